// net_kernels.h -- argument blocks and launchers of the proposal-network kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mvs {
namespace net {

// conv3x3 over NHWC bf16; input = concat(x0 [N][H][W][c0], upsample2(x1 [N][H1][W1][c1]))
struct ConvArgs {
    const uint16_t* x0;
    int c0;
    const uint16_t* x1;
    int c1;
    int H, W, H1, W1;
    const uint16_t* wt;    // [9][cout][c0 + c1] bf16
    const float* bias;     // [cout] or NULL
    int cout;
    int relu;
    int mode;              // 0: bf16 store (+bias, ReLU); 1: fp32 accumulate (data gradient)
    uint16_t* y;           // mode 0: [N][H][W][cout]
    int split;             // mode 1: channels < split -> yf0 (+=), others -> atomicAdd into half-res yf1
    float* yf0;
    float* yf1;
};

struct WgradArgs {
    const uint16_t* x0;
    int c0;
    const uint16_t* x1;
    int c1;
    int N, H, W, H1, W1;
    const uint16_t* dy;    // [N][H][W][cout] bf16 (ReLU-masked)
    int cout;
    int64_t chunk;         // pixels per workgroup
    float* dw;             // [9][c0 + c1][cout] fp32 (accumulated)
};

struct HeadArgs {
    const uint16_t* feat;  // [N][H][W][32]
    int64_t npix;
    int H, W;              // padded size
    int h_valid, w_valid;  // image size inside the padded tensor
    const float* w;        // [5][32] weights then [5] bias
    float* pred;           // [N][h_valid][w_valid][5] or NULL
    const float* target;   // [N][h_valid][w_valid][5]: inv (normalised), normal (3), valid
    float* gfeat;          // [N][H][W][32] fp32 gradient of feat (NULL = inference)
    float* gw;             // [5*32 + 5]
    float* loss;           // [4]: sum L1 depth, sum normal, sum BCE, valid count
    float w_depth, w_normal, w_conf, conf_tol;
    float scale_v;         // depth and normal terms: 1 / valid pixels of the batch
    float scale_c;         // confidence term: 1 / pixels of the batch
};

struct AdamArgs {
    float* p;
    float* g;
    float* m;
    float* v;
    int64_t n;
    float lr, b1, b2, eps, c1, c2, wd, gscale;
};

hipError_t launch_conv(const ConvArgs& a, int N, hipStream_t st);
hipError_t launch_wgrad(const WgradArgs& a, hipStream_t st);
hipError_t launch_bgrad(const uint16_t* dy, int64_t npix, int cout, float* db, hipStream_t st);
hipError_t launch_relu_bwd(const float* g, const uint16_t* y, int64_t n, uint16_t* out, hipStream_t st);
hipError_t launch_pool_fwd(const uint16_t* x, int N, int H, int W, int C, uint16_t* y, hipStream_t st);
hipError_t launch_pool_bwd(const float* gy, int N, int H, int W, int C, float* gx, hipStream_t st);
hipError_t launch_zero(float* p, int64_t n, hipStream_t st);
hipError_t launch_head(const HeadArgs& a, hipStream_t st);
hipError_t launch_adam(const AdamArgs& a, hipStream_t st);
hipError_t launch_wprep(const float* w, int cin, int cout, uint16_t* wf, uint16_t* wb, hipStream_t st);

}  // namespace net
}  // namespace mvs
