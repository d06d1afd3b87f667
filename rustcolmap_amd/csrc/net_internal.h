// net_internal.h -- the proposal network object shared by net.hip and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mvs {
namespace net {

constexpr int NLAYER = 18;

int layer_cin(int l);
int layer_cout(int l);

struct Net {
    // parameters: per layer [9][cin][cout] weights then [cout] bias; head [5][32] + [5]
    int64_t w_off[NLAYER], b_off[NLAYER], head_off = 0, n_params = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t* wf[NLAYER] = {nullptr};   // [9][cout][cin] bf16 (forward)
    uint16_t* wb[NLAYER] = {nullptr};   // [9][cin][cout] bf16, taps flipped (data gradient)
    int step = 0;
    hipStream_t stream = nullptr;
    // activations of the current batch (23 tensors, see net.hip)
    int N = 0, H = 0, W = 0, cap_N = 0, cap_H = 0, cap_W = 0;
    bool grads_alloc = false;
    uint16_t* act[23] = {nullptr};
    float* gact[23] = {nullptr};
    uint16_t* dyp = nullptr;
    float* pred = nullptr;     // [N][h][w][5]
    float* target = nullptr;   // [N][h][w][5]
    float* loss = nullptr;     // [4]

    Net();
    hipError_t init(uint64_t seed, hipStream_t st);
    hipError_t refresh_weights();
    hipError_t ensure(int N, int H, int W, bool train);
    uint16_t* input() { return act[0]; }
    hipError_t forward();
    // n_valid: valid target pixels of the batch (normalises the depth and
    // normal terms; the confidence term is a mean over all pixels)
    hipError_t head(bool train, int h_valid, int w_valid, float* pred_out, const float* tgt, float conf_tol,
                    int64_t n_valid);
    hipError_t backward();
    hipError_t adam(float lr);
    void release_acts();
    void release();
};

}  // namespace net
}  // namespace mvs
