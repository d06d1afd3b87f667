// net.hip -- host side of the learned plane-proposal network.
//
// Network: 5-level U-Net of 3x3 convolutions (channels 32-64-96-128-160),
// two convolutions per level, 2x2 average pooling down, nearest upsampling
// and skip concatenation up, 1x1 head to (inverse depth, normal, confidence).
// Input: 32-plane census sweep volume over 4 sources + the gray level.
// Forward, backward and Adam run on the device (net_kernels.hip); fp32
// master weights and moments, bf16 weights and activations, fp32
// accumulation and fp32 activation gradients.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "net_kernels.h"
#include "net_internal.h"

namespace mvs {
namespace net {

// --------------------------------------------------------------------------
// layer table
// --------------------------------------------------------------------------

// activation tensors
enum Act { IN = 0, A0, E0, P0, A1, E1, P1, A2, E2, P2, A3, E3, P3, A4, E4, B3, D3, B2, D2, B1, D1, B0, D0, NACT };

struct ActInfo {
    int ch, level;
};
static const ActInfo ACTS[NACT] = {
    {64, 0},  {32, 0},  {32, 0}, {32, 1},  {64, 1},  {64, 1},  {64, 2},  {96, 2},  {96, 2},  {96, 3},  {128, 3}, {128, 3},
    {128, 4}, {160, 4}, {160, 4}, {128, 3}, {128, 3}, {96, 2}, {96, 2}, {64, 1}, {64, 1}, {32, 0}, {32, 0}};

struct LayerDef {
    int x0, x1, out;   // activation ids (x1 = -1: none)
};
static const LayerDef LAYERS[NLAYER] = {
    {IN, -1, A0}, {A0, -1, E0}, {P0, -1, A1}, {A1, -1, E1}, {P1, -1, A2}, {A2, -1, E2},
    {P2, -1, A3}, {A3, -1, E3}, {P3, -1, A4}, {A4, -1, E4}, {E3, E4, B3}, {B3, -1, D3},
    {E2, D3, B2}, {B2, -1, D2}, {E1, D2, B1}, {B1, -1, D1}, {E0, D1, B0}, {B0, -1, D0}};
// pooled tensors: P_l = pool(E_{l})
static const int POOL_SRC[4][2] = {{E0, P0}, {E1, P1}, {E2, P2}, {E3, P3}};

int layer_cin(int l) { return ACTS[LAYERS[l].x0].ch + (LAYERS[l].x1 >= 0 ? ACTS[LAYERS[l].x1].ch : 0); }
int layer_cout(int l) { return ACTS[LAYERS[l].out].ch; }

Net::Net() {
    int64_t off = 0;
    for (int l = 0; l < NLAYER; l++) {
        w_off[l] = off;
        off += (int64_t)9 * layer_cin(l) * layer_cout(l);
        b_off[l] = off;
        off += layer_cout(l);
    }
    head_off = off;
    off += 5 * 32 + 5;
    n_params = off;
}

static hipError_t dalloc(void** p, size_t bytes) { return hipMalloc(p, bytes ? bytes : 16); }

hipError_t Net::init(uint64_t seed, hipStream_t st) {
    stream = st;
    hipError_t e;
    if ((e = dalloc((void**)&p, n_params * 4)) != hipSuccess) return e;
    if ((e = dalloc((void**)&g, n_params * 4)) != hipSuccess) return e;
    if ((e = dalloc((void**)&m, n_params * 4)) != hipSuccess) return e;
    if ((e = dalloc((void**)&v, n_params * 4)) != hipSuccess) return e;
    for (int l = 0; l < NLAYER; l++) {
        const size_t nw = (size_t)9 * layer_cin(l) * layer_cout(l);
        if ((e = dalloc((void**)&wf[l], nw * 2)) != hipSuccess) return e;
        if ((e = dalloc((void**)&wb[l], nw * 2)) != hipSuccess) return e;
    }
    // He-normal init from a seeded generator; input channels beyond 33 of
    // the first layer carry zeros in every input, so their weights never matter
    std::vector<float> h((size_t)n_params, 0.0f);
    std::mt19937_64 gen(seed);
    for (int l = 0; l < NLAYER; l++) {
        const int cin = layer_cin(l), cout = layer_cout(l);
        const int fan = 9 * (l == 0 ? 33 : cin);
        std::normal_distribution<float> nd(0.0f, sqrtf(2.0f / (float)fan));
        for (int64_t i = 0; i < (int64_t)9 * cin * cout; i++) h[w_off[l] + i] = nd(gen);
    }
    std::normal_distribution<float> nh(0.0f, 0.05f);
    for (int i = 0; i < 5 * 32; i++) h[head_off + i] = nh(gen);
    if ((e = hipMemcpyAsync(p, h.data(), n_params * 4, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    if ((e = launch_zero(g, n_params, st)) != hipSuccess) return e;
    if ((e = launch_zero(m, n_params, st)) != hipSuccess) return e;
    if ((e = launch_zero(v, n_params, st)) != hipSuccess) return e;
    step = 0;
    return refresh_weights();
}

hipError_t Net::refresh_weights() {
    for (int l = 0; l < NLAYER; l++) {
        hipError_t e = launch_wprep(p + w_off[l], layer_cin(l), layer_cout(l), wf[l], wb[l], stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

void Net::release_acts() {
    for (int i = 0; i < NACT; i++) {
        if (act[i]) (void)hipFree(act[i]);
        if (gact[i]) (void)hipFree(gact[i]);
        act[i] = nullptr;
        gact[i] = nullptr;
    }
    if (dyp) (void)hipFree(dyp);
    if (pred) (void)hipFree(pred);
    if (target) (void)hipFree(target);
    if (loss) (void)hipFree(loss);
    dyp = nullptr;
    pred = nullptr;
    target = nullptr;
    loss = nullptr;
    cap_N = cap_H = cap_W = 0;
    grads_alloc = false;
}

void Net::release() {
    release_acts();
    void* ps[] = {p, g, m, v};
    for (void* q : ps)
        if (q) (void)hipFree(q);
    p = g = m = v = nullptr;
    for (int l = 0; l < NLAYER; l++) {
        if (wf[l]) (void)hipFree(wf[l]);
        if (wb[l]) (void)hipFree(wb[l]);
        wf[l] = wb[l] = nullptr;
    }
}

// Buffers for a batch of N padded images of H x W (multiples of 16).
hipError_t Net::ensure(int N_, int H_, int W_, bool train) {
    // tensors are [N][H][W][C]: a smaller batch of the same size uses a
    // prefix of the buffers, so only a larger batch, another image size or
    // the first training call allocates
    if (N_ <= cap_N && H_ == cap_H && W_ == cap_W && (!train || grads_alloc)) {
        N = N_;
        H = H_;
        W = W_;
        return hipSuccess;
    }
    release_acts();
    N = cap_N = N_;
    H = cap_H = H_;
    W = cap_W = W_;
    hipError_t e;
    size_t max_elems = 0;
    for (int i = 0; i < NACT; i++) {
        const size_t n = (size_t)N * (H >> ACTS[i].level) * (W >> ACTS[i].level) * ACTS[i].ch;
        max_elems = std::max(max_elems, n);
        if ((e = dalloc((void**)&act[i], n * 2)) != hipSuccess) return e;
        if (train && i != IN) {
            if ((e = dalloc((void**)&gact[i], n * 4)) != hipSuccess) return e;
        }
    }
    if (train && (e = dalloc((void**)&dyp, max_elems * 2)) != hipSuccess) return e;
    if ((e = dalloc((void**)&pred, (size_t)N * H * W * 5 * 4)) != hipSuccess) return e;
    if ((e = dalloc((void**)&target, (size_t)N * H * W * 5 * 4)) != hipSuccess) return e;
    if ((e = dalloc((void**)&loss, 64)) != hipSuccess) return e;
    grads_alloc = train;
    return hipSuccess;
}

hipError_t Net::forward() {
    hipError_t e;
    for (int l = 0; l < NLAYER; l++) {
        // the pooled inputs of the encoder levels
        for (int k = 0; k < 4; k++) {
            if (LAYERS[l].x0 == POOL_SRC[k][1]) {
                const int src = POOL_SRC[k][0], lev = ACTS[src].level;
                if ((e = launch_pool_fwd(act[src], N, H >> lev, W >> lev, ACTS[src].ch, act[POOL_SRC[k][1]], stream)) !=
                    hipSuccess)
                    return e;
            }
        }
        const LayerDef& L = LAYERS[l];
        const int lev = ACTS[L.out].level;
        ConvArgs a{};
        a.x0 = act[L.x0];
        a.c0 = ACTS[L.x0].ch;
        a.x1 = L.x1 >= 0 ? act[L.x1] : nullptr;
        a.c1 = L.x1 >= 0 ? ACTS[L.x1].ch : 0;
        a.H = H >> lev;
        a.W = W >> lev;
        a.H1 = H >> (lev + 1);
        a.W1 = W >> (lev + 1);
        a.wt = wf[l];
        a.bias = p + b_off[l];
        a.cout = layer_cout(l);
        a.relu = 1;
        a.mode = 0;
        a.y = act[L.out];
        if ((e = launch_conv(a, N, stream)) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t Net::head(bool train, int h_valid, int w_valid, float* pred_out, const float* tgt, float conf_tol,
                     int64_t n_valid) {
    HeadArgs a{};
    a.feat = act[D0];
    a.npix = (int64_t)N * H * W;
    a.H = H;
    a.W = W;
    a.h_valid = h_valid;
    a.w_valid = w_valid;
    a.w = p + head_off;
    a.pred = pred_out;
    a.target = tgt;
    a.gfeat = train ? gact[D0] : nullptr;
    a.gw = g + head_off;
    a.loss = loss;
    a.w_depth = 1.0f;
    a.w_normal = 0.5f;
    a.w_conf = 0.25f;
    a.scale_c = 1.0f / (float)((int64_t)N * h_valid * w_valid);
    a.scale_v = 1.0f / (float)(n_valid > 0 ? n_valid : 1);
    a.conf_tol = conf_tol;
    if (train) {
        hipError_t e = launch_zero(loss, 4, stream);
        if (e != hipSuccess) return e;
    }
    return launch_head(a, stream);
}

hipError_t Net::backward() {
    hipError_t e;
    for (int i = 0; i < NACT; i++) {
        if (i == IN || i == D0) continue;
        const size_t n = (size_t)N * (H >> ACTS[i].level) * (W >> ACTS[i].level) * ACTS[i].ch;
        if ((e = launch_zero(gact[i], (int64_t)n, stream)) != hipSuccess) return e;
    }
    for (int l = NLAYER - 1; l >= 0; l--) {
        const LayerDef& L = LAYERS[l];
        const int lev = ACTS[L.out].level;
        const int Hl = H >> lev, Wl = W >> lev;
        const int cout = layer_cout(l);
        const int64_t npix = (int64_t)N * Hl * Wl;
        if ((e = launch_relu_bwd(gact[L.out], act[L.out], npix * cout, dyp, stream)) != hipSuccess) return e;
        WgradArgs wa{};
        wa.x0 = act[L.x0];
        wa.c0 = ACTS[L.x0].ch;
        wa.x1 = L.x1 >= 0 ? act[L.x1] : nullptr;
        wa.c1 = L.x1 >= 0 ? ACTS[L.x1].ch : 0;
        wa.N = N;
        wa.H = Hl;
        wa.W = Wl;
        wa.H1 = Hl >> 1;
        wa.W1 = Wl >> 1;
        wa.dy = dyp;
        wa.cout = cout;
        wa.chunk = std::max<int64_t>(1024, ((npix / 256) + 31) / 32 * 32);
        wa.dw = g + w_off[l];
        if ((e = launch_wgrad(wa, stream)) != hipSuccess) return e;
        if ((e = launch_bgrad(dyp, npix, cout, g + b_off[l], stream)) != hipSuccess) return e;
        if (l == 0) break;
        ConvArgs a{};
        a.x0 = dyp;
        a.c0 = cout;
        a.x1 = nullptr;
        a.c1 = 0;
        a.H = Hl;
        a.W = Wl;
        a.H1 = Hl >> 1;
        a.W1 = Wl >> 1;
        a.wt = wb[l];
        a.bias = nullptr;
        a.cout = layer_cin(l);
        a.relu = 0;
        a.mode = 1;
        a.split = ACTS[L.x0].ch;
        a.yf0 = gact[L.x0];
        a.yf1 = L.x1 >= 0 ? gact[L.x1] : nullptr;
        if ((e = launch_conv(a, N, stream)) != hipSuccess) return e;
        // pooled input: route its gradient to the pre-pool tensor
        for (int k = 0; k < 4; k++) {
            if (L.x0 == POOL_SRC[k][1]) {
                const int src = POOL_SRC[k][0], sl = ACTS[src].level;
                if ((e = launch_pool_bwd(gact[L.x0], N, H >> sl, W >> sl, ACTS[src].ch, gact[src], stream)) !=
                    hipSuccess)
                    return e;
            }
        }
    }
    return hipSuccess;
}

hipError_t Net::adam(float lr) {
    step++;
    AdamArgs a{};
    a.p = p;
    a.g = g;
    a.m = m;
    a.v = v;
    a.n = n_params;
    a.lr = lr;
    a.b1 = 0.9f;
    a.b2 = 0.999f;
    a.eps = 1e-8f;
    a.c1 = 1.0f - powf(0.9f, (float)step);
    a.c2 = 1.0f - powf(0.999f, (float)step);
    a.wd = 0.0f;
    a.gscale = 1.0f;
    hipError_t e = launch_adam(a, stream);
    if (e != hipSuccess) return e;
    return refresh_weights();
}

}  // namespace net
}  // namespace mvs
