// stereo_kernels.h -- host-visible launchers of the stereo and fusion kernels.
#pragma once
#include "mvs_device.h"

namespace mvs {

// One view of an upload: census + rays are produced for it.
struct PrepView {
    const uint8_t* gray;
    u64* census;
    float2* rays;
    int32_t w, h;
    double f, cx, cy, k;
    int32_t tiles_x, tile_off;
};

// Outputs of the sweep-volume kernel for one job (either may be NULL).
struct VolOut {
    uint16_t* vol;   // [pix][32] bf16 network features
    float* volf;     // [pix][nplanes] fp32 plane costs
};

// Launch shape of the stereo kernels (picked per shape class by the tuner).  The default is the shape the tuner
// picks at 1080p (8x8 tiles, persistent passes, evaluator 3).
struct TileCfg {
    int tw = 8, th = 8;
    int persistent = 1;   // 1: persistent grid over an atomic tile queue; 2: the same with one wave per tile,
                          // its pixels in rounds of 64 ordered by the previous pass's cost (evaluator 3)
    int eval = 3;         // plane_cost evaluator: 0 = reference loop order, 1 = hoisted geometry, 2 = + no-divide
                          // abort, 3/4 = per-tap geometry, exact fast divide, strengthened abort, shared-cost init
};

int prepare_tiles(int w, int h, int& tiles_x);
hipError_t launch_prepare(const PrepView* d_views, int nviews, int nblocks, hipStream_t st);
hipError_t launch_fill_u64(u64* p, int n, u64 v, hipStream_t st);
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t st);
hipError_t launch_init(const TileCfg& cfg, const JobDesc* jobs, const SrcDesc* srcs, const int4* tiles, int ntiles,
                       int nh, u64* counter, hipStream_t st);
hipError_t launch_pass(const TileCfg& cfg, const JobDesc* jobs, const SrcDesc* srcs, const int4* tiles, int ntiles,
                       int par, u64 salt, float sig_inv, float sig_n, u64* counter, int* queue, int grid_persistent,
                       hipStream_t st);
hipError_t launch_gate(const TileCfg& cfg, const JobDesc* jobs, const int4* tiles, int ntiles, int par,
                       float cost_thr, hipStream_t st);
int persistent_blocks_per_cu(const TileCfg& cfg);
hipError_t launch_volume(const JobDesc* jobs, const SrcDesc* srcs, const int4* tiles, int ntiles, int nplanes,
                         int stride, const VolOut* outs, hipStream_t st);
// launch shapes the learned initialisation kernel is built for (tile
// origins must be even so a pixel's coarse pixel lies in its tile)
bool proposal_supported(const TileCfg& cfg);
// prop == nullptr: an accepted proposal replaces the plain sweep at its
// pixel; otherwise (room for one float4 per pixel) it is an extra hypothesis
// that every pixel's plain sweep competes with by cost.
hipError_t launch_proposal_init(const TileCfg& cfg, const JobDesc* jobs, const SrcDesc* srcs, const int4* tiles,
                                int ntiles, int nh, const float* const* preds, int stride, float conf_thr,
                                u64* counter, u64* accepted, u64* list, float4* prop, hipStream_t st);
hipError_t launch_assemble(const uint16_t* const* vols, const uint8_t* const* grays, int n, int h, int w, int Hp,
                           int Wp, int stride, int wfull, uint16_t* out, hipStream_t st);
hipError_t launch_debug_cost(const JobDesc& J, const SrcDesc* S, const float* q, int n, float* cost, int* ham,
                             int* ninv, float* peak, hipStream_t st);
// input stage (image.hip)
hipError_t launch_resize(const uint8_t* src, int channels, int w, int h, int mw, int mh, float* tmp, uint8_t* out,
                         hipStream_t st);
hipError_t launch_luma(const uint8_t* rgb, int64_t n, uint8_t* gray, hipStream_t st);
hipError_t launch_debug_cvt(const float* x, int n, int* out, hipStream_t st);
hipError_t launch_debug_div(const float* a, const float* b, int n, float* out, hipStream_t st);
hipError_t launch_debug_rng(const int* xy, const u64* salt, int n, int ndraw, u64* seed_out, float* uni,
                            hipStream_t st);

// fusion (fuse.hip)
struct FuseScratch;
struct FuseArgs {
    const FuseView* d_views;   // nviews
    int nviews;
    const int32_t* d_order;    // concatenated fusion orders
    int64_t n_pix;             // candidate index space (pixels of views with depth)
    int min_consistent;
    double depth_diff;
    double cell;
};
// Runs fusion; points (first-insertion order) are written to a device buffer
// owned by the scratch; returns the counts through n_cand / n_pts.
hipError_t run_fusion(FuseScratch*& scratch, const FuseArgs& a, const FuseView* h_views, hipStream_t st,
                      int64_t* n_cand, int64_t* n_pts, const void** d_points, int* limit_err);
void free_fuse_scratch(FuseScratch* s);

}  // namespace mvs
