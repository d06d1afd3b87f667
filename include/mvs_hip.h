/*
 * mvs_hip.h -- C ABI of the MI355X (gfx950) dense multi-view-stereo library.
 *
 * Drop-in boundary for the reference `mvs` crate's dense path.  Every entry
 * point names the reference item it replaces (paths relative to the
 * reference repository root).  Conventions:
 *   - plain C types and pointers only; every buffer crossing the boundary is
 *     a HOST buffer owned by the caller (no device pointers cross it);
 *   - return value 0 = ok, < 0 = error (MVS_ERR_*); the message of the last
 *     failure on the calling thread is available from mvs_last_error();
 *   - no C++ exception crosses the boundary;
 *   - one context = one device + one HIP stream; calls on one context are
 *     serialised internally, so the reference's concurrent per-view calls
 *     (mvs/src/lib.rs:351-374, rayon) are safe; mvs_estimate_depth_batch is
 *     the intended replacement for that loop.
 */
#ifndef MVS_HIP_H
#define MVS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MVS_OK 0
#define MVS_ERR_ARG (-1)     /* invalid argument (sizes, indices, NULL) */
#define MVS_ERR_HIP (-2)     /* HIP runtime failure */
#define MVS_ERR_NOMEM (-3)   /* device or host allocation failed */
#define MVS_ERR_STATE (-4)   /* call out of order (e.g. no views uploaded) */
#define MVS_ERR_LIMIT (-5)   /* input beyond a documented limit */

typedef struct mvs_ctx mvs_ctx;

/* MvsOptions (mvs/src/lib.rs:33-54); defaults from lib.rs:56-70. */
typedef struct mvs_options {
    uint32_t max_dim;         /* 900 (input stage; views arrive at working resolution) */
    uint32_t num_sources;     /* 3 */
    uint32_t sweeps;          /* 4: passes = 2 * max(sweeps, 1) */
    uint32_t num_hyp;         /* 48: plane-sweep hypotheses, max(num_hyp, 2) used */
    uint32_t min_consistent;  /* 2 */
    uint32_t save_depth_vis;  /* 0 */
    double depth_diff;        /* 0.03 */
    double voxel_rel;         /* 0.0015 */
    float cost_thr;           /* 0.38 */
    uint32_t reserved;        /* 0 */
} mvs_options_t;

/* One registered view at working resolution: the fields of ViewData
 * (mvs/src/lib.rs:86-101) that the dense path reads.  `gray` is the u8
 * luma the reference divides by 255 (lib.rs:193-196); the census and the
 * per-pixel rays are computed on the device.  Cam / Pose as in
 * sfm-lite/src/geom.rs:13-53 (SIMPLE_RADIAL, world->camera, R row-major). */
typedef struct mvs_view {
    int32_t w, h;
    double f, cx, cy, k;
    double R[9];
    double t[3];
    const uint8_t *gray;      /* w*h */
    const uint8_t *rgb;       /* 3*w*h, or NULL = gray replicated (to_rgb8 of luma) */
    const char *name;         /* ViewData.name (image file name), for the log lines and the depth_vis file
                                 name; NULL = "view<index>" */
} mvs_view_t;

/* One reference view of a batched stereo run. */
typedef struct mvs_job {
    int32_t ref;              /* view index */
    int32_t k;                /* number of sources */
    const int32_t *srcs;      /* k view indices */
    double zmin, zmax;        /* depth range (depth_range_for, lib.rs:253-282) */
    float *depth;             /* optional host outputs (w*h, w*h, 3*w*h); NULL = keep on device */
    float *cost;
    float *normals;
} mvs_job_t;

/* Fused point: (x, y, z, [r, g, b]) of fuse_depth_maps (mvs/src/fuse.rs:71). */
typedef struct mvs_point {
    double x, y, z;
    uint8_t r, g, b, pad[5];
} mvs_point_t;

/* Work counters and device timings of the last stereo / fusion call. */
typedef struct mvs_stats {
    int64_t n_eval_init;      /* plane_cost evaluations in the init sweep */
    int64_t n_eval_pass;      /* plane_cost evaluations over all passes */
    int64_t n_pixels;         /* reference pixels processed */
    int64_t taps_per_eval_sum;/* sum over jobs of pixels * 25 * K (for byte accounting) */
    double ms_prepare;        /* census + rays kernels of the last upload */
    double ms_init;
    double ms_pass;           /* all passes */
    double ms_gate;
    double ms_fuse;
    int64_t n_candidates;     /* fusion candidates of the last fusion */
    int64_t n_points;
    int32_t n_pass_launches;
    int32_t graph_used;
    double ms_learned_volume; /* learned init: sweep volume (part of ms_init) */
    double ms_learned_net;    /* learned init: network forward (part of ms_init) */
    int32_t tile_w, tile_h;   /* launch shape of the last stereo call */
    int32_t persistent;       /* 1 = persistent tile queue for the passes, 2 = its sorted form */
    int32_t evaluator;        /* plane_cost evaluator (see mvs_autotune) */
} mvs_stats_t;

/* MvsOptions::default (mvs/src/lib.rs:56-70). */
void mvs_options_default(mvs_options_t *opts);

/* Thread-local message of the last failed call on this thread ("" if none). */
const char *mvs_last_error(void);

/* Context: device index (as HIP numbers them) and one stream. */
int mvs_ctx_create(int device, mvs_ctx **out);
void mvs_ctx_destroy(mvs_ctx *ctx);

/* The `log: &(dyn Fn(&str) + Sync)` argument of run_mvs / load_views /
 * fuse_depth_maps (mvs/src/lib.rs:148, 299; mvs/src/fuse.rs:70), as a
 * context-level callback: fn(line, user) receives each progress line the
 * reference writes (lib.rs:212-216, 344-347, 362-371, 376-380, 391,
 * 509-515; fuse.rs:89-92, 170-174) from mvs_run_mvs, mvs_fuse and
 * mvs_fuse_depth_maps, on the calling thread.  fn = NULL switches it off. */
typedef void (*mvs_log_fn)(const char *line, void *user);
int mvs_set_log(mvs_ctx *ctx, mvs_log_fn fn, void *user);
/* Host-only probe: `format!("{:?}", x)` of an f64 as the log lines print
 * the depth range (lib.rs:362-371), into out (cap bytes, NUL-terminated). */
int mvs_debug_rust_f64(double x, char *out, int32_t cap);

/* census_map (mvs/src/lib.rs:119-142): 7x7 census, bit (dy+3)*7+(dx+3),
 * clamp-to-edge, set when the neighbour is darker.  Host in, host out. */
int mvs_census(mvs_ctx *ctx, const uint8_t *gray, int32_t w, int32_t h, uint64_t *out);

/* No counterpart in the reference: mvs_census for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and result as
 * mvs_census and the same kernel, but the device staging (image, codes,
 * launch descriptor) lives in one file-scope device array per GPU, set up
 * when a context is created on that GPU, so the call allocates nothing.
 * Images above MVS_NOALLOC_MAX_PIXELS pixels get MVS_ERR_LIMIT.  Every GPU
 * the library runs on holds the array (about 36 MiB) until the process
 * ends; the calls on one GPU share it and are serialised by the library,
 * and each call rewrites everything it reads back. */
#define MVS_NOALLOC_MAX_PIXELS (1 << 22)
int mvs_census_noalloc(mvs_ctx *ctx, const uint8_t *gray, int32_t w, int32_t h, uint64_t *out);

/* load_views (mvs/src/lib.rs:144-218) after decode/resize: uploads n views,
 * replacing any previous set, and computes each view's census and rays on
 * the device (one launch for all views).  A view holds at most 2^29 pixels
 * and 2^24 - 1 on a side (MVS_ERR_LIMIT beyond: census offsets are 32-bit
 * byte offsets). */
int mvs_upload_views(mvs_ctx *ctx, const mvs_view_t *views, int32_t n);

/* ViewData.census of an uploaded view (w*h u64). */
int mvs_view_census(mvs_ctx *ctx, int32_t view, uint64_t *out);

/* stereo::estimate_depth (mvs/src/stereo.rs:234-394) for one reference
 * view: DepthMap.depth (0 = invalid) and DepthMap.cost, plus the accepted
 * unit normals (3 floats/px, camera frame; NULL to skip) that the reference
 * keeps internal (stereo.rs:24-31). */
int mvs_estimate_depth(mvs_ctx *ctx, int32_t ref, const int32_t *srcs, int32_t k, double zmin, double zmax,
                       const mvs_options_t *opts, float *depth, float *cost, float *normals);

/* As mvs_estimate_depth, plus: stop after `max_passes` passes (< 0 = all;
 * the decay schedule still uses 2*sweeps), the pre-gate plane buffer
 * (5 floats/px: nx, ny, nz, z, cost; NULL to skip) and per-stage evaluation
 * counts n_eval[0] = init, n_eval[1+p] = pass p (NULL to skip). */
int mvs_estimate_depth_ex(mvs_ctx *ctx, int32_t ref, const int32_t *srcs, int32_t k, double zmin, double zmax,
                          const mvs_options_t *opts, int32_t max_passes, float *depth, float *cost,
                          float *normals, float *planes, int64_t *n_eval);

/* The per-view stereo loop of run_mvs (mvs/src/lib.rs:351-374) for all
 * jobs at once: one launch per stage covers every job (uneven sizes and K),
 * and the stage schedule is replayed from a captured HIP graph.  Results
 * stay resident for mvs_fuse and are copied to the jobs' host pointers
 * where those are non-NULL.  `mode` bits 0-1: 0 = tuned schedule,
 * 1 = one launch per job and stage (unbatched), 2 = batched tile grid,
 * 3 = batched persistent tile queue (either form; a tuned tile-grid shape
 * becomes its persistent form, a tuned sorted shape becomes the default
 * shape under mode 2); bit 2 (+4): launch directly instead
 * of replaying the graph, with per-stage device timings in mvs_stats. */
int mvs_estimate_depth_batch(mvs_ctx *ctx, const mvs_job_t *jobs, int32_t njobs, const mvs_options_t *opts,
                             int32_t mode);

/* Launch-shape autotuner (no counterpart in the reference): runs the batch
 * with every candidate (tile 16x16 / 32x8 / 16x8 / 8x8; evaluator 0
 * reference order, 1 hoisted geometry, 2 divide-free abort, 3 / 4 per-tap
 * geometry with the scaling-free exact divide and the strengthened abort,
 * source-0 rows 1-2 batched or not; tile grid, persistent queue, or
 * persistent queue with one wave per tile whose pixels run in rounds
 * ordered by the previous pass's cost: persistent flag 2) for
 * `passes` passes, and records the fastest
 * for the batch's shape class (longest
 * side bucket x K bucket); later batches of that class use it.  Results
 * are identical for every candidate.  best (optional, 4 ints) receives
 * tile width, tile height, persistent flag, evaluator; cand_ms / cand_cfg
 * (optional, room for 32 entries) the estimated full-run time of every
 * candidate and its 4 parameters, n_cand their number. */
int mvs_autotune(mvs_ctx *ctx, const mvs_job_t *jobs, int32_t njobs, const mvs_options_t *opts, int32_t passes,
                 int32_t *best, float *cand_ms, int32_t *cand_cfg, int32_t *n_cand);

/* Download the resident depth / cost / normals of a view computed by the
 * last batch (NULL pointers skip). */
int mvs_get_depth(mvs_ctx *ctx, int32_t view, float *depth, float *cost, float *normals);

/* fuse::fuse_depth_maps (mvs/src/fuse.rs:64-215) over the depth maps held
 * on the device from the last stereo calls.  order_off has nviews+1
 * entries; order[order_off[i] .. order_off[i+1]) is fusion_order[i]
 * (lib.rs:398-418).  `has_depth[i]` != 0 marks views with a depth map
 * (Option::Some).  `cell` is the absolute voxel size (see mvs_fusion_cell).
 * Points come back in first-insertion order of their voxel (view-major
 * raster order of the first candidate); the reference's HashMap order is
 * unspecified.  *pts is allocated by the library: free with
 * mvs_free_points. */
int mvs_fuse(mvs_ctx *ctx, const int32_t *has_depth, const int32_t *order_off, const int32_t *order,
             double cell, const mvs_options_t *opts, mvs_point_t **pts, int64_t *n);

/* As mvs_fuse with caller-provided host depth maps (NULL entry = None),
 * the exact signature shape of fuse_depth_maps. */
int mvs_fuse_depth_maps(mvs_ctx *ctx, const float *const *depth_maps, const int32_t *order_off,
                        const int32_t *order, double cell, const mvs_options_t *opts, mvs_point_t **pts,
                        int64_t *n);

void mvs_free_points(mvs_point_t *pts);

/* Robust voxel cell of fuse_depth_maps (mvs/src/fuse.rs:72-88): per-axis
 * 2 %/98 % extent of the sparse points -> diag * voxel_rel (>= 1e-9).
 * Returns < 0 when n == 0 (the reference panics there). */
double mvs_fusion_cell(const double *xyz, int64_t n, double voxel_rel);
/* The robust sparse diagonal itself (fuse.rs:72-87), for the cell's log
 * line (fuse.rs:89-92).  Returns < 0 when n == 0. */
double mvs_fusion_diag(const double *xyz, int64_t n);

/* write_ply_binary (mvs/src/fuse.rs:221-239). */
int mvs_write_ply(const char *path, const mvs_point_t *pts, int64_t n);

/* dense_web.json of run_mvs (mvs/src/lib.rs:437-485): outliers beyond the
 * robust radius dropped, at most 400000 points kept by a fixed stride,
 * positions with 4 decimals.  Writes nothing for n = 0.  Host only. */
int mvs_write_dense_web_json(const char *path, const mvs_point_t *pts, int64_t n);

/* Counters and device timings of the last call on this context. */
int mvs_get_stats(mvs_ctx *ctx, mvs_stats_t *out);

/* Instrumentation for parity tests: evaluates plane_cost
 * (mvs/src/stereo.rs:158-228) on the device for n (pixel, plane, hint)
 * queries of reference view `ref` with sources srcs.  q holds 7 floats per
 * query: px, py, nx, ny, nz, d, hint.  Outputs per query: the returned
 * cost, the integer Hamming sum and invalid-tap count over the full
 * window, and the peak partial average (max over (source, row) prefixes). */
int mvs_debug_plane_cost(mvs_ctx *ctx, int32_t ref, const int32_t *srcs, int32_t k, const float *q, int32_t n,
                         float *cost, int32_t *ham, int32_t *ninv, float *peak);

/* Instrumentation: `(x) as i32` as the stereo kernels evaluate it on the
 * device (truncate toward zero, saturate, NaN -> 0), for n floats. */
int mvs_debug_cvt(mvs_ctx *ctx, const float *x, int32_t n, int32_t *out);

/* Input stage of load_views (mvs/src/lib.rs:185-197) on the device: a
 * decoded u8 image (channels 1 = luma or 3 = RGB, ow x oh, row-major
 * interleaved) resized to mw x mh as the image crate's
 * resize_exact(..., Triangle) does (a copy at equal size), then its luma
 * (gray_out, mw*mh) and RGB (rgb_out, mw*mh*3; gray replicated for a
 * one-channel image), as to_luma8 / to_rgb8.  Either output may be NULL. */
int mvs_resize_image(mvs_ctx *ctx, const uint8_t *src, int32_t channels, int32_t ow, int32_t oh, int32_t mw,
                     int32_t mh, uint8_t *gray_out, uint8_t *rgb_out);

/* Test probe: out[i] = a[i] / b[i] through the stereo kernels' scaling-free
 * division sequence, which is IEEE division when b and 1/b are normal, a / b
 * is normal, |a| >= 2^-104 and |a / b| < 2^96. */
int mvs_debug_div(mvs_ctx *ctx, const float *a, const float *b, int32_t n, float *out);

/* Instrumentation: one plain device-to-device copy of `bytes` (a multiple
 * of 16) by the library's copy kernel; a reference for traffic counters. */
int mvs_debug_copy(mvs_ctx *ctx, int64_t bytes);

/* Per-pixel RNG of estimate_depth (stereo.rs:47-78) evaluated on the
 * device: for n (x, y, salt) triples, the seed and the first `ndraw`
 * uniform() values of its stream. */
int mvs_debug_rng(mvs_ctx *ctx, const int32_t *xy, const uint64_t *salt, int32_t n, int32_t ndraw,
                  uint64_t *seed_out, float *uniform_out);

/* Instrumentation: passes pass_lo .. pass_hi-1 of a run of n_passes passes
 * (the pass loop of stereo.rs:271-371, decay schedule of stereo.rs:275-277)
 * for reference view `ref`, started from caller-given planes instead of the
 * init sweep: planes holds 5 floats per pixel (nx, ny, nz, z, cost) and is
 * overwritten with the planes after the last pass; n_eval (optional) gets
 * the evaluation count of each pass run.  Uses the context's tuned launch
 * shape. */
int mvs_debug_run_passes(mvs_ctx *ctx, int32_t ref, const int32_t *srcs, int32_t k, double zmin, double zmax,
                         int32_t pass_lo, int32_t pass_hi, int32_t n_passes, float *planes, int64_t *n_eval);

/* ------------------------------------------------------------------------
 * run_mvs orchestration (mvs/src/lib.rs:294-517)
 * ------------------------------------------------------------------------ */

/* run_mvs's host scheduling from the sparse reconstruction: co-visibility
 * (lib.rs:224-243), stereo sources = top num_sources views with >= 10
 * shared points (lib.rs:326-342; equal counts keep ascending view index),
 * depth range per view from the 2 / 98 % percentiles of its sparse depths
 * with the global fallback (lib.rs:245-282), fusion order (lib.rs:398-418).
 * Sparse point p has xyz[3p..3p+3) and track views
 * track_view[track_off[p] .. track_off[p+1]) (view indices; < 0 ignored).
 * Outputs: src_off[n+1] / srcs[n * num_sources], zrange[2n],
 * fusion_off[n+1] / fusion[n * n], covis[n * n] (optional).
 * MVS_ERR_STATE when npoints < 8 (lib.rs:310-315). */
int mvs_plan(const mvs_view_t *views, int32_t nviews, const double *xyz, const int32_t *track_off,
             const int32_t *track_view, int64_t npoints, const mvs_options_t *opts, int32_t *src_off, int32_t *srcs,
             double *zrange, int32_t *fusion_off, int32_t *fusion, uint32_t *covis);

typedef struct mvs_run_stats {   /* MvsStats (lib.rs:72-79) + timings */
    int32_t views, views_with_depth;
    int64_t dense_points;
    int64_t n_candidates;
    double cell;
    double stereo_ms, fuse_ms;   /* device time */
    int64_t *per_view_valid;     /* MvsStats.per_view_valid: set by the caller to an array of nviews
                                    entries (valid pixels per view, 0 without a depth map), or NULL */
} mvs_run_stats_t;

/* run_mvs for views at working resolution: plan, batched stereo of every
 * view with sources, fusion with the robust cell, and (out_dir != NULL)
 * <out_dir>/dense/dense.ply, dense_web.json and mvs_stats.json; with
 * opts->save_depth_vis also <out_dir>/dense/depth_vis/<stem>_depth.png of
 * every depth map (save_depth_vis, lib.rs:383-392, fuse.rs:241-269: the
 * same 8-bit grayscale pixels, this library's own PNG encoding).  Progress
 * lines go to the context's log callback (mvs_set_log); the per-view line's
 * time is the batched stereo stage's, which every view spans. */
int mvs_run_mvs(mvs_ctx *ctx, const mvs_view_t *views, int32_t nviews, const double *xyz, const int32_t *track_off,
                const int32_t *track_view, int64_t npoints, const mvs_options_t *opts, const char *out_dir,
                mvs_run_stats_t *stats);

/* ------------------------------------------------------------------------
 * Learned plane proposals (no counterpart in the reference: an additional
 * initialisation for estimate_depth's plane sweep, stereo.rs:247-269).
 * A 5-level U-Net (3x3 convs, channels 32-64-96-128-160, bf16 on MFMA with
 * fp32 accumulation) maps the gray level plus a 32-plane census sweep volume
 * over 4 sources to normalised inverse depth, unit normal and confidence.
 * The volume and the network run on the coarse grid of every
 * MVS_NET_STRIDE-th pixel: coarse pixel (cx, cy) is pixel (S*cx, S*cy) with
 * its full-resolution census window; a w x h view has a
 * ceil(w/S) x ceil(h/S) coarse grid.
 * ------------------------------------------------------------------------ */
#define MVS_NET_STRIDE 2
typedef struct mvs_net mvs_net;

/* Network with seeded He-normal weights, bound to the context's device.
 * Destroying the context first frees the network's device memory; the
 * handle then only accepts mvs_net_destroy (other calls: MVS_ERR_STATE). */
int mvs_net_create(mvs_ctx *ctx, uint64_t seed, mvs_net **out);
void mvs_net_destroy(mvs_net *net);
int64_t mvs_net_num_params(mvs_net *net);
/* fp32 master weights (mvs_net_num_params floats). */
int mvs_net_get_params(mvs_net *net, float *out);
int mvs_net_set_params(mvs_net *net, const float *in);

/* 32-plane fronto-parallel sweep volume of one job (its k sources) on the
 * grid of every stride-th pixel (stride 1 or 2): for plane h = 0..31,
 * inv = 1/zmax + (1/zmin - 1/zmax) * h / 31, the census cost of the plane
 * (0, 0, 1, 1/inv) over the 5x5 window at pixel (stride*cx, stride*cy) and
 * the k sources -- plane_cost's Hamming sum plus 48 per invalid tap, with no
 * abort, summed in integers -- divided by 48 * 25 * k;
 * out = ceil(w/stride) * ceil(h/stride) * 32 floats. */
int mvs_sweep_volume(mvs_ctx *ctx, const mvs_job_t *job, int32_t stride, float *out);

/* Adds training samples: for each job (a view of the context, usually 4
 * sources), the coarse sweep volume is computed and kept on the device with
 * the host target on the coarse grid (cw*ch*5 floats: normalised inverse
 * depth (1/z - 1/zmax) / (1/zmin - 1/zmax), unit normal, validity 0/1).
 * Returns the sample count (>= 0) or an error. */
int mvs_net_add_samples(mvs_net *net, const mvs_job_t *jobs, int32_t njobs, const float *const *targets);
int mvs_net_clear_samples(mvs_net *net);

/* One Adam step on samples idx[0..n) (same image size): forward, loss
 * (mean over valid pixels of L1 inverse depth + 0.5 (1 - cos) normal, plus
 * 0.25 x mean over all pixels of the BCE of the confidence against "within
 * 2 % of the target"), backward.  loss4
 * (optional): mean depth L1, mean normal term, mean BCE, valid fraction. */
int mvs_net_train_step(mvs_net *net, const int32_t *idx, int32_t n, float lr, float *loss4);

/* Prediction for one job on the coarse grid: cw*ch*5 floats (inverse depth
 * normalised to the job's range, unit normal, confidence). */
int mvs_net_predict(mvs_net *net, const mvs_job_t *job, float *pred);

/* Gradient probe for tests: forward + loss + backward on a host batch
 * (input n*Hp*Wp*64 floats, rounded to bf16; target n*h*w*5) without an
 * optimizer step.  grads: num_params floats; pred: n*h*w*5; loss4 raw sums. */
int mvs_net_debug_grads(mvs_net *net, const float *input, const float *target, int32_t n, int32_t Hp, int32_t Wp,
                        int32_t h, int32_t w, float *grads, float *pred, float *loss4);

/* estimate_depth with the learned initialisation: coarse sweep volume over
 * the first vol_k sources, network, then the plain plane sweep of every
 * pixel with one more hypothesis where the network is confident
 * (confidence >= conf_thr and both rays face the plane, n . r >= 0.1): the
 * plane proposed at the pixel's coarse pixel, passing through the predicted
 * depth on that pixel's ray, kept when its plane_cost is strictly lower;
 * then 2 * opts->sweeps passes and the gate as in
 * mvs_estimate_depth_batch.  Results stay resident and are copied to the
 * jobs' host pointers where non-NULL.  n_accepted (optional): pixels with
 * an accepted proposal. */
int mvs_estimate_depth_learned(mvs_ctx *ctx, mvs_net *net, const mvs_job_t *jobs, int32_t njobs,
                               const mvs_options_t *opts, int32_t vol_k, float conf_thr, int64_t *n_accepted);
/* mode 1 (what mvs_estimate_depth_learned runs): an accepted proposal is
 * an extra hypothesis -- every pixel runs the plain sweep and the proposal
 * replaces its result only when its plane_cost (in full, without the
 * sweep's early abort) is strictly lower (a 49th hypothesis); mode 0: an
 * accepted proposal replaces the plain sweep at its
 * pixel (no sweep there). */
int mvs_estimate_depth_learned_ex(mvs_ctx *ctx, mvs_net *net, const mvs_job_t *jobs, int32_t njobs,
                                  const mvs_options_t *opts, int32_t vol_k, float conf_thr, int32_t mode,
                                  int64_t *n_accepted);

/* ---- SfM descriptor matching (sfm-lite/src/matching.rs:22-98) ----------
 * Descriptors are 128 f32 each, row-major (SiftFrame.desc, sift.rs).  A
 * match (i, j) of sets a and b: j is i's nearest neighbour in b and i is j's
 * nearest in a (squared L2 summed in f32 in component order, first index
 * on ties), both second-nearest distances are finite, d1(i) <= 0.9,
 * d1(i) <= ratio * d2(i) and d1(j) <= ratio * d2(j) (ratio rounded to f32).
 * Matches are listed in increasing i.  Output slots: pair p owns
 * min(n_a, n_b) consecutive entries of out_a / out_b starting at the sum of
 * min(n_a, n_b) over the pairs before it; out_n[p] receives its count. */
typedef struct mvs_match_stats {   /* the last batch */
    int64_t pairs, filtered_pairs, distances, matches;
    int64_t overflow_items;            /* rows / columns scanned in full after the filter */
    int64_t reserved_i[3];
    double exact_ms;                   /* exact tiles + merge (pairs outside the filter) */
    double tile_ms, collect_ms, recheck_ms;   /* filtered search: MFMA tile pass, candidates, exact recheck */
    double select_ms;                  /* mutual + ratio tests, compaction */
    double host_launch_ms;             /* host: argument checks, pair lists, launches */
    double host_wait_ms;               /* host: waiting for the device */
    double host_out_ms;                /* host: match lists copied into the caller's slots */
    double reserved[4];
} mvs_match_stats_t;

/* match_pair(fa, fb, ratio) (matching.rs:22): host sets of na and nb
 * descriptors; out_a / out_b need min(na, nb) entries (may be NULL to get
 * the count only); *out_n = number of matches. */
int mvs_match_pair(mvs_ctx *ctx, const float *desc_a, int32_t na, const float *desc_b, int32_t nb, double ratio,
                   uint32_t *out_a, uint32_t *out_b, int64_t *out_n);
/* Frames of run_sfm (lib.rs:140-161) made resident: counts[f] descriptors
 * of frame f, concatenated in frame order.  Replaces earlier frames. */
int mvs_upload_descriptors(mvs_ctx *ctx, const float *desc, const int32_t *counts, int32_t nframes);
/* The job-pair loop of run_sfm (lib.rs:197-224) for the matching step:
 * pairs[2p], pairs[2p+1] = frame indices (i, j) of pair p, matched as
 * match_pair(frames[i], frames[j], ratio), all pairs in one batch. */
int mvs_match_frames(mvs_ctx *ctx, const int32_t *pairs, int32_t npairs, double ratio, uint32_t *out_a,
                     uint32_t *out_b, int64_t *out_n);
int mvs_get_match_stats(mvs_ctx *ctx, mvs_match_stats_t *out);
/* Search used for pairs of two sets whose nonzero components all lie in
 * [2^-40, 2^40]: 0 (default) = filtered (bf16x3 MFMA estimate with a proven
 * error bound, exact f32 recheck of every possible nearest / second-nearest
 * descriptor); 1 = exact f32 tiles for every pair.  Other sets always take
 * the exact tiles.  Both give the reference's result bit for bit. */
int mvs_match_set_mode(mvs_ctx *ctx, int32_t mode);
/* Check of the filter's error bound on two safe sets: *max_rel = max over
 * (i, j) of |exact - estimate| / (|a_i|^2 + |b_j|^2) (the filter assumes
 * <= 2^-11). */
int mvs_debug_match_bound(mvs_ctx *ctx, const float *desc_a, int32_t na, const float *desc_b, int32_t nb,
                          float *max_rel);

#ifdef __cplusplus
}
#endif

#endif /* MVS_HIP_H */
