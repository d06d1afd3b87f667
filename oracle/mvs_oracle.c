/*
 * mvs_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, CPU restatement of the reference `mvs` crate's dense path
 * (census patch-match stereo + geometric-consistency fusion), used as the
 * parity checker for the HIP library and as the timed CPU baseline.  Only
 * tests, the smoke check and the benchmark's CPU-baseline leg may load it;
 * the product path never calls into it.
 *
 * Followed line by line (reference paths relative to the reference root):
 *   mvs/src/lib.rs:119-142     census_map                 -> orc_census
 *   mvs/src/stereo.rs:47-78    pixel_rng/next_u64/uniform/gauss
 *   mvs/src/stereo.rs:114-152  RefWorkspace::new          -> orc_ws_new
 *   mvs/src/stereo.rs:158-228  plane_cost                 -> orc_plane_cost
 *   mvs/src/stereo.rs:234-394  estimate_depth             -> orc_estimate_depth
 *   mvs/src/fuse.rs:32-62      sample_depth               -> orc_sample_depth
 *   mvs/src/fuse.rs:64-215     fuse_depth_maps            -> orc_fuse
 *   mvs/src/fuse.rs:72-88      robust voxel cell          -> orc_fusion_cell
 *   mvs/src/fuse.rs:221-239    write_ply_binary           -> orc_write_ply
 *   sfm-lite/src/geom.rs:27-42 Cam::normalize             -> orc_normalize
 *   sfm-lite/src/geom.rs:63-76 Pose::project              -> orc_project
 *   sfm-lite/src/geom.rs:82-96 quat_to_rot                -> orc_quat_to_rot
 *   sfm-lite/src/matching.rs:22-98 match_pair             -> orc_match_pair
 *
 * Float semantics follow Rust: no FMA contraction (build with
 * -ffp-contract=off), correctly rounded / and sqrt, f32 literals where the
 * reference computes in f32, left-to-right evaluation made explicit, and
 * Rust `as` casts (truncate, saturate, NaN -> 0) written out.  nalgebra's
 * 3x3 products are left-to-right dot products (its gemv accumulates column
 * by column).
 *
 * Parity pins (what checks this file, beyond the GPU comparison) are listed
 * in the test modules tests/test_oracle_*.py.  The multi-pass trajectory as
 * a whole, nalgebra's f64 operation order and HashMap output order can only
 * be pinned by running the Rust reference, which cannot be built here.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Boundary types (same layout as include/mvs_hip.h)                         */
/* ------------------------------------------------------------------------ */

typedef struct {
    int32_t w, h;
    double f, cx, cy, k;     /* Cam (geom.rs:13-21) at working resolution */
    double R[9];             /* Pose.r, world->cam, row-major */
    double t[3];             /* Pose.t */
    const uint8_t *gray;     /* w*h luma */
    const uint8_t *rgb;      /* 3*w*h, or NULL => gray replicated */
    const uint64_t *census;  /* w*h, from orc_census */
} orc_view;

typedef struct {            /* MvsOptions, lib.rs:33-54 */
    uint32_t max_dim, num_sources, sweeps, num_hyp, min_consistent, save_depth_vis;
    double depth_diff, voxel_rel;
    float cost_thr;
    uint32_t reserved;
} orc_options;

typedef struct {
    double x, y, z;
    uint8_t r, g, b, pad[5];
} orc_point;

/* ------------------------------------------------------------------------ */
/* Rust cast helpers                                                         */
/* ------------------------------------------------------------------------ */

static inline int32_t rs_f32_to_i32(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
static inline int32_t rs_f64_to_i32(double x) {
    if (x != x) return 0;
    if (x >= 2147483648.0) return INT32_MAX;
    if (x <= -2147483648.0) return INT32_MIN;
    return (int32_t)x;
}
static inline int64_t rs_f64_to_i64(double x) {
    if (x != x) return 0;
    if (x >= 9223372036854775808.0) return INT64_MAX;
    if (x <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}
static inline uint64_t rs_f64_to_usize(double x) {
    if (x != x || x <= 0.0) return 0;
    if (x >= 18446744073709551616.0) return UINT64_MAX;
    return (uint64_t)x;
}
static inline uint8_t rs_f64_to_u8(double x) {
    if (x != x || x <= 0.0) return 0;
    if (x >= 255.0) return 255;
    return (uint8_t)x;
}
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ------------------------------------------------------------------------ */
/* Parallel-for over items (rayon stand-in: rows or views)                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    void (*fn)(void *ctx, int item);
    void *ctx;
    int n;
    int next;
} pf_job;

static void *pf_worker(void *arg) {
    pf_job *j = (pf_job *)arg;
    for (;;) {
        int i = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
        if (i >= j->n) break;
        j->fn(j->ctx, i);
    }
    return NULL;
}

static void parallel_for(int n, int nthreads, void (*fn)(void *, int), void *ctx) {
    pf_job job = {fn, ctx, n, 0};
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (nthreads > n) nthreads = n > 0 ? n : 1;
    pthread_t th[256];
    int started = 0;
    for (int t = 1; t < nthreads; t++) {
        if (pthread_create(&th[started], NULL, pf_worker, &job) == 0) started++;
    }
    pf_worker(&job);
    for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* Census (lib.rs:119-142).  gray[nb] < gray[c] on f32 = u8/255 is          */
/* order-equivalent to comparing the u8 values.                              */
/* ------------------------------------------------------------------------ */

ORC_API void orc_census(const uint8_t *gray, int w, int h, uint64_t *out) {
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            uint8_t c = gray[y * w + x];
            uint64_t bits = 0;
            for (int dy = -3; dy <= 3; dy++) {
                for (int dx = -3; dx <= 3; dx++) {
                    if (dx == 0 && dy == 0) continue;
                    int xx = clampi(x + dx, 0, w - 1);
                    int yy = clampi(y + dy, 0, h - 1);
                    int bit = (dy + 3) * 7 + (dx + 3);
                    if (gray[yy * w + xx] < c) bits |= (uint64_t)1 << bit;
                }
            }
            out[y * w + x] = bits;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* RNG (stereo.rs:47-78)                                                     */
/* ------------------------------------------------------------------------ */

ORC_API uint64_t orc_pixel_rng(int32_t x, int32_t y, uint64_t salt) {
    /* `x as u64` sign-extends an i32 */
    uint64_t s = ((uint64_t)(int64_t)x * 0x9E3779B97F4A7C15ULL) ^
                 ((uint64_t)(int64_t)y * 0xC2B2AE3D27D4EB4FULL) ^
                 (salt * 0x165667B19E3779F9ULL);
    s ^= s >> 30;
    s *= 0xBF58476D1CE4E5B9ULL;
    s ^= s >> 27;
    s *= 0x94D049BB133111EBULL;
    return s ^ (s >> 31);
}

ORC_API uint64_t orc_next_u64(uint64_t *s) {
    uint64_t x = *s;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    *s = x;
    return x;
}

ORC_API float orc_uniform(uint64_t *s) {
    return (float)(orc_next_u64(s) >> 40) / (float)(1ULL << 24);
}

ORC_API float orc_gauss(uint64_t *s) {
    float u1 = orc_uniform(s); /* Rust evaluates the three draws left to right */
    float u2 = orc_uniform(s);
    float u3 = orc_uniform(s);
    return ((u1 + u2) + u3) - 1.5f;
}

/* ------------------------------------------------------------------------ */
/* Camera / pose (geom.rs:27-42, 63-76, 82-96)                               */
/* ------------------------------------------------------------------------ */

ORC_API void orc_normalize(double f, double cx, double cy, double k, double u, double v, double out[2]) {
    double xn = (u - cx) / f;
    double yn = (v - cy) / f;
    for (int it = 0; it < 2; it++) {
        double r2 = xn * xn + yn * yn;
        double d = 1.0 + k * r2;
        if (fabs(d) < 1e-12) break;
        xn = (u - cx) / (f * d);
        yn = (v - cy) / (f * d);
    }
    out[0] = xn;
    out[1] = yn;
}

/* x_cam = R x + t with nalgebra's left-to-right row dot products. */
static inline void mat_vec(const double R[9], const double x[3], double out[3]) {
    for (int i = 0; i < 3; i++) out[i] = (R[3 * i] * x[0] + R[3 * i + 1] * x[1]) + R[3 * i + 2] * x[2];
}
static inline void mat_t_vec(const double R[9], const double x[3], double out[3]) {
    for (int i = 0; i < 3; i++) out[i] = (R[i] * x[0] + R[3 + i] * x[1]) + R[6 + i] * x[2];
}

ORC_API int orc_project(const double R[9], const double t[3], double f, double cx, double cy, double k,
                        const double X[3], double out[2]) {
    double xc[3];
    mat_vec(R, X, xc);
    for (int i = 0; i < 3; i++) xc[i] = xc[i] + t[i];
    double z = xc[2];
    if (z <= 1e-9) return 0;
    double xn = xc[0] / z, yn = xc[1] / z;
    double r2 = xn * xn + yn * yn;
    double d = 1.0 + k * r2;
    out[0] = f * xn * d + cx;
    out[1] = f * yn * d + cy;
    return 1;
}

ORC_API void orc_quat_to_rot(const double q[4], double R[9]) {
    double w = q[0], x = q[1], y = q[2], z = q[3];
    double n = w * w + x * x + y * y + z * z;
    if (fabs(n) - 1.0 > 1e-12) {
        double s = sqrt(n);
        w /= s; x /= s; y /= s; z /= s;
    }
    R[0] = 1.0 - 2.0 * (y * y + z * z); R[1] = 2.0 * (x * y - w * z); R[2] = 2.0 * (x * z + w * y);
    R[3] = 2.0 * (x * y + w * z); R[4] = 1.0 - 2.0 * (x * x + z * z); R[5] = 2.0 * (y * z - w * x);
    R[6] = 2.0 * (x * z - w * y); R[7] = 2.0 * (y * z + w * x); R[8] = 1.0 - 2.0 * (x * x + y * y);
}

/* ------------------------------------------------------------------------ */
/* Workspace (stereo.rs:84-152)                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    const uint64_t *census;
    int w, h;
    float f, cx, cy, k;
    float rs[9];
    float ts[3];
} orc_src;

typedef struct {
    int w, h;
    const uint64_t *census;
    float *rx, *ry;
    int nsrc;
    orc_src *src;
    float zmin, zmax;
} orc_ws;

ORC_API orc_ws *orc_ws_new(const orc_view *views, int nviews, int ref, const int *srcs, int k,
                           double zmin, double zmax) {
    if (ref < 0 || ref >= nviews || k < 1) return NULL;
    const orc_view *rv = &views[ref];
    orc_ws *ws = (orc_ws *)calloc(1, sizeof(orc_ws));
    size_t n = (size_t)rv->w * rv->h;
    ws->w = rv->w;
    ws->h = rv->h;
    ws->census = rv->census;
    ws->rx = (float *)malloc(n * sizeof(float));
    ws->ry = (float *)malloc(n * sizeof(float));
    for (int y = 0; y < rv->h; y++)
        for (int x = 0; x < rv->w; x++) {
            double c[2];
            orc_normalize(rv->f, rv->cx, rv->cy, rv->k, (double)x, (double)y, c);
            ws->rx[(size_t)y * rv->w + x] = (float)c[0];
            ws->ry[(size_t)y * rv->w + x] = (float)c[1];
        }
    ws->nsrc = k;
    ws->src = (orc_src *)calloc((size_t)k, sizeof(orc_src));
    for (int s = 0; s < k; s++) {
        const orc_view *sv = &views[srcs[s]];
        double rs64[9], ts64[3], rt_t[3];
        /* rs64 = R_s * R_r^T ; element (i,j) = sum_m Rs[i,m] * Rr[j,m] */
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                rs64[3 * i + j] = (sv->R[3 * i] * rv->R[3 * j] + sv->R[3 * i + 1] * rv->R[3 * j + 1]) +
                                  sv->R[3 * i + 2] * rv->R[3 * j + 2];
        mat_vec(rs64, rv->t, rt_t);
        for (int i = 0; i < 3; i++) ts64[i] = sv->t[i] - rt_t[i];
        orc_src *o = &ws->src[s];
        o->census = sv->census;
        o->w = sv->w;
        o->h = sv->h;
        o->f = (float)sv->f;
        o->cx = (float)sv->cx;
        o->cy = (float)sv->cy;
        o->k = (float)sv->k;
        for (int i = 0; i < 9; i++) o->rs[i] = (float)rs64[i];
        for (int i = 0; i < 3; i++) o->ts[i] = (float)ts64[i];
    }
    ws->zmin = (float)zmin;
    ws->zmax = (float)zmax;
    return ws;
}

ORC_API void orc_ws_free(orc_ws *ws) {
    if (!ws) return;
    free(ws->rx);
    free(ws->ry);
    free(ws->src);
    free(ws);
}

/* Copy out the f32 workspace parameters (rays and per-source warp). */
ORC_API void orc_ws_export(const orc_ws *ws, float *rx, float *ry, float *src_params /* k*16 */) {
    size_t n = (size_t)ws->w * ws->h;
    if (rx) memcpy(rx, ws->rx, n * sizeof(float));
    if (ry) memcpy(ry, ws->ry, n * sizeof(float));
    if (src_params)
        for (int s = 0; s < ws->nsrc; s++) {
            float *p = src_params + 16 * s;
            const orc_src *o = &ws->src[s];
            for (int i = 0; i < 9; i++) p[i] = o->rs[i];
            for (int i = 0; i < 3; i++) p[9 + i] = o->ts[i];
            p[12] = o->f; p[13] = o->cx; p[14] = o->cy; p[15] = o->k;
        }
}

/* ------------------------------------------------------------------------ */
/* plane_cost (stereo.rs:158-228)                                            */
/* ------------------------------------------------------------------------ */

typedef struct {
    int32_t ham_sum;    /* integer Hamming total over valid taps (full window) */
    int32_t n_invalid;  /* taps that added 1.0 */
    float full_sum;     /* f32 running sum without the abort */
    float peak;         /* max over (source, row) prefixes of partial/denominator */
    int32_t aborted;    /* the reference's early abort fired */
    float result;       /* value plane_cost returns */
} orc_eval_info;

/* Evaluates the plane; when `info` is given the whole window is evaluated
 * (the returned value is identical: the abort fires iff peak > threshold). */
ORC_API float orc_plane_cost(const orc_ws *ws, int px, int py, float nx, float ny, float nz, float d,
                             float hint, orc_eval_info *info) {
    const int w = ws->w, h = ws->h;
    const float nsrc = (float)ws->nsrc;
    const float span = (float)(2 * 2 + 1);
    const float total = span * span * nsrc;
    const float thr = hint + 0.12f;
    float sum = 0.0f;
    int aborted = 0;
    int32_t ham = 0, ninv = 0;
    float peak = -INFINITY;
    for (int si = 0; si < ws->nsrc && !(aborted && !info); si++) {
        const orc_src *s = &ws->src[si];
        const int sw = s->w, sh = s->h;
        const float f = s->f, cx = s->cx, cy = s->cy, k = s->k;
        const float *rs = s->rs, *ts = s->ts;
        for (int dy = -2; dy <= 2; dy++) {
            const int ycl = clampi(py + dy, 0, h - 1);
            for (int dx = -2; dx <= 2; dx++) {
                const int xcl = clampi(px + dx, 0, w - 1);
                const size_t j = (size_t)ycl * w + xcl;
                const float xo = ws->rx[j], yo = ws->ry[j];
                const float denom = (nx * xo + ny * yo) + nz;
                if (denom < 0.05f) { sum += 1.0f; ninv++; continue; }
                const float z = d / denom;
                const float xr = z * xo, yr = z * yo;
                const float xs = ((rs[0] * xr + rs[1] * yr) + rs[2] * z) + ts[0];
                const float ys = ((rs[3] * xr + rs[4] * yr) + rs[5] * z) + ts[1];
                const float zs = ((rs[6] * xr + rs[7] * yr) + rs[8] * z) + ts[2];
                if (zs < 1e-3f) { sum += 1.0f; ninv++; continue; }
                const float xn = xs / zs, yn = ys / zs;
                const float dist = 1.0f + k * (xn * xn + yn * yn);
                const int32_t u = rs_f32_to_i32(f * xn * dist + cx);
                const int32_t v = rs_f32_to_i32(f * yn * dist + cy);
                if (u < 0 || v < 0 || u >= sw || v >= sh) { sum += 1.0f; ninv++; continue; }
                const uint64_t csrc = s->census[(size_t)v * sw + u];
                const uint64_t cref = ws->census[j];
                const int pc = __builtin_popcountll(cref ^ csrc);
                ham += pc;
                sum += (float)pc / 48.0f;
            }
            const float rows_done = (float)(dy + 2 + 1);
            const float part = sum / (rows_done * span * (float)(si + 1));
            if (part > peak) peak = part;
            if (part > thr) {
                aborted = 1;
                if (!info) break;
            }
        }
    }
    const float res = aborted ? 1e9f : sum / total;
    if (info) {
        info->ham_sum = ham;
        info->n_invalid = ninv;
        info->full_sum = sum;
        info->peak = peak;
        info->aborted = aborted;
        info->result = res;
    }
    return res;
}

/* ------------------------------------------------------------------------ */
/* estimate_depth (stereo.rs:234-394)                                        */
/* ------------------------------------------------------------------------ */

typedef struct { float nx, ny, nz, z, cost; } orc_plane;

typedef struct {
    const orc_ws *ws;
    orc_plane *prev, *next;
    float inv_min, inv_max;
    int nh;
    uint64_t pass;
    float sig_inv, sig_n;
    int64_t n_eval;
    int y_lo;          /* first row of the computed band */
    int variant;       /* 0 = the reference; >0 = a planted mistake (orc_run_passes) */
    float *dpl;        /* variants 5, 6: the d_plane each pixel's plane was costed with */
    uint8_t *dok;      /* ... where known (fronto-parallel or accepted in this run) */
} est_ctx;

/* Variants of orc_run_passes, tests only: the planted mistakes (1-4) that
 * the pass pins reject, and the device's skipping of repeated neighbour
 * planes (5, 6), which must change no plane.  A neighbour candidate whose
 * (n, d_plane) bits equal those of the current plane with the d_plane it was
 * costed with (z for (+0, +0, 1, z); for other planes known only once
 * accepted in this run), or of the first neighbour evaluated at the pixel,
 * is not evaluated; 5 counts it as an evaluation (as the device does), 6
 * does not (so the tests can see how many were skipped). */
enum {
    ORC_PASS_REFERENCE = 0,
    ORC_PASS_IN_PLACE = 1,        /* neighbours and cur read from the buffer being written */
    ORC_PASS_DRAW_ALL = 2,        /* all 12 uniforms drawn before the inv_new / zcand test */
    ORC_PASS_PERTURB_BEST = 3,    /* perturbations centred on the best plane so far */
    ORC_PASS_STORE_ZCAND = 4,     /* the accepted candidate's zcand stored instead of bd / a_here */
    ORC_PASS_SKIP_REPEATS = 5,    /* repeated neighbour planes counted, not evaluated */
    ORC_PASS_SKIP_UNCOUNTED = 6,  /* the same, not counted */
};

static uint32_t f32_bits(float v) {
    uint32_t u;
    memcpy(&u, &v, 4);
    return u;
}

static void init_row(void *c, int item) {
    est_ctx *e = (est_ctx *)c;
    const int y = e->y_lo + item;
    const orc_ws *ws = e->ws;
    int64_t ne = 0;
    for (int x = 0; x < ws->w; x++) {
        float best_c = INFINITY, best_d = 0.0f;
        for (int hi = 0; hi < e->nh; hi++) {
            const float inv = e->inv_min + (e->inv_max - e->inv_min) * (float)hi / (float)(e->nh - 1);
            const float d = 1.0f / inv;
            const float cst = orc_plane_cost(ws, x, y, 0.0f, 0.0f, 1.0f, d, best_c + 0.02f, NULL);
            ne++;
            if (cst < best_c) { best_c = cst; best_d = d; }
        }
        orc_plane p = {0.0f, 0.0f, 1.0f, best_d, fminf(best_c, 1.0f)};
        e->prev[(size_t)y * ws->w + x] = p;
    }
    __atomic_fetch_add(&e->n_eval, ne, __ATOMIC_RELAXED);
}

static void pass_row(void *c, int item) {
    est_ctx *e = (est_ctx *)c;
    const int y = e->y_lo + item;
    const orc_ws *ws = e->ws;
    const int w = ws->w, h = ws->h;
    const float *rx = ws->rx, *ry = ws->ry;
    const float zlo = ws->zmin * 0.5f, zhi = ws->zmax * 2.0f;
    int64_t ne = 0;
    static const int NB[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    const int var = e->variant;
    /* the reference reads every plane from the previous buffer */
    const orc_plane *src = var == ORC_PASS_IN_PLACE ? e->next : e->prev;
    for (int x = 0; x < w; x++) {
        const size_t i = (size_t)y * w + x;
        uint64_t seed = orc_pixel_rng(x, y, e->pass + 7);
        const orc_plane cur = src[i];
        float best_c = cur.cost;
        float bn0 = cur.nx, bn1 = cur.ny, bn2 = cur.nz;
        float bd = cur.z * ((bn0 * rx[i] + bn1 * ry[i]) + bn2);
        float bz = cur.z;   /* zcand of the best candidate (planted mistake 4 only) */
        /* repeat skipping (variants 5, 6): the planes a repeat is compared with */
        const int skip = var == ORC_PASS_SKIP_REPEATS || var == ORC_PASS_SKIP_UNCOUNTED;
        const int cur_known = skip && e->dok[i];
        const uint32_t c0 = f32_bits(cur.nx), c1 = f32_bits(cur.ny), c2 = f32_bits(cur.nz),
                       c3 = skip ? f32_bits(e->dpl[i]) : 0u;
        uint32_t f0 = 0u, f1 = 0u, f2 = 0u, f3 = 0u;
        int have = 0;
        for (int q = 0; q < 4; q++) {
            const int xx = x + NB[q][0], yy = y + NB[q][1];
            if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
            const size_t j = (size_t)yy * w + xx;
            const orc_plane nb = src[j];
            const float a_nb = (nb.nx * rx[j] + nb.ny * ry[j]) + nb.nz;
            if (a_nb < 0.1f) continue;
            const float d_plane = nb.z * a_nb;
            const float a_here = (nb.nx * rx[i] + nb.ny * ry[i]) + nb.nz;
            if (a_here < 0.1f) continue;
            const float zcand = d_plane / a_here;
            if (!(zlo <= zcand && zcand <= zhi)) continue;
            if (skip) {
                const uint32_t u0 = f32_bits(nb.nx), u1 = f32_bits(nb.ny), u2 = f32_bits(nb.nz), u3 = f32_bits(d_plane);
                if ((cur_known && u0 == c0 && u1 == c1 && u2 == c2 && u3 == c3) ||
                    (have && u0 == f0 && u1 == f1 && u2 == f2 && u3 == f3)) {
                    if (var == ORC_PASS_SKIP_REPEATS) ne++;
                    continue;
                }
                if (!have) { f0 = u0; f1 = u1; f2 = u2; f3 = u3; have = 1; }
            }
            const float cst = orc_plane_cost(ws, x, y, nb.nx, nb.ny, nb.nz, d_plane, best_c + 0.02f, NULL);
            ne++;
            if (cst < best_c) { best_c = cst; bn0 = nb.nx; bn1 = nb.ny; bn2 = nb.nz; bd = d_plane; bz = zcand; }
        }
        for (int r = 0; r < 3; r++) {
            /* the reference perturbs cur; planted mistake 3 perturbs the best plane so far */
            float cz = cur.z, cnx = cur.nx, cny = cur.ny, cnz = cur.nz;
            if (var == ORC_PASS_PERTURB_BEST) {
                cnx = bn0; cny = bn1; cnz = bn2;
                cz = bd / ((bn0 * rx[i] + bn1 * ry[i]) + bn2);
            }
            const float inv = 1.0f / fmaxf(cz, 1e-6f);
            const float g0 = orc_gauss(&seed);
            float g1 = 0.0f, g2 = 0.0f, g3 = 0.0f;
            if (var == ORC_PASS_DRAW_ALL) {
                g1 = orc_gauss(&seed);
                g2 = orc_gauss(&seed);
                g3 = orc_gauss(&seed);
            }
            const float inv_new = inv + e->sig_inv * (e->inv_max - e->inv_min) * g0;
            if (inv_new <= 0.0f) continue;
            const float zcand = 1.0f / inv_new;
            if (!(zlo <= zcand && zcand <= zhi)) continue;
            if (var != ORC_PASS_DRAW_ALL) g1 = orc_gauss(&seed);
            float nnx = cnx + e->sig_n * g1;
            if (var != ORC_PASS_DRAW_ALL) g2 = orc_gauss(&seed);
            float nny = cny + e->sig_n * g2;
            if (var != ORC_PASS_DRAW_ALL) g3 = orc_gauss(&seed);
            float nnz = cnz + e->sig_n * g3;
            const float len = sqrtf((nnx * nnx + nny * nny) + nnz * nnz);
            if (len < 1e-6f) continue;
            nnx = nnx / len; nny = nny / len; nnz = nnz / len;
            const float a_here = (nnx * rx[i] + nny * ry[i]) + nnz;
            if (a_here < 0.1f) continue;
            const float d_plane = zcand * a_here;
            const float cst = orc_plane_cost(ws, x, y, nnx, nny, nnz, d_plane, best_c + 0.02f, NULL);
            ne++;
            if (cst < best_c) { best_c = cst; bn0 = nnx; bn1 = nny; bn2 = nnz; bd = d_plane; bz = zcand; }
        }
        if (best_c < cur.cost) {
            const float a_here = (bn0 * rx[i] + bn1 * ry[i]) + bn2;
            if (a_here > 0.05f) {
                orc_plane p = {bn0, bn1, bn2, var == ORC_PASS_STORE_ZCAND ? bz : bd / a_here, fminf(best_c, 1.0f)};
                e->next[i] = p;
                if (skip) { e->dpl[i] = bd; e->dok[i] = 1; }   /* this pixel's own entry */
            }
        }
    }
    __atomic_fetch_add(&e->n_eval, ne, __ATOMIC_RELAXED);
}

/* Decay schedule (stereo.rs:275-277), computed with libm powf as Rust does. */
ORC_API void orc_schedule(uint32_t pass, uint32_t n_passes, float *sig_inv, float *sig_n) {
    const float t = (float)pass / (float)n_passes;
    *sig_inv = 0.02f * powf(0.28f, t * 4.0f);
    *sig_n = 0.30f * powf(0.42f, t * 4.0f);
}

/*
 * Per-view estimate.  `max_passes` < 0 runs all 2*sweeps passes; a smaller
 * value stops early (the schedule still uses the full pass count).
 * `planes` (5 floats/px, pre-gate), `normals` (3 floats/px) and `n_eval`
 * (1 + n_passes counters: init, then each pass) are optional outputs.
 *
 * Row band: only rows [row_lo, row_hi) are computed (row_hi <= 0 = all).
 * Window taps still read the whole image; planes outside the band stay
 * zero.  A pass reads rows y-1..y+1 of the previous pass only, so after P
 * passes the rows at distance >= P from the band edges (that are not image
 * edges) equal the full-image result exactly.
 */
ORC_API int orc_estimate_depth_band(const orc_view *views, int nviews, int ref, const int *srcs, int k,
                                    double zmin, double zmax, const orc_options *opts, int max_passes,
                                    int nthreads, int row_lo, int row_hi, float *depth, float *cost,
                                    float *normals, float *planes, int64_t *n_eval) {
    orc_ws *ws = orc_ws_new(views, nviews, ref, srcs, k, zmin, zmax);
    if (!ws) return -1;
    const int w = ws->w, h = ws->h;
    const size_t n = (size_t)w * h;
    orc_plane *prev = (orc_plane *)calloc(n, sizeof(orc_plane));
    orc_plane *next = (orc_plane *)calloc(n, sizeof(orc_plane));
    est_ctx e;
    memset(&e, 0, sizeof(e));
    e.ws = ws;
    e.prev = prev;
    e.inv_min = 1.0f / ws->zmax;
    e.inv_max = 1.0f / ws->zmin;
    e.nh = (int)(opts->num_hyp < 2 ? 2 : opts->num_hyp);
    if (row_hi <= 0 || row_hi > h) row_hi = h;
    if (row_lo < 0) row_lo = 0;
    if (row_lo >= row_hi) { free(prev); free(next); orc_ws_free(ws); return -1; }
    e.y_lo = row_lo;
    parallel_for(row_hi - row_lo, nthreads, init_row, &e);
    if (n_eval) n_eval[0] = e.n_eval;

    const uint32_t n_passes = (opts->sweeps < 1 ? 1 : opts->sweeps) * 2;
    const uint32_t run = (max_passes < 0 || (uint32_t)max_passes > n_passes) ? n_passes : (uint32_t)max_passes;
    for (uint32_t pass = 0; pass < run; pass++) {
        orc_schedule(pass, n_passes, &e.sig_inv, &e.sig_n);
        memcpy(next, prev, n * sizeof(orc_plane));
        e.prev = prev;
        e.next = next;
        e.pass = pass;
        e.n_eval = 0;
        parallel_for(row_hi - row_lo, nthreads, pass_row, &e);
        if (n_eval) n_eval[1 + pass] = e.n_eval;
        orc_plane *tmp = prev; prev = next; next = tmp;
    }
    if (planes)
        for (size_t i = 0; i < n; i++) {
            planes[5 * i + 0] = prev[i].nx; planes[5 * i + 1] = prev[i].ny; planes[5 * i + 2] = prev[i].nz;
            planes[5 * i + 3] = prev[i].z; planes[5 * i + 4] = prev[i].cost;
        }
    /* texture gate + cost threshold + depth clamp (stereo.rs:378-388) */
    for (size_t i = 0; i < n; i++) {
        const int tex = __builtin_popcountll(ws->census[i]);
        const int textured = tex >= 4 && tex <= 44;
        const int ok = textured && prev[i].cost < opts->cost_thr && prev[i].z > ws->zmin * 0.5f &&
                       prev[i].z < ws->zmax * 2.0f;
        if (!ok) prev[i].z = 0.0f;
        if (depth) depth[i] = prev[i].z;
        if (cost) cost[i] = prev[i].cost;
        if (normals) { normals[3 * i] = prev[i].nx; normals[3 * i + 1] = prev[i].ny; normals[3 * i + 2] = prev[i].nz; }
    }
    free(prev);
    free(next);
    orc_ws_free(ws);
    return 0;
}

/*
 * Passes pass_lo .. pass_hi-1 of a run of n_passes passes (stereo.rs:271-371),
 * started from caller-given planes (5 floats/px: nx, ny, nz, z, cost) that
 * stand in for the init sweep; `planes` is updated in place.  n_eval gets
 * one counter per pass run.  `variant` 0 is the reference; the others are
 * the planted mistakes of the enum above, for the tests that show the pass
 * pins reject them (ORC_PASS_IN_PLACE is only meaningful with nthreads 1).
 */
ORC_API int orc_run_passes(const orc_view *views, int nviews, int ref, const int *srcs, int k, double zmin,
                           double zmax, int pass_lo, int pass_hi, int n_passes, int nthreads, int variant,
                           float *planes, int64_t *n_eval) {
    if (pass_lo < 0 || pass_hi < pass_lo || n_passes < 1 || !planes) return -1;
    orc_ws *ws = orc_ws_new(views, nviews, ref, srcs, k, zmin, zmax);
    if (!ws) return -1;
    const size_t n = (size_t)ws->w * ws->h;
    orc_plane *prev = (orc_plane *)malloc(n * sizeof(orc_plane));
    orc_plane *next = (orc_plane *)malloc(n * sizeof(orc_plane));
    for (size_t i = 0; i < n; i++) {
        orc_plane p = {planes[5 * i], planes[5 * i + 1], planes[5 * i + 2], planes[5 * i + 3], planes[5 * i + 4]};
        prev[i] = p;
    }
    est_ctx e;
    memset(&e, 0, sizeof(e));
    e.ws = ws;
    e.inv_min = 1.0f / ws->zmax;
    e.inv_max = 1.0f / ws->zmin;
    e.variant = variant;
    if (variant == ORC_PASS_SKIP_REPEATS || variant == ORC_PASS_SKIP_UNCOUNTED) {
        /* known d_plane: z for the fronto-parallel planes given; none for others */
        e.dpl = (float *)malloc(n * sizeof(float));
        e.dok = (uint8_t *)malloc(n);
        for (size_t i = 0; i < n; i++) {
            e.dpl[i] = prev[i].z;
            e.dok[i] = f32_bits(prev[i].nx) == 0u && f32_bits(prev[i].ny) == 0u && f32_bits(prev[i].nz) == f32_bits(1.0f);
        }
    }
    for (int pass = pass_lo; pass < pass_hi; pass++) {
        orc_schedule((uint32_t)pass, (uint32_t)n_passes, &e.sig_inv, &e.sig_n);
        memcpy(next, prev, n * sizeof(orc_plane));
        e.prev = prev;
        e.next = next;
        e.pass = (uint64_t)pass;
        e.n_eval = 0;
        parallel_for(ws->h, nthreads, pass_row, &e);
        if (n_eval) n_eval[pass - pass_lo] = e.n_eval;
        orc_plane *tmp = prev; prev = next; next = tmp;
    }
    for (size_t i = 0; i < n; i++) {
        planes[5 * i + 0] = prev[i].nx; planes[5 * i + 1] = prev[i].ny; planes[5 * i + 2] = prev[i].nz;
        planes[5 * i + 3] = prev[i].z; planes[5 * i + 4] = prev[i].cost;
    }
    free(prev);
    free(next);
    free(e.dpl);
    free(e.dok);
    orc_ws_free(ws);
    return 0;
}

ORC_API int orc_estimate_depth(const orc_view *views, int nviews, int ref, const int *srcs, int k,
                               double zmin, double zmax, const orc_options *opts, int max_passes,
                               int nthreads, float *depth, float *cost, float *normals, float *planes,
                               int64_t *n_eval) {
    return orc_estimate_depth_band(views, nviews, ref, srcs, k, zmin, zmax, opts, max_passes, nthreads, 0, 0,
                                   depth, cost, normals, planes, n_eval);
}

/* ------------------------------------------------------------------------ */
/* Fusion (fuse.rs:32-215)                                                   */
/* ------------------------------------------------------------------------ */

ORC_API double orc_sample_depth(const float *depth, int w, int h, double u, double v) {
    const int32_t u0 = rs_f64_to_i32(floor(u));
    const int32_t v0 = rs_f64_to_i32(floor(v));
    if (u0 < 0 || v0 < 0 || (int64_t)u0 + 1 >= w || (int64_t)v0 + 1 >= h) return 0.0;
    const float fx = (float)(u - (double)u0);
    const float fy = (float)(v - (double)v0);
    const float d[4] = {depth[(size_t)v0 * w + u0], depth[(size_t)v0 * w + u0 + 1],
                        depth[(size_t)(v0 + 1) * w + u0], depth[(size_t)(v0 + 1) * w + u0 + 1]};
    const float wt[4] = {(1.0f - fx) * (1.0f - fy), fx * (1.0f - fy), (1.0f - fx) * fy, fx * fy};
    float acc = 0.0f, wsum = 0.0f;
    for (int q = 0; q < 4; q++)
        if (d[q] > 0.0f) { acc += d[q] * wt[q]; wsum += wt[q]; }
    return wsum > 1e-6f ? (double)(acc / wsum) : 0.0;
}

static int cmp_f64(const void *p, const void *q) {
    const double a = *(const double *)p, b = *(const double *)q;
    return a < b ? -1 : (a > b ? 1 : 0);
}

/* fuse.rs:72-88 robust per-axis 2/98 % extent -> cell. */
ORC_API double orc_fusion_cell(const double *xyz, int64_t n, double voxel_rel) {
    if (n <= 0) return -1.0;
    double *ax = (double *)malloc((size_t)n * sizeof(double));
    double ext[3];
    for (int a = 0; a < 3; a++) {
        for (int64_t i = 0; i < n; i++) ax[i] = xyz[3 * i + a];
        /* values only: tie order of the sort does not matter */
        qsort(ax, (size_t)n, sizeof(double), cmp_f64);
        uint64_t il = rs_f64_to_usize((double)n * 0.02), ih = rs_f64_to_usize((double)n * 0.98);
        if (il > (uint64_t)(n - 1)) il = (uint64_t)(n - 1);
        if (ih > (uint64_t)(n - 1)) ih = (uint64_t)(n - 1);
        ext[a] = ax[ih] - ax[il];
    }
    free(ax);
    const double diag = sqrt(ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2]);
    const double cell = diag * voxel_rel;
    return cell > 1e-9 ? cell : 1e-9;
}

typedef struct {
    const orc_view *views;
    const float *const *depth;
    const int *order_off, *order;
    int min_consistent;
    double depth_diff;
    double **cand;     /* per view: 6 doubles per candidate + (as double) pixel index */
    int64_t *ncand;
} fuse_ctx;

static void fuse_view(void *c, int ri) {
    fuse_ctx *e = (fuse_ctx *)c;
    e->ncand[ri] = 0;
    e->cand[ri] = NULL;
    const float *dm = e->depth[ri];
    if (!dm) return;
    const orc_view *v = &e->views[ri];
    const int w = v->w, h = v->h;
    int64_t cap = 1024, nc = 0;
    double *out = (double *)malloc((size_t)cap * 7 * sizeof(double));
    const int nord = e->order_off[ri + 1] - e->order_off[ri];
    const int *ord = e->order + e->order_off[ri];
    double *cons = (double *)malloc((size_t)(nord > 0 ? nord : 1) * 3 * sizeof(double));
    for (int py = 0; py < h; py++)
        for (int px = 0; px < w; px++) {
            const size_t i = (size_t)py * w + px;
            const double zr = (double)dm[i];
            if (zr <= 0.0) continue;
            double nrm[2];
            orc_normalize(v->f, v->cx, v->cy, v->k, (double)px, (double)py, nrm);
            const double x_r[3] = {nrm[0] * zr, nrm[1] * zr, 1.0 * zr};
            const double dx[3] = {x_r[0] - v->t[0], x_r[1] - v->t[1], x_r[2] - v->t[2]};
            double x_w[3];
            mat_t_vec(v->R, dx, x_w);
            int ncons = 0;
            for (int q = 0; q < nord; q++) {
                const int si = ord[q];
                if (!e->depth[si]) continue; /* the reference only lists views with a depth map */
                const orc_view *sv = &e->views[si];
                double x_s[3];
                mat_vec(sv->R, x_w, x_s);
                for (int a = 0; a < 3; a++) x_s[a] = x_s[a] + sv->t[a];
                if (x_s[2] <= 1e-6) continue;
                double uv[2];
                if (!orc_project(sv->R, sv->t, sv->f, sv->cx, sv->cy, sv->k, x_w, uv)) continue;
                const double zs = orc_sample_depth(e->depth[si], sv->w, sv->h, uv[0], uv[1]);
                if (zs <= 0.0) continue;
                const double rel = fabs((x_s[2] - zs) / x_s[2]);
                if (rel < e->depth_diff) {
                    double ns[2];
                    orc_normalize(sv->f, sv->cx, sv->cy, sv->k, uv[0], uv[1], ns);
                    const double xs2[3] = {ns[0] * zs, ns[1] * zs, 1.0 * zs};
                    const double d2[3] = {xs2[0] - sv->t[0], xs2[1] - sv->t[1], xs2[2] - sv->t[2]};
                    mat_t_vec(sv->R, d2, &cons[3 * ncons]);
                    ncons++;
                }
            }
            if (ncons < e->min_consistent) continue;
            double acc[3] = {x_w[0], x_w[1], x_w[2]};
            for (int q = 0; q < ncons; q++)
                for (int a = 0; a < 3; a++) acc[a] = acc[a] + cons[3 * q + a];
            const double m = (double)(ncons + 1);
            for (int a = 0; a < 3; a++) acc[a] = acc[a] / m;
            if (nc == cap) { cap *= 2; out = (double *)realloc(out, (size_t)cap * 7 * sizeof(double)); }
            double *o = out + 7 * nc;
            o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[2];
            if (v->rgb) { o[3] = v->rgb[3 * i]; o[4] = v->rgb[3 * i + 1]; o[5] = v->rgb[3 * i + 2]; }
            else { o[3] = o[4] = o[5] = v->gray[i]; }
            o[6] = (double)i;
            nc++;
        }
    free(cons);
    e->cand[ri] = out;
    e->ncand[ri] = nc;
}

typedef struct {
    int64_t key[3];
    double x, y, z;
    uint16_t r, g, b;
    uint32_t n;
    int used;
} vox_entry;

static inline uint64_t key_hash(const int64_t k[3]) {
    uint64_t hsh = (uint64_t)k[0] * 0x9E3779B97F4A7C15ULL;
    hsh ^= (uint64_t)k[1] * 0xC2B2AE3D27D4EB4FULL + (hsh << 6) + (hsh >> 2);
    hsh ^= (uint64_t)k[2] * 0x165667B19E3779F9ULL + (hsh << 6) + (hsh >> 2);
    hsh ^= hsh >> 29;
    return hsh;
}

/*
 * Runs fusion.  Points are returned in first-insertion order of their voxel
 * (view-major raster order of the first candidate), a deterministic order;
 * the reference's HashMap order is random, so outputs compare as sets keyed
 * by voxel.  *pts_out / *cand_out are malloc'ed (free with orc_free).
 * cand_out (optional) receives 8 doubles per candidate:
 * x, y, z, r, g, b, view, pixel.
 */
ORC_API int64_t orc_fuse(const orc_view *views, int nviews, const float *const *depth, const int *order_off,
                         const int *order, int min_consistent, double depth_diff, double cell, int nthreads,
                         orc_point **pts_out, double **cand_out, int64_t *n_cand_out) {
    fuse_ctx e;
    e.views = views;
    e.depth = depth;
    e.order_off = order_off;
    e.order = order;
    e.min_consistent = min_consistent;
    e.depth_diff = depth_diff;
    e.cand = (double **)calloc((size_t)nviews, sizeof(double *));
    e.ncand = (int64_t *)calloc((size_t)nviews, sizeof(int64_t));
    parallel_for(nviews, nthreads, fuse_view, &e);
    int64_t total = 0;
    for (int v = 0; v < nviews; v++) total += e.ncand[v];
    if (n_cand_out) *n_cand_out = total;
    if (cand_out) {
        double *co = (double *)malloc((size_t)(total > 0 ? total : 1) * 8 * sizeof(double));
        int64_t o = 0;
        for (int v = 0; v < nviews; v++)
            for (int64_t c = 0; c < e.ncand[v]; c++, o++) {
                for (int a = 0; a < 6; a++) co[8 * o + a] = e.cand[v][7 * c + a];
                co[8 * o + 6] = v;
                co[8 * o + 7] = e.cand[v][7 * c + 6];
            }
        *cand_out = co;
    }
    /* voxel hash (fuse.rs:176-202), sequential view-major merge */
    uint64_t cap = 16;
    while (cap < (uint64_t)(2 * total + 2)) cap <<= 1;
    vox_entry *tab = (vox_entry *)calloc(cap, sizeof(vox_entry));
    int64_t *slot_order = (int64_t *)malloc((size_t)(total > 0 ? total : 1) * sizeof(int64_t));
    int64_t nslots = 0;
    for (int v = 0; v < nviews; v++)
        for (int64_t c = 0; c < e.ncand[v]; c++) {
            const double *p = e.cand[v] + 7 * c;
            const int64_t key[3] = {rs_f64_to_i64(floor(p[0] / cell)), rs_f64_to_i64(floor(p[1] / cell)),
                                    rs_f64_to_i64(floor(p[2] / cell))};
            uint64_t hsh = key_hash(key) & (cap - 1);
            for (;;) {
                vox_entry *ent = &tab[hsh];
                if (!ent->used) {
                    ent->used = 1;
                    ent->key[0] = key[0]; ent->key[1] = key[1]; ent->key[2] = key[2];
                    slot_order[nslots++] = (int64_t)hsh;
                    break;
                }
                if (ent->key[0] == key[0] && ent->key[1] == key[1] && ent->key[2] == key[2]) break;
                hsh = (hsh + 1) & (cap - 1);
            }
            vox_entry *ent = &tab[hsh];
            ent->x = ent->x + p[0];
            ent->y = ent->y + p[1];
            ent->z = ent->z + p[2];
            ent->r = (uint16_t)(ent->r + (uint16_t)p[3]);
            ent->g = (uint16_t)(ent->g + (uint16_t)p[4]);
            ent->b = (uint16_t)(ent->b + (uint16_t)p[5]);
            ent->n += 1;
        }
    orc_point *pts = (orc_point *)calloc((size_t)(nslots > 0 ? nslots : 1), sizeof(orc_point));
    for (int64_t s = 0; s < nslots; s++) {
        const vox_entry *ent = &tab[slot_order[s]];
        const double n = (double)ent->n;
        pts[s].x = ent->x / n;
        pts[s].y = ent->y / n;
        pts[s].z = ent->z / n;
        pts[s].r = rs_f64_to_u8(round((double)ent->r / n));
        pts[s].g = rs_f64_to_u8(round((double)ent->g / n));
        pts[s].b = rs_f64_to_u8(round((double)ent->b / n));
    }
    free(tab);
    free(slot_order);
    for (int v = 0; v < nviews; v++) free(e.cand[v]);
    free(e.cand);
    free(e.ncand);
    *pts_out = pts;
    return nslots;
}

ORC_API void orc_free(void *p) { free(p); }

/* fuse.rs:221-239: binary little-endian PLY, 15 B per vertex. */
ORC_API int orc_write_ply(const char *path, const orc_point *pts, int64_t n) {
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    fprintf(fp,
            "ply\nformat binary_little_endian 1.0\ncomment rustcolmap dense MVS fusion\n"
            "element vertex %lld\nproperty float x\nproperty float y\nproperty float z\n"
            "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n",
            (long long)n);
    for (int64_t i = 0; i < n; i++) {
        float xyz[3] = {(float)pts[i].x, (float)pts[i].y, (float)pts[i].z};
        uint8_t c[3] = {pts[i].r, pts[i].g, pts[i].b};
        fwrite(xyz, sizeof(float), 3, fp);
        fwrite(c, 1, 3, fp);
    }
    return fclose(fp) == 0 ? 0 : -1;
}

/* ------------------------------------------------------------------------ */
/* match_pair (sfm-lite/src/matching.rs:22-98)                               */
/* ------------------------------------------------------------------------ */

/* Best and second-best squared distance of descriptor `d` against `set`:
 * s = sum over k of (d[k] - e[k])^2 accumulated in f32 in k order; a strictly
 * smaller s replaces the best (so the first of equal minima is kept) and
 * demotes it to second, otherwise a strictly smaller s than second replaces
 * second.  idx stays UINT32_MAX for an empty set. */
static void best_two(const float *d, const float *set, int n, uint32_t *idx, float *best, float *second) {
    float b = INFINITY, s2 = INFINITY;
    uint32_t bi = UINT32_MAX;
    for (int j = 0; j < n; j++) {
        const float *e = set + (size_t)j * 128;
        float s = 0.0f;
        for (int k = 0; k < 128; k++) {
            const float t = d[k] - e[k];
            s += t * t;
        }
        if (s < b) {
            s2 = b;
            b = s;
            bi = (uint32_t)j;
        } else if (s < s2) {
            s2 = s;
        }
    }
    *idx = bi;
    *best = b;
    *second = s2;
}

/* Mutual-best matches with the symmetric ratio test, in increasing index of
 * `a`; returns their number (out_a / out_b need min(na, nb) entries). */
ORC_API int64_t orc_match_pair(const float *da, int na, const float *db, int nb, double ratio_f64, int32_t *out_a,
                               int32_t *out_b) {
    if (na <= 0 || nb <= 0) return 0;
    uint32_t *fi = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)na);
    float *f1 = (float *)malloc(sizeof(float) * (size_t)na), *f2 = (float *)malloc(sizeof(float) * (size_t)na);
    uint32_t *bi = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nb);
    float *b1 = (float *)malloc(sizeof(float) * (size_t)nb), *b2 = (float *)malloc(sizeof(float) * (size_t)nb);
    for (int i = 0; i < na; i++) best_two(da + (size_t)i * 128, db, nb, &fi[i], &f1[i], &f2[i]);
    for (int j = 0; j < nb; j++) best_two(db + (size_t)j * 128, da, na, &bi[j], &b1[j], &b2[j]);
    const float ratio = (float)ratio_f64;
    int64_t n = 0;
    for (int i = 0; i < na; i++) {
        const uint32_t j = fi[i];
        if (j == UINT32_MAX || !isfinite(f2[i])) continue;
        if (f1[i] > 0.9f || f1[i] > ratio * f2[i]) continue;
        if (bi[j] != (uint32_t)i) continue;
        if (!isfinite(b2[j]) || b1[j] > ratio * b2[j]) continue;
        out_a[n] = i;
        out_b[n] = (int32_t)j;
        n++;
    }
    free(fi); free(f1); free(f2); free(bi); free(b1); free(b2);
    return n;
}
