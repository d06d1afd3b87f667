// capi.hip -- context, host orchestration and the C ABI (include/mvs_hip.h).
//
// Host-side duties kept from the reference:
//   * RefWorkspace::new warp parameters (mvs/src/stereo.rs:126-149) in f64
//     with nalgebra's left-to-right products, cast to f32;
//   * the decay schedule (stereo.rs:275-277) with libm powf;
//   * the stage schedule of estimate_depth (init, 2*sweeps passes, gate),
//     recorded once per batch shape into a HIP graph and replayed;
//   * fusion inputs (fusion order, depth-map presence) and the robust voxel
//     cell (fuse.rs:72-88); PLY writing (fuse.rs:221-239).
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mvs_hip.h"
#include "net_internal.h"
#include "match_kernels.h"
#include "net_kernels.h"
#include "stereo_kernels.h"
#include "tuner.h"

using namespace mvs;

// --------------------------------------------------------------------------
// error plumbing
// --------------------------------------------------------------------------

static thread_local std::string g_last_error;

static int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HCHK(call)                                                                                       \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return fail(MVS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));                 \
    } while (0)

// --------------------------------------------------------------------------
// context
// --------------------------------------------------------------------------

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t need(size_t n) {
        if (p && n <= cap) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc((void**)&p, (n ? n : 1) * sizeof(T));
        if (e == hipSuccess) cap = n;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Pinned host staging (hipHostMalloc), grown on demand.
template <class T>
struct HBuf {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t need(size_t n) {
        if (p && n <= cap) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipHostMalloc((void**)&p, (n ? n : 1) * sizeof(T), hipHostMallocDefault);
        if (e == hipSuccess) cap = n;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct ViewState {
    int w = 0, h = 0;
    double f = 0, cx = 0, cy = 0, k = 0, R[9] = {0}, t[3] = {0};
    bool has_rgb = false;
    DBuf<uint8_t> gray, rgb;
    DBuf<u64> census;
    DBuf<float2> rays;
    // stereo state (allocated when the view is a reference view)
    DBuf<float4> plane[2];
    DBuf<float> dplane;     // d_plane of the stored plane where it is not fronto-parallel (JobDesc.dplane)
    DBuf<float> cost[2];
    DBuf<float> depth, cost_out, normal;
    DBuf<uint8_t> work;   // per-pixel pass cost of the sorted pass kernel (scheduling only)
    bool has_depth = false;
    int64_t n_valid = 0;
    void release() {
        gray.release(); rgb.release(); census.release(); rays.release();
        plane[0].release(); plane[1].release(); cost[0].release(); cost[1].release(); dplane.release();
        depth.release(); cost_out.release(); normal.release(); work.release();
        has_depth = false;
    }
};

struct GraphCache {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<uint64_t> sig;   // batch signature the graph was captured for
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        sig.clear();
    }
};

// descriptor matching: resident frames and the batch buffers
struct MatchSets {
    DBuf<float> desc;                  // [total][128]
    DBuf<float> norm;                  // squared norms
    DBuf<uint16_t> split;              // bf16 split: hi [rows][128], then lo [rows][128]
    size_t rows = 0;
    const uint16_t* hi() const { return split.p; }
    const uint16_t* lo() const { return split.p + rows * mt::DIM; }
    std::vector<int64_t> off;          // first row of each set
    std::vector<int32_t> n;            // rows per set
    std::vector<int> safe;             // set may use the filtered search
    void release() {
        desc.release(); norm.release(); split.release();
        rows = 0;
        off.clear(); n.clear(); safe.clear();
    }
};

struct MatchState {
    MatchSets frames;                  // mvs_upload_descriptors
    MatchSets tmp;                     // the two sets of one mvs_match_pair call
    int mode = 0;                      // 0: filtered where safe, 1: exact tiles only
    DBuf<int64_t> set_off;
    DBuf<int> bad;
    DBuf<mt::PairDesc> pairs;
    DBuf<mt::Best2> rpart, cpart, fwd, bwd;
    DBuf<mt::TileMeta> rmeta, cmeta;
    DBuf<uint4> rmask, cmask;
    DBuf<uint32_t> rcnt, ccnt, rcand, ccand, ovf, ovf_n;
    DBuf<uint32_t> oa, ob, da, db;
    DBuf<int64_t> counts, prefix;
    HBuf<mt::PairDesc> h_pairs;        // pinned: pair lists up
    HBuf<int64_t> h_counts;            // pinned: counts (+ overflow count) down
    HBuf<uint32_t> h_out;              // pinned: dense match lists down
    mvs_match_stats_t stats;
    void release() {
        frames.release(); tmp.release(); set_off.release(); bad.release();
        pairs.release(); rpart.release(); cpart.release(); fwd.release(); bwd.release();
        rmeta.release(); cmeta.release(); rmask.release(); cmask.release();
        rcnt.release(); ccnt.release(); rcand.release(); ccand.release(); ovf.release(); ovf_n.release();
        oa.release(); ob.release(); da.release(); db.release(); counts.release(); prefix.release();
        h_pairs.release(); h_counts.release(); h_out.release();
    }
};

struct mvs_net;
static void orphan_net(mvs_net* nt);

struct mvs_ctx {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::vector<ViewState> views;
    DBuf<PrepView> prep;
    // batch descriptors
    DBuf<JobDesc> jobs;
    DBuf<SrcDesc> srcs;
    DBuf<int4> tiles;
    DBuf<u64> counters;   // [0] init, [1 + p] pass p
    DBuf<int> queues;     // persistent-kernel queue heads, one per pass
    GraphCache graph;
    hipEvent_t ev[8] = {nullptr};
    // fusion
    FuseScratch* fuse = nullptr;
    DBuf<FuseView> fviews;
    DBuf<int32_t> forder;
    DBuf<float> tmp_depth;   // host depth maps uploaded for mvs_fuse_depth_maps
    // debug
    DBuf<uint8_t> dbg;
    mvs_stats_t stats;
    Tuner tuner;
    std::vector<mvs_net*> nets;   // proposal networks bound to this context
    MatchState match;
    mvs_log_fn log_fn = nullptr;  // the reference's `log` callback (mvs_set_log)
    void* log_user = nullptr;
    std::vector<std::string> names;   // view names of the last upload (log lines, depth_vis files)
};

static int set_device(mvs_ctx* c) {
    HCHK(hipSetDevice(c->device));
    return MVS_OK;
}

// --------------------------------------------------------------------------
// log lines (the reference's `log` callback) and their Rust number formats
// --------------------------------------------------------------------------

static void emit(mvs_ctx* c, const std::string& line) {
    if (c->log_fn) c->log_fn(line.c_str(), c->log_user);
}

static std::string strf(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return std::string(buf);
}

// `{:?}` of an f64: the shortest digits that read back as x (of the
// candidates with that many digits, the nearest), in positional notation for
// 1e-4 <= |x| < 1e16 (with ".0" when integral), else as <d>[.<ddd>]e<exp>.
static std::string rust_f64_debug(double x) {
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
    if (x == 0.0) return std::signbit(x) ? "-0.0" : "0.0";
    char buf[64];
    std::string digits;
    int exp10 = 0;
    for (int p = 1; p <= 17 && digits.empty(); p++) {
        // glibc's %.*e is correctly rounded: v is the nearest p-digit decimal to x.  When v does not read
        // back as x, a neighbour in the last digit still can (the rounding interval of a power of two is
        // narrower below it than above it).
        snprintf(buf, sizeof(buf), "%.*e", p - 1, fabs(x));
        std::string m;
        for (const char* q = buf; *q && *q != 'e'; q++)
            if (*q != '.') m += *q;
        const int e = atoi(strchr(buf, 'e') + 1);
        const long long v = atoll(m.c_str());
        const long long order[3] = {v, v + 1, v - 1};
        for (long long cand : order) {
            if (cand <= 0) continue;
            char tmp[64];
            snprintf(tmp, sizeof(tmp), "%llde%d", cand, e - (p - 1));
            if (strtod(tmp, nullptr) != fabs(x)) continue;
            const std::string cs = std::to_string(cand);
            exp10 = e + ((int)cs.size() - p);   // a carry (9..9 + 1) adds a digit
            digits = cs;
            break;
        }
    }
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    const std::string sign = x < 0 ? "-" : "";
    const double ax = fabs(x);
    if (ax < 1e-4 || ax >= 1e16) {
        std::string r = sign + digits.substr(0, 1);
        if (digits.size() > 1) r += "." + digits.substr(1);
        return r + "e" + std::to_string(exp10);
    }
    const int point = exp10 + 1;   // digits before the decimal point
    if (point <= 0) return sign + "0." + std::string(-point, '0') + digits;
    if (point >= (int)digits.size()) return sign + digits + std::string(point - digits.size(), '0') + ".0";
    return sign + digits.substr(0, point) + "." + digits.substr(point);
}

// Path::file_stem of a view name (the depth_vis file name, fuse.rs:262-266)
static std::string file_stem(std::string name) {
    while (name.size() > 1 && name.back() == '/') name.pop_back();   // "dir/" names "dir"
    const size_t sl = name.find_last_of('/');
    const std::string file = sl == std::string::npos ? name : name.substr(sl + 1);
    if (file.empty() || file == "." || file == "..") return "view";   // no file name
    const size_t dot = file.find_last_of('.');
    if (dot == std::string::npos || dot == 0) return file;          // "a", ".hidden"
    return file.substr(0, dot);                                      // "a.b.png" -> "a.b"
}

// save_depth_vis pixels (fuse.rs:243-260): 2 % / 98 % of the valid depths,
// near = bright, invalid = 0, f32 arithmetic, `as u8` truncation.
static std::vector<uint8_t> depth_vis_u8(const float* d, size_t n) {
    std::vector<float> ds;
    for (size_t i = 0; i < n; i++)
        if (d[i] > 0.0f) ds.push_back(d[i]);
    std::sort(ds.begin(), ds.end());
    float lo = 0.0f, hi = 1.0f;
    if (ds.size() >= 32) {
        lo = ds[(size_t)((double)ds.size() * 0.02)];
        hi = ds[(size_t)((double)ds.size() * 0.98) - 1];
    }
    const float range = std::max(hi - lo, 1e-9f);
    std::vector<uint8_t> out(n);
    for (size_t i = 0; i < n; i++) {
        const float z = d[i];
        if (!(z > 0.0f)) {
            out[i] = 0;
            continue;
        }
        float t = (z - lo) / range;
        t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
        const float v = 255.0f * (1.0f - t);
        out[i] = (uint8_t)(v >= 255.0f ? 255 : (v > 0.0f ? (int)v : 0));
    }
    return out;
}

// 8-bit grayscale PNG (filter type 0 rows, zlib stream)
static bool write_png_gray8(const std::string& path, const uint8_t* img, int w, int h) {
    std::vector<uint8_t> raw((size_t)h * (w + 1));
    for (int y = 0; y < h; y++) {
        raw[(size_t)y * (w + 1)] = 0;
        memcpy(&raw[(size_t)y * (w + 1) + 1], img + (size_t)y * w, w);
    }
    uLongf zlen = compressBound(raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), raw.size(), 6) != Z_OK) return false;
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    auto be32 = [](uint32_t v, uint8_t* o) { o[0] = v >> 24; o[1] = v >> 16; o[2] = v >> 8; o[3] = v; };
    auto chunk = [&](const char* tag, const uint8_t* data, uint32_t len) {
        uint8_t hdr[8];
        be32(len, hdr);
        memcpy(hdr + 4, tag, 4);
        uLong crc = crc32(0L, (const Bytef*)tag, 4);
        if (len) crc = crc32(crc, data, len);
        uint8_t tail[4];
        be32((uint32_t)crc, tail);
        fwrite(hdr, 1, 8, f);
        if (len) fwrite(data, 1, len, f);
        fwrite(tail, 1, 4, f);
    };
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    fwrite(sig, 1, 8, f);
    uint8_t ihdr[13];
    be32((uint32_t)w, ihdr);
    be32((uint32_t)h, ihdr + 4);
    ihdr[8] = 8; ihdr[9] = 0; ihdr[10] = 0; ihdr[11] = 0; ihdr[12] = 0;
    chunk("IHDR", ihdr, 13);
    chunk("IDAT", z.data(), (uint32_t)zlen);
    chunk("IEND", nullptr, 0);
    return fclose(f) == 0;
}

// --------------------------------------------------------------------------
// allocation-free census: staging in one file-scope device array per GPU
// --------------------------------------------------------------------------
//
// Layout of `scratch` (every part 256-byte aligned): the u8 image, the u64
// census codes and the one-view PrepView descriptor of the prepare kernel.
// In host code `scratch` names a host-side stand-in, never device memory:
// host code uses only the per-device address from hipGetSymbolAddress.
static constexpr size_t NA_GRAY_OFF = 0;
static constexpr size_t NA_CENSUS_OFF = NA_GRAY_OFF + (((size_t)MVS_NOALLOC_MAX_PIXELS + 255) & ~(size_t)255);
static constexpr size_t NA_DESC_OFF = NA_CENSUS_OFF + (size_t)MVS_NOALLOC_MAX_PIXELS * sizeof(u64);
static constexpr size_t NA_BYTES = NA_DESC_OFF + ((sizeof(PrepView) + 255) & ~(size_t)255);
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

static constexpr int NA_MAX_DEVICES = 64;
struct NoAllocSlot {
    std::mutex mu;            // serialises the calls on one device (they share its scratch)
    bool looked_up = false;
    hipError_t err = hipSuccess;
    unsigned char* ptr = nullptr;
};
static NoAllocSlot g_noalloc[NA_MAX_DEVICES];
static std::mutex g_noalloc_init;

// Once per device, with that device current: the device address of its copy
// of `scratch`.  Called from mvs_ctx_create.
static void noalloc_lookup(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_noalloc_init);
    NoAllocSlot& s = g_noalloc[device];
    if (s.looked_up) return;
    void* p = nullptr;
    s.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    s.ptr = s.err == hipSuccess ? (unsigned char*)p : nullptr;
    s.looked_up = true;
}

extern "C" {

void mvs_options_default(mvs_options_t* o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->max_dim = 900;
    o->num_sources = 3;
    o->sweeps = 4;
    o->num_hyp = 48;
    o->min_consistent = 2;
    o->depth_diff = 0.03;
    o->cost_thr = 0.38f;
    o->voxel_rel = 0.0015;
    o->save_depth_vis = 0;
}

const char* mvs_last_error(void) { return g_last_error.c_str(); }

int mvs_ctx_create(int device, mvs_ctx** out) {
    if (!out) return fail(MVS_ERR_ARG, "mvs_ctx_create: out is NULL");
    *out = nullptr;
    try {
        int n = 0;
        HCHK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) return fail(MVS_ERR_ARG, "mvs_ctx_create: no such device");
        mvs_ctx* c = new mvs_ctx();
        c->device = device;
        if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return fail(MVS_ERR_HIP, "mvs_ctx_create: stream creation failed");
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->n_cu = prop.multiProcessorCount;
        noalloc_lookup(device);
        for (auto& e : c->ev) (void)hipEventCreate(&e);
        memset(&c->stats, 0, sizeof(c->stats));
        *out = c;
        return MVS_OK;
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_ctx_create: ") + e.what());
    }
}

void mvs_ctx_destroy(mvs_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    // networks still alive lose their device memory and their context; a
    // later mvs_net_destroy only frees the handle
    for (mvs_net* nt : c->nets) orphan_net(nt);
    c->nets.clear();
    c->graph.release();
    for (auto& v : c->views) v.release();
    c->prep.release(); c->jobs.release(); c->srcs.release(); c->tiles.release();
    c->counters.release(); c->queues.release(); c->fviews.release(); c->forder.release();
    c->tmp_depth.release(); c->dbg.release();
    c->match.release();
    free_fuse_scratch(c->fuse);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

int mvs_debug_rust_f64(double x, char* out, int32_t cap) {
    if (!out || cap < 1) return fail(MVS_ERR_ARG, "mvs_debug_rust_f64: no output buffer");
    const std::string r = rust_f64_debug(x);
    if ((int32_t)r.size() + 1 > cap) return fail(MVS_ERR_LIMIT, "mvs_debug_rust_f64: buffer too small");
    memcpy(out, r.c_str(), r.size() + 1);
    return MVS_OK;
}

int mvs_set_log(mvs_ctx* c, mvs_log_fn fn, void* user) {
    if (!c) return fail(MVS_ERR_ARG, "mvs_set_log: NULL context");
    std::lock_guard<std::mutex> lk(c->mu);
    c->log_fn = fn;
    c->log_user = user;
    return MVS_OK;
}

// --------------------------------------------------------------------------
// census / upload
// --------------------------------------------------------------------------

int mvs_census(mvs_ctx* c, const uint8_t* gray, int32_t w, int32_t h, uint64_t* out) {
    if (!c || !gray || !out || w <= 0 || h <= 0) return fail(MVS_ERR_ARG, "mvs_census: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int r = set_device(c)) return r;
    const size_t n = (size_t)w * h;
    DBuf<uint8_t> g;
    DBuf<u64> cen;
    DBuf<PrepView> pv;
    if (g.need(n) != hipSuccess || cen.need(n) != hipSuccess || pv.need(1) != hipSuccess) {
        g.release(); cen.release(); pv.release();
        return fail(MVS_ERR_NOMEM, "mvs_census: device allocation failed");
    }
    PrepView v{};
    v.gray = g.p;
    v.census = cen.p;
    v.rays = nullptr;
    v.w = w;
    v.h = h;
    const int nb = prepare_tiles(w, h, v.tiles_x);
    v.tile_off = 0;
    hipError_t e = hipMemcpyAsync(g.p, gray, n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pv.p, &v, sizeof(v), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_prepare(pv.p, 1, nb, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, cen.p, n * sizeof(u64), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    g.release(); cen.release(); pv.release();
    if (e != hipSuccess) return fail(MVS_ERR_HIP, std::string("mvs_census: ") + hipGetErrorString(e));
    return MVS_OK;
}

int mvs_census_noalloc(mvs_ctx* c, const uint8_t* gray, int32_t w, int32_t h, uint64_t* out) {
    if (!c || !gray || !out || w <= 0 || h <= 0) return fail(MVS_ERR_ARG, "mvs_census_noalloc: bad arguments");
    if ((int64_t)w * h > (int64_t)MVS_NOALLOC_MAX_PIXELS)
        return fail(MVS_ERR_LIMIT, "mvs_census_noalloc: image has more than MVS_NOALLOC_MAX_PIXELS (" +
                                       std::to_string((long long)MVS_NOALLOC_MAX_PIXELS) + ") pixels");
    if (c->device < 0 || c->device >= NA_MAX_DEVICES)
        return fail(MVS_ERR_LIMIT, "mvs_census_noalloc: device index beyond the scratch table");
    NoAllocSlot& slot = g_noalloc[c->device];
    if (!slot.ptr)
        return fail(MVS_ERR_HIP, std::string("mvs_census_noalloc: no scratch address for this device: ") +
                                     hipGetErrorString(slot.looked_up ? slot.err : hipErrorNotInitialized));
    std::lock_guard<std::mutex> lk(c->mu);
    std::lock_guard<std::mutex> lk2(slot.mu);
    if (int r = set_device(c)) return r;
    const size_t n = (size_t)w * h;
    uint8_t* d_gray = slot.ptr + NA_GRAY_OFF;
    u64* d_census = (u64*)(slot.ptr + NA_CENSUS_OFF);
    PrepView* d_desc = (PrepView*)(slot.ptr + NA_DESC_OFF);
    PrepView v{};
    v.gray = d_gray;
    v.census = d_census;
    v.rays = nullptr;
    v.w = w;
    v.h = h;
    const int nb = prepare_tiles(w, h, v.tiles_x);
    v.tile_off = 0;
    // every census code of the image is written by the kernel: nothing to clear
    hipError_t e = hipMemcpyAsync(d_gray, gray, n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_desc, &v, sizeof(v), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = launch_prepare(d_desc, 1, nb, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_census, n * sizeof(u64), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(MVS_ERR_HIP, std::string("mvs_census_noalloc: ") + hipGetErrorString(e));
    return MVS_OK;
}

int mvs_upload_views(mvs_ctx* c, const mvs_view_t* views, int32_t n) {
    if (!c || !views || n < 1) return fail(MVS_ERR_ARG, "mvs_upload_views: bad arguments");
    for (int i = 0; i < n; i++) {
        if (views[i].w <= 0 || views[i].h <= 0 || !views[i].gray)
            return fail(MVS_ERR_ARG, "mvs_upload_views: view " + std::to_string(i) + " has no image");
        // the stereo kernels address a census by 32-bit byte offsets (8 B per
        // pixel) computed with 24-bit multiplies
        if ((int64_t)views[i].w * views[i].h > (int64_t)1 << 29 || views[i].w >= (1 << 24) || views[i].h >= (1 << 24))
            return fail(MVS_ERR_LIMIT, "mvs_upload_views: view larger than 2^29 pixels or 2^24 - 1 on a side");
        if (!(views[i].f != 0.0)) return fail(MVS_ERR_ARG, "mvs_upload_views: focal length is 0");
    }
    std::lock_guard<std::mutex> lk(c->mu);
    if (int r = set_device(c)) return r;
    HCHK(hipStreamSynchronize(c->stream));
    c->graph.release();
    if ((int)c->views.size() > n) {
        for (size_t i = n; i < c->views.size(); i++) c->views[i].release();
    }
    c->views.resize(n);
    c->names.assign(n, std::string());
    for (int i = 0; i < n; i++) c->names[i] = views[i].name ? views[i].name : "view" + std::to_string(i);
    std::vector<PrepView> pv(n);
    int tile_off = 0;
    for (int i = 0; i < n; i++) {
        const mvs_view_t& in = views[i];
        ViewState& v = c->views[i];
        const size_t np = (size_t)in.w * in.h;
        if (v.w != in.w || v.h != in.h) {
            v.release();
        }
        v.w = in.w; v.h = in.h; v.f = in.f; v.cx = in.cx; v.cy = in.cy; v.k = in.k;
        memcpy(v.R, in.R, sizeof(v.R));
        memcpy(v.t, in.t, sizeof(v.t));
        v.has_depth = false;
        v.has_rgb = in.rgb != nullptr;
        HCHK(v.gray.need(np));
        HCHK(v.census.need(np));
        HCHK(v.rays.need(np));
        HCHK(hipMemcpyAsync(v.gray.p, in.gray, np, hipMemcpyHostToDevice, c->stream));
        if (in.rgb) {
            HCHK(v.rgb.need(3 * np));
            HCHK(hipMemcpyAsync(v.rgb.p, in.rgb, 3 * np, hipMemcpyHostToDevice, c->stream));
        }
        PrepView& p = pv[i];
        p.gray = v.gray.p;
        p.census = v.census.p;
        p.rays = v.rays.p;
        p.w = in.w; p.h = in.h; p.f = in.f; p.cx = in.cx; p.cy = in.cy; p.k = in.k;
        const int nt = prepare_tiles(in.w, in.h, p.tiles_x);
        p.tile_off = tile_off;
        tile_off += nt;
    }
    HCHK(c->prep.need(n));
    HCHK(hipMemcpyAsync(c->prep.p, pv.data(), n * sizeof(PrepView), hipMemcpyHostToDevice, c->stream));
    HCHK(hipEventRecord(c->ev[0], c->stream));
    HCHK(launch_prepare(c->prep.p, n, tile_off, c->stream));
    HCHK(hipEventRecord(c->ev[1], c->stream));
    HCHK(hipStreamSynchronize(c->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    c->stats.ms_prepare = ms;
    return MVS_OK;
}

int mvs_view_census(mvs_ctx* c, int32_t view, uint64_t* out) {
    if (!c || !out) return fail(MVS_ERR_ARG, "mvs_view_census: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (view < 0 || view >= (int)c->views.size()) return fail(MVS_ERR_ARG, "mvs_view_census: no such view");
    if (int r = set_device(c)) return r;
    ViewState& v = c->views[view];
    HCHK(hipMemcpyAsync(out, v.census.p, (size_t)v.w * v.h * sizeof(u64), hipMemcpyDeviceToHost, c->stream));
    HCHK(hipStreamSynchronize(c->stream));
    return MVS_OK;
}

}  // extern "C"

// --------------------------------------------------------------------------
// stereo batch
// --------------------------------------------------------------------------

// SourceCtx of stereo.rs:127-148: rs = R_s R_r^T, ts = t_s - rs t_r in f64.
static SrcDesc make_src(const ViewState& rv, const ViewState& sv) {
    double rs64[9], ts64[3], rtt[3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            rs64[3 * i + j] = (sv.R[3 * i] * rv.R[3 * j] + sv.R[3 * i + 1] * rv.R[3 * j + 1]) + sv.R[3 * i + 2] * rv.R[3 * j + 2];
    for (int i = 0; i < 3; i++) rtt[i] = (rs64[3 * i] * rv.t[0] + rs64[3 * i + 1] * rv.t[1]) + rs64[3 * i + 2] * rv.t[2];
    for (int i = 0; i < 3; i++) ts64[i] = sv.t[i] - rtt[i];
    SrcDesc s{};
    s.census = sv.census.p;
    s.w = sv.w;
    s.h = sv.h;
    for (int i = 0; i < 9; i++) s.rs[i] = (float)rs64[i];
    for (int i = 0; i < 3; i++) s.ts[i] = (float)ts64[i];
    s.f = (float)sv.f;
    s.cx = (float)sv.cx;
    s.cy = (float)sv.cy;
    s.k = (float)sv.k;
    return s;
}

struct BatchPlan;
static int counter_words(const BatchPlan& P);

struct BatchPlan {
    std::vector<JobDesc> jobs;
    std::vector<SrcDesc> srcs;
    std::vector<int4> tiles;
    std::vector<uint64_t> sig;
    TileCfg cfg;
    int n_passes = 0, run_passes = 0, nh = 0;
    int grid_p = 0;   // persistent grid size
    int64_t n_pixels = 0, taps = 0;
};

// counters buffer: [0] init evals, [1 + p] pass p evals, then the persistent
// queue heads (one int per pass), then 4 spare words, then one valid-pixel
// count per job
static int valid_base(const BatchPlan& P) { return 1 + P.n_passes + (P.n_passes + 1) / 2 + 1 + 4; }
static int counter_words(const BatchPlan& P) { return valid_base(P) + (int)P.jobs.size(); }

static int plan_batch(mvs_ctx* c, const mvs_job_t* jobs, int njobs, const mvs_options_t* o, int mode, int max_passes,
                      BatchPlan& P) {
    if (c->views.empty()) return fail(MVS_ERR_STATE, "estimate_depth: no views uploaded");
    std::vector<char> seen(c->views.size(), 0);
    P.nh = (int)std::max<uint32_t>(o->num_hyp, 2);
    P.n_passes = (int)std::max<uint32_t>(o->sweeps, 1) * 2;
    P.run_passes = (max_passes < 0 || max_passes > P.n_passes) ? P.n_passes : max_passes;
    // launch shape: tuned per shape class unless the mode pins it
    int maxw = 0, maxh = 0, maxk = 0;
    int64_t total_px = 0;
    for (int j = 0; j < njobs; j++) {
        const mvs_job_t& jb = jobs[j];
        if (jb.ref < 0 || jb.ref >= (int)c->views.size()) return fail(MVS_ERR_ARG, "estimate_depth: bad ref index");
        if (seen[jb.ref]) return fail(MVS_ERR_ARG, "estimate_depth: a view is the reference of two jobs");
        seen[jb.ref] = 1;
        if (jb.k < 1 || !jb.srcs) return fail(MVS_ERR_ARG, "estimate_depth: job needs >= 1 source");
        for (int s = 0; s < jb.k; s++)
            if (jb.srcs[s] < 0 || jb.srcs[s] >= (int)c->views.size())
                return fail(MVS_ERR_ARG, "estimate_depth: bad source index");
        if (!(jb.zmin > 0.0) || !(jb.zmax > 0.0)) return fail(MVS_ERR_ARG, "estimate_depth: zmin/zmax must be > 0");
        maxw = std::max(maxw, c->views[jb.ref].w);
        maxh = std::max(maxh, c->views[jb.ref].h);
        maxk = std::max(maxk, jb.k);
        total_px += (int64_t)c->views[jb.ref].w * c->views[jb.ref].h;
    }
    P.cfg = c->tuner.pick(maxw, maxh, maxk, total_px);
    if ((mode & 3) == 2) {
        if (P.cfg.persistent == 2) P.cfg = TileCfg();   // the sorted shapes have no tile-grid form
        P.cfg.persistent = 0;
    }
    if ((mode & 3) == 3 && P.cfg.persistent == 0) P.cfg.persistent = 1;
    const int TW = P.cfg.tw, TH = P.cfg.th;
    int tile_off = 0;
    for (int j = 0; j < njobs; j++) {
        const mvs_job_t& jb = jobs[j];
        ViewState& rv = c->views[jb.ref];
        const size_t np = (size_t)rv.w * rv.h;
        for (int b = 0; b < 2; b++) {
            HCHK(rv.plane[b].need(np));
            HCHK(rv.cost[b].need(np));
        }
        HCHK(rv.dplane.need(np));
        HCHK(rv.depth.need(np));
        HCHK(rv.cost_out.need(np));
        HCHK(rv.normal.need(3 * np));
        if (P.cfg.persistent == 2 && (!rv.work.p || rv.work.cap < np)) {
            // any contents give the same results; zeros keep the first pass's order spatial
            HCHK(rv.work.need(np));
            HCHK(hipMemsetAsync(rv.work.p, 0, np, c->stream));
        }
        JobDesc J{};
        J.census = rv.census.p;
        J.rays = rv.rays.p;
        J.plane[0] = rv.plane[0].p;
        J.plane[1] = rv.plane[1].p;
        J.cost[0] = rv.cost[0].p;
        J.cost[1] = rv.cost[1].p;
        J.dplane = rv.dplane.p;
        J.depth_out = rv.depth.p;
        J.cost_out = rv.cost_out.p;
        J.normal_out = rv.normal.p;
        J.work = P.cfg.persistent == 2 ? rv.work.p : nullptr;
        J.w = rv.w;
        J.h = rv.h;
        J.nsrc = jb.k;
        J.src_off = (int)P.srcs.size();
        J.zmin = (float)jb.zmin;
        J.zmax = (float)jb.zmax;
        volatile float one = 1.0f;
        J.inv_min = one / J.zmax;
        J.inv_max = one / J.zmin;
        // pinhole fast path of the third evaluator: dist = 1 + 0 * r2 is 1
        // exactly while r2 stays finite; with |rs| <= 1, denom >= 0.05,
        // zs >= 1e-3 and |d| <= 3 zmax (|ray| + 1), r2 is finite whenever
        // 120 zmax (|ray| + 1)^2 + |ts| < 1e16, ensured by these bounds
        {
            double ray_max = 0.0;
            if (rv.k == 0.0 && rv.f != 0.0) {
                const double cand[4] = {fabs(rv.cx), fabs(rv.w - 1 - rv.cx), fabs(rv.cy), fabs(rv.h - 1 - rv.cy)};
                for (double cv : cand) ray_max = std::max(ray_max, cv / fabs(rv.f));
            }
            bool ok = rv.k == 0.0 && rv.f != 0.0 && J.zmax < 1e6f && ray_max < 1e3;
            for (int s2 = 0; s2 < jb.k && ok; s2++) {
                const SrcDesc sd = make_src(rv, c->views[jb.srcs[s2]]);
                for (int i = 0; i < 3; i++) ok = ok && fabsf(sd.ts[i]) < 1e12f;
            }
            // bit 1: every valid tap's divides stay inside the domain where the
            // scaling-free division sequence is IEEE division (zmin >= 1e-6
            // bounds the plane distance below; the bounds above bound it and
            // the projected coordinates above)
            J.pinhole_ok = ok ? (J.zmin >= 1e-6f ? 3 : 1) : 0;
        }
        // cleared by mvs_debug_run_passes, whose planes and costs are the
        // caller's; MVS_COSTS_EXACT=0 clears it everywhere (tests: the repeat
        // rule that relies on it must change no result)
        {
            const char* env = getenv("MVS_COSTS_EXACT");
            J.costs_exact = (env && env[0] == '0') ? 0 : 1;
        }
        J.tiles_x = (rv.w + TW - 1) / TW;
        J.tiles_y = (rv.h + TH - 1) / TH;
        J.tile_off = tile_off;
        for (int s = 0; s < jb.k; s++) P.srcs.push_back(make_src(rv, c->views[jb.srcs[s]]));
        for (int ty = 0; ty < J.tiles_y; ty++)
            for (int tx = 0; tx < J.tiles_x; tx++) P.tiles.push_back(make_int4(j, tx * TW, ty * TH, 0));
        tile_off += J.tiles_x * J.tiles_y;
        P.jobs.push_back(J);
        P.n_pixels += (int64_t)np;
        P.taps += (int64_t)np * 25 * jb.k;
        // signature: everything the captured launch sequence depends on
        P.sig.push_back((uint64_t)jb.ref);
        P.sig.push_back((uint64_t)(uintptr_t)rv.plane[0].p);
        P.sig.push_back((uint64_t)jb.k);
        for (int s = 0; s < jb.k; s++) P.sig.push_back((uint64_t)jb.srcs[s]);
        uint64_t zz;
        memcpy(&zz, &jb.zmin, 8);
        P.sig.push_back(zz);
        memcpy(&zz, &jb.zmax, 8);
        P.sig.push_back(zz);
    }
    P.sig.push_back((uint64_t)P.nh);
    P.sig.push_back((uint64_t)P.n_passes);
    P.sig.push_back((uint64_t)P.run_passes);
    uint32_t thr_bits;
    memcpy(&thr_bits, &o->cost_thr, 4);
    P.sig.push_back(thr_bits);
    P.sig.push_back((uint64_t)(P.cfg.tw * 10000 + P.cfg.th * 100 + P.cfg.persistent * 10 + P.cfg.eval));
    P.grid_p = std::min((int)P.tiles.size(), c->n_cu * persistent_blocks_per_cu(P.cfg));
    P.sig.push_back((uint64_t)P.grid_p);
    P.sig.push_back(P.jobs.empty() ? 0u : (uint64_t)P.jobs[0].costs_exact);
    P.sig.push_back((uint64_t)(uintptr_t)c->jobs.p);
    return MVS_OK;
}

// Records the stage sequence of a batch on c->stream (captured into a graph
// by the caller when requested).  Counters and queue heads are cleared by a
// fill kernel, never by a memset node.
static hipError_t enqueue_batch(mvs_ctx* c, const BatchPlan& P, float cost_thr, bool timed) {
    hipStream_t st = c->stream;
    const int ntiles = (int)P.tiles.size();
    const int ncnt = 1 + P.n_passes;
    hipError_t e = launch_fill_u64(c->counters.p, counter_words(P), 0ULL, st);
    if (e != hipSuccess) return e;
    if (timed) (void)hipEventRecord(c->ev[0], st);
    e = launch_init(P.cfg, c->jobs.p, c->srcs.p, c->tiles.p, ntiles, P.nh, c->counters.p, st);
    if (e != hipSuccess) return e;
    if (timed) (void)hipEventRecord(c->ev[1], st);
    for (int p = 0; p < P.run_passes; p++) {
        const float t = (float)p / (float)P.n_passes;
        const float sig_inv = 0.02f * powf(0.28f, t * 4.0f);
        const float sig_n = 0.30f * powf(0.42f, t * 4.0f);
        e = launch_pass(P.cfg, c->jobs.p, c->srcs.p, c->tiles.p, ntiles, p & 1, (u64)p + 7, sig_inv, sig_n,
                        c->counters.p + 1 + p, ((int*)(c->counters.p + ncnt)) + p, P.grid_p, st);
        if (e != hipSuccess) return e;
    }
    if (timed) (void)hipEventRecord(c->ev[2], st);
    e = launch_gate(P.cfg, c->jobs.p, c->tiles.p, ntiles, P.run_passes & 1, cost_thr, st);
    if (timed) (void)hipEventRecord(c->ev[3], st);
    return e;
}

static int run_batch(mvs_ctx* c, const mvs_job_t* jobs, int njobs, const mvs_options_t* o, int mode, int max_passes,
                     float* planes_out, int64_t* n_eval_out) {
    BatchPlan P;
    if (int r = plan_batch(c, jobs, njobs, o, mode, max_passes, P)) return r;
    const int ncnt = 1 + P.n_passes;
    const int sched = mode & 3;
    const bool use_graph = sched != 1 && !(mode & 4);
    bool reuse = use_graph && c->graph.exec && c->graph.sig == P.sig;
    if (!reuse) {
        c->graph.release();
        HCHK(c->jobs.need(P.jobs.size()));
        HCHK(c->srcs.need(P.srcs.size()));
        HCHK(c->tiles.need(P.tiles.size()));
        HCHK(c->counters.need(counter_words(P)));
        for (size_t j = 0; j < P.jobs.size(); j++) P.jobs[j].valid_count = c->counters.p + valid_base(P) + j;
        // the job array pointer is part of the signature
        P.sig.back() = (uint64_t)(uintptr_t)c->jobs.p;
        HCHK(hipMemcpyAsync(c->jobs.p, P.jobs.data(), P.jobs.size() * sizeof(JobDesc), hipMemcpyHostToDevice, c->stream));
        HCHK(hipMemcpyAsync(c->srcs.p, P.srcs.data(), P.srcs.size() * sizeof(SrcDesc), hipMemcpyHostToDevice, c->stream));
        HCHK(hipMemcpyAsync(c->tiles.p, P.tiles.data(), P.tiles.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
        HCHK(hipStreamSynchronize(c->stream));
    }
    if (sched == 1) {
        // unbatched schedule: every stage launched per job (same kernels)
        float ms_init = 0, ms_pass = 0, ms_gate = 0;
        HCHK(launch_fill_u64(c->counters.p, counter_words(P), 0ULL, c->stream));
        int toff = 0;
        for (int j = 0; j < njobs; j++) {
            const int nt = P.jobs[j].tiles_x * P.jobs[j].tiles_y;
            HCHK(hipEventRecord(c->ev[0], c->stream));
            HCHK(launch_init(P.cfg, c->jobs.p, c->srcs.p, c->tiles.p + toff, nt, P.nh, c->counters.p, c->stream));
            HCHK(hipEventRecord(c->ev[1], c->stream));
            const int grid_p = std::min(nt, c->n_cu * persistent_blocks_per_cu(P.cfg));
            for (int p = 0; p < P.run_passes; p++) {
                const float t = (float)p / (float)P.n_passes;
                const float sig_inv = 0.02f * powf(0.28f, t * 4.0f);
                const float sig_n = 0.30f * powf(0.42f, t * 4.0f);
                HCHK(launch_fill_u64(c->counters.p + ncnt, (P.n_passes + 1) / 2 + 1, 0ULL, c->stream));
                HCHK(launch_pass(P.cfg, c->jobs.p, c->srcs.p, c->tiles.p + toff, nt, p & 1, (u64)p + 7, sig_inv, sig_n,
                                 c->counters.p + 1 + p, ((int*)(c->counters.p + ncnt)) + p, grid_p, c->stream));
            }
            HCHK(hipEventRecord(c->ev[2], c->stream));
            HCHK(launch_gate(P.cfg, c->jobs.p, c->tiles.p + toff, nt, P.run_passes & 1, o->cost_thr, c->stream));
            HCHK(hipEventRecord(c->ev[3], c->stream));
            HCHK(hipStreamSynchronize(c->stream));
            float a = 0, b = 0, d = 0;
            (void)hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
            (void)hipEventElapsedTime(&b, c->ev[1], c->ev[2]);
            (void)hipEventElapsedTime(&d, c->ev[2], c->ev[3]);
            ms_init += a; ms_pass += b; ms_gate += d;
            toff += nt;
        }
        c->stats.ms_init = ms_init;
        c->stats.ms_pass = ms_pass;
        c->stats.ms_gate = ms_gate;
        c->stats.graph_used = 0;
    } else if (!use_graph) {
        // batched schedule launched directly, with per-stage events
        HCHK(enqueue_batch(c, P, o->cost_thr, true));
        HCHK(hipStreamSynchronize(c->stream));
        float a = 0, b = 0, d = 0;
        (void)hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&b, c->ev[1], c->ev[2]);
        (void)hipEventElapsedTime(&d, c->ev[2], c->ev[3]);
        c->stats.ms_init = a;
        c->stats.ms_pass = b;
        c->stats.ms_gate = d;
        c->stats.graph_used = 0;
    } else {
        if (!reuse) {
            HCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = enqueue_batch(c, P, o->cost_thr, false);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(c->stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(MVS_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
            }
            hipGraphExec_t ex = nullptr;
            hipError_t e3 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (e3 != hipSuccess) {
                (void)hipGraphDestroy(g);
                return fail(MVS_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e3));
            }
            c->graph.graph = g;
            c->graph.exec = ex;
            c->graph.sig = P.sig;
        }
        HCHK(hipEventRecord(c->ev[4], c->stream));
        HCHK(hipGraphLaunch(c->graph.exec, c->stream));
        HCHK(hipEventRecord(c->ev[5], c->stream));
        HCHK(hipStreamSynchronize(c->stream));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev[4], c->ev[5]);
        // per-stage split of the graph replay is not observable from events
        // inside it; the total is reported as passes + init.
        c->stats.ms_init = 0;
        c->stats.ms_pass = ms;
        c->stats.ms_gate = 0;
        c->stats.graph_used = 1;
    }
    std::vector<u64> cnt(counter_words(P));
    HCHK(hipMemcpy(cnt.data(), c->counters.p, cnt.size() * sizeof(u64), hipMemcpyDeviceToHost));
    for (int j = 0; j < njobs; j++) c->views[jobs[j].ref].n_valid = (int64_t)cnt[valid_base(P) + j];
    c->stats.n_eval_init = (int64_t)cnt[0];
    c->stats.n_eval_pass = 0;
    for (int p = 0; p < P.run_passes; p++) c->stats.n_eval_pass += (int64_t)cnt[1 + p];
    c->stats.n_pixels = P.n_pixels;
    c->stats.taps_per_eval_sum = P.taps;
    c->stats.n_pass_launches = P.run_passes;
    c->stats.tile_w = P.cfg.tw;
    c->stats.tile_h = P.cfg.th;
    c->stats.persistent = P.cfg.persistent;
    c->stats.evaluator = P.cfg.eval;
    if (n_eval_out) {
        for (int p = 0; p < 1 + P.run_passes; p++) n_eval_out[p] = (int64_t)cnt[p];
    }
    for (int j = 0; j < njobs; j++) {
        ViewState& rv = c->views[jobs[j].ref];
        const size_t np = (size_t)rv.w * rv.h;
        rv.has_depth = true;
        if (jobs[j].depth) HCHK(hipMemcpyAsync(jobs[j].depth, rv.depth.p, np * 4, hipMemcpyDeviceToHost, c->stream));
        if (jobs[j].cost) HCHK(hipMemcpyAsync(jobs[j].cost, rv.cost_out.p, np * 4, hipMemcpyDeviceToHost, c->stream));
        if (jobs[j].normals)
            HCHK(hipMemcpyAsync(jobs[j].normals, rv.normal.p, 3 * np * 4, hipMemcpyDeviceToHost, c->stream));
    }
    if (planes_out && njobs == 1) {
        ViewState& rv = c->views[jobs[0].ref];
        const size_t np = (size_t)rv.w * rv.h;
        const int par = P.run_passes & 1;
        std::vector<float4> pl(np);
        std::vector<float> co(np);
        HCHK(hipMemcpyAsync(pl.data(), rv.plane[par].p, np * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        HCHK(hipMemcpyAsync(co.data(), rv.cost[par].p, np * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HCHK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < np; i++) {
            planes_out[5 * i] = pl[i].x;
            planes_out[5 * i + 1] = pl[i].y;
            planes_out[5 * i + 2] = pl[i].z;
            planes_out[5 * i + 3] = pl[i].w;
            planes_out[5 * i + 4] = co[i];
        }
    }
    HCHK(hipStreamSynchronize(c->stream));
    return MVS_OK;
}

extern "C" {

int mvs_estimate_depth_ex(mvs_ctx* c, int32_t ref, const int32_t* srcs, int32_t k, double zmin, double zmax,
                          const mvs_options_t* opts, int32_t max_passes, float* depth, float* cost, float* normals,
                          float* planes, int64_t* n_eval) {
    if (!c || !opts) return fail(MVS_ERR_ARG, "mvs_estimate_depth: NULL context or options");
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        mvs_job_t jb{};
        jb.ref = ref;
        jb.k = k;
        jb.srcs = srcs;
        jb.zmin = zmin;
        jb.zmax = zmax;
        jb.depth = depth;
        jb.cost = cost;
        jb.normals = normals;
        return run_batch(c, &jb, 1, opts, 0, max_passes, planes, n_eval);
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_estimate_depth: ") + e.what());
    }
}

int mvs_estimate_depth(mvs_ctx* c, int32_t ref, const int32_t* srcs, int32_t k, double zmin, double zmax,
                       const mvs_options_t* opts, float* depth, float* cost, float* normals) {
    return mvs_estimate_depth_ex(c, ref, srcs, k, zmin, zmax, opts, -1, depth, cost, normals, nullptr, nullptr);
}

int mvs_estimate_depth_batch(mvs_ctx* c, const mvs_job_t* jobs, int32_t njobs, const mvs_options_t* opts,
                             int32_t mode) {
    if (!c || !opts || !jobs || njobs < 1) return fail(MVS_ERR_ARG, "mvs_estimate_depth_batch: bad arguments");
    if (mode < 0 || mode > 7) return fail(MVS_ERR_ARG, "mvs_estimate_depth_batch: mode must be 0..7");
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        return run_batch(c, jobs, njobs, opts, mode, -1, nullptr, nullptr);
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_estimate_depth_batch: ") + e.what());
    }
}

int mvs_debug_run_passes(mvs_ctx* c, int32_t ref, const int32_t* srcs, int32_t k, double zmin, double zmax,
                         int32_t pass_lo, int32_t pass_hi, int32_t n_passes, float* planes, int64_t* n_eval) {
    if (!c || !srcs || !planes) return fail(MVS_ERR_ARG, "mvs_debug_run_passes: NULL argument");
    if (pass_lo < 0 || pass_hi < pass_lo || n_passes < 1 || pass_hi - pass_lo > n_passes)
        return fail(MVS_ERR_ARG, "mvs_debug_run_passes: need 0 <= pass_lo <= pass_hi <= pass_lo + n_passes");
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        mvs_job_t jb{};
        jb.ref = ref;
        jb.k = k;
        jb.srcs = srcs;
        jb.zmin = zmin;
        jb.zmax = zmax;
        mvs_options_t o;
        mvs_options_default(&o);
        o.sweeps = (uint32_t)((n_passes + 1) / 2);   // counter room for n_passes passes
        BatchPlan P;
        if (int r = plan_batch(c, &jb, 1, &o, 0, -1, P)) return r;
        // the descriptor buffers are rewritten: a cached graph must not replay them
        c->graph.release();
        HCHK(c->jobs.need(1));
        HCHK(c->srcs.need(P.srcs.size()));
        HCHK(c->tiles.need(P.tiles.size()));
        HCHK(c->counters.need(counter_words(P)));
        P.jobs[0].valid_count = c->counters.p + valid_base(P);
        // planted planes carry whatever costs the caller chose
        P.jobs[0].costs_exact = 0;
        HCHK(hipMemcpyAsync(c->jobs.p, P.jobs.data(), sizeof(JobDesc), hipMemcpyHostToDevice, c->stream));
        HCHK(hipMemcpyAsync(c->srcs.p, P.srcs.data(), P.srcs.size() * sizeof(SrcDesc), hipMemcpyHostToDevice,
                            c->stream));
        HCHK(hipMemcpyAsync(c->tiles.p, P.tiles.data(), P.tiles.size() * sizeof(int4), hipMemcpyHostToDevice,
                            c->stream));
        ViewState& rv = c->views[ref];
        const size_t np = (size_t)rv.w * rv.h;
        std::vector<float4> pl(np);
        std::vector<float> co(np);
        for (size_t i = 0; i < np; i++) {
            pl[i] = make_float4(planes[5 * i], planes[5 * i + 1], planes[5 * i + 2], planes[5 * i + 3]);
            co[i] = planes[5 * i + 4];
        }
        const int par0 = pass_lo & 1;
        HCHK(hipMemcpyAsync(rv.plane[par0].p, pl.data(), np * sizeof(float4), hipMemcpyHostToDevice, c->stream));
        HCHK(hipMemcpyAsync(rv.cost[par0].p, co.data(), np * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HCHK(launch_fill_u64(c->counters.p, counter_words(P), 0ULL, c->stream));
        const int ncnt = 1 + P.n_passes;
        for (int p = pass_lo; p < pass_hi; p++) {
            const float t = (float)p / (float)n_passes;
            const float sig_inv = 0.02f * powf(0.28f, t * 4.0f);
            const float sig_n = 0.30f * powf(0.42f, t * 4.0f);
            const int q = p - pass_lo;
            HCHK(launch_pass(P.cfg, c->jobs.p, c->srcs.p, c->tiles.p, (int)P.tiles.size(), p & 1, (u64)p + 7, sig_inv,
                             sig_n, c->counters.p + 1 + q, ((int*)(c->counters.p + ncnt)) + q, P.grid_p, c->stream));
        }
        const int par1 = pass_hi & 1;
        HCHK(hipMemcpyAsync(pl.data(), rv.plane[par1].p, np * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        HCHK(hipMemcpyAsync(co.data(), rv.cost[par1].p, np * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        std::vector<u64> cnt(counter_words(P));
        HCHK(hipMemcpyAsync(cnt.data(), c->counters.p, cnt.size() * sizeof(u64), hipMemcpyDeviceToHost, c->stream));
        HCHK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < np; i++) {
            planes[5 * i] = pl[i].x;
            planes[5 * i + 1] = pl[i].y;
            planes[5 * i + 2] = pl[i].z;
            planes[5 * i + 3] = pl[i].w;
            planes[5 * i + 4] = co[i];
        }
        if (n_eval)
            for (int q = 0; q < pass_hi - pass_lo; q++) n_eval[q] = (int64_t)cnt[1 + q];
        return MVS_OK;
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_debug_run_passes: ") + e.what());
    }
}

int mvs_autotune(mvs_ctx* c, const mvs_job_t* jobs, int32_t njobs, const mvs_options_t* opts, int32_t passes,
                 int32_t* best, float* cand_ms, int32_t* cand_cfg, int32_t* n_cand) {
    if (!c || !opts || !jobs || njobs < 1) return fail(MVS_ERR_ARG, "mvs_autotune: bad arguments");
    if (passes < 1) passes = 2;
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        // in order of preference (the sorted pass shapes first, 16x16 the
        // fastest of them at 1080p): a later candidate must beat the best so
        // far by 2 %, so a near-tie keeps the preferred shape
        static const int cand[][4] = {{16, 16, 2, 3}, {16, 8, 2, 3}, {24, 8, 2, 3},
                                      {16, 16, 0, 1}, {16, 8, 0, 1}, {8, 8, 0, 1},  {8, 8, 1, 1},  {16, 16, 0, 2},
                                      {16, 8, 0, 2},  {8, 8, 0, 2},  {8, 8, 1, 2},  {16, 16, 0, 0}, {16, 8, 0, 0},
                                      {16, 16, 0, 3}, {16, 16, 0, 4}, {16, 8, 0, 3}, {16, 8, 0, 4}, {8, 8, 0, 3},
                                      {8, 8, 0, 4},   {8, 8, 1, 3},  {8, 8, 1, 4},  {32, 8, 0, 3}, {16, 16, 1, 4}};
        const int ncand = (int)(sizeof(cand) / sizeof(cand[0]));
        const int n_passes = (int)std::max<uint32_t>(opts->sweeps, 1) * 2;
        double best_t = 1e300;
        TileCfg best_cfg;
        int maxw = 0, maxh = 0, maxk = 0;
        for (int j = 0; j < njobs; j++) {
            if (jobs[j].ref < 0 || jobs[j].ref >= (int)c->views.size()) return fail(MVS_ERR_ARG, "mvs_autotune: bad ref");
            maxw = std::max(maxw, c->views[jobs[j].ref].w);
            maxh = std::max(maxh, c->views[jobs[j].ref].h);
            maxk = std::max(maxk, jobs[j].k);
        }
        for (int i = 0; i < ncand; i++) {
            TileCfg cf;
            cf.tw = cand[i][0];
            cf.th = cand[i][1];
            cf.persistent = cand[i][2];
            cf.eval = cand[i][3];
            c->tuner.forced = true;
            c->tuner.force_cfg = cf;
            double t = 0;
            for (int rep = 0; rep < 2; rep++) {
                const int r = run_batch(c, jobs, njobs, opts, 4, passes, nullptr, nullptr);
                if (r) {
                    c->tuner.forced = false;
                    return r;
                }
                t = c->stats.ms_init + c->stats.ms_pass * (double)n_passes / (double)passes;
            }
            if (cand_ms) cand_ms[i] = (float)t;
            if (cand_cfg) {
                cand_cfg[4 * i] = cf.tw;
                cand_cfg[4 * i + 1] = cf.th;
                cand_cfg[4 * i + 2] = cf.persistent;
                cand_cfg[4 * i + 3] = cf.eval;
            }
            // a later candidate must win by 2 % to displace an earlier one, so
            // near-ties resolve the same way run after run (the one-pass
            // estimates of one candidate vary by 1-3 % from run to run)
            if (t < best_t * 0.98) {
                best_t = t;
                best_cfg = cf;
            }
        }
        if (n_cand) *n_cand = ncand;
        c->tuner.forced = false;
        c->tuner.table[Tuner::shape_class(maxw, maxh, maxk)] = best_cfg;
        c->graph.release();
        if (best) {
            best[0] = best_cfg.tw;
            best[1] = best_cfg.th;
            best[2] = best_cfg.persistent;
            best[3] = best_cfg.eval;
        }
        return MVS_OK;
    } catch (const std::exception& e) {
        c->tuner.forced = false;
        return fail(MVS_ERR_NOMEM, std::string("mvs_autotune: ") + e.what());
    }
}

int mvs_get_depth(mvs_ctx* c, int32_t view, float* depth, float* cost, float* normals) {
    if (!c) return fail(MVS_ERR_ARG, "mvs_get_depth: NULL context");
    std::lock_guard<std::mutex> lk(c->mu);
    if (view < 0 || view >= (int)c->views.size()) return fail(MVS_ERR_ARG, "mvs_get_depth: no such view");
    ViewState& v = c->views[view];
    if (!v.has_depth) return fail(MVS_ERR_STATE, "mvs_get_depth: view has no depth map");
    if (int r = set_device(c)) return r;
    const size_t np = (size_t)v.w * v.h;
    if (depth) HCHK(hipMemcpyAsync(depth, v.depth.p, np * 4, hipMemcpyDeviceToHost, c->stream));
    if (cost) HCHK(hipMemcpyAsync(cost, v.cost_out.p, np * 4, hipMemcpyDeviceToHost, c->stream));
    if (normals) HCHK(hipMemcpyAsync(normals, v.normal.p, 3 * np * 4, hipMemcpyDeviceToHost, c->stream));
    HCHK(hipStreamSynchronize(c->stream));
    return MVS_OK;
}

// --------------------------------------------------------------------------
// fusion
// --------------------------------------------------------------------------

static int fuse_common(mvs_ctx* c, const std::vector<const float*>& dmaps, const int32_t* order_off,
                       const int32_t* order, double cell, const mvs_options_t* o, mvs_point_t** pts, int64_t* n) {
    const int nv = (int)c->views.size();
    if (!order_off || !pts || !n) return fail(MVS_ERR_ARG, "fuse: NULL argument");
    if (!(cell > 0.0)) return fail(MVS_ERR_ARG, "fuse: cell must be > 0");
    *pts = nullptr;
    *n = 0;
    if (order_off[0] != 0) return fail(MVS_ERR_ARG, "fuse: order_off[0] must be 0");
    const int total_ord = order_off[nv];
    for (int i = 0; i < nv; i++)
        if (order_off[i + 1] < order_off[i]) return fail(MVS_ERR_ARG, "fuse: order_off must be non-decreasing");
    for (int q = 0; q < total_ord; q++)
        if (order[q] < 0 || order[q] >= nv) return fail(MVS_ERR_ARG, "fuse: bad view index in order");
    std::vector<FuseView> fv(nv);
    int64_t off = 0;
    for (int i = 0; i < nv; i++) {
        const ViewState& v = c->views[i];
        FuseView& f = fv[i];
        f.depth = dmaps[i];
        f.rgb = v.has_rgb ? v.rgb.p : nullptr;
        f.gray = v.gray.p;
        f.w = v.w;
        f.h = v.h;
        f.f = v.f; f.cx = v.cx; f.cy = v.cy; f.k = v.k;
        memcpy(f.R, v.R, sizeof(f.R));
        memcpy(f.t, v.t, sizeof(f.t));
        f.pix_off = off;
        f.ord_off = order_off[i];
        f.ord_n = order_off[i + 1] - order_off[i];
        if (f.depth) off += (int64_t)v.w * v.h;
    }
    HCHK(c->fviews.need(nv));
    HCHK(c->forder.need(total_ord > 0 ? total_ord : 1));
    HCHK(hipMemcpyAsync(c->fviews.p, fv.data(), nv * sizeof(FuseView), hipMemcpyHostToDevice, c->stream));
    if (total_ord > 0)
        HCHK(hipMemcpyAsync(c->forder.p, order, total_ord * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    FuseArgs a;
    a.d_views = c->fviews.p;
    a.nviews = nv;
    a.d_order = c->forder.p;
    a.n_pix = off;
    a.min_consistent = (int)o->min_consistent;
    a.depth_diff = o->depth_diff;
    a.cell = cell;
    int64_t ncand = 0, npts = 0;
    const void* dp = nullptr;
    int lim = 0;
    HCHK(hipEventRecord(c->ev[6], c->stream));
    HCHK(run_fusion(c->fuse, a, fv.data(), c->stream, &ncand, &npts, &dp, &lim));
    HCHK(hipEventRecord(c->ev[7], c->stream));
    HCHK(hipStreamSynchronize(c->stream));
    if (lim == 1) return fail(MVS_ERR_LIMIT, "fuse: more than 2^31-1 depth pixels");
    if (lim == 2) return fail(MVS_ERR_LIMIT, "fuse: voxel grid extent exceeds 2^21 cells on an axis");
    if (lim == 3) return fail(MVS_ERR_HIP, "fuse: voxel hash table overflow");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[6], c->ev[7]);
    c->stats.ms_fuse = ms;
    c->stats.n_candidates = ncand;
    c->stats.n_points = npts;
    {
        int ndm = 0;
        for (const float* d : dmaps) ndm += d != nullptr;
        emit(c, strf("mvs: %lld consistency-passed candidates from %d depth maps", (long long)ncand, ndm));
    }
    mvs_point_t* out = (mvs_point_t*)malloc((size_t)(npts > 0 ? npts : 1) * sizeof(mvs_point_t));
    if (!out) return fail(MVS_ERR_NOMEM, "fuse: host allocation failed");
    if (npts > 0) {
        hipError_t e = hipMemcpyAsync(out, dp, (size_t)npts * sizeof(mvs_point_t), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            free(out);
            return fail(MVS_ERR_HIP, std::string("fuse: ") + hipGetErrorString(e));
        }
    }
    *pts = out;
    *n = npts;
    return MVS_OK;
}

int mvs_fuse(mvs_ctx* c, const int32_t* has_depth, const int32_t* order_off, const int32_t* order, double cell,
             const mvs_options_t* opts, mvs_point_t** pts, int64_t* n) {
    if (!c || !opts || !has_depth) return fail(MVS_ERR_ARG, "mvs_fuse: NULL argument");
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        std::vector<const float*> dm(c->views.size(), nullptr);
        for (size_t i = 0; i < c->views.size(); i++) {
            if (has_depth[i]) {
                if (!c->views[i].has_depth) return fail(MVS_ERR_STATE, "mvs_fuse: view " + std::to_string(i) + " has no depth map");
                dm[i] = c->views[i].depth.p;
            }
        }
        return fuse_common(c, dm, order_off, order, cell, opts, pts, n);
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_fuse: ") + e.what());
    }
}

int mvs_fuse_depth_maps(mvs_ctx* c, const float* const* depth_maps, const int32_t* order_off, const int32_t* order,
                        double cell, const mvs_options_t* opts, mvs_point_t** pts, int64_t* n) {
    if (!c || !opts || !depth_maps) return fail(MVS_ERR_ARG, "mvs_fuse_depth_maps: NULL argument");
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        const int nv = (int)c->views.size();
        size_t total = 0;
        for (int i = 0; i < nv; i++)
            if (depth_maps[i]) total += (size_t)c->views[i].w * c->views[i].h;
        HCHK(c->tmp_depth.need(total));
        std::vector<const float*> dm(nv, nullptr);
        size_t off = 0;
        for (int i = 0; i < nv; i++) {
            if (!depth_maps[i]) continue;
            const size_t np = (size_t)c->views[i].w * c->views[i].h;
            HCHK(hipMemcpyAsync(c->tmp_depth.p + off, depth_maps[i], np * 4, hipMemcpyHostToDevice, c->stream));
            dm[i] = c->tmp_depth.p + off;
            off += np;
        }
        return fuse_common(c, dm, order_off, order, cell, opts, pts, n);
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_fuse_depth_maps: ") + e.what());
    }
}

void mvs_free_points(mvs_point_t* p) { free(p); }

}  // extern "C"

// Robust sparse diagonal of fuse_depth_maps (fuse.rs:72-88).
static double robust_diag(const double* xyz, int64_t n) {
    std::vector<double> ax((size_t)n);
    double ext[3];
    for (int a = 0; a < 3; a++) {
        for (int64_t i = 0; i < n; i++) ax[i] = xyz[3 * i + a];
        std::sort(ax.begin(), ax.end());
        // `(len as f64 * q) as usize`, clamped to len - 1
        const double fl = (double)n * 0.02, fh = (double)n * 0.98;
        uint64_t il = fl > 0 ? (uint64_t)fl : 0, ih = fh > 0 ? (uint64_t)fh : 0;
        il = std::min<uint64_t>(il, (uint64_t)n - 1);
        ih = std::min<uint64_t>(ih, (uint64_t)n - 1);
        ext[a] = ax[ih] - ax[il];
    }
    return sqrt(ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2]);
}

extern "C" {

double mvs_fusion_diag(const double* xyz, int64_t n) {
    if (!xyz || n <= 0) return -1.0;
    return robust_diag(xyz, n);
}

double mvs_fusion_cell(const double* xyz, int64_t n, double voxel_rel) {
    if (!xyz || n <= 0) return -1.0;
    const double cell = robust_diag(xyz, n) * voxel_rel;
    return cell > 1e-9 ? cell : 1e-9;
}

// Decimated JSON copy of a fused cloud for web viewers, as run_mvs writes it
// beside dense.ply (mvs/src/lib.rs:437-485): centroid by f64 sums in point
// order, distances to it, median and 95th-percentile distance of the sorted
// list, keep radius med + 6 max(|p95 - med|, max(med, 1e-9) / 2), then every
// ceil(kept / 400000)-th kept point; positions to 4 decimals, colours as
// "[r, g,b]".  No file for an empty cloud.
int mvs_write_dense_web_json(const char* path, const mvs_point_t* pts, int64_t n) {
    if (!path || (n > 0 && !pts) || n < 0) return fail(MVS_ERR_ARG, "mvs_write_dense_web_json: bad arguments");
    if (n == 0) return MVS_OK;   // the reference writes no file for an empty cloud
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int64_t i = 0; i < n; i++) {
        sx += pts[i].x;
        sy += pts[i].y;
        sz += pts[i].z;
    }
    const double cx = sx / (double)n, cy = sy / (double)n, cz = sz / (double)n;
    std::vector<double> dist((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        const double dx = pts[i].x - cx, dy = pts[i].y - cy, dz = pts[i].z - cz;
        dist[(size_t)i] = sqrt(dx * dx + dy * dy + dz * dz);
    }
    std::vector<double> sorted(dist);
    std::sort(sorted.begin(), sorted.end());
    const double med = sorted[(size_t)n / 2];
    const size_t i95 = std::min((size_t)((double)n * 0.95), (size_t)n - 1);
    const double p95 = sorted[i95];
    const double r_keep = med + 6.0 * std::max(fabs(p95 - med), std::max(med, 1e-9) * 0.5);
    std::vector<int64_t> kept;
    kept.reserve((size_t)n);
    for (int64_t i = 0; i < n; i++)
        if (dist[(size_t)i] <= r_keep) kept.push_back(i);
    const size_t cap = 400000;
    const size_t stride = (kept.size() + cap - 1) / cap;
    std::vector<int64_t> web;
    for (size_t k = 0; k < kept.size(); k += stride) web.push_back(kept[k]);
    std::string s;
    s.reserve(web.size() * 48 + 64);
    char buf[256];
    s += "{\"n\":" + std::to_string(web.size()) + ",\"pos\":[";
    for (size_t k = 0; k < web.size(); k++) {
        const mvs_point_t& p = pts[web[k]];
        snprintf(buf, sizeof buf, "%s[%.4f,%.4f,%.4f]", k ? "," : "", p.x, p.y, p.z);
        s += buf;
    }
    s += "],\"col\":[";
    for (size_t k = 0; k < web.size(); k++) {
        const mvs_point_t& p = pts[web[k]];
        snprintf(buf, sizeof buf, "%s[%u, %u,%u]", k ? "," : "", (unsigned)p.r, (unsigned)p.g, (unsigned)p.b);
        s += buf;
    }
    s += "]}";
    FILE* f = fopen(path, "wb");
    if (!f) return fail(MVS_ERR_ARG, std::string("mvs_write_dense_web_json: cannot create ") + path);
    const bool ok = fwrite(s.data(), 1, s.size(), f) == s.size();
    if (fclose(f) != 0 || !ok) return fail(MVS_ERR_ARG, std::string("mvs_write_dense_web_json: write failed: ") + path);
    return MVS_OK;
}

int mvs_write_ply(const char* path, const mvs_point_t* pts, int64_t n) {
    if (!path || (n > 0 && !pts)) return fail(MVS_ERR_ARG, "mvs_write_ply: bad arguments");
    FILE* fp = fopen(path, "wb");
    if (!fp) return fail(MVS_ERR_ARG, std::string("mvs_write_ply: cannot create ") + path);
    fprintf(fp,
            "ply\nformat binary_little_endian 1.0\ncomment rustcolmap dense MVS fusion\n"
            "element vertex %lld\nproperty float x\nproperty float y\nproperty float z\n"
            "property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n",
            (long long)n);
    std::vector<uint8_t> buf;
    buf.reserve((size_t)std::min<int64_t>(n, 1 << 20) * 15);
    for (int64_t i = 0; i < n; i++) {
        const float xyz[3] = {(float)pts[i].x, (float)pts[i].y, (float)pts[i].z};
        const size_t o = buf.size();
        buf.resize(o + 15);
        memcpy(&buf[o], xyz, 12);
        buf[o + 12] = pts[i].r;
        buf[o + 13] = pts[i].g;
        buf[o + 14] = pts[i].b;
        if (buf.size() >= ((size_t)1 << 20) * 15) {
            fwrite(buf.data(), 1, buf.size(), fp);
            buf.clear();
        }
    }
    if (!buf.empty()) fwrite(buf.data(), 1, buf.size(), fp);
    if (fclose(fp) != 0) return fail(MVS_ERR_ARG, "mvs_write_ply: write failed");
    return MVS_OK;
}

int mvs_get_stats(mvs_ctx* c, mvs_stats_t* out) {
    if (!c || !out) return fail(MVS_ERR_ARG, "mvs_get_stats: NULL argument");
    std::lock_guard<std::mutex> lk(c->mu);
    *out = c->stats;
    return MVS_OK;
}

// --------------------------------------------------------------------------
// instrumentation
// --------------------------------------------------------------------------

int mvs_debug_plane_cost(mvs_ctx* c, int32_t ref, const int32_t* srcs, int32_t k, const float* q, int32_t n,
                         float* cost, int32_t* ham, int32_t* ninv, float* peak) {
    if (!c || !srcs || k < 1 || !q || n < 0 || !cost || !ham || !ninv || !peak)
        return fail(MVS_ERR_ARG, "mvs_debug_plane_cost: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (ref < 0 || ref >= (int)c->views.size()) return fail(MVS_ERR_ARG, "mvs_debug_plane_cost: bad ref");
    for (int s = 0; s < k; s++)
        if (srcs[s] < 0 || srcs[s] >= (int)c->views.size()) return fail(MVS_ERR_ARG, "mvs_debug_plane_cost: bad source");
    const ViewState& rv = c->views[ref];
    for (int i = 0; i < n; i++) {
        const int px = (int)q[7 * i], py = (int)q[7 * i + 1];
        if (px < 0 || py < 0 || px >= rv.w || py >= rv.h) return fail(MVS_ERR_ARG, "mvs_debug_plane_cost: pixel outside view");
    }
    if (int r = set_device(c)) return r;
    std::vector<SrcDesc> S;
    for (int s = 0; s < k; s++) S.push_back(make_src(rv, c->views[srcs[s]]));
    JobDesc J{};
    J.census = rv.census.p;
    J.rays = rv.rays.p;
    J.w = rv.w;
    J.h = rv.h;
    J.nsrc = k;
    const size_t bytes = S.size() * sizeof(SrcDesc) + (size_t)n * (7 * 4 + 16) + 256;
    HCHK(c->dbg.need(bytes));
    uint8_t* base = c->dbg.p;
    SrcDesc* dS = (SrcDesc*)base;
    size_t o = (S.size() * sizeof(SrcDesc) + 255) & ~(size_t)255;
    float* dq = (float*)(base + o);
    o += (size_t)n * 7 * 4;
    float* dc = (float*)(base + o);
    o += (size_t)n * 4;
    int* dh = (int*)(base + o);
    o += (size_t)n * 4;
    int* di = (int*)(base + o);
    o += (size_t)n * 4;
    float* dp = (float*)(base + o);
    HCHK(hipMemcpyAsync(dS, S.data(), S.size() * sizeof(SrcDesc), hipMemcpyHostToDevice, c->stream));
    HCHK(hipMemcpyAsync(dq, q, (size_t)n * 7 * 4, hipMemcpyHostToDevice, c->stream));
    HCHK(launch_debug_cost(J, dS, dq, n, dc, dh, di, dp, c->stream));
    HCHK(hipMemcpyAsync(cost, dc, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HCHK(hipMemcpyAsync(ham, dh, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HCHK(hipMemcpyAsync(ninv, di, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HCHK(hipMemcpyAsync(peak, dp, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HCHK(hipStreamSynchronize(c->stream));
    return MVS_OK;
}

int mvs_debug_copy(mvs_ctx* c, int64_t bytes) {
    if (!c || bytes <= 0 || (bytes & 15)) return fail(MVS_ERR_ARG, "mvs_debug_copy: bytes must be a positive multiple of 16");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int r = set_device(c)) return r;
    DBuf<uint8_t> a, b;
    if (a.need((size_t)bytes) != hipSuccess || b.need((size_t)bytes) != hipSuccess) {
        a.release();
        b.release();
        return fail(MVS_ERR_NOMEM, "mvs_debug_copy: allocation failed");
    }
    hipError_t e = hipMemsetAsync(a.p, 1, (size_t)bytes, c->stream);
    if (e == hipSuccess) e = launch_copy(a.p, b.p, (size_t)bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    a.release();
    b.release();
    if (e != hipSuccess) return fail(MVS_ERR_HIP, std::string("mvs_debug_copy: ") + hipGetErrorString(e));
    return MVS_OK;
}

int mvs_debug_cvt(mvs_ctx* c, const float* x, int32_t n, int32_t* out) {
    if (!c || !x || !out || n < 0) return fail(MVS_ERR_ARG, "mvs_debug_cvt: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int r = set_device(c)) return r;
    HCHK(c->dbg.need((size_t)n * 8 + 64));
    float* dx = (float*)c->dbg.p;
    int* dout = (int*)(c->dbg.p + (((size_t)n * 4 + 15) & ~(size_t)15));
    HCHK(hipMemcpyAsync(dx, x, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    HCHK(launch_debug_cvt(dx, n, dout, c->stream));
    HCHK(hipMemcpyAsync(out, dout, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HCHK(hipStreamSynchronize(c->stream));
    return MVS_OK;
}

int mvs_resize_image(mvs_ctx* c, const uint8_t* src, int32_t channels, int32_t ow, int32_t oh, int32_t mw,
                     int32_t mh, uint8_t* gray_out, uint8_t* rgb_out) {
    if (!c || !src || (channels != 1 && channels != 3) || ow < 1 || oh < 1 || mw < 1 || mh < 1 ||
        (!gray_out && !rgb_out))
        return fail(MVS_ERR_ARG, "mvs_resize_image: bad arguments");
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        const size_t nin = (size_t)ow * oh * channels, nout = (size_t)mw * mh * channels, npx = (size_t)mw * mh;
        DBuf<uint8_t> dsrc, dout, dgray;
        DBuf<float> dtmp;
        hipError_t e = dsrc.need(nin);
        if (e == hipSuccess) e = dout.need(nout);
        if (e == hipSuccess) e = dgray.need(npx);
        if (e == hipSuccess && (mw != ow || mh != oh)) e = dtmp.need((size_t)ow * mh * channels);
        if (e == hipSuccess) e = hipMemcpyAsync(dsrc.p, src, nin, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) {
            if (mw == ow && mh == oh)   // resize_exact to the same size is a copy
                e = hipMemcpyAsync(dout.p, dsrc.p, nin, hipMemcpyDeviceToDevice, c->stream);
            else
                e = launch_resize(dsrc.p, channels, ow, oh, mw, mh, dtmp.p, dout.p, c->stream);
        }
        if (e == hipSuccess && channels == 3 && gray_out) e = launch_luma(dout.p, (int64_t)npx, dgray.p, c->stream);
        if (e == hipSuccess && gray_out)
            e = hipMemcpyAsync(gray_out, channels == 3 ? dgray.p : dout.p, npx, hipMemcpyDeviceToHost, c->stream);
        std::vector<uint8_t> host;
        if (e == hipSuccess && rgb_out) {
            if (channels == 3) {
                e = hipMemcpyAsync(rgb_out, dout.p, nout, hipMemcpyDeviceToHost, c->stream);
            } else {
                host.resize(npx);
                e = hipMemcpyAsync(host.data(), dout.p, npx, hipMemcpyDeviceToHost, c->stream);
            }
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess && rgb_out && channels == 1)
            for (size_t q = 0; q < npx; q++) rgb_out[3 * q] = rgb_out[3 * q + 1] = rgb_out[3 * q + 2] = host[q];
        dsrc.release();
        dout.release();
        dgray.release();
        dtmp.release();
        if (e != hipSuccess) return fail(MVS_ERR_HIP, std::string("mvs_resize_image: ") + hipGetErrorString(e));
        return MVS_OK;
    } catch (const std::exception& ex) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_resize_image: ") + ex.what());
    }
}

int mvs_debug_div(mvs_ctx* c, const float* a, const float* b, int32_t n, float* out) {
    if (!c || !a || !b || !out || n < 0) return fail(MVS_ERR_ARG, "mvs_debug_div: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int r = set_device(c)) return r;
    const size_t nb = (((size_t)n * 4 + 255) & ~(size_t)255);
    HCHK(c->dbg.need(3 * nb + 64));
    float* da = (float*)c->dbg.p;
    float* db = (float*)(c->dbg.p + nb);
    float* dout = (float*)(c->dbg.p + 2 * nb);
    HCHK(hipMemcpyAsync(da, a, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    HCHK(hipMemcpyAsync(db, b, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    HCHK(launch_debug_div(da, db, n, dout, c->stream));
    HCHK(hipMemcpyAsync(out, dout, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HCHK(hipStreamSynchronize(c->stream));
    return MVS_OK;
}

int mvs_debug_rng(mvs_ctx* c, const int32_t* xy, const uint64_t* salt, int32_t n, int32_t ndraw, uint64_t* seed_out,
                  float* uniform_out) {
    if (!c || !xy || !salt || n < 0 || ndraw < 0 || !seed_out || (ndraw > 0 && !uniform_out))
        return fail(MVS_ERR_ARG, "mvs_debug_rng: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int r = set_device(c)) return r;
    const size_t bytes = (size_t)n * (8 + 8 + 8) + (size_t)n * ndraw * 4 + 64;
    HCHK(c->dbg.need(bytes));
    uint8_t* base = c->dbg.p;
    int* dxy = (int*)base;
    u64* ds = (u64*)(base + (((size_t)n * 8 + 7) & ~(size_t)7));
    u64* dseed = ds + n;
    float* du = (float*)(dseed + n);
    HCHK(hipMemcpyAsync(dxy, xy, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    HCHK(hipMemcpyAsync(ds, salt, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    HCHK(launch_debug_rng(dxy, ds, n, ndraw, dseed, du, c->stream));
    HCHK(hipMemcpyAsync(seed_out, dseed, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    if (ndraw > 0) HCHK(hipMemcpyAsync(uniform_out, du, (size_t)n * ndraw * 4, hipMemcpyDeviceToHost, c->stream));
    HCHK(hipStreamSynchronize(c->stream));
    return MVS_OK;
}

}  // extern "C"

// --------------------------------------------------------------------------
// learned plane proposals
// --------------------------------------------------------------------------

// One training sample on the network's coarse grid (h x w = the view's
// size divided by MVS_NET_STRIDE, rounded up); the gray image stays at the
// view's full width wf.
struct NetSample {
    int view = 0, h = 0, w = 0, wf = 0;
    int64_t n_valid = 0;    // valid target pixels
    DBuf<uint16_t> vol;     // [h*w][32] bf16 sweep features
    DBuf<uint8_t> gray;     // full-resolution gray
    DBuf<float> target;     // [h*w][5]
};

static inline int coarse(int v) { return (v + MVS_NET_STRIDE - 1) / MVS_NET_STRIDE; }

struct mvs_net {
    mvs_ctx* ctx = nullptr;
    net::Net net;
    std::vector<NetSample*> samples;
    DBuf<uint16_t> vol_tmp;
    DBuf<float> volf_tmp;
    DBuf<VolOut> vol_outs;
    DBuf<uintptr_t> ptrs;     // device pointer tables (vol, gray, pred)
    DBuf<float> preds;
    DBuf<u64> cnt;
    DBuf<u64> acc_list;       // accepted proposals (job, y, x) of the learned initialisation
    DBuf<float4> acc_plane;   // their planes (extra-hypothesis mode)
};

static inline int round16(int v) { return (v + 15) & ~15; }

// Frees every device buffer of a network (its context's device current).
static void release_net(mvs_net* nt) {
    for (NetSample* s : nt->samples) {
        s->vol.release();
        s->gray.release();
        s->target.release();
        delete s;
    }
    nt->samples.clear();
    nt->net.release();
    nt->vol_tmp.release();
    nt->volf_tmp.release();
    nt->vol_outs.release();
    nt->ptrs.release();
    nt->preds.release();
    nt->cnt.release();
    nt->acc_list.release();
    nt->acc_plane.release();
}

static void orphan_net(mvs_net* nt) {
    release_net(nt);
    nt->ctx = nullptr;
}

#define NET_CTX(nt, name)                                                              \
    do {                                                                               \
        if (!(nt)) return fail(MVS_ERR_ARG, name ": NULL network");                    \
        if (!(nt)->ctx) return fail(MVS_ERR_STATE, name ": the network's context is gone"); \
    } while (0)

static hipError_t launch_zero_grads(mvs_net* nt) { return net::launch_zero(nt->net.g, nt->net.n_params, nt->ctx->stream); }

// Plans `jobs` with 16x16 tiles (the sweep-volume and proposal kernels' shape).
static int plan_16(mvs_ctx* c, const mvs_job_t* jobs, int njobs, const mvs_options_t* o, BatchPlan& P, int eval) {
    c->tuner.forced = true;
    c->tuner.force_cfg = TileCfg();
    c->tuner.force_cfg.tw = 16;
    c->tuner.force_cfg.th = 16;
    c->tuner.force_cfg.persistent = 0;
    c->tuner.force_cfg.eval = eval;
    const int r = plan_batch(c, jobs, njobs, o, 4, -1, P);
    c->tuner.forced = false;
    return r;
}

static int upload_plan(mvs_ctx* c, BatchPlan& P) {
    for (size_t j = 0; j < P.jobs.size(); j++) P.jobs[j].valid_count = nullptr;
    c->graph.release();
    HCHK(c->jobs.need(P.jobs.size()));
    HCHK(c->srcs.need(P.srcs.size()));
    HCHK(c->tiles.need(P.tiles.size()));
    HCHK(c->counters.need(counter_words(P)));
    HCHK(hipMemcpyAsync(c->jobs.p, P.jobs.data(), P.jobs.size() * sizeof(JobDesc), hipMemcpyHostToDevice, c->stream));
    HCHK(hipMemcpyAsync(c->srcs.p, P.srcs.data(), P.srcs.size() * sizeof(SrcDesc), hipMemcpyHostToDevice, c->stream));
    HCHK(hipMemcpyAsync(c->tiles.p, P.tiles.data(), P.tiles.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    return MVS_OK;
}

// Sweep volume of each job, on the grid of every stride-th pixel, into
// per-job outputs (device pointers).  Tiles are 16 x 16 coarse pixels.
static int run_volume(mvs_ctx* c, mvs_net* nt, const mvs_job_t* jobs, int njobs, const std::vector<VolOut>& outs,
                      int stride) {
    mvs_options_t o;
    mvs_options_default(&o);
    BatchPlan P;
    if (int r = plan_16(c, jobs, njobs, &o, P, 4)) return r;
    P.tiles.clear();
    for (int j = 0; j < njobs; j++) {
        const int cw = (P.jobs[j].w + stride - 1) / stride, ch = (P.jobs[j].h + stride - 1) / stride;
        for (int ty = 0; ty < (ch + 15) / 16; ty++)
            for (int tx = 0; tx < (cw + 15) / 16; tx++) P.tiles.push_back(make_int4(j, tx * 16, ty * 16, 0));
    }
    if (int r = upload_plan(c, P)) return r;
    HCHK(nt->vol_outs.need(outs.size()));
    HCHK(hipMemcpyAsync(nt->vol_outs.p, outs.data(), outs.size() * sizeof(VolOut), hipMemcpyHostToDevice, c->stream));
    HCHK(launch_volume(c->jobs.p, c->srcs.p, c->tiles.p, (int)P.tiles.size(), 32, stride, nt->vol_outs.p, c->stream));
    return MVS_OK;
}

// Forward of the network over n same-size samples given device vol/gray
// pointers (coarse h x w grid, gray rows of full width wf); predictions
// written to dpred ([n][h][w][5]).
static int net_predict_batch(mvs_net* nt, const std::vector<const uint16_t*>& vols,
                             const std::vector<const uint8_t*>& grays, int h, int w, int wf, float* dpred, bool train,
                             const float* dtarget, float* loss4, int64_t n_valid = 0) {
    mvs_ctx* c = nt->ctx;
    const int n = (int)vols.size();
    const int Hp = round16(h), Wp = round16(w);
    HCHK(nt->net.ensure(n, Hp, Wp, train));
    HCHK(nt->ptrs.need(2 * n));
    std::vector<uintptr_t> tab(2 * n);
    for (int i = 0; i < n; i++) {
        tab[i] = (uintptr_t)vols[i];
        tab[n + i] = (uintptr_t)grays[i];
    }
    HCHK(hipMemcpyAsync(nt->ptrs.p, tab.data(), tab.size() * sizeof(uintptr_t), hipMemcpyHostToDevice, c->stream));
    HCHK(launch_assemble((const uint16_t* const*)nt->ptrs.p, (const uint8_t* const*)(nt->ptrs.p + n), n, h, w, Hp, Wp,
                         MVS_NET_STRIDE, wf, nt->net.input(), c->stream));
    HCHK(nt->net.forward());
    HCHK(nt->net.head(train, h, w, dpred, dtarget, 0.02f, n_valid));
    if (train) {
        HCHK(nt->net.backward());
        if (loss4) {
            HCHK(hipMemcpyAsync(loss4, nt->net.loss, 4 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        }
    }
    return MVS_OK;
}

extern "C" {

int mvs_net_create(mvs_ctx* c, uint64_t seed, mvs_net** out) {
    if (!c || !out) return fail(MVS_ERR_ARG, "mvs_net_create: NULL argument");
    *out = nullptr;
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        mvs_net* nt = new mvs_net();
        nt->ctx = c;
        hipError_t e = nt->net.init(seed, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            nt->net.release();
            delete nt;
            return fail(MVS_ERR_HIP, std::string("mvs_net_create: ") + hipGetErrorString(e));
        }
        c->nets.push_back(nt);
        *out = nt;
        return MVS_OK;
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_net_create: ") + e.what());
    }
}

void mvs_net_destroy(mvs_net* nt) {
    if (!nt) return;
    if (mvs_ctx* c = nt->ctx) {
        std::lock_guard<std::mutex> lk(c->mu);
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        release_net(nt);
        c->nets.erase(std::remove(c->nets.begin(), c->nets.end(), nt), c->nets.end());
    }
    delete nt;
}

int64_t mvs_net_num_params(mvs_net* nt) { return nt ? nt->net.n_params : -1; }

int mvs_net_get_params(mvs_net* nt, float* out) {
    NET_CTX(nt, "mvs_net_get_params");
    if (!out) return fail(MVS_ERR_ARG, "mvs_net_get_params: NULL argument");
    std::lock_guard<std::mutex> lk(nt->ctx->mu);
    if (int r = set_device(nt->ctx)) return r;
    HCHK(hipMemcpyAsync(out, nt->net.p, nt->net.n_params * 4, hipMemcpyDeviceToHost, nt->ctx->stream));
    HCHK(hipStreamSynchronize(nt->ctx->stream));
    return MVS_OK;
}

int mvs_net_set_params(mvs_net* nt, const float* in) {
    NET_CTX(nt, "mvs_net_set_params");
    if (!in) return fail(MVS_ERR_ARG, "mvs_net_set_params: NULL argument");
    std::lock_guard<std::mutex> lk(nt->ctx->mu);
    if (int r = set_device(nt->ctx)) return r;
    HCHK(hipMemcpyAsync(nt->net.p, in, nt->net.n_params * 4, hipMemcpyHostToDevice, nt->ctx->stream));
    HCHK(nt->net.refresh_weights());
    HCHK(hipStreamSynchronize(nt->ctx->stream));
    return MVS_OK;
}

int mvs_sweep_volume(mvs_ctx* c, const mvs_job_t* job, int32_t stride, float* out) {
    if (!c || !job || !out) return fail(MVS_ERR_ARG, "mvs_sweep_volume: NULL argument");
    if (stride != 1 && stride != 2) return fail(MVS_ERR_ARG, "mvs_sweep_volume: stride must be 1 or 2");
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        if (job->ref < 0 || job->ref >= (int)c->views.size()) return fail(MVS_ERR_ARG, "mvs_sweep_volume: bad ref");
        const ViewState& v = c->views[job->ref];
        const size_t np = (size_t)((v.w + stride - 1) / stride) * ((v.h + stride - 1) / stride);
        mvs_net tmp;
        tmp.ctx = c;
        DBuf<float> vf;
        HCHK(vf.need(np * 32));
        std::vector<VolOut> outs(1);
        outs[0].vol = nullptr;
        outs[0].volf = vf.p;
        int r = run_volume(c, &tmp, job, 1, outs, stride);
        if (r == MVS_OK) {
            hipError_t e = hipMemcpyAsync(out, vf.p, np * 32 * 4, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) r = fail(MVS_ERR_HIP, std::string("mvs_sweep_volume: ") + hipGetErrorString(e));
        }
        vf.release();
        tmp.vol_outs.release();
        return r;
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_sweep_volume: ") + e.what());
    }
}

int mvs_net_add_samples(mvs_net* nt, const mvs_job_t* jobs, int32_t njobs, const float* const* targets) {
    NET_CTX(nt, "mvs_net_add_samples");
    if (!jobs || njobs < 1 || !targets) return fail(MVS_ERR_ARG, "mvs_net_add_samples: bad arguments");
    mvs_ctx* c = nt->ctx;
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        std::vector<VolOut> outs(njobs);
        std::vector<NetSample*> fresh;
        for (int j = 0; j < njobs; j++) {
            if (jobs[j].ref < 0 || jobs[j].ref >= (int)c->views.size())
                return fail(MVS_ERR_ARG, "mvs_net_add_samples: bad ref");
            const ViewState& v = c->views[jobs[j].ref];
            NetSample* s = new NetSample();
            s->view = jobs[j].ref;
            s->h = coarse(v.h);
            s->w = coarse(v.w);
            s->wf = v.w;
            const size_t np = (size_t)s->w * s->h;
            HCHK(s->vol.need(np * 32));
            HCHK(s->gray.need((size_t)v.w * v.h));
            HCHK(s->target.need(np * 5));
            HCHK(hipMemcpyAsync(s->gray.p, v.gray.p, (size_t)v.w * v.h, hipMemcpyDeviceToDevice, c->stream));
            HCHK(hipMemcpyAsync(s->target.p, targets[j], np * 5 * 4, hipMemcpyHostToDevice, c->stream));
            for (size_t q = 0; q < np; q++) s->n_valid += targets[j][5 * q + 4] > 0.5f ? 1 : 0;
            outs[j].vol = s->vol.p;
            outs[j].volf = nullptr;
            fresh.push_back(s);
        }
        if (int r = run_volume(c, nt, jobs, njobs, outs, MVS_NET_STRIDE)) return r;
        HCHK(hipStreamSynchronize(c->stream));
        for (NetSample* s : fresh) nt->samples.push_back(s);
        return (int)nt->samples.size();
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_net_add_samples: ") + e.what());
    }
}

int mvs_net_clear_samples(mvs_net* nt) {
    NET_CTX(nt, "mvs_net_clear_samples");
    std::lock_guard<std::mutex> lk(nt->ctx->mu);
    (void)hipSetDevice(nt->ctx->device);
    (void)hipStreamSynchronize(nt->ctx->stream);
    for (NetSample* s : nt->samples) {
        s->vol.release();
        s->gray.release();
        s->target.release();
        delete s;
    }
    nt->samples.clear();
    return MVS_OK;
}

int mvs_net_train_step(mvs_net* nt, const int32_t* idx, int32_t n, float lr, float* loss4) {
    NET_CTX(nt, "mvs_net_train_step");
    if (!idx || n < 1) return fail(MVS_ERR_ARG, "mvs_net_train_step: bad arguments");
    mvs_ctx* c = nt->ctx;
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        std::vector<const uint16_t*> vols;
        std::vector<const uint8_t*> grays;
        int h = -1, w = -1, wf = -1;
        int64_t n_valid = 0;
        for (int i = 0; i < n; i++) {
            if (idx[i] < 0 || idx[i] >= (int)nt->samples.size()) return fail(MVS_ERR_ARG, "train_step: bad sample");
            NetSample* s = nt->samples[idx[i]];
            if (h < 0) {
                h = s->h;
                w = s->w;
                wf = s->wf;
            } else if (s->h != h || s->w != w || s->wf != wf) {
                return fail(MVS_ERR_ARG, "train_step: samples of one batch must share their size");
            }
            vols.push_back(s->vol.p);
            grays.push_back(s->gray.p);
            n_valid += s->n_valid;
        }
        const size_t np = (size_t)h * w;
        HCHK(nt->net.ensure(n, round16(h), round16(w), true));
        for (int i = 0; i < n; i++)
            HCHK(hipMemcpyAsync(nt->net.target + (size_t)i * np * 5, nt->samples[idx[i]]->target.p, np * 5 * 4,
                                hipMemcpyDeviceToDevice, c->stream));
        float l4[4] = {0, 0, 0, 0};
        if (int r = net_predict_batch(nt, vols, grays, h, w, wf, nullptr, true, nt->net.target, l4, n_valid))
            return r;
        HCHK(nt->net.adam(lr));
        HCHK(hipStreamSynchronize(c->stream));
        if (loss4) {
            const float npx = (float)((double)n * np);
            loss4[0] = l4[0] / std::max(l4[3], 1.0f);   // mean L1 (normalised inverse depth) over valid px
            loss4[1] = l4[1] / std::max(l4[3], 1.0f);   // mean (1 - cos) over valid px
            loss4[2] = l4[2] / npx;                     // mean confidence BCE
            loss4[3] = l4[3] / npx;                     // valid fraction
        }
        return MVS_OK;
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_net_train_step: ") + e.what());
    }
}

int mvs_net_predict(mvs_net* nt, const mvs_job_t* job, float* pred) {
    NET_CTX(nt, "mvs_net_predict");
    if (!job || !pred) return fail(MVS_ERR_ARG, "mvs_net_predict: NULL argument");
    mvs_ctx* c = nt->ctx;
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        if (job->ref < 0 || job->ref >= (int)c->views.size()) return fail(MVS_ERR_ARG, "mvs_net_predict: bad ref");
        const ViewState& v = c->views[job->ref];
        const size_t np = (size_t)coarse(v.w) * coarse(v.h);
        HCHK(nt->vol_tmp.need(np * 32));
        HCHK(nt->preds.need(np * 5));
        std::vector<VolOut> outs(1);
        outs[0].vol = nt->vol_tmp.p;
        outs[0].volf = nullptr;
        if (int r = run_volume(c, nt, job, 1, outs, MVS_NET_STRIDE)) return r;
        std::vector<const uint16_t*> vols(1, nt->vol_tmp.p);
        std::vector<const uint8_t*> grays(1, v.gray.p);
        if (int r = net_predict_batch(nt, vols, grays, coarse(v.h), coarse(v.w), v.w, nt->preds.p, false, nullptr,
                                      nullptr))
            return r;
        HCHK(hipMemcpyAsync(pred, nt->preds.p, np * 5 * 4, hipMemcpyDeviceToHost, c->stream));
        HCHK(hipStreamSynchronize(c->stream));
        return MVS_OK;
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_net_predict: ") + e.what());
    }
}

int mvs_net_debug_grads(mvs_net* nt, const float* input, const float* target, int32_t n, int32_t Hp, int32_t Wp,
                        int32_t h, int32_t w, float* grads, float* pred, float* loss4) {
    NET_CTX(nt, "mvs_net_debug_grads");
    if (!input || !target || n < 1 || Hp % 16 || Wp % 16 || h > Hp || w > Wp)
        return fail(MVS_ERR_ARG, "mvs_net_debug_grads: bad arguments");
    mvs_ctx* c = nt->ctx;
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        HCHK(nt->net.ensure(n, Hp, Wp, true));
        const size_t ni = (size_t)n * Hp * Wp * 64;
        std::vector<uint16_t> hb(ni);
        for (size_t i = 0; i < ni; i++) {
            uint32_t u;
            memcpy(&u, &input[i], 4);
            u += 0x7fffu + ((u >> 16) & 1u);
            hb[i] = (uint16_t)(u >> 16);
        }
        HCHK(hipMemcpyAsync(nt->net.input(), hb.data(), ni * 2, hipMemcpyHostToDevice, c->stream));
        HCHK(hipMemcpyAsync(nt->net.target, target, (size_t)n * h * w * 5 * 4, hipMemcpyHostToDevice, c->stream));
        HCHK(launch_zero_grads(nt));
        HCHK(nt->net.forward());
        int64_t n_valid = 0;
        for (size_t q = 0; q < (size_t)n * h * w; q++) n_valid += target[5 * q + 4] > 0.5f ? 1 : 0;
        HCHK(nt->net.head(true, h, w, nt->net.pred, nt->net.target, 0.02f, n_valid));
        HCHK(nt->net.backward());
        if (grads) HCHK(hipMemcpyAsync(grads, nt->net.g, nt->net.n_params * 4, hipMemcpyDeviceToHost, c->stream));
        if (pred) HCHK(hipMemcpyAsync(pred, nt->net.pred, (size_t)n * h * w * 5 * 4, hipMemcpyDeviceToHost, c->stream));
        if (loss4) HCHK(hipMemcpyAsync(loss4, nt->net.loss, 16, hipMemcpyDeviceToHost, c->stream));
        HCHK(launch_zero_grads(nt));
        HCHK(hipStreamSynchronize(c->stream));
        return MVS_OK;
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_net_debug_grads: ") + e.what());
    }
}

int mvs_estimate_depth_learned(mvs_ctx* c, mvs_net* nt, const mvs_job_t* jobs, int32_t njobs,
                               const mvs_options_t* opts, int32_t vol_k, float conf_thr, int64_t* n_accepted) {
    return mvs_estimate_depth_learned_ex(c, nt, jobs, njobs, opts, vol_k, conf_thr, 1, n_accepted);
}

int mvs_estimate_depth_learned_ex(mvs_ctx* c, mvs_net* nt, const mvs_job_t* jobs, int32_t njobs,
                                  const mvs_options_t* opts, int32_t vol_k, float conf_thr, int32_t mode,
                                  int64_t* n_accepted) {
    if (!c || !nt || !jobs || njobs < 1 || !opts || vol_k < 1 || mode < 0 || mode > 1)
        return fail(MVS_ERR_ARG, "mvs_estimate_depth_learned: bad arguments");
    if (nt->ctx != c) return fail(MVS_ERR_STATE, "mvs_estimate_depth_learned: the network belongs to another context");
    try {
        std::lock_guard<std::mutex> lk(c->mu);
        if (int r = set_device(c)) return r;
        // 1. sweep volumes over the first vol_k sources of every job
        std::vector<mvs_job_t> vj(jobs, jobs + njobs);
        size_t total = 0;
        std::vector<size_t> off(njobs);
        for (int j = 0; j < njobs; j++) {
            if (jobs[j].ref < 0 || jobs[j].ref >= (int)c->views.size())
                return fail(MVS_ERR_ARG, "mvs_estimate_depth_learned: bad ref");
            vj[j].k = std::min(jobs[j].k, vol_k);
            off[j] = total;
            total += (size_t)coarse(c->views[jobs[j].ref].w) * coarse(c->views[jobs[j].ref].h);
        }
        HCHK(nt->vol_tmp.need(total * 32));
        HCHK(nt->preds.need(total * 5));
        HCHK(hipEventRecord(c->ev[0], c->stream));
        std::vector<VolOut> outs(njobs);
        for (int j = 0; j < njobs; j++) {
            outs[j].vol = nt->vol_tmp.p + off[j] * 32;
            outs[j].volf = nullptr;
        }
        if (int r = run_volume(c, nt, vj.data(), njobs, outs, MVS_NET_STRIDE)) return r;
        HCHK(hipEventRecord(c->ev[1], c->stream));
        // 2. network over same-size groups, in chunks that bound activation memory
        std::vector<int> order(njobs);
        for (int j = 0; j < njobs; j++) order[j] = j;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            const ViewState &va = c->views[jobs[a].ref], &vb = c->views[jobs[b].ref];
            return va.w != vb.w ? va.w < vb.w : va.h < vb.h;
        });
        for (size_t i = 0; i < order.size();) {
            const ViewState& v0 = c->views[jobs[order[i]].ref];
            const int cw0 = coarse(v0.w), ch0 = coarse(v0.h);
            const int64_t px = (int64_t)round16(cw0) * round16(ch0);
            const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(16, ((int64_t)24 << 20) / px));
            std::vector<const uint16_t*> vols;
            std::vector<const uint8_t*> grays;
            size_t k = i;
            while (k < order.size() && (int)vols.size() < chunk) {
                const ViewState& vk = c->views[jobs[order[k]].ref];
                if (vk.w != v0.w || vk.h != v0.h) break;
                vols.push_back(nt->vol_tmp.p + off[order[k]] * 32);
                grays.push_back(vk.gray.p);
                ++k;
            }
            // predictions of a chunk land contiguously; jobs of one chunk are consecutive in `order`
            DBuf<float> chunk_pred;
            float* dst = nt->preds.p + off[order[i]] * 5;
            bool contiguous = true;
            const size_t cnp = (size_t)cw0 * ch0;
            for (size_t q = i; q < k; q++)
                if (off[order[q]] != off[order[i]] + (q - i) * cnp) contiguous = false;
            if (!contiguous) {
                HCHK(chunk_pred.need((k - i) * cnp * 5));
                dst = chunk_pred.p;
            }
            if (int r = net_predict_batch(nt, vols, grays, ch0, cw0, v0.w, dst, false, nullptr, nullptr)) return r;
            if (!contiguous) {
                for (size_t q = i; q < k; q++)
                    HCHK(hipMemcpyAsync(nt->preds.p + off[order[q]] * 5, dst + (q - i) * cnp * 5, cnp * 5 * 4,
                                        hipMemcpyDeviceToDevice, c->stream));
                HCHK(hipStreamSynchronize(c->stream));
                chunk_pred.release();
            }
            i = k;
        }
        HCHK(hipEventRecord(c->ev[2], c->stream));
        // 3. proposal init with the full source sets, then the passes and the
        // gate, in the launch shape the tuner holds for this batch (16x16
        // with the fourth evaluator where the proposal kernel lacks it)
        BatchPlan P;
        if (int r = plan_batch(c, jobs, njobs, opts, 0, -1, P)) return r;
        if (!proposal_supported(P.cfg))
            if (int r = plan_16(c, jobs, njobs, opts, P, 4)) return r;
        if (int r = upload_plan(c, P)) return r;
        const int ncnt = 1 + P.n_passes;
        HCHK(nt->ptrs.need(njobs));
        std::vector<uintptr_t> tab(njobs);
        for (int j = 0; j < njobs; j++) tab[j] = (uintptr_t)(nt->preds.p + off[j] * 5);
        HCHK(hipMemcpyAsync(nt->ptrs.p, tab.data(), njobs * sizeof(uintptr_t), hipMemcpyHostToDevice, c->stream));
        HCHK(launch_fill_u64(c->counters.p, counter_words(P), 0ULL, c->stream));
        u64* acc_counter = c->counters.p + valid_base(P) - 4;
        HCHK(nt->acc_list.need((size_t)P.n_pixels));
        if (mode == 1) HCHK(nt->acc_plane.need((size_t)P.n_pixels));
        HCHK(launch_proposal_init(P.cfg, c->jobs.p, c->srcs.p, c->tiles.p, (int)P.tiles.size(), P.nh,
                                  (const float* const*)nt->ptrs.p, MVS_NET_STRIDE, conf_thr, c->counters.p,
                                  acc_counter, nt->acc_list.p, mode == 1 ? nt->acc_plane.p : nullptr, c->stream));
        HCHK(hipEventRecord(c->ev[3], c->stream));
        for (int p = 0; p < P.run_passes; p++) {
            const float t = (float)p / (float)P.n_passes;
            const float sig_inv = 0.02f * powf(0.28f, t * 4.0f);
            const float sig_n = 0.30f * powf(0.42f, t * 4.0f);
            HCHK(launch_pass(P.cfg, c->jobs.p, c->srcs.p, c->tiles.p, (int)P.tiles.size(), p & 1, (u64)p + 7, sig_inv,
                             sig_n, c->counters.p + 1 + p, ((int*)(c->counters.p + ncnt)) + p, P.grid_p, c->stream));
        }
        HCHK(hipEventRecord(c->ev[4], c->stream));
        HCHK(launch_gate(P.cfg, c->jobs.p, c->tiles.p, (int)P.tiles.size(), P.run_passes & 1, opts->cost_thr,
                         c->stream));
        HCHK(hipEventRecord(c->ev[5], c->stream));
        HCHK(hipStreamSynchronize(c->stream));
        float t_vol = 0, t_net = 0, t_init = 0, t_pass = 0, t_gate = 0;
        (void)hipEventElapsedTime(&t_vol, c->ev[0], c->ev[1]);
        (void)hipEventElapsedTime(&t_net, c->ev[1], c->ev[2]);
        (void)hipEventElapsedTime(&t_init, c->ev[2], c->ev[3]);
        (void)hipEventElapsedTime(&t_pass, c->ev[3], c->ev[4]);
        (void)hipEventElapsedTime(&t_gate, c->ev[4], c->ev[5]);
        std::vector<u64> cnt(ncnt);
        HCHK(hipMemcpy(cnt.data(), c->counters.p, ncnt * sizeof(u64), hipMemcpyDeviceToHost));
        u64 acc = 0;
        HCHK(hipMemcpy(&acc, acc_counter, sizeof(u64), hipMemcpyDeviceToHost));
        c->stats.n_eval_init = (int64_t)cnt[0];
        c->stats.n_eval_pass = 0;
        for (int p = 0; p < P.run_passes; p++) c->stats.n_eval_pass += (int64_t)cnt[1 + p];
        c->stats.n_pixels = P.n_pixels;
        c->stats.taps_per_eval_sum = P.taps;
        c->stats.ms_init = t_vol + t_net + t_init;
        c->stats.ms_pass = t_pass;
        c->stats.ms_gate = t_gate;
        c->stats.ms_learned_volume = t_vol;
        c->stats.ms_learned_net = t_net;
        c->stats.n_pass_launches = P.run_passes;
        c->stats.graph_used = 0;
        c->stats.tile_w = P.cfg.tw;
        c->stats.tile_h = P.cfg.th;
        c->stats.persistent = P.cfg.persistent;
        c->stats.evaluator = P.cfg.eval;
        if (n_accepted) *n_accepted = (int64_t)acc;
        for (int j = 0; j < njobs; j++) {
            ViewState& rv = c->views[jobs[j].ref];
            const size_t np = (size_t)rv.w * rv.h;
            rv.has_depth = true;
            if (jobs[j].depth) HCHK(hipMemcpyAsync(jobs[j].depth, rv.depth.p, np * 4, hipMemcpyDeviceToHost, c->stream));
            if (jobs[j].cost) HCHK(hipMemcpyAsync(jobs[j].cost, rv.cost_out.p, np * 4, hipMemcpyDeviceToHost, c->stream));
            if (jobs[j].normals)
                HCHK(hipMemcpyAsync(jobs[j].normals, rv.normal.p, 3 * np * 4, hipMemcpyDeviceToHost, c->stream));
        }
        HCHK(hipStreamSynchronize(c->stream));
        return MVS_OK;
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_estimate_depth_learned: ") + e.what());
    }
}

}  // extern "C"

// --------------------------------------------------------------------------
// run_mvs (mvs/src/lib.rs:294-517) without image decoding
// --------------------------------------------------------------------------

extern "C" int mvs_plan(const mvs_view_t* views, int32_t nviews, const double* xyz, const int32_t* track_off,
                        const int32_t* track_view, int64_t npoints, const mvs_options_t* opts, int32_t* src_off,
                        int32_t* srcs, double* zrange, int32_t* fusion_off, int32_t* fusion, uint32_t* covis);

extern "C" int mvs_run_mvs(mvs_ctx* c, const mvs_view_t* views, int32_t nviews, const double* xyz,
                           const int32_t* track_off, const int32_t* track_view, int64_t npoints,
                           const mvs_options_t* opts, const char* out_dir, mvs_run_stats_t* st) {
    if (!c || !views || nviews < 2 || !opts) return fail(MVS_ERR_ARG, "mvs_run_mvs: need at least 2 views");
    if (npoints < 8)
        return fail(MVS_ERR_STATE, "sparse reconstruction too small for MVS (" + std::to_string(npoints) +
                                       " points, need >= 8)");
    try {
        const auto t_all = std::chrono::steady_clock::now();
        const int n = nviews;
        std::vector<int32_t> src_off(n + 1), srcs((size_t)n * std::max<uint32_t>(opts->num_sources, 1));
        std::vector<double> zr(2 * n);
        std::vector<int32_t> fus_off(n + 1), fus((size_t)n * n);
        int r = mvs_plan(views, n, xyz, track_off, track_view, npoints, opts, src_off.data(), srcs.data(), zr.data(),
                         fus_off.data(), fus.data(), nullptr);
        if (r) return fail(r, "mvs_run_mvs: planning failed");
        if ((r = mvs_upload_views(c, views, n))) return r;
        emit(c, strf("mvs: loaded %d views at working resolution (max_dim %u)", n, opts->max_dim));
        std::vector<mvs_job_t> jobs;
        for (int i = 0; i < n; i++) {
            if (src_off[i + 1] == src_off[i]) continue;   // no source view: no depth map (lib.rs:355-357)
            mvs_job_t j{};
            j.ref = i;
            j.k = src_off[i + 1] - src_off[i];
            j.srcs = srcs.data() + src_off[i];
            j.zmin = zr[2 * i];
            j.zmax = zr[2 * i + 1];
            jobs.push_back(j);
        }
        emit(c, strf("mvs: %d views with >=1 source view (K = %u)", (int)jobs.size(), opts->num_sources));
        mvs_run_stats_t rs{};
        rs.per_view_valid = st ? st->per_view_valid : nullptr;
        rs.views = n;
        rs.views_with_depth = (int32_t)jobs.size();
        double stereo_ms = 0.0;
        if (!jobs.empty()) {
            if ((r = mvs_estimate_depth_batch(c, jobs.data(), (int32_t)jobs.size(), opts, 4))) return r;
            stereo_ms = c->stats.ms_init + c->stats.ms_pass + c->stats.ms_gate;
        }
        std::vector<int32_t> has(n, 0);
        for (auto& j : jobs) has[j.ref] = 1;
        std::vector<int64_t> valid(n, 0);
        {
            std::lock_guard<std::mutex> lk(c->mu);
            for (int i = 0; i < n; i++) valid[i] = has[i] ? c->views[i].n_valid : 0;
        }
        // every view of the batch spans the whole batched stereo stage
        const float stereo_s = (float)(stereo_ms / 1000.0);
        for (auto& j : jobs)
            emit(c, "mvs: [" + c->names[j.ref] + "] depth " + rust_f64_debug(j.zmin) + "-" + rust_f64_debug(j.zmax) +
                        strf(" (%d srcs) valid %lld/%lld px in %.1fs", j.k, (long long)valid[j.ref],
                             (long long)views[j.ref].w * views[j.ref].h, (double)stereo_s));
        emit(c, strf("mvs: stereo done in %.1fs (%d depth maps)", (double)stereo_s, (int)jobs.size()));
        const std::string dense = out_dir ? std::string(out_dir) + "/dense" : std::string();
        if (out_dir) {
            (void)mkdir(out_dir, 0755);
            (void)mkdir(dense.c_str(), 0755);
        }
        if (opts->save_depth_vis && out_dir) {
            const std::string vis = dense + "/depth_vis";
            if (mkdir(vis.c_str(), 0755) != 0 && errno != EEXIST)
                return fail(MVS_ERR_ARG, "mvs_run_mvs: cannot create " + vis);
            for (auto& j : jobs) {
                const size_t np = (size_t)views[j.ref].w * views[j.ref].h;
                std::vector<float> d(np);
                if ((r = mvs_get_depth(c, j.ref, d.data(), nullptr, nullptr))) return r;
                const std::vector<uint8_t> img = depth_vis_u8(d.data(), np);
                // a failed image write is skipped, as the reference ignores it (lib.rs:387)
                (void)write_png_gray8(vis + "/" + file_stem(c->names[j.ref]) + "_depth.png", img.data(),
                                      views[j.ref].w, views[j.ref].h);
            }
            emit(c, "mvs: depth visualizations saved");
        }
        const double diag = robust_diag(xyz, npoints);
        const double cell = mvs_fusion_cell(xyz, npoints, opts->voxel_rel);
        emit(c, strf("mvs: fusion voxel cell %.4f (robust sparse diag %.2f)", cell, diag));
        mvs_point_t* pts = nullptr;
        int64_t npts = 0;
        if ((r = mvs_fuse(c, has.data(), fus_off.data(), fus.data(), cell, opts, &pts, &npts))) return r;
        rs.dense_points = npts;
        rs.cell = cell;
        rs.stereo_ms = stereo_ms;
        rs.fuse_ms = c->stats.ms_fuse;
        rs.n_candidates = c->stats.n_candidates;
        if (rs.per_view_valid)
            for (int i = 0; i < n; i++) rs.per_view_valid[i] = valid[i];
        std::string ply;
        if (out_dir) {
            ply = dense + "/dense.ply";
            r = mvs_write_ply(ply.c_str(), pts, npts);
            if (r == MVS_OK) r = mvs_write_dense_web_json((dense + "/dense_web.json").c_str(), pts, npts);
            if (r) {
                mvs_free_points(pts);
                return r;
            }
            FILE* f = fopen((dense + "/mvs_stats.json").c_str(), "w");
            if (f) {
                // same keys as lib.rs:503-507; seconds of device time here
                fprintf(f,
                        "{\"views\": %d, \"views_with_depth\": %d, \"dense_points\": %lld, \"stereo_s\": %.1f, "
                        "\"fuse_s\": %.1f, \"elapsed_s\": %.1f}",
                        rs.views, rs.views_with_depth, (long long)npts, stereo_ms / 1000.0, rs.fuse_ms / 1000.0,
                        (stereo_ms + rs.fuse_ms) / 1000.0);
                fclose(f);
            }
        }
        mvs_free_points(pts);
        const float total_s =
            std::chrono::duration<float>(std::chrono::steady_clock::now() - t_all).count();
        emit(c, strf("mvs: fused %lld dense points -> %s in %.1fs (total %.1fs)", (long long)npts,
                     out_dir ? ply.c_str() : "(not written)", (double)(float)(rs.fuse_ms / 1000.0),
                     (double)total_s));
        if (st) *st = rs;
        return MVS_OK;
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_run_mvs: ") + e.what());
    }
}

// --------------------------------------------------------------------------
// descriptor matching (sfm-lite/src/matching.rs:22-98, lib.rs:197-224)
// --------------------------------------------------------------------------

// Upload sets to S (host descriptors, concatenated) and prepare norms, the
// bf16 split and the per-set safety flags.
static int match_prepare(mvs_ctx* c, MatchSets& S, const float* desc, const int32_t* counts, int32_t nsets,
                         const char* who) {
    MatchState& M = c->match;
    int64_t total = 0;
    for (int32_t f = 0; f < nsets; f++) {
        if (counts[f] < 0) return fail(MVS_ERR_ARG, std::string(who) + ": negative descriptor count");
        total += counts[f];
    }
    if (total > 0 && !desc) return fail(MVS_ERR_ARG, std::string(who) + ": NULL descriptors");
    S.off.assign((size_t)nsets, 0);
    S.n.assign(counts, counts + nsets);
    S.safe.assign((size_t)nsets, 0);
    int64_t o = 0;
    for (int32_t f = 0; f < nsets; f++) {
        S.off[(size_t)f] = o;
        o += counts[f];
    }
    const size_t rows = (size_t)std::max<int64_t>(total, 1);
    HCHK(S.desc.need(rows * mt::DIM));
    HCHK(S.norm.need(rows));
    HCHK(S.split.need(2 * rows * mt::DIM));
    S.rows = rows;
    HCHK(M.set_off.need((size_t)std::max(nsets, 1)));
    HCHK(M.bad.need((size_t)std::max(nsets, 1)));
    hipStream_t st = c->stream;
    if (total > 0)
        HCHK(hipMemcpyAsync(S.desc.p, desc, sizeof(float) * mt::DIM * (size_t)total, hipMemcpyHostToDevice, st));
    if (nsets > 0) {
        HCHK(hipMemcpyAsync(M.set_off.p, S.off.data(), sizeof(int64_t) * (size_t)nsets, hipMemcpyHostToDevice, st));
        HCHK(hipMemsetAsync(M.bad.p, 0, sizeof(int) * (size_t)nsets, st));
    }
    HCHK(mt::launch_prep(S.desc.p, total, M.set_off.p, nsets, S.norm.p, S.split.p, S.split.p + rows * mt::DIM,
                         M.bad.p, st));
    std::vector<int> bad((size_t)nsets);
    if (nsets > 0) HCHK(hipMemcpyAsync(bad.data(), M.bad.p, sizeof(int) * (size_t)nsets, hipMemcpyDeviceToHost, st));
    HCHK(hipStreamSynchronize(st));
    for (int32_t f = 0; f < nsets; f++) S.safe[(size_t)f] = bad[(size_t)f] == 0;
    return MVS_OK;
}

// One batch over the sets of S; pairs index the sets.  Pairs of two safe
// sets take the filtered search (unless mode == 1), the rest the exact
// tiles; both write fwd / bwd, then one selection pass lists the matches.
static int match_batch(mvs_ctx* c, MatchSets& S, const int32_t* pairs, int32_t npairs, double ratio,
                       uint32_t* out_a, uint32_t* out_b, int64_t* out_n, const char* who) {
    const auto t_start = std::chrono::steady_clock::now();
    MatchState& M = c->match;
    std::vector<mt::PairDesc> all((size_t)npairs), ex, fl;
    int64_t fr = 0, fc = 0, out = 0;
    for (int32_t p = 0; p < npairs; p++) {
        const int32_t ia = pairs[2 * p], ib = pairs[2 * p + 1];
        if (ia < 0 || ib < 0 || ia >= (int32_t)S.n.size() || ib >= (int32_t)S.n.size())
            return fail(MVS_ERR_ARG, std::string(who) + ": pair index out of range");
        mt::PairDesc d = {};
        d.a = S.off[(size_t)ia];
        d.b = S.off[(size_t)ib];
        d.na = S.n[(size_t)ia];
        d.nb = S.n[(size_t)ib];
        d.frow = fr;
        d.fcol = fc;
        d.out = out;
        fr += d.na;
        fc += d.nb;
        out += std::min(d.na, d.nb);
        all[(size_t)p] = d;
        if (d.na == 0 || d.nb == 0) continue;   // no matches, nothing to search
        const bool filt = M.mode == 0 && S.safe[(size_t)ia] && S.safe[(size_t)ib];
        (filt ? fl : ex).push_back(d);
    }
    // launch-local numbering: tiles, partials, rows and columns
    auto number = [](std::vector<mt::PairDesc>& v, int tile, int64_t& tiles, int64_t& rp, int64_t& cp, int64_t& lr,
                     int64_t& lc) {
        tiles = rp = cp = lr = lc = 0;
        for (auto& d : v) {
            d.ta = (d.na + tile - 1) / tile;
            d.tb = (d.nb + tile - 1) / tile;
            d.tile0 = tiles;
            d.rpart = rp;
            d.cpart = cp;
            d.lrow = lr;
            d.lcol = lc;
            tiles += (int64_t)d.ta * d.tb;
            rp += (int64_t)d.tb * d.na;
            cp += (int64_t)d.ta * d.nb;
            lr += d.na;
            lc += d.nb;
        }
    };
    int64_t xt, xrp, xcp, xr, xc, ft, frp, fcp, frr, fcc;
    number(ex, mt::TILE, xt, xrp, xcp, xr, xc);
    number(fl, mt::FTILE, ft, frp, fcp, frr, fcc);
    const size_t nex = ex.size(), nfl = fl.size();
    if (frr + fcc >= (int64_t)0x7fffffff) return fail(MVS_ERR_LIMIT, std::string(who) + ": batch too large");
    HCHK(M.pairs.need((size_t)npairs + nex + nfl + 1));
    HCHK(M.rpart.need((size_t)std::max<int64_t>(xrp, 1)));
    HCHK(M.cpart.need((size_t)std::max<int64_t>(xcp, 1)));
    HCHK(M.fwd.need((size_t)std::max<int64_t>(fr, 1)));
    HCHK(M.bwd.need((size_t)std::max<int64_t>(fc, 1)));
    HCHK(M.rmeta.need((size_t)std::max<int64_t>(frp, 1)));
    HCHK(M.cmeta.need((size_t)std::max<int64_t>(fcp, 1)));
    HCHK(M.rmask.need((size_t)std::max<int64_t>(frp, 1)));
    HCHK(M.cmask.need((size_t)std::max<int64_t>(fcp, 1)));
    HCHK(M.rcnt.need((size_t)std::max<int64_t>(frr, 1)));
    HCHK(M.ccnt.need((size_t)std::max<int64_t>(fcc, 1)));
    HCHK(M.rcand.need((size_t)std::max<int64_t>(frr, 1) * mt::CAP));
    HCHK(M.ccand.need((size_t)std::max<int64_t>(fcc, 1) * mt::CAP));
    HCHK(M.ovf.need((size_t)std::max<int64_t>(frr + fcc, 1)));
    HCHK(M.ovf_n.need(1));
    HCHK(M.oa.need((size_t)std::max<int64_t>(out, 1)));
    HCHK(M.ob.need((size_t)std::max<int64_t>(out, 1)));
    HCHK(M.da.need((size_t)std::max<int64_t>(out, 1)));
    HCHK(M.db.need((size_t)std::max<int64_t>(out, 1)));
    HCHK(M.counts.need((size_t)std::max(npairs, 1)));
    HCHK(M.prefix.need((size_t)std::max(npairs, 1)));
    mt::PairDesc* d_all = M.pairs.p;
    mt::PairDesc* d_ex = M.pairs.p + npairs;
    mt::PairDesc* d_fl = d_ex + nex;
    hipStream_t st = c->stream;
    const size_t nlist = (size_t)npairs + nex + nfl;
    HCHK(M.h_pairs.need(nlist + 1));
    HCHK(M.h_counts.need((size_t)npairs + 1));
    std::copy(all.begin(), all.end(), M.h_pairs.p);
    std::copy(ex.begin(), ex.end(), M.h_pairs.p + npairs);
    std::copy(fl.begin(), fl.end(), M.h_pairs.p + npairs + nex);
    if (nlist > 0) HCHK(hipMemcpyAsync(d_all, M.h_pairs.p, sizeof(mt::PairDesc) * nlist, hipMemcpyHostToDevice, st));
    mt::FiltBufs fb = {M.rmeta.p, M.rmask.p, M.cmeta.p, M.cmask.p, M.rcnt.p,
                       M.ccnt.p,  M.rcand.p, M.ccand.p, M.ovf.p,   M.ovf_n.p};
    HCHK(hipEventRecord(c->ev[0], st));
    HCHK(mt::launch_dist_tiles(S.desc.p, d_ex, (int)nex, xt, M.rpart.p, M.cpart.p, st));
    HCHK(mt::launch_merge_partials(d_ex, (int)nex, xr, xc, M.rpart.p, M.cpart.p, M.fwd.p, M.bwd.p, st));
    HCHK(hipEventRecord(c->ev[1], st));
    HCHK(mt::launch_filtered(S.desc.p, S.norm.p, S.hi(), S.lo(), d_fl, (int)nfl, ft, frr, fcc, fb, M.fwd.p, M.bwd.p,
                             st, c->ev + 2));
    if (ft <= 0)
        for (int e = 2; e < 4; e++) HCHK(hipEventRecord(c->ev[e], st));
    HCHK(hipEventRecord(c->ev[4], st));
    HCHK(mt::launch_select(d_all, npairs, M.fwd.p, M.bwd.p, (float)ratio, M.oa.p, M.ob.p, M.counts.p, st));
    HCHK(hipEventRecord(c->ev[5], st));
    HCHK(mt::launch_pack(d_all, npairs, M.counts.p, M.prefix.p, M.oa.p, M.ob.p, M.da.p, M.db.p, st));
    const auto t_launched = std::chrono::steady_clock::now();
    int64_t* cnt = M.h_counts.p;
    uint32_t novf = 0;
    if (npairs > 0)
        HCHK(hipMemcpyAsync(cnt, M.counts.p, sizeof(int64_t) * (size_t)npairs, hipMemcpyDeviceToHost, st));
    if (ft > 0) HCHK(hipMemcpyAsync(&cnt[npairs], M.ovf_n.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HCHK(hipStreamSynchronize(st));
    if (ft > 0) memcpy(&novf, &cnt[npairs], sizeof(uint32_t));
    const auto t_synced = std::chrono::steady_clock::now();
    // the dense lists come back in one copy per side and are spread into
    // the callers' slots (pair p at all[p].out)
    int64_t nm = 0;
    for (int32_t p = 0; p < npairs; p++) {
        out_n[p] = cnt[p];
        nm += cnt[p];
    }
    if (nm > 0 && (out_a || out_b)) {
        HCHK(M.h_out.need(2 * (size_t)nm));
        uint32_t* ha = M.h_out.p;
        uint32_t* hb = M.h_out.p + nm;
        if (out_a) HCHK(hipMemcpyAsync(ha, M.da.p, sizeof(uint32_t) * (size_t)nm, hipMemcpyDeviceToHost, st));
        if (out_b) HCHK(hipMemcpyAsync(hb, M.db.p, sizeof(uint32_t) * (size_t)nm, hipMemcpyDeviceToHost, st));
        HCHK(hipStreamSynchronize(st));
        int64_t q = 0;
        for (int32_t p = 0; p < npairs; p++) {
            const int64_t n = cnt[p], o = all[(size_t)p].out;
            if (n > 0 && out_a) memcpy(out_a + o, ha + q, sizeof(uint32_t) * (size_t)n);
            if (n > 0 && out_b) memcpy(out_b + o, hb + q, sizeof(uint32_t) * (size_t)n);
            q += n;
        }
    }
    const auto t_done = std::chrono::steady_clock::now();
    float t[5] = {0, 0, 0, 0, 0};
    for (int e = 0; e < 5; e++) (void)hipEventElapsedTime(&t[e], c->ev[e], c->ev[e + 1]);
    mvs_match_stats_t& R = M.stats;
    memset(&R, 0, sizeof(R));
    R.pairs = npairs;
    R.filtered_pairs = (int64_t)nfl;
    for (const auto& q : all) R.distances += (int64_t)q.na * q.nb;
    R.matches = nm;
    R.host_launch_ms = std::chrono::duration<double, std::milli>(t_launched - t_start).count();
    R.host_wait_ms = std::chrono::duration<double, std::milli>(t_synced - t_launched).count();
    R.host_out_ms = std::chrono::duration<double, std::milli>(t_done - t_synced).count();
    R.overflow_items = novf;
    R.exact_ms = t[0];
    R.tile_ms = t[1];
    R.collect_ms = t[2];
    R.recheck_ms = t[3];
    R.select_ms = t[4];
    return MVS_OK;
}

extern "C" {

int mvs_upload_descriptors(mvs_ctx* c, const float* desc, const int32_t* counts, int32_t nframes) {
    if (!c || nframes < 0 || (nframes > 0 && !counts)) return fail(MVS_ERR_ARG, "mvs_upload_descriptors: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int r = set_device(c)) return r;
    try {
        return match_prepare(c, c->match.frames, desc, counts, nframes, "mvs_upload_descriptors");
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_upload_descriptors: ") + e.what());
    }
}

int mvs_match_frames(mvs_ctx* c, const int32_t* pairs, int32_t npairs, double ratio, uint32_t* out_a,
                     uint32_t* out_b, int64_t* out_n) {
    if (!c || npairs < 0 || (npairs > 0 && (!pairs || !out_n)))
        return fail(MVS_ERR_ARG, "mvs_match_frames: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int r = set_device(c)) return r;
    try {
        return match_batch(c, c->match.frames, pairs, npairs, ratio, out_a, out_b, out_n, "mvs_match_frames");
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_match_frames: ") + e.what());
    }
}

int mvs_match_pair(mvs_ctx* c, const float* desc_a, int32_t na, const float* desc_b, int32_t nb, double ratio,
                   uint32_t* out_a, uint32_t* out_b, int64_t* out_n) {
    if (!c || na < 0 || nb < 0 || (na > 0 && !desc_a) || (nb > 0 && !desc_b) || !out_n)
        return fail(MVS_ERR_ARG, "mvs_match_pair: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int r = set_device(c)) return r;
    try {
        // the two sets side by side in one host buffer
        std::vector<float> both(((size_t)na + (size_t)nb) * mt::DIM);
        if (na > 0) memcpy(both.data(), desc_a, sizeof(float) * mt::DIM * (size_t)na);
        if (nb > 0) memcpy(both.data() + (size_t)na * mt::DIM, desc_b, sizeof(float) * mt::DIM * (size_t)nb);
        const int32_t counts[2] = {na, nb};
        if (int r = match_prepare(c, c->match.tmp, both.data(), counts, 2, "mvs_match_pair")) return r;
        const int32_t pr[2] = {0, 1};
        return match_batch(c, c->match.tmp, pr, 1, ratio, out_a, out_b, out_n, "mvs_match_pair");
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_match_pair: ") + e.what());
    }
}

int mvs_match_set_mode(mvs_ctx* c, int32_t mode) {
    if (!c || mode < 0 || mode > 1) return fail(MVS_ERR_ARG, "mvs_match_set_mode: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    c->match.mode = mode;
    return MVS_OK;
}

int mvs_get_match_stats(mvs_ctx* c, mvs_match_stats_t* out) {
    if (!c || !out) return fail(MVS_ERR_ARG, "mvs_get_match_stats: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    *out = c->match.stats;
    return MVS_OK;
}

int mvs_debug_match_bound(mvs_ctx* c, const float* desc_a, int32_t na, const float* desc_b, int32_t nb,
                          float* max_rel) {
    if (!c || na <= 0 || nb <= 0 || !desc_a || !desc_b || !max_rel)
        return fail(MVS_ERR_ARG, "mvs_debug_match_bound: bad arguments");
    std::lock_guard<std::mutex> lk(c->mu);
    if (int r = set_device(c)) return r;
    MatchState& M = c->match;
    try {
        std::vector<float> both(((size_t)na + (size_t)nb) * mt::DIM);
        memcpy(both.data(), desc_a, sizeof(float) * mt::DIM * (size_t)na);
        memcpy(both.data() + (size_t)na * mt::DIM, desc_b, sizeof(float) * mt::DIM * (size_t)nb);
        const int32_t counts[2] = {na, nb};
        if (int r = match_prepare(c, M.tmp, both.data(), counts, 2, "mvs_debug_match_bound")) return r;
        if (!M.tmp.safe[0] || !M.tmp.safe[1])
            return fail(MVS_ERR_ARG, "mvs_debug_match_bound: components outside {0} U [2^-40, 2^40]");
        mt::PairDesc d = {};
        d.a = 0;
        d.b = na;
        d.na = na;
        d.nb = nb;
        d.ta = (na + mt::FTILE - 1) / mt::FTILE;
        d.tb = (nb + mt::FTILE - 1) / mt::FTILE;
        HCHK(M.pairs.need(1));
        HCHK(M.ovf_n.need(1));
        HCHK(hipMemcpyAsync(M.pairs.p, &d, sizeof(d), hipMemcpyHostToDevice, c->stream));
        HCHK(hipMemsetAsync(M.ovf_n.p, 0, sizeof(uint32_t), c->stream));
        mt::FiltBufs fb = {};
        fb.ovf_n = M.ovf_n.p;
        HCHK(mt::launch_filter_check(M.tmp.desc.p, M.tmp.norm.p, M.tmp.hi(), M.tmp.lo(), M.pairs.p, 1,
                                     (int64_t)d.ta * d.tb, fb, c->stream));
        uint32_t bitsv = 0;
        HCHK(hipMemcpyAsync(&bitsv, M.ovf_n.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HCHK(hipStreamSynchronize(c->stream));
        memcpy(max_rel, &bitsv, sizeof(float));
        return MVS_OK;
    } catch (const std::exception& e) {
        return fail(MVS_ERR_NOMEM, std::string("mvs_debug_match_bound: ") + e.what());
    }
}

}  // extern "C"
