// plan.cpp -- host scheduling of run_mvs: co-visibility, per-view depth
// ranges, stereo sources and fusion order (mvs/src/lib.rs:224-282,
// 304-347, 398-418), as native code feeding the batched stereo schedule.
//
// Tie order: the reference sorts candidates with sort_unstable_by on the
// shared-point count only (lib.rs:335, 407), so equal counts come out in an
// order its algorithm does not define.  Here equal counts keep ascending
// view index (a stable sort), which is one of the orders the reference may
// produce.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/mvs_hip.h"

namespace {

// percentile (lib.rs:245-251): sorted[round((len - 1) * q) clamped]
double percentile(const std::vector<double>& s, double q) {
    if (s.empty()) return 0.0;
    double idx = std::round(((double)s.size() - 1.0) * q);
    idx = std::min(std::max(idx, 0.0), (double)s.size() - 1.0);
    return s[(size_t)idx];
}

}  // namespace

extern "C" {

int mvs_plan(const mvs_view_t* views, int32_t nviews, const double* xyz, const int32_t* track_off,
             const int32_t* track_view, int64_t npoints, const mvs_options_t* opts, int32_t* src_off, int32_t* srcs,
             double* zrange, int32_t* fusion_off, int32_t* fusion, uint32_t* covis) {
    if (!views || nviews < 2 || !opts || (npoints > 0 && (!xyz || !track_off || !track_view)))
        return MVS_ERR_ARG;
    if (npoints < 8) return MVS_ERR_STATE;   // lib.rs:310-315
    const int n = nviews;
    // global depths: |p| of every sparse point, sorted (lib.rs:304-309)
    std::vector<double> global((size_t)npoints);
    for (int64_t i = 0; i < npoints; i++) {
        const double* p = xyz + 3 * i;
        global[i] = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    }
    std::sort(global.begin(), global.end());
    // co-visibility (lib.rs:224-243): per point, deduplicated views of its track
    std::vector<uint32_t> cov((size_t)n * n, 0);
    std::vector<int32_t> ids;
    for (int64_t i = 0; i < npoints; i++) {
        ids.assign(track_view + track_off[i], track_view + track_off[i + 1]);
        ids.erase(std::remove_if(ids.begin(), ids.end(), [&](int v) { return v < 0 || v >= n; }), ids.end());
        std::sort(ids.begin(), ids.end());
        ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
        for (int a : ids)
            for (int b : ids)
                if (a != b) cov[(size_t)a * n + b] += 1;
    }
    if (covis) memcpy(covis, cov.data(), cov.size() * sizeof(uint32_t));
    // sources: top num_sources views with >= 10 shared points (lib.rs:326-342)
    std::vector<std::vector<int>> sources(n);
    int pos = 0;
    for (int i = 0; i < n; i++) {
        std::vector<std::pair<uint32_t, int>> order;
        for (int j = 0; j < n; j++)
            if (cov[(size_t)i * n + j] >= 10) order.push_back({cov[(size_t)i * n + j], j});
        std::stable_sort(order.begin(), order.end(), [](const std::pair<uint32_t, int>& a,
                                                        const std::pair<uint32_t, int>& b) { return a.first > b.first; });
        for (size_t q = 0; q < order.size() && q < opts->num_sources; q++) sources[i].push_back(order[q].second);
        src_off[i] = pos;
        for (int j : sources[i]) srcs[pos++] = j;
    }
    src_off[n] = pos;
    // depth ranges (lib.rs:253-282)
    for (int i = 0; i < n; i++) {
        const mvs_view_t& v = views[i];
        std::vector<double> ds;
        for (int64_t p = 0; p < npoints; p++) {
            bool in_track = false;
            for (int q = track_off[p]; q < track_off[p + 1]; q++)
                if (track_view[q] == i) {
                    in_track = true;
                    break;
                }
            if (!in_track) continue;
            const double* X = xyz + 3 * p;
            const double zc = ((v.R[6] * X[0] + v.R[7] * X[1]) + v.R[8] * X[2]) + v.t[2];
            if (zc > 1e-6) ds.push_back(zc);
        }
        const std::vector<double>* src = &global;
        if (ds.size() >= 8) {
            std::sort(ds.begin(), ds.end());
            src = &ds;
        }
        double lo = 0.0, hi = 0.0;
        if (src->size() >= 8) {
            lo = percentile(*src, 0.02);
            hi = percentile(*src, 0.98);
        }
        if (hi <= lo * 1.05 || lo <= 0.0) {
            const double glo = percentile(global, 0.01), ghi = percentile(global, 0.99);
            zrange[2 * i] = std::max(glo * 0.7, 1e-3);
            zrange[2 * i + 1] = ghi * 1.4;
        } else {
            zrange[2 * i] = lo * 0.8;
            zrange[2 * i + 1] = hi * 1.25;
        }
    }
    // fusion order (lib.rs:398-418): co-visible views (>= 10) that have a
    // depth map (= have sources), by count, then stereo sources moved first
    // in their source order
    pos = 0;
    for (int i = 0; i < n; i++) {
        std::vector<std::pair<uint32_t, int>> order;
        for (int j = 0; j < n; j++)
            if (cov[(size_t)i * n + j] >= 10 && !sources[j].empty()) order.push_back({cov[(size_t)i * n + j], j});
        std::stable_sort(order.begin(), order.end(), [](const std::pair<uint32_t, int>& a,
                                                        const std::pair<uint32_t, int>& b) { return a.first > b.first; });
        std::vector<int> ord;
        for (auto& o : order) ord.push_back(o.second);
        const std::vector<int>& s = sources[i];
        auto key = [&](int j) -> size_t {
            auto it = std::find(s.begin(), s.end(), j);
            return it == s.end() ? (size_t)-1 : (size_t)(it - s.begin());
        };
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key(a) < key(b); });
        fusion_off[i] = pos;
        for (int j : ord) fusion[pos++] = j;
    }
    fusion_off[n] = pos;
    return MVS_OK;
}

}  // extern "C"
