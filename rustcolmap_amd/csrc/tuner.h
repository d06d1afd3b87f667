// tuner.h -- launch-shape table for the stereo stages.
//
// A shape class is (longest image side bucket, K bucket).  The table starts
// from built-in defaults and is overwritten by mvs_autotune, which times the
// candidate tile shapes / schedules on the live device for the batch it is
// given.  MVS_TILE="WxH[p]" in the environment pins a shape (p = persistent).
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <utility>

#include "stereo_kernels.h"

namespace mvs {

struct Tuner {
    std::map<std::pair<int, int>, TileCfg> table;
    bool forced = false;
    TileCfg force_cfg;

    static std::pair<int, int> shape_class(int maxw, int maxh, int maxk) {
        const int side = maxw > maxh ? maxw : maxh;
        const int sb = side <= 512 ? 0 : (side <= 1024 ? 1 : (side <= 2048 ? 2 : 3));
        const int kb = maxk <= 4 ? 0 : (maxk <= 8 ? 1 : 2);
        return {sb, kb};
    }

    // "WxH[p|s][r|v|w|x]": tile shape, p = persistent queue, s = persistent
    // queue with one wave per tile in rounds sorted by the previous pass's
    // cost, evaluator r = reference order, v = divide-free abort, w =
    // per-tap geometry, exact scaling-free divide and strengthened abort
    // with the source-0 rows 1-2 batched, x = as w row by row (default:
    // hoisted geometry)
    static bool parse(const char* s, TileCfg& c) {
        int w = 0, h = 0, used = 0;
        if (sscanf(s, "%dx%d%n", &w, &h, &used) < 2) return false;
        c.tw = w;
        c.th = h;
        c.persistent = strchr(s + used, 's') ? 2 : strchr(s + used, 'p') ? 1 : 0;
        const char* t = s + used;
        c.eval = strchr(t, 'r') ? 0 : strchr(t, 'v') ? 2 : strchr(t, 'w') ? 3 : strchr(t, 'x') ? 4 : 1;
        return true;
    }

    // Built-in default, used until mvs_autotune has timed the shape class:
    // 16x16 tiles in the sorted form once the batch gives every wave slot of
    // the machine a few such tiles (at 6 waves x 4 SIMDs x 256 CUs, 16 Mpx is
    // about 10 tiles per slot), else 8x8 tiles on the persistent queue, whose
    // shorter tiles keep the end of each launch short.  Measured pass time,
    // sorted 24x8 against 8x8: -5 % on 20 x 1920x1080, -6.6 % on 100 x
    // 1920x1080, -9.3 % on 64 views of 640x480 .. 2560x1440, +29 % on 10 x
    // 320x240; on 20 x 1920x1080 the three sorted shapes are within 3 %,
    // 16x16 the fastest (16 launches: 168 ms against 172 at 16x8 and 173 at
    // 24x8), as it also is in the tuner's one-pass estimates run after run.
    static TileCfg default_cfg(int64_t total_px) {
        TileCfg c;   // 8x8, persistent, evaluator 3
        if (total_px >= ((int64_t)16 << 20)) {
            c.tw = 16;
            c.th = 16;
            c.persistent = 2;
        }
        return c;
    }

    TileCfg pick(int maxw, int maxh, int maxk, int64_t total_px) const {
        if (forced) return force_cfg;
        const char* env = getenv("MVS_TILE");
        TileCfg c;
        if (env && parse(env, c)) return c;
        auto it = table.find(shape_class(maxw, maxh, maxk));
        if (it != table.end()) return it->second;
        return default_cfg(total_px);
    }
};

}  // namespace mvs
