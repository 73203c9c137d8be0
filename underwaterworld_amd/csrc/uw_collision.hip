// Collision-triangle lookup: Chunk::tris_around (underwater_world/src/chunk.rs:315-342) over the
// per-triangle cell records that the emit kernels write (tri_cells). Triangles are emitted in cell
// scan order, so a chunk's tri_cells are non-decreasing and every (x, y) row of the query box is
// one contiguous triangle range, found by two binary searches instead of the reference's
// per-cell HashMap lookups.
#include <algorithm>
#include <cmath>

#include "uw_internal.h"

namespace {

// Rust's `f32 as i32`: truncation toward zero, saturating, NaN -> 0.
int64_t rust_f32_to_i32(float v) {
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return -2147483647 - 1;
    return int64_t(v);
}

}  // namespace

extern "C" {

int uw_tris_around(const uint32_t* tri_cells, uint32_t ntris, int32_t n, float px, float py, float pz, int32_t range,
                   uint32_t* ids, uint32_t cap, uint32_t* count) {
    if (!count || n < 1 || n > UW_MAX_RESOLUTION || (ntris > 0 && !tri_cells) || (cap > 0 && !ids))
        return UW_ERR_INVALID_ARG;
    *count = 0;
    const float pct[3] = {px, py, pz};
    int64_t lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        const int64_t m = rust_f32_to_i32(std::floor(pct[k] * float(n)));  // chunk.rs:316-318
        lo[k] = std::max<int64_t>(m - range, 0);                            // chunk.rs:320-322
        hi[k] = std::min<int64_t>(m + range, n);                            // chunk.rs:324-326
        hi[k] = std::min<int64_t>(hi[k], n - 1);  // cell n has no entry in the reference's map
    }
    const uint32_t* end = tri_cells + ntris;
    uint64_t total = 0;
    for (int64_t x = lo[0]; x <= hi[0]; ++x)
        for (int64_t y = lo[1]; y <= hi[1]; ++y) {
            if (lo[2] > hi[2]) continue;
            const uint32_t g0 = uint32_t((x * n + y) * n + lo[2]), g1 = uint32_t((x * n + y) * n + hi[2]);
            const uint32_t* a = std::lower_bound(tri_cells, end, g0);
            const uint32_t* b = std::upper_bound(a, end, g1);
            for (const uint32_t* t = a; t < b; ++t, ++total)
                if (total < cap) ids[total] = uint32_t(t - tri_cells);
        }
    *count = uint32_t(total);
    return total > cap ? UW_ERR_OVERFLOW : UW_OK;
}

}  // extern "C"
