// Device-side marching-cubes helpers: cube cases, edge ownership and the reference's vertex order.
//
// Restates Chunk::build_mesh (underwater_world/src/chunk.rs:135-264):
//   - cells are scanned x -> y -> z (linear index (x n + y) n + z), corners in the order of
//     chunk.rs:144-153, bit i of the case set iff rho(corner i) < iso (strict, chunk.rs:159);
//   - each triangle-table entry names an edge (a -> b as listed in marching_table.rs:1-14),
//     lerped at t = (iso - rho_a) / (rho_b - rho_a) in f32 (chunk.rs:178-213);
//   - vertices are deduplicated by the ORDERED corner pair (a, b) and numbered in first-seen
//     order (chunk.rs:233-243).
//
// Edge ownership. A directed edge record (a, b) is shared by two cells (x and z edges, which
// differ in y) or four cells (y edges, which differ in x and z). The tables emit every edge whose
// endpoints straddle the isovalue, so every sharer of an active record emits it, and the record
// is first seen in the sharer that comes first in scan order: its OWNER. For cell (x, y, z):
//   e4 e5 e6 e7 e10      always owned
//   e0 e1 e2 e3          owned iff y == 0   (else by (x, y-1, z) as e4 e5 e6 e7)
//   e8                   owned iff x == 0 and z == 0
//   e9                   owned iff z == 0   (else by (x, y, z-1) as e10)
//   e11                  owned iff x == 0   (else by (x-1, y, z) as e10)
// Hence the reference's vertex id of a record is
//   (owned active edges of all cells before the owner) + (rank of the edge among the owner's
//   owned edges in first-appearance order),
// which every cell can compute locally from a per-cell prefix sum.
#pragma once

#include "uw_internal.h"

// Corner offsets bit0 = dx, bit1 = dy, bit2 = dz (chunk.rs:144-153).
#define UW_CORNER_OFS_PACKED 0x67324510u /* nibbles: corner 0..7 -> 0,1,5,4,2,3,7,6 */
#define UW_EDGE_A_PACKED 0x321076543210ull /* nibbles: edge 0..11 -> start corner */
#define UW_EDGE_B_PACKED 0x765447650321ull /* nibbles: edge 0..11 -> end corner   */

__device__ __forceinline__ uint32_t uw_corner_ofs(uint32_t i) { return (UW_CORNER_OFS_PACKED >> (4 * i)) & 15u; }
__device__ __forceinline__ uint32_t uw_edge_a(uint32_t e) { return uint32_t(UW_EDGE_A_PACKED >> (4 * e)) & 15u; }
__device__ __forceinline__ uint32_t uw_edge_b(uint32_t e) { return uint32_t(UW_EDGE_B_PACKED >> (4 * e)) & 15u; }

__device__ __forceinline__ uint32_t uw_owned_mask(int32_t x, int32_t y, int32_t z) {
    uint32_t m = 0x4F0u;  // e4 e5 e6 e7 e10
    if (y == 0) m |= 0x00Fu;
    if (x == 0 && z == 0) m |= 1u << 8;
    if (z == 0) m |= 1u << 9;
    if (x == 0) m |= 1u << 11;
    return m;
}

// Owner of the record that cell (x, y, z) sees as edge e, when the cell does not own it.
__device__ __forceinline__ void uw_edge_owner(int32_t x, int32_t y, int32_t z, uint32_t e, int32_t* ox,
                                              int32_t* oy, int32_t* oz, uint32_t* oe) {
    *ox = x;
    *oy = y;
    *oz = z;
    if (e < 4) {
        *oy = y - 1;
        *oe = e + 4;
    } else if (e == 8) {
        if (x > 0 && z > 0) { *ox = x - 1; *oz = z - 1; *oe = 10; }
        else if (x > 0) { *ox = x - 1; *oe = 9; }
        else { *oz = z - 1; *oe = 11; }
    } else if (e == 9) {
        *oz = z - 1;
        *oe = 10;
    } else {  // e == 11
        *ox = x - 1;
        *oe = 10;
    }
}

// Decode a chunk-local cell index g = (x n + y) n + z.
__device__ __forceinline__ void uw_cell_xyz(uint32_t g, uw_divider dn, uw_divider dnn, int32_t* x, int32_t* y,
                                            int32_t* z) {
    const uint32_t xq = uw_div(g, dnn);
    const uint32_t r = g - xq * dnn.d;
    const uint32_t yq = uw_div(r, dn);
    *x = int32_t(xq);
    *y = int32_t(yq);
    *z = int32_t(r - yq * dn.d);
}

__device__ __forceinline__ uint32_t uw_case_of(const float r[8], float iso) {
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) c |= (r[i] < iso ? 1u : 0u) << i;
    return c;
}

// Vertex colour of the reference's terrain (chunk.rs:215-222), f32 in the reference's order:
// hue from the world height (create_mix_ratio over world.rs MIN_Z..MAX_Z = -2..2 in chunk units,
// util.rs:93-95), value from the edge's end corner index mod 3, hsv_to_rgb (util.rs:122-153,
// saturation 0.6, base value 0.4, chunk.rs:14-17), then to_srgb (util.rs:106-111).
__device__ __forceinline__ void uw_vertex_color(float world_z, uint32_t corner_b, float extent, float rgb[3]) {
    const float ratio = world_z / extent;
    const float mix = (ratio - (-2.0f)) / (2.0f - (-2.0f));
    const float value = 0.4f + float(corner_b % 3u) / 9.0f;
    float hue = -150.0f + (60.0f - (-150.0f)) * mix;
    hue = fmodf(hue, 360.0f);
    if (hue < 0.0f) hue = hue + 360.0f;
    const float c = value * 0.6f;
    const float h = hue / 60.0f;
    const float x = c * (1.0f - fabsf(fmodf(h, 2.0f) - 1.0f));
    const float m = value - c;
    float r, g, b;
    if (0.0f <= h && h < 1.0f) { r = c; g = x; b = 0.0f; }
    else if (1.0f <= h && h < 2.0f) { r = x; g = c; b = 0.0f; }
    else if (2.0f <= h && h < 3.0f) { r = 0.0f; g = c; b = x; }
    else if (3.0f <= h && h < 4.0f) { r = 0.0f; g = x; b = c; }
    else if (4.0f <= h && h < 5.0f) { r = x; g = 0.0f; b = c; }
    else { r = c; g = 0.0f; b = x; }
    const float lin[3] = {(r + m) * 255.0f, (g + m) * 255.0f, (b + m) * 255.0f};
#pragma unroll
    for (int i = 0; i < 3; ++i) rgb[i] = powf((lin[i] / 255.0f + 0.055f) / 1.055f, 2.4f);
}

// Collision-triangle normal (util::Tri::new, util.rs:12-20, with cgmath's Vector3 cross, dot and
// magnitude): n = (b - a) x (c - a); n / |n| unless |n| == 0, then n itself. f32, the reference's
// operation order: |n| = sqrt((x x + y y) + z z).
__device__ __forceinline__ void uw_tri_normal(const float a[3], const float b[3], const float c[3], float out[3]) {
    const float ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    const float vx = c[0] - a[0], vy = c[1] - a[1], vz = c[2] - a[2];
    const float nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
    const float mag = sqrtf(nx * nx + ny * ny + nz * nz);
    if (mag == 0.0f) {
        out[0] = nx; out[1] = ny; out[2] = nz;
    } else {
        out[0] = nx / mag; out[1] = ny / mag; out[2] = nz / mag;
    }
}

// Copies the first-seen rank masks (uw_tables::before, 6144 bytes) into a 16-byte aligned LDS
// array with 16-byte loads, every load of a thread issued before its first store (a loop of
// 2-byte loads, each waited for before its store, costs one memory round trip per element).
__device__ __forceinline__ void uw_lds_load_before(uint16_t* dst, const uw_tables* tab, int tid, int nthr) {
    constexpr int kWords = 256 * 12 * 2 / 16;
    const uint4* src = reinterpret_cast<const uint4*>(&tab->before[0][0]);
    uint4* d = reinterpret_cast<uint4*>(dst);
    const int e0 = tid, e1 = tid + nthr, e2 = tid + 2 * nthr;
    uint4 v0 = make_uint4(0, 0, 0, 0), v1 = v0, v2 = v0;
    if (e0 < kWords) v0 = src[e0];
    if (e1 < kWords) v1 = src[e1];
    if (e2 < kWords) v2 = src[e2];
    if (e0 < kWords) d[e0] = v0;
    if (e1 < kWords) d[e1] = v1;
    if (e2 < kWords) d[e2] = v2;
    for (int e = tid + 3 * nthr; e < kWords; e += nthr) d[e] = src[e];
}

// Per-launch output options shared by the emit kernels.
struct uw_emit_opts {
    uint32_t vstride;     // floats per vertex: 3 (xyz) or 6 (xyz + sRGB colour)
    uint32_t ind16;       // 1: indices are written as u16 (chunk.rs:243)
    float* tri_normals;   // optional, 3 floats per triangle
    uint32_t* tri_cells;  // optional, 1 u32 per triangle
};

__host__ inline uw_emit_opts uw_make_emit_opts(const uw_mesh_out* out) {
    uw_emit_opts o;
    o.vstride = out->vertex_format == UW_VERTEX_XYZ_RGB ? 6u : 3u;
    o.ind16 = out->index_format == UW_INDEX_U16 ? 1u : 0u;
    o.tri_normals = out->tri_normals;
    o.tri_cells = out->tri_cells;
    return o;
}

__device__ __forceinline__ void uw_store_index(uint32_t* inds, uint32_t ind16, uint64_t at, uint32_t id) {
    if (ind16) reinterpret_cast<uint16_t*>(inds)[at] = uint16_t(id);
    else inds[at] = id;
}

// World position of the lerp on edge e of cell (x, y, z) (chunk.rs:185-229) from the densities ra,
// rb of its start and end corners, f32 with the reference's operation order (the file is compiled
// with -ffp-contract=off).
__device__ __forceinline__ void uw_edge_vertex_ab(uint32_t e, int32_t x, int32_t y, int32_t z, float ra, float rb,
                                                  float iso, float step, float ox, float oy, float oz, float out[3]) {
    const uint32_t ca = uw_corner_ofs(uw_edge_a(e)), cb = uw_corner_ofs(uw_edge_b(e));
    const float t = (iso - ra) / (rb - ra);
    const float sax = float(x + int32_t(ca & 1u)) * step, sbx = float(x + int32_t(cb & 1u)) * step;
    const float say = float(y + int32_t((ca >> 1) & 1u)) * step, sby = float(y + int32_t((cb >> 1) & 1u)) * step;
    const float saz = float(z + int32_t((ca >> 2) & 1u)) * step, sbz = float(z + int32_t((cb >> 2) & 1u)) * step;
    out[0] = (sax + t * (sbx - sax)) + ox;
    out[1] = (say + t * (sby - say)) + oy;
    out[2] = (saz + t * (sbz - saz)) + oz;
}

// The same from the cell's eight corner densities r (corner order of chunk.rs:144-153).
__device__ __forceinline__ void uw_edge_vertex(uint32_t e, int32_t x, int32_t y, int32_t z, const float r[8],
                                               float iso, float step, float ox, float oy, float oz, float out[3]) {
    uw_edge_vertex_ab(e, x, y, z, r[uw_edge_a(e)], r[uw_edge_b(e)], iso, step, ox, oy, oz, out);
}
