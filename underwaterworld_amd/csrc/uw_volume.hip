// Large single volume: ONE chunk of n cells per axis (n up to UW_VOLUME_MAX_RESOLUTION) meshed as
// one grid with one index space -- Chunk::new(pos) + build_full (chunk.rs:89-103,266-313) for a
// chunk too large for the per-chunk paths. Output equals the reference's builder run on that
// chunk: densities of chunk.rs:105-129, cases of chunk.rs:141-164, lerps of chunk.rs:178-231 and
// first-seen vertex numbering over the whole grid (chunk.rs:233-243).
//
// Kernels (one launch each, no host synchronisation between them):
//   k_vol_tables   per-axis octave entries of the whole grid (x hashed), fade derivatives, the z
//                  records of the line evaluator, and the launch's counters reset
//   k_vol_density  densities of every sample into HBM; one lane per z-line, lanes of a wave on
//                  consecutive lines, so each sample's store is one coalesced row
//   k_vol_count    per 32^3 tile: cube cases and, per row segment (x, y, 32-cell z range), the
//                  owned vertex records and triangles
//   k_vol_scan_*   exclusive scan of the row-segment counts in scan order (x, y, z range): the
//                  first vertex id and first triangle of every row segment; totals and the record
//   k_vol_emit     per tile with a one-cell halo on its low sides: vertex ids from the row bases
//                  plus a within-row prefix (the halo gives the ids of records owned by cells of
//                  neighbouring tiles), vertices, indices, optional tape / colour / collision
//
// HBM layout: densities f32 [(x L + z) L + y] (L = n + 1, y fastest: a wave's store covers 64
// consecutive y); row-segment counts u64 [(x n + y) T + tz] (T = ceil(n / 32)) holding
// vertex records | triangles << 32, scanned in place.
#include <hip/hip_runtime.h>
#include <string.h>

#include "uw_mesh.h"
#include "uw_noise.h"
#include "uw_wave.h"

namespace {

constexpr int kVolBlock = 256;
constexpr int kTile = 32;           // cells per tile axis
constexpr int kTileS = kTile + 2;   // samples per tile axis with the low halo: 34
constexpr int kPitch = kTileS + 1;  // LDS row pitch of a sample plane (odd: no bank aliasing along z)
constexpr int kScanPer = 16;        // row segments per thread in the scan's block passes
constexpr int kScanBlock = kVolBlock * kScanPer;

// Counters of one build (device): early-blank bits, active cells.
struct VolCounters {
    uint32_t below;
    uint32_t active;
};

struct VolArgs {
    uw_field f;
    const uw_tables* tab;
    int32_t c[3];      // chunk coordinate
    int32_t n, L, T;   // cells, samples per axis, tiles per axis
    float step;
    const uw_axis* axis;  // [3][O][L] (x hashed)
    const float* dfade;   // [3][O][L] fade derivative 30 d^2 (d - 1)^2
    const uint32_t* zrec; // L + 1 z records (uw_zrec_words(O) words each)
    float* dens;          // [(x L + z) L + y]
    unsigned long long* rows;  // [(x n + y) T + tz]
    unsigned long long* bsum;  // per scan block
    VolCounters* cnt;
    uw_chunk_info* info;
    uw_totals* totals;
    float* verts;
    uint32_t* inds;
    uw_vertex_tape* tape;
    uw_emit_opts eo;
    uint64_t vcap, icap;
    int32_t zseg, zlen;  // z segments per line in the density kernel
    int32_t nx_tab;      // x entries per workgroup in the density kernel's LDS table
};

// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kVolBlock) k_vol_tables(VolArgs a, uw_axis* axis, float* dfade, uint32_t* zrec) {
    const int32_t O = a.f.octaves, L = a.L;
    const int64_t nent = int64_t(3) * O * L;
    const int64_t g = int64_t(blockIdx.x) * kVolBlock + threadIdx.x;
    if (g == 0) {
        VolCounters z = {0u, 0u};
        *a.cnt = z;
    }
    if (g < nent) {
        const int32_t ax = int32_t(g / (int64_t(O) * L));
        const int32_t r = int32_t(g - int64_t(ax) * O * L);
        const int32_t k = r / L, i = r - k * L;
        const double p = uw_sample_coord(a.f, ax, a.c[ax], i, a.step);
        uw_axis e = uw_axis_of(p, a.f.octave_freq[k]);
        if (ax == 0) {
            const uint32_t lo = e.bits & 255u, hi = (e.bits >> 8) & 255u;
            e.bits = uint32_t(a.tab->perm[lo]) | (uint32_t(a.tab->perm[hi]) << 8);
        }
        axis[g] = e;
        const double pk = p * a.f.octave_freq[k];
        const double d = pk - floor(pk);
        dfade[g] = float(30.0 * d * d * (d - 1.0) * (d - 1.0));
    } else if (g < nent + L + 1) {
        const int32_t i = int32_t(g - nent);
        uw_fill_zrec(a.f, a.c[2], L, a.step, i, zrec + size_t(i) * uw_zrec_words(O));
    }
}

// ---------------------------------------------------------------------------------------------
// Densities: lines t = x L + y, one per lane, workgroup b takes lines [256 b, 256 b + 256) over
// z segment s. The line's y entries are copied per lane into LDS, its x entries per distinct x of
// the workgroup.
template <int OCT>
__global__ void __launch_bounds__(kVolBlock) k_vol_density(VolArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int32_t O = uw_octaves<OCT>::count(a.f);
    const int32_t L = a.L;
    uint8_t* GC = smem;  // gradient code of each permutation slot (uw_gcode)
    uint8_t* P = GC + 256;
    uw_axis* AYs = reinterpret_cast<uw_axis*>(P + 256);  // [O][256] per line
    uw_axis* AXs = AYs + O * kVolBlock;                  // [O][nx_tab] per distinct x
    __shared__ uint32_t s_bits;
    const int tid = threadIdx.x;
    const int32_t seg = int32_t(blockIdx.x % uint32_t(a.zseg));
    const int64_t blk = int64_t(blockIdx.x / uint32_t(a.zseg));
    const int64_t lines = int64_t(L) * L;
    const int64_t t0 = blk * kVolBlock;
    const int64_t t = t0 + tid;
    const bool valid = t < lines;
    const int32_t xfirst = int32_t(t0 / L);
    const int32_t x = valid ? int32_t(t / L) : xfirst, y = valid ? int32_t(t - int64_t(x) * L) : 0;
    if (tid == 0) s_bits = 0u;
    for (int i = tid; i < 256; i += kVolBlock) {
        const uint8_t p = a.tab->perm[i];
        P[i] = p;
        GC[i] = uw_gcode_of(p);
    }
    for (int k = 0; k < O; ++k) AYs[k * kVolBlock + tid] = a.axis[(size_t(O) + k) * L + y];
    for (int e = tid; e < O * a.nx_tab; e += kVolBlock) {
        const int k = e / a.nx_tab, j = e - k * a.nx_tab;
        const int32_t xx = min(xfirst + j, L - 1);  // entries past the grid are never read
        AXs[e] = a.axis[size_t(k) * L + xx];
    }
    __syncthreads();
    const int32_t z0 = seg * a.zlen, z1 = min(L, z0 + a.zlen);
    float vmin = INFINITY, vmax = -INFINITY;
    if (z0 < z1) {
        const int32_t rw = OCT > 0 ? uw_zrec_words(OCT) : uw_zrec_words(O);
        const uw_axis* ax = AXs + (x - xfirst);
        const uw_axis* ay = AYs + tid;
        const int32_t nxt = a.nx_tab;
        float* out = a.dens + (size_t(x) * L + z0) * L + y;
        const size_t zstride = size_t(L);
        uw_noise_zline<OCT, true>(
            a.f, [=](int k) { return ax[k * nxt]; }, [=](int k) { return ay[k * kVolBlock]; }, P, uw_gcode{GC},
            a.zrec + size_t(z0) * rw, rw, z1 - z0, vmin, vmax, [=](int32_t iz, float rho) {
                if (valid) out[size_t(iz) * zstride] = rho;
            });
    }
    const int bits = valid ? uw_below_bits(vmin, vmax, a.f.iso) : 0;
    if (bits) atomicOr(&s_bits, uint32_t(bits));
    __syncthreads();
    if (tid == 0 && s_bits) atomicOr(&a.cnt->below, s_bits);
}

// ---------------------------------------------------------------------------------------------
// Tile helpers. A tile covers cells [32 t, min(32 t + 32, n)) per axis; the emit kernel adds the
// cell plane / row / column just below it (when it exists) as a halo.

// Inclusive scan within each 32-lane half of a wave (rows 0-1 and rows 2-3 scanned separately).
__device__ __forceinline__ uint32_t uw_half_incl_scan(uint32_t v) {
    v += uw_dpp_u32<UW_DPP_ROW_SHR(1), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_SHR(2), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_SHR(4), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_SHR(8), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_BCAST15, 0xA>(v);
    return v;
}

// Sample plane x, y in [ys, ys + sy), z in [zs, zs + sz), through registers (kPlaneRegs floats per
// thread) into an LDS plane [z - zs][y - ys] of pitch kPitch: the loads are issued one plane ahead
// and their registers stored into LDS when the plane's slot is free.
constexpr int kPlaneRegs = (kTileS * kTileS + kVolBlock - 1) / kVolBlock;
__device__ __forceinline__ void fetch_plane(const VolArgs& a, int32_t x, int32_t ys, int32_t sy, int32_t zs, int32_t sz,
                                            float (&r)[kPlaneRegs]) {
    const float* src = a.dens + size_t(x) * a.L * a.L;
#pragma unroll
    for (int i = 0; i < kPlaneRegs; ++i) {
        const int e = int(threadIdx.x) + i * kVolBlock;
        const int zz = e / sy, yy = e - zz * sy;
        r[i] = e < sy * sz ? src[size_t(zs + zz) * a.L + ys + yy] : 0.0f;
    }
}
__device__ __forceinline__ void put_plane(int32_t sy, int32_t sz, const float (&r)[kPlaneRegs], float* dst) {
#pragma unroll
    for (int i = 0; i < kPlaneRegs; ++i) {
        const int e = int(threadIdx.x) + i * kVolBlock;
        const int zz = e / sy, yy = e - zz * sy;
        if (e < sy * sz) dst[zz * kPitch + yy] = r[i];
    }
}
// Corner densities of cell (x, y, z) from the LDS sample planes p0 (plane x) and p1 (plane x + 1)
// whose origin is (ys, zs).
__device__ __forceinline__ void tile_corners(const float* p0, const float* p1, int32_t y, int32_t z, int32_t ys,
                                             int32_t zs, float r[8]) {
    const int o = (z - zs) * kPitch + (y - ys);
    r[0] = p0[o];
    r[1] = p1[o];
    r[2] = p1[o + kPitch];
    r[3] = p0[o + kPitch];
    r[4] = p0[o + 1];
    r[5] = p1[o + 1];
    r[6] = p1[o + kPitch + 1];
    r[7] = p0[o + kPitch + 1];
}

// Workgroup b takes x-slab (b % slabs) of tile b / slabs: cell planes [X0, X1) of `planes` each.
__device__ __forceinline__ void tile_of(const VolArgs& a, int32_t planes, int32_t* X0, int32_t* Y0, int32_t* Z0,
                                        int32_t* X1, int32_t* Y1, int32_t* Z1, int32_t* tz) {
    const int32_t T = a.T;
    const int32_t slabs = kTile / planes;
    const int32_t b = int32_t(blockIdx.x) / slabs, sl = int32_t(blockIdx.x) - b * slabs;
    const int32_t tx = b / (T * T), ty = (b / T) % T;
    *tz = b % T;
    *X0 = tx * kTile + sl * planes;
    *Y0 = ty * kTile;
    *Z0 = *tz * kTile;
    *X1 = min(min(a.n, tx * kTile + kTile), *X0 + planes);
    *Y1 = min(a.n, *Y0 + kTile);
    *Z1 = min(a.n, *Z0 + kTile);
}

constexpr int32_t kCountPlanes = 8;  // cell planes per count workgroup (a quarter of a tile)
// Cell planes emitted per emit workgroup (plus one halo plane): 2 / 4 / 8 / 16 measured 1.70 /
// 1.44 / 1.34 / 1.33 ms on the 512^3 grid (the halo plane's classification is repeated once per slab;
// 16 leaves the LDS at the four-workgroup limit).
constexpr int32_t kEmitPlanes = 8;

// Per-row-segment counts of one x-slab of a tile.
__global__ void __launch_bounds__(kVolBlock) k_vol_count(VolArgs a) {
    __shared__ float SP[2][kTileS * kPitch];
    __shared__ uint32_t LUT[256];
    __shared__ uint32_t s_act;
    int32_t X0, Y0, Z0, X1, Y1, Z1, tz;
    tile_of(a, kCountPlanes, &X0, &Y0, &Z0, &X1, &Y1, &Z1, &tz);
    if (X0 >= X1) return;  // a slab past the grid's last plane (whole workgroup)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < 256; i += kVolBlock) LUT[i] = uint32_t(a.tab->used[i]) | (uint32_t(a.tab->tri[i] & 15u) << 16);
    if (tid == 0) s_act = 0u;
    const int32_t sy = Y1 - Y0 + 1, sz = Z1 - Z0 + 1;
    // sample plane x + 2 is fetched into registers while plane x is classified
    float pre[kPlaneRegs];
    fetch_plane(a, X0, Y0, sy, Z0, sz, pre);
    put_plane(sy, sz, pre, SP[X0 & 1]);
    fetch_plane(a, X0 + 1, Y0, sy, Z0, sz, pre);
    const float iso = a.f.iso;
    uint32_t act = 0;
    for (int32_t x = X0; x < X1; ++x) {
        put_plane(sy, sz, pre, SP[(x + 1) & 1]);
        if (x + 2 <= X1) fetch_plane(a, x + 2, Y0, sy, Z0, sz, pre);  // block-uniform
        __syncthreads();
        const float* p0 = SP[x & 1];
        const float* p1 = SP[(x + 1) & 1];
        // rows y (two per wave, 32 z cells each)
        for (int32_t q = wave; 2 * q < Y1 - Y0; q += kVolBlock / 64) {
            const int32_t y = Y0 + 2 * q + (lane >> 5), z = Z0 + (lane & 31);
            uint32_t v = 0;
            if (y < Y1 && z < Z1) {
                float r[8];
                tile_corners(p0, p1, y, z, Y0, Z0, r);
                const uint32_t cs = uw_case_of(r, iso);
                const uint32_t w = LUT[cs];
                const uint32_t nv = __popc(w & 0xFFFu & uw_owned_mask(x, y, z));
                const uint32_t nt = w >> 16;
                const uint32_t ac = (cs != 0u && cs != 255u) ? 1u : 0u;
                v = nv | (nt << 10) | (ac << 20);
            }
            const uint32_t inc = uw_half_incl_scan(v);
            if ((lane & 31) == 31 && y < Y1) {
                a.rows[(size_t(x) * a.n + y) * a.T + tz] =
                    (unsigned long long)(inc & 0x3FFu) | ((unsigned long long)((inc >> 10) & 0x3FFu) << 32);
                act += inc >> 20;
            }
        }
        __syncthreads();  // plane x is read no more; its slot takes plane x + 2
    }
    if (act) atomicAdd(&s_act, act);
    __syncthreads();
    if (tid == 0 && s_act) atomicAdd(&a.cnt->active, s_act);
}

// ---------------------------------------------------------------------------------------------
// Exclusive scan of the row-segment counts (u64 pairs that never carry between halves: both
// totals stay below 2^32 for n <= UW_VOLUME_MAX_RESOLUTION).
__device__ __forceinline__ unsigned long long vol_block_scan(unsigned long long v, unsigned long long* total) {
    __shared__ unsigned long long ws[kVolBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long inc = uw_wave_incl_scan(v);
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    unsigned long long base = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < kVolBlock / 64; ++w) {
        base += w < wave ? ws[w] : 0ull;
        sum += ws[w];
    }
    __syncthreads();
    *total = sum;
    return base + inc - v;
}

__global__ void __launch_bounds__(kVolBlock) k_vol_scan_reduce(const unsigned long long* __restrict__ rows,
                                                               uint64_t nrows, unsigned long long* __restrict__ bsum) {
    const uint64_t b0 = uint64_t(blockIdx.x) * kScanBlock;
    unsigned long long s = 0;
#pragma unroll 4
    for (int j = 0; j < kScanPer; ++j) {
        const uint64_t i = b0 + uint64_t(j) * kVolBlock + threadIdx.x;
        if (i < nrows) s += rows[i];
    }
    unsigned long long tot;
    (void)vol_block_scan(s, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// One workgroup: exclusive scan of the block sums; the totals, the chunk record and the overflow
// flag (the emit kernel reads them).
__global__ void __launch_bounds__(kVolBlock) k_vol_scan_blocks(VolArgs a, uint32_t nblocks) {
    unsigned long long carry = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += kVolBlock) {
        const uint32_t b = b0 + threadIdx.x;
        const unsigned long long v = b < nblocks ? a.bsum[b] : 0ull;
        unsigned long long tot;
        const unsigned long long ex = vol_block_scan(v, &tot);
        if (b < nblocks) a.bsum[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        const VolCounters cn = *a.cnt;
        const uint64_t nv = carry & 0xFFFFFFFFull, ni = 3ull * (carry >> 32);
        uint32_t flags = 0;
        if (!(cn.below & UW_BELOW_NOT_ABOVE)) flags = UW_CHUNK_BLANK | UW_CHUNK_EARLY_SKIP;
        else if (ni == 0) flags = UW_CHUNK_BLANK;
        const bool ovf = nv > a.vcap || ni > a.icap;
        if (ovf) flags |= UW_CHUNK_OVERFLOW;
        if (a.eo.ind16 && nv > 65535u) flags |= UW_CHUNK_INDEX_WRAP;
        uw_chunk_info r = {flags, uint32_t(nv), uint32_t(ni), cn.active, 0, 0};
        *a.info = r;
        uw_totals t;
        t.n_verts = nv;
        t.n_inds = ni;
        t.n_nonblank = ni > 0 ? 1u : 0u;
        t.overflow = ovf ? 1u : 0u;
        *a.totals = t;
    }
}

__global__ void __launch_bounds__(kVolBlock) k_vol_scan_apply(unsigned long long* __restrict__ rows, uint64_t nrows,
                                                              const unsigned long long* __restrict__ bsum) {
    const uint64_t b0 = uint64_t(blockIdx.x) * kScanBlock;
    unsigned long long carry = bsum[blockIdx.x];
    for (int j = 0; j < kScanPer; ++j) {
        const uint64_t i = b0 + uint64_t(j) * kVolBlock + threadIdx.x;
        const unsigned long long v = i < nrows ? rows[i] : 0ull;
        unsigned long long tot;
        const unsigned long long ex = vol_block_scan(v, &tot);
        if (i < nrows) rows[i] = carry + ex;
        carry += tot;
    }
}

// ---------------------------------------------------------------------------------------------
// Emission of one x-slab of a tile (cell planes [X0, X1)). Cell planes x = xs .. X1 - 1 (xs = X0 - 1
// when a plane below the slab exists: its cells are classified for owner lookups only); in each
// plane the rows y = ys .. Y1 - 1 and the cells z = zs .. Z1 - 1, likewise with halo row / column.
// Per emitted plane the active cells are compacted into a list (LDS atomics per row, any order:
// every id and index position comes from the row bases), then one lane per listed cell writes its
// vertices and one lane per triangle its three index entries, so no lane idles on an empty cell or
// waits for the cell with the most triangles. The slab's row bases are read into LDS once, and each
// sample plane is fetched into registers while the plane before it is processed, so no memory
// round trip sits between two planes.
constexpr int kT2C = 2048;  // triangles of a plane taken one per lane (more: one lane per cell)

__global__ void __launch_bounds__(kVolBlock) k_vol_emit(VolArgs a) {
    __shared__ float SP[2][kTileS * kPitch];
    // first owned record of the cell, relative to its row segment's base (the halo column's cells,
    // which end the segment below, hold minus their own records)
    __shared__ int16_t CB[2][(kTile + 1) * (kTile + 1)];
    __shared__ unsigned long long RB[(kEmitPlanes + 1) * (kTile + 1)];  // row bases of planes xs .. X1 - 1
    __shared__ uint8_t CC[2][(kTile + 1) * (kTile + 1)];   // cube case
    __shared__ uint32_t LIST[kTile * kTile];  // active cells of the plane: (y - Y0) 32 + z - Z0 | triangle prefix << 10
    __shared__ uint32_t LIB[kTile * kTile];   // index position of each listed cell's first index entry
    __shared__ uint16_t T2C[kT2C];            // plane-local triangle -> list slot
    __shared__ uint32_t LUT[256];
    __shared__ __align__(16) uint16_t BEF[256 * 12];
    __shared__ uint32_t s_cnt;  // listed cells | triangles << 16
    const uw_totals tot = *a.totals;
    if (tot.overflow || tot.n_inds == 0) return;
    int32_t X0, Y0, Z0, X1, Y1, Z1, tz;
    tile_of(a, kEmitPlanes, &X0, &Y0, &Z0, &X1, &Y1, &Z1, &tz);
    if (X0 >= X1) return;  // a slab past the grid's last plane (whole workgroup)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < 256; i += kVolBlock) LUT[i] = uint32_t(a.tab->used[i]) | (uint32_t(a.tab->tri[i] & 15u) << 16);
    uw_lds_load_before(BEF, a.tab, tid, kVolBlock);
    const int32_t xs = X0 > 0 ? X0 - 1 : X0, ys = Y0 > 0 ? Y0 - 1 : Y0, zs = Z0 > 0 ? Z0 - 1 : Z0;
    const int32_t ny = Y1 - ys, nz = Z1 - zs;  // cells per plane row / column, halo included
    const int32_t sy = ny + 1, sz = nz + 1;
    const float iso = a.f.iso;
    const float step = a.step;
    const float ox = float(a.f.extent * a.c[0]), oy = float(a.f.extent * a.c[1]), oz = float(a.f.extent * a.c[2]);
    const uint64_t vlim = tot.n_verts, ilim = tot.n_inds;
    const int32_t L = a.L, n = a.n;
    const uw_emit_opts& eo = a.eo;
    const bool coll = eo.tri_normals || eo.tri_cells;
    const uint64_t* __restrict__ tri_rows = a.tab->tri;
    for (int32_t e = tid; e < (X1 - xs) * ny; e += kVolBlock) {
        const int32_t pl = e / ny, r = e - pl * ny;
        RB[pl * (kTile + 1) + r] = a.rows[(size_t(xs + pl) * n + ys + r) * a.T + tz];
    }
    // row base of row r of cell plane x
    auto rbase = [&](int32_t x, int32_t r) -> unsigned long long { return RB[(x - xs) * (kTile + 1) + r]; };
    float pre[kPlaneRegs];
    fetch_plane(a, xs, ys, sy, zs, sz, pre);
    put_plane(sy, sz, pre, SP[xs & 1]);
    fetch_plane(a, xs + 1, ys, sy, zs, sz, pre);
    for (int32_t x = xs; x < X1; ++x) {
        put_plane(sy, sz, pre, SP[(x + 1) & 1]);
        if (x + 2 <= X1) fetch_plane(a, x + 2, ys, sy, zs, sz, pre);  // block-uniform
        if (tid == 0) s_cnt = 0u;
        __syncthreads();
        const float* p0 = SP[x & 1];
        const float* p1 = SP[(x + 1) & 1];
        int16_t* cb = CB[x & 1];
        uint8_t* cc = CC[x & 1];
        // classification, vertex and index bases of the plane's rows (cells z >= Z0 of a row lie
        // in the row segment (x, y, tz); the halo cell z = Z0 - 1 is the last of segment tz - 1),
        // and the list of the emitted cells
        for (int32_t q = wave; 2 * q < ny; q += kVolBlock / 64) {
            const int32_t r = 2 * q + (lane >> 5);
            const int32_t y = ys + r, z = Z0 + (lane & 31);
            const bool in = y < Y1 && z < Z1;
            uint32_t v = 0, cs = 0;
            if (in) {
                float rr[8];
                tile_corners(p0, p1, y, z, ys, zs, rr);
                cs = uw_case_of(rr, iso);
                const uint32_t w = LUT[cs];
                const uint32_t nt = w >> 16;
                // records (< 2^10), triangles (< 2^10) and active cells (< 2^6) of a row of 32
                v = __popc(w & 0xFFFu & uw_owned_mask(x, y, z)) | (nt << 10) | ((nt ? 1u : 0u) << 20);
            }
            const uint32_t inc = uw_half_incl_scan(v);
            const bool emitted = x >= X0 && y >= Y0 && y < Y1;
            uint32_t old = 0;
            if ((lane & 31) == 31 && emitted && (inc >> 20))
                old = atomicAdd(&s_cnt, (inc >> 20) | (((inc >> 10) & 0x3FFu) << 16));
            const uint32_t lbase = (lane & 32) ? uint32_t(__builtin_amdgcn_readlane(int(old), 63))
                                               : uint32_t(__builtin_amdgcn_readlane(int(old), 31));
            if (in) {
                const unsigned long long rb = rbase(x, r);
                const uint32_t ex = inc - v;
                const int o = r * nz + (z - zs);
                cb[o] = int16_t(ex & 0x3FFu);
                cc[o] = uint8_t(cs);
                if (emitted && (v >> 20)) {
                    const uint32_t slot = (lbase & 0xFFFFu) + (ex >> 20);
                    LIST[slot] = uint32_t((y - Y0) * kTile + (z - Z0)) | (((lbase >> 16) + ((ex >> 10) & 0x3FFu)) << 10);
                    LIB[slot] = 3u * (uint32_t(rb >> 32) + ((ex >> 10) & 0x3FFu));
                }
            }
        }
        if (zs < Z0) {
            for (int32_t r = tid; r < ny; r += kVolBlock) {
                const int32_t y = ys + r, z = zs;
                float rr[8];
                tile_corners(p0, p1, y, z, ys, zs, rr);
                const uint32_t cs = uw_case_of(rr, iso);
                const uint32_t nv = __popc(LUT[cs] & 0xFFFu & uw_owned_mask(x, y, z));
                cb[r * nz] = int16_t(-int32_t(nv));
                cc[r * nz] = uint8_t(cs);
            }
        }
        __syncthreads();
        if (x >= X0) {
            const uint32_t cnt = s_cnt;
            const uint32_t na = cnt & 0xFFFFu, ntp = cnt >> 16;
            const bool flat = ntp <= uint32_t(kT2C);  // block-uniform
            // record id of edge ed as seen from cell (y, z) of plane x: its own rank, or its owner's
            auto edge_id = [&](int32_t y, int32_t z, uint32_t cs, uint32_t owned, uint32_t myv, uint32_t ed) -> uint32_t {
                if (owned & (1u << ed)) return myv + __popc(uint32_t(BEF[cs * 12 + ed]) & owned);
                int32_t qx, qy, qz;
                uint32_t qe;
                uw_edge_owner(x, y, z, ed, &qx, &qy, &qz, &qe);
                const int qo = (qy - ys) * nz + (qz - zs);
                const uint32_t qc = CC[qx & 1][qo];
                return uint32_t(rbase(qx, qy - ys)) + uint32_t(int32_t(CB[qx & 1][qo])) +
                       __popc(uint32_t(BEF[qc * 12 + qe]) & uw_owned_mask(qx, qy, qz));
            };
            // index entries (and collision records) of triangle j of a listed cell
            auto emit_tri = [&](int32_t y, int32_t z, uint32_t cs, uint32_t owned, uint32_t myv, uint64_t row, uint32_t j,
                                uint64_t gi, const float* r) {
                float q[3][3];
#pragma unroll
                for (uint32_t k = 0; k < 3; ++k) {
                    const uint32_t ed = uint32_t(row >> (4 + 4 * (3 * j + k))) & 15u;
                    const uint32_t id = edge_id(y, z, cs, owned, myv, ed);
                    if (gi + k < ilim) uw_store_index(a.inds, eo.ind16, gi + k, id);
                    if (coll) uw_edge_vertex(ed, x, y, z, r, iso, step, ox, oy, oz, q[k]);
                }
                if (coll && gi + 2 < ilim) {
                    const uint64_t t = gi / 3;
                    if (eo.tri_normals) {
                        float nm[3];
                        uw_tri_normal(q[0], q[1], q[2], nm);
                        eo.tri_normals[3 * t] = nm[0];
                        eo.tri_normals[3 * t + 1] = nm[1];
                        eo.tri_normals[3 * t + 2] = nm[2];
                    }
                    if (eo.tri_cells) eo.tri_cells[t] = uint32_t((uint64_t(x) * n + y) * n + z);
                }
            };
            for (uint32_t k = tid; k < na; k += kVolBlock) {
                const uint32_t ent = LIST[k];
                const int32_t y = Y0 + int32_t((ent & 0x3FFu) >> 5), z = Z0 + int32_t(ent & 31u);
                const int o = (y - ys) * nz + (z - zs);
                const uint32_t cs = cc[o];
                float r[8];
                tile_corners(p0, p1, y, z, ys, zs, r);
                const uint32_t owned = uw_owned_mask(x, y, z);
                const uint32_t myv = uint32_t(rbase(x, y - ys)) + uint32_t(int32_t(cb[o]));
                for (uint32_t vm = LUT[cs] & 0xFFFu & owned; vm; vm &= vm - 1u) {
                    const uint32_t ed = uint32_t(__builtin_ctz(vm));
                    const uint64_t id = uint64_t(myv) + __popc(uint32_t(BEF[cs * 12 + ed]) & owned);
                    if (id >= vlim) continue;
                    float p[3];
                    uw_edge_vertex(ed, x, y, z, r, iso, step, ox, oy, oz, p);
                    float* dst = a.verts + eo.vstride * id;
                    dst[0] = p[0];
                    dst[1] = p[1];
                    dst[2] = p[2];
                    if (eo.vstride == 6) uw_vertex_color(p[2], uw_edge_b(ed), float(a.f.extent), dst + 3);
                    if (a.tape) {
                        const uint32_t ia = uw_edge_a(ed), ib = uw_edge_b(ed);
                        const uint32_t ca = uw_corner_ofs(ia), cbo = uw_corner_ofs(ib);
                        uw_vertex_tape tv;
                        tv.a = uint32_t((uint64_t(x + int32_t(ca & 1u)) * L + y + int32_t((ca >> 1) & 1u)) * L + z +
                                        int32_t((ca >> 2) & 1u));
                        tv.b = uint32_t((uint64_t(x + int32_t(cbo & 1u)) * L + y + int32_t((cbo >> 1) & 1u)) * L + z +
                                        int32_t((cbo >> 2) & 1u));
                        tv.rho_a = r[ia];
                        tv.rho_b = r[ib];
                        a.tape[id] = tv;
                    }
                }
                const uint32_t ntri = LUT[cs] >> 16;
                if (flat) {
                    const uint32_t tp = ent >> 10;
                    for (uint32_t i = 0; i < ntri; ++i) T2C[tp + i] = uint16_t(k);
                } else {
                    const uint64_t row = tri_rows[cs];
                    for (uint32_t j = 0; j < ntri; ++j) emit_tri(y, z, cs, owned, myv, row, j, uint64_t(LIB[k]) + 3 * j, r);
                }
            }
            if (flat) {
                __syncthreads();  // the triangle map is complete
                for (uint32_t t = tid; t < ntp; t += kVolBlock) {
                    const uint32_t k = T2C[t];
                    const uint32_t ent = LIST[k];
                    const int32_t y = Y0 + int32_t((ent & 0x3FFu) >> 5), z = Z0 + int32_t(ent & 31u);
                    const int o = (y - ys) * nz + (z - zs);
                    const uint32_t cs = cc[o];
                    const uint32_t j = t - (ent >> 10);
                    float r[8];
                    if (coll) tile_corners(p0, p1, y, z, ys, zs, r);
                    emit_tri(y, z, cs, uw_owned_mask(x, y, z), uint32_t(rbase(x, y - ys)) + uint32_t(int32_t(cb[o])),
                             tri_rows[cs], j, uint64_t(LIB[k]) + 3 * j, r);
                }
            }
        }
        __syncthreads();  // plane x - 1's cells, sample plane x and the plane's lists are read no more
    }
}

// ---------------------------------------------------------------------------------------------
// Backward over the volume's vertices (the adjoint of uw_adjoint.hip for one grid): the octave
// values and gradients at each vertex's endpoints are recomputed from the per-axis tables of the
// grid (read through the cache; 3 O L entries). Per-workgroup parameter sums in f64, added in
// workgroup order by k_vol_backward_sum.
constexpr int kVolBwdParts = 19;  // UW_MAX_OCTAVES + 3

template <int OMAX>
__global__ void __launch_bounds__(kVolBlock) k_vol_backward(VolArgs a, const uw_chunk_info* __restrict__ info,
                                                            const uw_vertex_tape* __restrict__ tape,
                                                            const float* __restrict__ grad_verts, float* grad_dens,
                                                            int want_params, double* __restrict__ part) {
    __shared__ float G3[3 * 256];
    __shared__ uint8_t P[256];
    __shared__ double red[kVolBlock / 64][OMAX + 3];
    const int tid = threadIdx.x;
    for (int i = tid; i < 256; i += kVolBlock) {
        const uint8_t p = a.tab->perm[i];
        P[i] = p;
        const float4 g = uw_grad_vec(p & 15u);
        G3[3 * i] = g.x;
        G3[3 * i + 1] = g.y;
        G3[3 * i + 2] = g.z;
    }
    __syncthreads();
    const uw_chunk_info ci = *info;
    const bool skip = (ci.flags & (UW_CHUNK_OVERFLOW | UW_CHUNK_INVALID)) != 0;
    const int32_t O = a.f.octaves, L = a.L;
    const uint32_t LL = uint32_t(L) * uint32_t(L);
    const float iso = a.f.iso;
    const double inv_sum = 1.0 / double(a.f.amp_sum);
    const double slope = a.f.bias_kind == UW_BIAS_LINEAR ? double(a.f.bias_a) : 0.0;
    const uw_axis* AXg = a.axis;
    const uw_axis* AYg = a.axis + size_t(O) * L;
    const uw_axis* AZg = a.axis + 2 * size_t(O) * L;
    const float* DFx = a.dfade;
    const float* DFy = a.dfade + size_t(O) * L;
    const float* DFz = a.dfade + 2 * size_t(O) * L;
    double acc[OMAX + 3];
#pragma unroll
    for (int k = 0; k < OMAX + 3; ++k) acc[k] = 0.0;
    const uint64_t nv = skip ? 0 : ci.n_verts;
    for (uint64_t v = uint64_t(blockIdx.x) * kVolBlock + tid; v < nv; v += uint64_t(gridDim.x) * kVolBlock) {
        const uw_vertex_tape tv = tape[v];
        const int64_t diff = int64_t(tv.b) - int64_t(tv.a);
        const int axis = (diff == 1 || diff == -1) ? 2 : ((diff == L || diff == -L) ? 1 : 0);
        const float dir = diff > 0 ? a.step : -a.step;
        const float gv = grad_verts[3 * v + axis] * dir;
        const float D = tv.rho_b - tv.rho_a;
        const float inv_d2 = 1.0f / (D * D);
        const float w[2] = {gv * (iso - tv.rho_b) * inv_d2, gv * (tv.rho_a - iso) * inv_d2};
        const uint32_t s[2] = {tv.a, tv.b};
        if (grad_dens) {
            unsafeAtomicAdd(grad_dens + s[0], w[0]);
            unsafeAtomicAdd(grad_dens + s[1], w[1]);
        }
        if (!want_params) continue;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const uint32_t xx = s[e] / LL, rem = s[e] - xx * LL;
            const uint32_t yy = rem / uint32_t(L), zz = rem - yy * uint32_t(L);
            float nk[OMAX];
            float total = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
            for (int k = 0; k < OMAX; ++k) {
                if (k < O) {
                    float g[3];
                    const size_t kl = size_t(k) * L;
                    nk[k] = uw_octave_grad(AXg[kl + xx], AYg[kl + yy], AZg[kl + zz], DFx[kl + xx], DFy[kl + yy],
                                           DFz[kl + zz], P, G3, g);
                    const float af = a.f.amp[k] * float(a.f.octave_freq[k]);
                    total = fmaf(a.f.amp[k], nk[k], total);
                    gx = fmaf(af, g[0], gx);
                    gy = fmaf(af, g[1], gy);
                    gz = fmaf(af, g[2], gz);
                }
            }
            const double we = double(w[e]);
            const double fbm = double(total) * inv_sum;
#pragma unroll
            for (int k = 0; k < OMAX; ++k)
                if (k < O) acc[k] += we * (double(nk[k]) - fbm) * inv_sum;
            const double fq = a.f.frequency;
            acc[OMAX + 0] += we * fq * double(gx) * inv_sum;
            acc[OMAX + 1] += we * fq * double(gy) * inv_sum;
            acc[OMAX + 2] += we * fq * (double(gz) * inv_sum + slope);
        }
    }
    if (!want_params) return;
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int k = 0; k < OMAX + 3; ++k) {
        double vsum = acc[k];
        for (int d = 32; d > 0; d >>= 1) vsum += __shfl_down(vsum, d, 64);
        if (lane == 0) red[wave][k] = vsum;
    }
    __syncthreads();
    if (tid < OMAX + 3) {
        double t = 0.0;
        for (int w2 = 0; w2 < kVolBlock / 64; ++w2) t += red[w2][tid];
        part[size_t(blockIdx.x) * kVolBwdParts + tid] = t;
    }
}

// One workgroup per parameter: each thread adds a stride of the workgroup sums, then the waves'
// totals are added in wave order (a fixed order, so runs agree). A single thread walking all
// workgroups' sums in turn waited on one load per sum: 0.64 ms for 2048 workgroups.
template <int OMAX>
__global__ void __launch_bounds__(kVolBlock) k_vol_backward_sum(const double* __restrict__ part, int32_t nblocks,
                                                                int32_t O, double* grad_amp, double* grad_off) {
    __shared__ double ws[kVolBlock / 64];
    const int k = int(blockIdx.x), tid = threadIdx.x;
    double t = 0.0;
    for (int32_t b = tid; b < nblocks; b += kVolBlock) t += part[size_t(b) * kVolBwdParts + k];
    t = uw_wave_sum_f64(t);
    if ((tid & 63) == 0) ws[tid >> 6] = t;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < kVolBlock / 64; ++w) s += ws[w];
        if (k < OMAX) {
            if (k < O && grad_amp) grad_amp[k] = s;
        } else if (grad_off) {
            grad_off[k - OMAX] = s;
        }
    }
}

// ---------------------------------------------------------------------------------------------
struct VolLayout {
    size_t axis_off, df_off, zrec_off, dens_off, rows_off, bsum_off, cnt_off, part_off, bytes;
    uint64_t nrows;
    uint32_t nscan;
};

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

VolLayout vol_layout(int32_t n, int32_t octaves, int32_t bwd_blocks) {
    VolLayout l;
    const size_t L = size_t(n) + 1;
    const size_t T = (size_t(n) + kTile - 1) / kTile;
    l.nrows = uint64_t(n) * uint64_t(n) * T;
    l.nscan = uint32_t((l.nrows + kScanBlock - 1) / kScanBlock);
    size_t o = 0;
    l.axis_off = o;
    o = align256(o + 3 * size_t(octaves) * L * sizeof(uw_axis));
    l.df_off = o;
    o = align256(o + 3 * size_t(octaves) * L * sizeof(float));
    l.zrec_off = o;
    o = align256(o + (L + 1) * size_t(uw_zrec_words(octaves)) * sizeof(uint32_t));
    l.dens_off = o;
    o = align256(o + L * L * L * sizeof(float));
    l.rows_off = o;
    o = align256(o + size_t(l.nrows) * sizeof(unsigned long long));
    l.bsum_off = o;
    o = align256(o + size_t(l.nscan) * sizeof(unsigned long long));
    l.cnt_off = o;
    o = align256(o + sizeof(VolCounters));
    l.part_off = o;
    o = align256(o + size_t(bwd_blocks) * kVolBwdParts * sizeof(double));
    l.bytes = o;
    return l;
}

int vol_backward_blocks(const uw_ctx* ctx) { return ctx->num_cus * 8; }

// Fills the launch arguments over the context's volume scratch (grown as needed).
int vol_prepare(uw_ctx* ctx, int32_t cx, int32_t cy, int32_t cz, int32_t n, VolArgs* a, VolLayout* lay) {
    const int32_t O = ctx->params.octaves;
    *lay = vol_layout(n, O, vol_backward_blocks(ctx));
    if (lay->bytes > ctx->vol_bytes) {
        (void)hipFree(ctx->d_vol);
        ctx->d_vol = nullptr;
        ctx->vol_bytes = 0;
        UW_HIP(ctx, hipMalloc(&ctx->d_vol, lay->bytes));
        ctx->vol_bytes = lay->bytes;
    }
    unsigned char* base = ctx->d_vol;
    memset(a, 0, sizeof(*a));
    a->f = uw_make_field(ctx->params);
    a->tab = ctx->d_tables;
    a->c[0] = cx;
    a->c[1] = cy;
    a->c[2] = cz;
    a->n = n;
    a->L = n + 1;
    a->T = (n + kTile - 1) / kTile;
    a->step = float(a->f.extent) / float(n);
    a->axis = reinterpret_cast<const uw_axis*>(base + lay->axis_off);
    a->dfade = reinterpret_cast<const float*>(base + lay->df_off);
    a->zrec = reinterpret_cast<const uint32_t*>(base + lay->zrec_off);
    a->dens = reinterpret_cast<float*>(base + lay->dens_off);
    a->rows = reinterpret_cast<unsigned long long*>(base + lay->rows_off);
    a->bsum = reinterpret_cast<unsigned long long*>(base + lay->bsum_off);
    a->cnt = reinterpret_cast<VolCounters*>(base + lay->cnt_off);
    const int32_t L = n + 1;
    a->zseg = L >= 256 ? (L + 127) / 128 : 1;
    a->zlen = (L + a->zseg - 1) / a->zseg;
    a->nx_tab = kVolBlock / L + 2;
    if (a->nx_tab > L) a->nx_tab = L;
    return UW_OK;
}

using VolDensityKernel = void (*)(VolArgs);
VolDensityKernel pick_vol_density(int octaves) {
    switch (octaves) {
        case 3: return k_vol_density<3>;
        case 4: return k_vol_density<4>;
        case 6: return k_vol_density<6>;
        case 8: return k_vol_density<8>;
        default: return k_vol_density<0>;
    }
}

// The tables kernel (per-axis entries, z records, counters reset); shared by forward and backward.
int vol_launch_tables(uw_ctx* ctx, VolArgs& a, const VolLayout& lay, hipStream_t st) {
    unsigned char* base = ctx->d_vol;
    const int64_t work = int64_t(3) * a.f.octaves * a.L + a.L + 1;
    if (int rc = uw_prof_start(ctx, st, UW_STAGE_VOLUME_TABLES)) return rc;
    hipLaunchKernelGGL(k_vol_tables, dim3(uint32_t((work + kVolBlock - 1) / kVolBlock)), dim3(kVolBlock), 0, st, a,
                       reinterpret_cast<uw_axis*>(base + lay.axis_off), reinterpret_cast<float*>(base + lay.df_off),
                       reinterpret_cast<uint32_t*>(base + lay.zrec_off));
    return uw_prof_end(ctx, st);
}

bool vol_args_ok(int32_t n, const uw_mesh_out* out) {
    return n >= 1 && n <= UW_VOLUME_MAX_RESOLUTION && out && out->info && out->totals && out->vertex_format <= 1 &&
           out->index_format <= 1 && (out->vert_cap == 0 || out->verts) && (out->ind_cap == 0 || out->inds);
}

}  // namespace

extern "C" {

int uw_build_volume(uw_ctx* ctx, int32_t cx, int32_t cy, int32_t cz, int32_t n, const uw_mesh_out* out,
                    uw_vertex_tape* tape, void* stream) {
    if (!ctx) return UW_ERR_INVALID_ARG;
    if (!vol_args_ok(n, out)) return uw_fail(ctx, UW_ERR_INVALID_ARG, "bad arguments (1 <= n <= UW_VOLUME_MAX_RESOLUTION)");
    UW_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (int rc = uw_refuse_capture(ctx, st)) return rc;
    VolArgs a;
    VolLayout lay;
    if (int rc = vol_prepare(ctx, cx, cy, cz, n, &a, &lay)) return rc;
    a.info = out->info;
    a.totals = out->totals;
    a.verts = out->verts;
    a.inds = out->inds;
    a.tape = tape;
    a.eo = uw_make_emit_opts(out);
    a.vcap = out->vert_cap;
    a.icap = out->ind_cap;
    if (int rc = vol_launch_tables(ctx, a, lay, st)) return rc;
    // densities
    const VolDensityKernel dk = pick_vol_density(a.f.octaves);
    const size_t lds = 256 + 256 + sizeof(uw_axis) * size_t(a.f.octaves) * (kVolBlock + a.nx_tab);
    UW_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(dk), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lds)));
    const uint64_t lines = uint64_t(a.L) * uint64_t(a.L);
    const uint64_t dblocks = (lines + kVolBlock - 1) / kVolBlock * uint64_t(a.zseg);
    if (int rc = uw_prof_start(ctx, st)) return rc;
    hipLaunchKernelGGL(dk, dim3(uint32_t(dblocks)), dim3(kVolBlock), lds, st, a);
    if (int rc = uw_prof_end(ctx, st)) return rc;
    const uint32_t tiles = uint32_t(a.T) * uint32_t(a.T) * uint32_t(a.T);
    if (int rc = uw_prof_start(ctx, st, UW_STAGE_VOLUME_COUNT)) return rc;
    hipLaunchKernelGGL(k_vol_count, dim3(tiles * uint32_t(kTile / kCountPlanes)), dim3(kVolBlock), 0, st, a);
    if (int rc = uw_prof_end(ctx, st)) return rc;
    if (int rc = uw_prof_start(ctx, st, UW_STAGE_VOLUME_SCAN)) return rc;
    hipLaunchKernelGGL(k_vol_scan_reduce, dim3(lay.nscan), dim3(kVolBlock), 0, st, a.rows, lay.nrows, a.bsum);
    hipLaunchKernelGGL(k_vol_scan_blocks, dim3(1), dim3(kVolBlock), 0, st, a, lay.nscan);
    hipLaunchKernelGGL(k_vol_scan_apply, dim3(lay.nscan), dim3(kVolBlock), 0, st, a.rows, lay.nrows, a.bsum);
    if (int rc = uw_prof_end(ctx, st)) return rc;
    if (int rc = uw_prof_start(ctx, st, UW_STAGE_VOLUME_EMIT)) return rc;
    hipLaunchKernelGGL(k_vol_emit, dim3(tiles * uint32_t(kTile / kEmitPlanes)), dim3(kVolBlock), 0, st, a);
    if (int rc = uw_prof_end(ctx, st)) return rc;
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}

int uw_volume_backward(uw_ctx* ctx, int32_t cx, int32_t cy, int32_t cz, int32_t n, const uw_chunk_info* info,
                       const uw_vertex_tape* tape, const float* grad_verts, float* grad_dens, double* grad_amp,
                       double* grad_off, void* stream) {
    if (!ctx) return UW_ERR_INVALID_ARG;
    if (n < 1 || n > UW_VOLUME_MAX_RESOLUTION || !info || !tape || !grad_verts)
        return uw_fail(ctx, UW_ERR_INVALID_ARG, "bad arguments");
    UW_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (int rc = uw_refuse_capture(ctx, st)) return rc;
    VolArgs a;
    VolLayout lay;
    if (int rc = vol_prepare(ctx, cx, cy, cz, n, &a, &lay)) return rc;
    if (int rc = vol_launch_tables(ctx, a, lay, st)) return rc;
    const int want = (grad_amp || grad_off) ? 1 : 0;
    const int blocks = vol_backward_blocks(ctx);
    double* part = reinterpret_cast<double*>(ctx->d_vol + lay.part_off);
    const int O = a.f.octaves;
    if (int rc = uw_prof_start(ctx, st)) return rc;
    if (O <= 4) {
        hipLaunchKernelGGL(k_vol_backward<4>, dim3(blocks), dim3(kVolBlock), 0, st, a, info, tape, grad_verts, grad_dens,
                           want, part);
        if (want) hipLaunchKernelGGL(k_vol_backward_sum<4>, dim3(4 + 3), dim3(kVolBlock), 0, st, part, blocks, O, grad_amp,
                                   grad_off);
    } else if (O <= 8) {
        hipLaunchKernelGGL(k_vol_backward<8>, dim3(blocks), dim3(kVolBlock), 0, st, a, info, tape, grad_verts, grad_dens,
                           want, part);
        if (want) hipLaunchKernelGGL(k_vol_backward_sum<8>, dim3(8 + 3), dim3(kVolBlock), 0, st, part, blocks, O, grad_amp,
                                   grad_off);
    } else {
        hipLaunchKernelGGL(k_vol_backward<16>, dim3(blocks), dim3(kVolBlock), 0, st, a, info, tape, grad_verts,
                           grad_dens, want, part);
        if (want) hipLaunchKernelGGL(k_vol_backward_sum<16>, dim3(16 + 3), dim3(kVolBlock), 0, st, part, blocks, O, grad_amp,
                                   grad_off);
    }
    if (int rc = uw_prof_end(ctx, st)) return rc;
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}

uint64_t uw_volume_scratch_bytes(const uw_ctx* ctx, int32_t n) {
    if (!ctx || n < 1 || n > UW_VOLUME_MAX_RESOLUTION) return 0;
    return vol_layout(n, ctx->params.octaves, vol_backward_blocks(ctx)).bytes;
}

}  // extern "C"
