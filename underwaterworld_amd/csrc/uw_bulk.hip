// Bulk path of the batched build: chunks of at least `bulk_min` cells per axis (the 24^3 and 32^3
// chunks of mixed-resolution and adjoint launches) leave the persistent queue. One workgroup per
// chunk would run a 24^3 chunk's density, classification and emission as a serial chain of x-slab
// phases; here every phase is spread over the device instead:
//
//   k_bulk_density  one lane per z-line (lanes on consecutive lines t = x L + y of one chunk),
//                   densities into a per-chunk HBM field [x][z][y], so each sample's store is a row
//   k_bulk_count    one workgroup per (chunk, x-slab of 8 or 4 cell planes): cube cases and, per row
//                   (x, y), the owned vertex records and triangles (chunk.rs:141-164 with the
//                   ownership rule)
//   k_bulk_scan     one wave per chunk, four per workgroup: the chunk's row bases (first-seen
//                   order within the chunk), its totals and record; the workgroup's chunks reserve
//                   their vertex and index ranges with one atomic on the launch's range counter
//                   (the persistent kernel's reservations use the same counter, so ranges never
//                   overlap)
//   k_bulk_emit     one workgroup per (chunk, x-slab of 1, 2 or 4 cell planes): classifies the slab's planes
//                   and the one below it (owners of shared records lie at most one plane back),
//                   then writes the planes' vertices, indices and, when requested, tape, colours
//                   and collision records
//
// The queue set-up (k_prep / k_place in uw_batched.hip) sorts chunks by resolution, largest first,
// so the bulk chunks are the first `nbig` queue entries; k_place publishes nbig in the queue state
// and starts the persistent kernel's queue after them. Output (vertex numbering, index order,
// records) is that of the persistent and plain paths, chunk by chunk.
#include <hip/hip_runtime.h>
#include <string.h>

#include "uw_mesh.h"
#include "uw_noise.h"
#include "uw_queue.h"
#include "uw_wave.h"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxN = UW_BATCH_MAX_RESOLUTION;
constexpr size_t kBulkScratchMax = size_t(8) << 30;  // 8 GiB: about 55,000 chunks at max_res 32
constexpr int kMaxL = kMaxN + 1;
// Distinct x values among a density workgroup's 256 consecutive lines t = x L + y of a chunk with
// L samples per axis: at most (256 - 1) / L + 2.
__host__ __device__ constexpr int bulk_nx_tab(int32_t L) { return (kBlock - 1) / L + 2; }

struct BulkArgs {
    uw_field f;
    const uw_tables* tab;
    const uw_task* tasks;
    uint32_t* qhead;            // queue state: bulk entry count, range and exit counters
    int32_t max_res;            // largest resolution of the launch (sizes the per-chunk slots)
    int32_t entries;            // chunks of the launch: count and emit workgroups per slab row
    int32_t Lm;                 // max_res + 1
    int32_t wpc;                // density workgroup items per chunk: ceil(Lm^2 / 256)
    int32_t nx_tab;             // x entries per octave in the density kernel's LDS (smallest bulk L)
    const unsigned char* ztab;  // per-chunk z tables written by k_prep
    int32_t zstride;
    int32_t packed;
    float* dens;                  // [j][Lm^3], chunk layout [x][z][y] with the chunk's own L
    unsigned long long* rows;     // [j][max_res^2]: per row (x n + y) records | triangles << 32
    uint32_t* cases;              // [j][max_res][32 x 32 bytes]: cube case of cell (x, y, z) at x, y 32 + z
    uint32_t* below;              // [j] early-blank bits, zeroed by k_place
    uint32_t* act;                // [j] active cells, zeroed by k_place
    unsigned long long* rec;      // [j][4]: vertex base, index base, vertices, indices (~0: nothing to emit)
    uw_chunk_info* info;
    uw_totals* totals;
    float* verts;
    uint32_t* inds;
    uw_vertex_tape* tape;
    uw_emit_opts eo;
    uint64_t vcap, icap;
};

__device__ __forceinline__ uint32_t bulk_count(const BulkArgs& a) { return *uw_queue_bulk_count(a.qhead); }

// Inclusive scan within each 32-lane half of a wave.
__device__ __forceinline__ uint32_t half_incl_scan(uint32_t v) {
    v += uw_dpp_u32<UW_DPP_ROW_SHR(1), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_SHR(2), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_SHR(4), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_SHR(8), 0xF>(v);
    v += uw_dpp_u32<UW_DPP_ROW_BCAST15, 0xA>(v);
    return v;
}

// Corner densities of cell (y, z) from two LDS sample planes [z][y] of pitch L (planes x, x + 1).
__device__ __forceinline__ void plane_corners(const float* p0, const float* p1, int32_t L, int32_t y, int32_t z,
                                              float r[8]) {
    const int o = z * L + y;
    r[0] = p0[o];
    r[1] = p1[o];
    r[2] = p1[o + L];
    r[3] = p0[o + L];
    r[4] = p0[o + 1];
    r[5] = p1[o + 1];
    r[6] = p1[o + L + 1];
    r[7] = p0[o + L + 1];
}

__device__ __forceinline__ void load_task(const BulkArgs& a, uint32_t j, int32_t* c, int32_t* n, int32_t* cx,
                                          int32_t* cy, int32_t* cz) {
    const int4* p = reinterpret_cast<const int4*>(a.tasks + j);
    const int4 r0 = p[0];
    const int4 r1 = p[1];
    *c = __builtin_amdgcn_readfirstlane(r0.x);
    *n = __builtin_amdgcn_readfirstlane(r0.y);
    *cx = __builtin_amdgcn_readfirstlane(r0.z);
    *cy = __builtin_amdgcn_readfirstlane(r0.w);
    *cz = __builtin_amdgcn_readfirstlane(r1.x);
}

// ---------------------------------------------------------------------------------------------
template <int OCT>
__global__ void __launch_bounds__(kBlock) k_bulk_density(BulkArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int32_t O = uw_octaves<OCT>::count(a.f);
    float* G3 = reinterpret_cast<float*>(smem);
    uint8_t* P = reinterpret_cast<uint8_t*>(G3 + 3 * 256);
    uw_axis* AYs = reinterpret_cast<uw_axis*>(P + 256);  // [O][256]: each line's y entries
    uw_axis* AXs = AYs + O * kBlock;                     // [O][nx_tab]: the workgroup's x entries
    __shared__ double s_freq[UW_MAX_OCTAVES];
    __shared__ uint32_t s_bits;
    const int tid = threadIdx.x;
    for (int i = tid; i < 256; i += kBlock) {
        const uint8_t p = a.tab->perm[i];
        P[i] = p;
        const float4 g = uw_grad_vec(p & 15u);
        G3[3 * i] = g.x;
        G3[3 * i + 1] = g.y;
        G3[3 * i + 2] = g.z;
    }
    if (tid < UW_MAX_OCTAVES) s_freq[tid] = a.f.octave_freq[tid];
    const uint32_t items = bulk_count(a) * uint32_t(a.wpc);
    const int32_t rw = OCT > 0 ? uw_zrec_words(OCT) : uw_zrec_words(O);
    const uint64_t slot = uint64_t(a.Lm) * a.Lm * a.Lm;
    for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
        const uint32_t j = item / uint32_t(a.wpc), wb = item - j * uint32_t(a.wpc);
        int32_t c, n, cx, cy, cz;
        load_task(a, j, &c, &n, &cx, &cy, &cz);
        if (n < 1 || n > a.max_res) continue;
        const int32_t L = n + 1, lines = L * L, t0 = int32_t(wb) * kBlock;
        if (t0 >= lines) continue;
        __syncthreads();  // the previous item's tables are read no more
        const int32_t t = t0 + tid;
        const bool valid = t < lines;
        const int32_t xfirst = t0 / L;
        const int32_t x = valid ? t / L : xfirst, y = valid ? t - x * L : 0;
        const int32_t nx = min(t0 + kBlock - 1, lines - 1) / L - xfirst + 1;
        const float step = float(a.f.extent) / float(n);
        if (tid == 0) s_bits = 0u;
        for (int k = 0; k < O; ++k) AYs[k * kBlock + tid] = uw_axis_of(uw_sample_coord(a.f, 1, cy, y, step), s_freq[k]);
        for (int e = tid; e < O * nx; e += kBlock) {
            const int k = e / nx, jx = e - k * nx;
            AXs[k * a.nx_tab + jx] = uw_hash_x(uw_axis_of(uw_sample_coord(a.f, 0, cx, xfirst + jx, step), s_freq[k]), P);
        }
        __syncthreads();
        const uw_ztab Zt = uw_ztab_of(a.ztab, c, O, a.zstride);
        const uw_axis* ax = AXs + (x - xfirst);
        const uw_axis* ay = AYs + tid;
        float* out = a.dens + uint64_t(j) * slot + (size_t(x) * L) * L + y;
        float vmin = INFINITY, vmax = -INFINITY;
        const int32_t nxt = a.nx_tab;
        // a wave with no line of the chunk (the last item of a chunk: 65 of 256 lines at 32^3, 113
        // at 24^3) skips the walk, leaving its SIMD to the other waves
        if (__any(valid))
            uw_noise_zline<OCT, false>(
                a.f, [=](int k) { return ax[k * nxt]; }, [=](int k) { return ay[k * kBlock]; }, P, G3, Zt.rec, rw, L,
                vmin, vmax, [=](int32_t iz, float rho) {
                    if (valid) out[iz * L] = rho;
                });
        const int bits = valid ? uw_below_bits(vmin, vmax, a.f.iso) : 0;
        if (bits) atomicOr(&s_bits, uint32_t(bits));
        __syncthreads();
        if (tid == 0 && s_bits) atomicOr(a.below + j, s_bits);
    }
}

// ---------------------------------------------------------------------------------------------
// Count and emit work on x-slabs: workgroup b takes chunk entry b / slabs and the cell planes
// [X0, X1) of slab b % slabs (one launch covers every chunk of the launch; workgroups past the bulk
// entries or the chunk's planes leave at once). Sample planes pass through a ring of two LDS
// planes; the next plane is fetched into registers while the current one is classified, so the
// slab's plane loads overlap its work.
// Cell planes per count workgroup: 8, or 4 when the launch has few chunks (the host's choice in
// uw_bulk_shape_for), so that short launches still spread over the device.
constexpr int kEmitBlock = 256;  // threads per emit workgroup (512 measured slower)

// One sample plane (up to kMaxL^2 floats) spread over the registers of a B-thread workgroup.
template <int B>
struct PlaneRegsT {
    static constexpr int kRegs = (kMaxL * kMaxL + B - 1) / B;
    float v[kRegs];
};
using PlaneRegs = PlaneRegsT<kBlock>;
template <int B>
__device__ __forceinline__ void plane_fetch(const float* __restrict__ src, int32_t LL, PlaneRegsT<B>& r) {
#pragma unroll
    for (int i = 0; i < PlaneRegsT<B>::kRegs; ++i) {
        const int e = int(threadIdx.x) + i * B;
        r.v[i] = e < LL ? src[e] : 0.0f;
    }
}
template <int B>
__device__ __forceinline__ void plane_put(float* dst, int32_t LL, const PlaneRegsT<B>& r) {
#pragma unroll
    for (int i = 0; i < PlaneRegsT<B>::kRegs; ++i) {
        const int e = int(threadIdx.x) + i * B;
        if (e < LL) dst[e] = r.v[i];
    }
}

struct BulkSlab {
    int32_t c, n, cx, cy, cz, X0, X1;
    uint32_t j;
};
// The workgroup's chunk entry and slab; false when it has nothing to do. The grid is sized for
// every chunk at the launch's largest resolution; live workgroups must not sit at a fixed stride
// among dead ones, since the dispatcher deals workgroup b to XCD b % 8 (and alternates halves of
// an XCD's CUs): 24^3 chunks laid out chunk-major in slots sized for 32^3 put all live emission
// workgroups of half their slabs on four of the eight XCDs, and within those on half of the CUs.
// Slab-major: workgroup b takes slab b / entries of entry b % entries, so the live workgroups of
// each slab row are one contiguous run. For the emission (whose cost follows the surface) and bulk
// chunks of one resolution (bulk entries are a prefix of the queue, sorted by resolution, largest
// first: its first and last entries tell), chunk-major with that resolution's slab count instead,
// gap-free and in queue order, so the chunks nearest the surface, the costly ones, start first and
// the launch does not end on them (adjoint launch of 256 chunks of 24^3: 39.9 -> 35.7 us).
template <bool BY_COST>
__device__ __forceinline__ bool bulk_slab(const BulkArgs& a, int32_t planes, BulkSlab* s) {
    const uint32_t nb = bulk_count(a);
    if (nb == 0) return false;
    const int32_t n_first = __builtin_amdgcn_readfirstlane(a.tasks[0].n);
    const int32_t n_last = __builtin_amdgcn_readfirstlane(a.tasks[nb - 1].n);
    uint32_t row;
    if (BY_COST && n_first == n_last && n_first >= 1 && n_first <= a.max_res) {
        const uint32_t slabs = uint32_t((n_first + planes - 1) / planes);
        s->j = blockIdx.x / slabs;
        row = blockIdx.x - s->j * slabs;
    } else {
        row = blockIdx.x / uint32_t(a.entries);
        s->j = blockIdx.x - row * uint32_t(a.entries);
    }
    if (s->j >= nb) return false;
    load_task(a, s->j, &s->c, &s->n, &s->cx, &s->cy, &s->cz);
    if (s->n < 1 || s->n > a.max_res) return false;
    s->X0 = int32_t(row) * planes;
    s->X1 = min(s->n, s->X0 + planes);
    return s->X0 < s->X1;
}

// Per-row counts of one x-slab of a chunk: rows[j][x n + y] = owned records | triangles << 32. The
// slab's CP + 1 sample planes are fetched into registers at once (one memory round trip
// per workgroup) and pass through a ring of two LDS planes.
template <int CP>
__global__ void __launch_bounds__(kBlock) k_bulk_count(BulkArgs a) {
    constexpr int kCountPlanes = CP;
    __shared__ float SP[2][kMaxL * kMaxL];
    __shared__ uint32_t LUT[256];
    BulkSlab s;
    if (!bulk_slab<false>(a, kCountPlanes, &s)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t n = s.n, L = n + 1, LL = L * L;
    const float* src = a.dens + uint64_t(s.j) * uint64_t(a.Lm) * a.Lm * a.Lm;
    PlaneRegs pr[kCountPlanes + 1];
#pragma unroll
    for (int p = 0; p <= kCountPlanes; ++p)
        if (s.X0 + p <= s.X1) plane_fetch(src + int64_t(s.X0 + p) * LL, LL, pr[p]);
    for (int i = tid; i < 256; i += kBlock) LUT[i] = uint32_t(a.tab->used[i]) | (uint32_t(a.tab->tri[i] & 15u) << 16);
    plane_put(SP[0], LL, pr[0]);
    plane_put(SP[1], LL, pr[1]);
    __syncthreads();
    unsigned long long* rows = a.rows + uint64_t(s.j) * a.max_res * a.max_res;
    const float iso = a.f.iso;
    uint32_t act = 0;
#pragma unroll
    for (int p = 0; p < kCountPlanes; ++p) {
        const int32_t x = s.X0 + p;
        if (x < s.X1) {
            const float* p0 = SP[p & 1];
            const float* p1 = SP[(p + 1) & 1];
            uint8_t* cplane = reinterpret_cast<uint8_t*>(a.cases) + (uint64_t(s.j) * a.max_res + x) * (kMaxN * kMaxN);
            for (int32_t q = wave; 2 * q < n; q += kBlock / 64) {
                const int32_t y = 2 * q + (lane >> 5), z = lane & 31;
                uint32_t v = 0;
                if (y < n && z < n) {
                    float r[8];
                    plane_corners(p0, p1, L, y, z, r);
                    const uint32_t cs = uw_case_of(r, iso);
                    cplane[y * kMaxN + z] = uint8_t(cs);  // for the emission, which then skips the corner loads
                    const uint32_t w = LUT[cs];
                    v = __popc(w & 0xFFFu & uw_owned_mask(x, y, z)) | ((w >> 16) << 10) |
                        ((cs != 0u && cs != 255u) ? (1u << 20) : 0u);
                }
                const uint32_t inc = half_incl_scan(v);
                if ((lane & 31) == 31 && y < n) {
                    rows[x * n + y] = (unsigned long long)(inc & 0x3FFu) | ((unsigned long long)((inc >> 10) & 0x3FFu) << 32);
                    act += inc >> 20;
                }
            }
            if (p + 2 <= kCountPlanes && x + 1 < s.X1) {
                __syncthreads();  // sample plane x is read no more: its slot takes plane x + 2
                plane_put(SP[p & 1], LL, pr[p + 2]);
                __syncthreads();
            }
        }
    }
    if (act) atomicAdd(a.act + s.j, act);
}

// ---------------------------------------------------------------------------------------------
// One wave per bulk entry, four entries per workgroup pass: the wave scans its chunk's rows (16
// per lane, then a DPP wave scan), and the workgroup's chunks reserve their vertex and index ranges
// with one atomic for their sum, each chunk's range following from its prefix among them. The
// range counter is the launch's one contended address (the memory-side atomic unit serialises it):
// one reservation per chunk had made this kernel 41 us for the mixed set's 4,096 bulk chunks.
__global__ void __launch_bounds__(kBlock) k_bulk_scan(BulkArgs a) {
    constexpr int W = kBlock / 64;
    constexpr int kPerLane = (kMaxN * kMaxN + 63) / 64;  // rows per lane
    __shared__ unsigned long long s_tot[W];  // per wave: its chunk's vertices | indices << 32 (0: none)
    __shared__ unsigned long long s_vo, s_io;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nbig = bulk_count(a);
    unsigned long long* resv = uw_queue_resv(a.qhead);
    for (uint32_t j0 = blockIdx.x * uint32_t(W); j0 < nbig; j0 += gridDim.x * uint32_t(W)) {  // block-uniform
        const uint32_t j = j0 + uint32_t(wave);
        int32_t c = 0, n = 0, cx, cy, cz;
        bool valid = false;
        unsigned long long tot = 0;
        if (j < nbig) {  // wave-uniform
            load_task(a, j, &c, &n, &cx, &cy, &cz);
            valid = n >= 1 && n <= a.max_res;
            if (valid) {
                const int32_t R = n * n;
                unsigned long long* rows = a.rows + uint64_t(j) * a.max_res * a.max_res;
                unsigned long long v[kPerLane];
                unsigned long long sum = 0;
#pragma unroll
                for (int k = 0; k < kPerLane; ++k) {
                    const int r = lane * kPerLane + k;
                    v[k] = r < R ? rows[r] : 0ull;
                    sum += v[k];
                }
                const unsigned long long inc = uw_wave_incl_scan(sum);
                unsigned long long ex = inc - sum;
#pragma unroll
                for (int k = 0; k < kPerLane; ++k) {
                    const int r = lane * kPerLane + k;
                    if (r < R) rows[r] = ex;
                    ex += v[k];
                }
                tot = uw_wave_last(inc);  // vertex records | triangles << 32
            }
        }
        const uint32_t nv = uint32_t(tot & 0xFFFFFFFFull), ni = 3u * uint32_t(tot >> 32);
        if (lane == 0) s_tot[wave] = ni > 0 ? ((unsigned long long)nv | ((unsigned long long)ni << 32)) : 0ull;
        __syncthreads();
        if (tid == 0) {
            unsigned long long all = 0;
            uint32_t nonblank = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                all += s_tot[w];
                nonblank += s_tot[w] ? 1u : 0u;
            }
            unsigned long long vo = 0, io = 0;
            if (all) {
                const uint32_t av = uint32_t(all & 0xFFFFFFFFull), ai = uint32_t(all >> 32);
                if (a.packed) {
                    const unsigned long long old = atomicAdd(resv, all);
                    vo = old & 0xFFFFFFFFull;
                    io = old >> 32;
                } else {
                    vo = atomicAdd(resv, (unsigned long long)av);
                    io = atomicAdd(resv + kResvIndsOffset, (unsigned long long)ai);
                }
                atomicAdd(resv + kExitOffset, (unsigned long long)nonblank << 32);  // non-blank chunks of the launch
            }
            s_vo = vo;
            s_io = io;
        }
        __syncthreads();
        if (j < nbig && lane == 0) {
            unsigned long long* rec = a.rec + 4 * uint64_t(j);
            if (!valid) {
                uw_chunk_info r = {UW_CHUNK_INVALID | UW_CHUNK_BLANK, 0, 0, 0, 0, 0};
                a.info[c] = r;
                rec[0] = ~0ull;
            } else {
                unsigned long long pre = 0;
                for (int w = 0; w < wave; ++w) pre += s_tot[w];
                uint32_t flags = 0;
                unsigned long long vo = 0, io = 0;
                if (!(a.below[j] & UW_BELOW_NOT_ABOVE)) flags = UW_CHUNK_BLANK | UW_CHUNK_EARLY_SKIP;
                else if (ni == 0) flags = UW_CHUNK_BLANK;
                if (ni > 0) {
                    vo = s_vo + (pre & 0xFFFFFFFFull);
                    io = s_io + (pre >> 32);
                    if (vo + nv > a.vcap || io + ni > a.icap) {
                        flags |= UW_CHUNK_OVERFLOW;
                        atomicOr(&a.totals->overflow, 1u);
                    }
                }
                if (a.eo.ind16 && nv > 65535u) flags |= UW_CHUNK_INDEX_WRAP;
                uw_chunk_info r = {flags, nv, ni, a.act[j], vo, io};
                a.info[c] = r;
                const bool emit = ni > 0 && !(flags & UW_CHUNK_OVERFLOW);
                rec[0] = emit ? vo : ~0ull;
                rec[1] = io;
                rec[2] = nv;
                rec[3] = ni;
            }
        }
        __syncthreads();  // s_tot, s_vo and s_io are written again in the next pass
    }
}

// ---------------------------------------------------------------------------------------------
// Emission of one x-slab of a chunk: cell planes xs = X0 - 1 (when X0 > 0: classified for owner
// lookups only) to X1 - 1. The slab's sample planes and row bases are fetched at once (one memory
// round trip per workgroup). Classifying an emitted plane also writes its vertex map (one entry per
// owned vertex record: cell and edge, at the record's id less the plane's first id) and triangle
// map (one entry per triangle: cell and triangle of the cell, at the triangle's number less the
// plane's first), so the emission that follows takes one vertex, then one triangle, per lane: no
// lane idles on a cell with fewer records than its neighbours'. Map positions come from the row
// bases of the chunk's scan, so no counter is shared. A plane with more records or triangles than
// a map holds falls back to one lane per cell.
constexpr int kMapCap = 1280;  // entries of each per-plane map (LDS)

template <int P>
// At most 96 VGPRs (the 4-plane form spills two): five workgroups per CU, which the 30 KB of LDS
// allows (mixed emission 0.383 -> 0.326 ms).
__global__ void __launch_bounds__(kEmitBlock, 5) k_bulk_emit(BulkArgs a) {
    constexpr int kEmitPlanes = P;
    __shared__ float SP[2][kMaxL * kMaxL];     // sample planes x, x + 1 ([z][y], pitch L)
    __shared__ uint16_t CB[2][kMaxN * kMaxN];  // first vertex id within its row: cells of planes x - 1, x
    __shared__ __align__(16) uint8_t CC[2][kMaxN * kMaxN];  // cube cases of the same cells (from k_bulk_count)
    __shared__ uint16_t V2C[kMapCap];          // cell y 32 + z | edge << 10
    __shared__ uint16_t T2C[kMapCap];          // cell y 32 + z | triangle of the cell << 10
    __shared__ uint64_t TRI[256];
    __shared__ uint32_t LUT[256];
    __shared__ __align__(16) uint16_t BEF[256 * 12];
    // row bases (vertex records | triangles << 32, chunk-local) of cell planes xs .. X1 - 1, then the
    // first row base of plane X1 (the chunk's totals when X1 is its last plane)
    __shared__ unsigned long long RB[(kEmitPlanes + 2) * kMaxN];
    BulkSlab s;
    if (!bulk_slab<true>(a, kEmitPlanes, &s)) return;
    const unsigned long long* rec = a.rec + 4 * uint64_t(s.j);
    const unsigned long long vo = rec[0];
    if (vo == ~0ull) return;  // nothing to emit (blank, overflow or invalid chunk)
    const unsigned long long io = rec[1], vlim = vo + rec[2], ilim = io + rec[3];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t n = s.n, L = n + 1, LL = L * L;
    const int32_t xs = s.X0 > 0 ? s.X0 - 1 : s.X0;
    const float* src = a.dens + uint64_t(s.j) * uint64_t(a.Lm) * a.Lm * a.Lm;
    const unsigned long long* rows = a.rows + uint64_t(s.j) * a.max_res * a.max_res;
    PlaneRegsT<kEmitBlock> pr[kEmitPlanes + 2];  // sample planes xs .. X1
#pragma unroll
    for (int p = 0; p < kEmitPlanes + 2; ++p)
        if (xs + p <= s.X1) plane_fetch(src + int64_t(xs + p) * LL, LL, pr[p]);
    // cube cases of cell planes xs .. X1 - 1, written by k_bulk_count (1 KB per plane, 4 per thread)
    static_assert(kEmitBlock * 4 == kMaxN * kMaxN, "one case word per thread and plane");
    const uint32_t* cw_src = a.cases + (uint64_t(s.j) * a.max_res) * (kMaxN * kMaxN / 4);
    uint32_t cw[kEmitPlanes + 1];
#pragma unroll
    for (int p = 0; p <= kEmitPlanes; ++p)
        cw[p] = xs + p < s.X1 ? cw_src[(xs + p) * (kMaxN * kMaxN / 4) + tid] : 0u;
    const int32_t nrb = (s.X1 - xs) * n;  // at most 5 planes of 32 rows: the last thread is free
    if (tid < nrb) RB[(tid / n) * kMaxN + tid % n] = rows[xs * n + tid];
    if (tid == kEmitBlock - 1)
        RB[(s.X1 - xs) * kMaxN] = s.X1 < n ? rows[s.X1 * n] : (rec[2] | ((rec[3] / 3u) << 32));
    for (int i = tid; i < 256; i += kEmitBlock) {
        const uint64_t row = a.tab->tri[i];
        TRI[i] = row;
        LUT[i] = uint32_t(a.tab->used[i]) | (uint32_t(row & 15u) << 16);
    }
    uw_lds_load_before(BEF, a.tab, tid, kEmitBlock);
    plane_put(SP[0], LL, pr[0]);
    plane_put(SP[1], LL, pr[1]);
    reinterpret_cast<uint32_t*>(CC[xs & 1])[tid] = cw[0];
    __syncthreads();
    const float iso = a.f.iso;
    const uw_emit_opts& eo = a.eo;
    const bool coll = eo.tri_normals || eo.tri_cells;
    const float step = float(a.f.extent) / float(n);
    const float ox = float(a.f.extent * s.cx), oy = float(a.f.extent * s.cy), oz = float(a.f.extent * s.cz);
    static_assert(kEmitPlanes <= 4, "the plane hand-over below covers slabs of up to 4 planes");
#pragma unroll 1
    for (int pl = 0; pl <= kEmitPlanes; ++pl) {
        const int32_t x = xs + pl;
        if (x >= s.X1) break;  // block-uniform
        const float* p0 = SP[pl & 1];
        const float* p1 = SP[(pl + 1) & 1];
        const bool emit_plane = x >= s.X0;
        const unsigned long long rb0 = RB[pl * kMaxN];  // the plane's first vertex id and triangle
        // classification of plane x; for an emitted plane, its vertex and triangle maps
        for (int32_t q = wave; 2 * q < n; q += kEmitBlock / 64) {
            const int32_t y = 2 * q + (lane >> 5), z = lane & 31;
            const bool in = y < n && z < n;
            uint32_t v = 0, cs = 0;
            if (in) {
                cs = CC[x & 1][y * kMaxN + z];
                const uint32_t w = LUT[cs];
                // records (< 2^10) and triangles of a row of 32
                v = __popc(w & 0xFFFu & uw_owned_mask(x, y, z)) | ((w >> 16) << 10);
            }
            const uint32_t inc = half_incl_scan(v);
            if (in) {
                const unsigned long long rb = RB[pl * kMaxN + y];
                const uint32_t ex = inc - v;
                const int o = y * kMaxN + z;
                CB[x & 1][o] = uint16_t(ex & 0x3FFu);
                if (emit_plane && (v >> 10)) {
                    const uint32_t owned = uw_owned_mask(x, y, z);
                    // map positions: the cell's first id and triangle, less the plane's first
                    const uint32_t vp = uint32_t(rb) - uint32_t(rb0) + (ex & 0x3FFu);
                    for (uint32_t vm = LUT[cs] & 0xFFFu & owned; vm; vm &= vm - 1u) {
                        const uint32_t ed = uint32_t(__builtin_ctz(vm));
                        const uint32_t rank = __popc(uint32_t(BEF[cs * 12 + ed]) & owned);
                        if (vp + rank < uint32_t(kMapCap)) V2C[vp + rank] = uint16_t(uint32_t(o) | (ed << 10));
                    }
                    const uint32_t tp = uint32_t(rb >> 32) - uint32_t(rb0 >> 32) + (ex >> 10), ntri = v >> 10;
                    for (uint32_t i = 0; i < ntri; ++i)
                        if (tp + i < uint32_t(kMapCap)) T2C[tp + i] = uint16_t(uint32_t(o) | (i << 10));
                }
            }
        }
        __syncthreads();
        if (emit_plane) {
            const unsigned long long rb1 = RB[(pl + 1) * kMaxN];  // the next plane's first id and triangle
            const uint32_t nvp = uint32_t(rb1) - uint32_t(rb0), ntp = uint32_t(rb1 >> 32) - uint32_t(rb0 >> 32);
            // one vertex record at its chunk-local id: corner densities of the edge's two ends
            auto put_vertex = [&](int32_t y, int32_t z, uint32_t ed, uint32_t idl, float ra, float rb) {
                const uint64_t id = vo + idl;
                if (id >= vlim) return;
                float p[3];
                uw_edge_vertex_ab(ed, x, y, z, ra, rb, iso, step, ox, oy, oz, p);
                float* dst = a.verts + eo.vstride * id;
                dst[0] = p[0];
                dst[1] = p[1];
                dst[2] = p[2];
                if (eo.vstride == 6) uw_vertex_color(p[2], uw_edge_b(ed), float(a.f.extent), dst + 3);
                if (a.tape) {
                    const uint32_t ca = uw_corner_ofs(uw_edge_a(ed)), cbo = uw_corner_ofs(uw_edge_b(ed));
                    uw_vertex_tape tv;
                    tv.a = uint32_t(((x + int32_t(ca & 1u)) * L + y + int32_t((ca >> 1) & 1u)) * L + z + int32_t((ca >> 2) & 1u));
                    tv.b = uint32_t(((x + int32_t(cbo & 1u)) * L + y + int32_t((cbo >> 1) & 1u)) * L + z +
                                    int32_t((cbo >> 2) & 1u));
                    tv.rho_a = ra;
                    tv.rho_b = rb;
                    a.tape[id] = tv;
                }
            };
            // density at corner c of cell (y, z) of plane x
            auto corner = [&](int32_t y, int32_t z, uint32_t c) -> float {
                const uint32_t d = uw_corner_ofs(c);
                const float* pp = (d & 1u) ? p1 : p0;
                return pp[(z + int32_t((d >> 2) & 1u)) * L + y + int32_t((d >> 1) & 1u)];
            };
            // chunk-local id of the record that cell (x, y, z) sees as edge ed: its own rank, or its owner's
            auto edge_id = [&](int32_t y, int32_t z, uint32_t cs, uint32_t owned, uint32_t myv, uint32_t ed) -> uint32_t {
                if (owned & (1u << ed)) return myv + __popc(uint32_t(BEF[cs * 12 + ed]) & owned);
                int32_t qx, qy, qz;
                uint32_t qe;
                uw_edge_owner(x, y, z, ed, &qx, &qy, &qz, &qe);
                const int qo = qy * kMaxN + qz;
                return uint32_t(RB[(qx - xs) * kMaxN + qy]) + CB[qx & 1][qo] +
                       __popc(uint32_t(BEF[CC[qx & 1][qo] * 12 + qe]) & uw_owned_mask(qx, qy, qz));
            };
            // index entries (and collision records) of triangle tj of an active cell; gi: its first entry
            auto emit_tri = [&](int32_t y, int32_t z, uint32_t cs, uint32_t tj, uint64_t gi) {
                const uint32_t owned = uw_owned_mask(x, y, z);
                const uint32_t myv = uint32_t(RB[pl * kMaxN + y]) + CB[x & 1][y * kMaxN + z];
                const uint64_t row = TRI[cs];
                float q[3][3];
#pragma unroll
                for (uint32_t k = 0; k < 3; ++k) {
                    const uint32_t ed = uint32_t(row >> (4 + 4 * (3 * tj + k))) & 15u;
                    const uint32_t id = edge_id(y, z, cs, owned, myv, ed);
                    if (gi + k < ilim) uw_store_index(a.inds, eo.ind16, gi + k, id);
                    if (coll)
                        uw_edge_vertex_ab(ed, x, y, z, corner(y, z, uw_edge_a(ed)), corner(y, z, uw_edge_b(ed)), iso, step,
                                          ox, oy, oz, q[k]);
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
                    if (eo.tri_cells) eo.tri_cells[t] = uint32_t((x * n + y) * n + z);
                }
            };
            if (nvp <= uint32_t(kMapCap)) {
                for (uint32_t t = tid; t < nvp; t += kEmitBlock) {
                    const uint32_t ent = V2C[t];
                    const int32_t o = int32_t(ent & 0x3FFu), y = o >> 5, z = o & 31;
                    const uint32_t ed = ent >> 10;
                    put_vertex(y, z, ed, uint32_t(rb0) + t, corner(y, z, uw_edge_a(ed)), corner(y, z, uw_edge_b(ed)));
                }
            } else {  // one lane per cell
                for (int32_t o = tid; o < n * kMaxN; o += kEmitBlock) {
                    const int32_t y = o >> 5, z = o & 31;
                    if (z >= n) continue;
                    const uint32_t cs = CC[x & 1][o];
                    const uint32_t owned = uw_owned_mask(x, y, z), myv = uint32_t(RB[pl * kMaxN + y]) + CB[x & 1][o];
                    for (uint32_t vm = LUT[cs] & 0xFFFu & owned; vm; vm &= vm - 1u) {
                        const uint32_t ed = uint32_t(__builtin_ctz(vm));
                        put_vertex(y, z, ed, myv + __popc(uint32_t(BEF[cs * 12 + ed]) & owned), corner(y, z, uw_edge_a(ed)),
                                   corner(y, z, uw_edge_b(ed)));
                    }
                }
            }
            if (ntp <= uint32_t(kMapCap)) {
                for (uint32_t t = tid; t < ntp; t += kEmitBlock) {
                    const uint32_t ent = T2C[t];
                    const int32_t o = int32_t(ent & 0x3FFu), y = o >> 5, z = o & 31;
                    const uint64_t gi = io + 3ull * (uint64_t(uint32_t(rb0 >> 32)) + t);
                    emit_tri(y, z, CC[x & 1][o], ent >> 10, gi);
                }
            } else {  // one lane per cell
                for (int32_t o = tid; o < n * kMaxN; o += kEmitBlock) {
                    const int32_t y = o >> 5, z = o & 31;
                    if (z >= n) continue;
                    const uint32_t cs = CC[x & 1][o];
                    const uint32_t ntri = LUT[cs] >> 16;
                    if (!ntri) continue;
                    // the cell's first triangle: the row's triangle base plus the triangles of the
                    // row's cells before it
                    uint32_t pre = 0;
                    for (int32_t zz = 0; zz < z; ++zz) pre += LUT[CC[x & 1][y * kMaxN + zz]] >> 16;
                    const uint64_t gi = io + 3u * (uint32_t(RB[pl * kMaxN + y] >> 32) + pre);
                    for (uint32_t tj = 0; tj < ntri; ++tj) emit_tri(y, z, cs, tj, gi + 3 * tj);
                }
            }
        }
        if (pl + 2 < kEmitPlanes + 2 && x + 1 < s.X1) {
            __syncthreads();  // sample plane x, the plane's maps and plane x - 1's cells are read no more
            // sample plane x + 2 and the cases of cell plane x + 1 (whose slot held plane x - 1's, read
            // no more); a switch keeps the register arrays' indices static in the rolled loop
            uint32_t* cnext = reinterpret_cast<uint32_t*>(CC[(x + 1) & 1]);
            switch (pl) {
                case 0: plane_put(SP[0], LL, pr[2 < kEmitPlanes + 2 ? 2 : 0]); cnext[tid] = cw[1 <= kEmitPlanes ? 1 : 0]; break;
                case 1: plane_put(SP[1], LL, pr[3 < kEmitPlanes + 2 ? 3 : 0]); cnext[tid] = cw[2 <= kEmitPlanes ? 2 : 0]; break;
                case 2: plane_put(SP[0], LL, pr[4 < kEmitPlanes + 2 ? 4 : 0]); cnext[tid] = cw[3 <= kEmitPlanes ? 3 : 0]; break;
                default: plane_put(SP[1], LL, pr[5 < kEmitPlanes + 2 ? 5 : 0]); cnext[tid] = cw[4 <= kEmitPlanes ? 4 : 0]; break;
            }
            __syncthreads();
        }
    }
}

using BulkDensityKernel = void (*)(BulkArgs);
BulkDensityKernel pick_bulk_density(int octaves) {
    switch (octaves) {
        case 3: return k_bulk_density<3>;
        case 4: return k_bulk_density<4>;
        case 6: return k_bulk_density<6>;
        case 8: return k_bulk_density<8>;
        default: return k_bulk_density<0>;
    }
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

int grid_for(uw_ctx* ctx, const void* k, size_t lds, int* grid, int block = kBlock) {
    int per_cu = 0;
    UW_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, lds));
    if (per_cu < 1) return uw_fail(ctx, UW_ERR_INVALID_ARG, "bulk kernel does not fit on a CU");
    *grid = per_cu * ctx->num_cus;
    return UW_OK;
}

}  // namespace

int uw_bulk_shape_for(uw_ctx* ctx, const uw_batch_tuning& t, int32_t nchunks, uw_bulk_shape* s) {
    memset(s, 0, sizeof(*s));
    s->on = t.bulk_min >= 1 && t.bulk_min <= t.max_res;
    if (!s->on) return UW_OK;
    s->max_res = t.max_res;
    const int32_t Lm = t.max_res + 1;
    s->wpc = (Lm * Lm + kBlock - 1) / kBlock;
    const int32_t O = ctx->params.octaves;
    s->nx_tab = bulk_nx_tab(t.bulk_min + 1);  // every bulk chunk has at least bulk_min + 1 samples per axis
    s->lds_density = 256 * 3 * sizeof(float) + 256 + sizeof(uw_axis) * size_t(O) * size_t(kBlock + s->nx_tab);
    const BulkDensityKernel kd = pick_bulk_density(O);
    UW_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(kd), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(s->lds_density)));
    if (int rc = grid_for(ctx, reinterpret_cast<const void*>(kd), s->lds_density, &s->grid_density)) return rc;
    if (int rc = grid_for(ctx, reinterpret_cast<const void*>(k_bulk_scan), 0, &s->grid_scan)) return rc;
    // count and emit: one workgroup per (chunk entry, x-slab) of every chunk of the launch
    const int64_t nc64 = nchunks > 0 ? nchunks : 1;
    s->entries = int32_t(nc64);
    int resident_count = 0;
    if (int rc = grid_for(ctx, reinterpret_cast<const void*>(k_bulk_count<8>), 0, &resident_count)) return rc;
    s->count_planes = nc64 * ((t.max_res + 7) / 8) < 2 * int64_t(resident_count) ? 4 : 8;
    const int64_t gc = nc64 * ((t.max_res + s->count_planes - 1) / s->count_planes);
    // emit: the deepest slab (fewest halo planes) whose workgroups still cover the resident slots
    // about three times over; a launch of few chunks gets shallow slabs, so more workgroups share
    // its chunks (the mixed set: 4 planes; the adjoint's 256 chunks: 2)
    int resident = 0;
    if (int rc = grid_for(ctx, reinterpret_cast<const void*>(k_bulk_emit<4>), 0, &resident, kEmitBlock)) return rc;
    s->emit_planes = 4;
    while (s->emit_planes > 1 && nc64 * ((t.max_res + s->emit_planes - 1) / s->emit_planes) < 3 * int64_t(resident))
        s->emit_planes /= 2;
    const int64_t ge = nc64 * ((t.max_res + s->emit_planes - 1) / s->emit_planes);
    if (ge > 0x7FFFFFFF) return uw_fail(ctx, UW_ERR_INVALID_ARG, "too many chunks for the bulk path");
    s->grid_count = int(gc);
    s->grid_emit = int(ge);
    // scratch for every chunk of the launch (the resolutions are on the device; a launch's bulk
    // chunks are at most all of them)
    const size_t nc = size_t(nchunks > 0 ? nchunks : 1);
    size_t o = 0;
    s->dens_off = o;
    o = align256(o + nc * size_t(Lm) * Lm * Lm * sizeof(float));
    s->rows_off = o;
    o = align256(o + nc * size_t(t.max_res) * t.max_res * sizeof(unsigned long long));
    s->cases_off = o;
    o = align256(o + nc * size_t(t.max_res) * kMaxN * kMaxN);
    s->below_off = o;
    o = align256(o + nc * sizeof(uint32_t));
    s->act_off = o;
    o = align256(o + nc * sizeof(uint32_t));
    s->rec_off = o;
    o = align256(o + nc * 4 * sizeof(unsigned long long));
    s->bytes = o;
    // a launch whose worst case would need more scratch than this keeps every chunk on the
    // persistent kernel (same output, no per-chunk density slots)
    if (s->bytes > kBulkScratchMax) {
        memset(s, 0, sizeof(*s));
        return UW_OK;
    }
    return UW_OK;
}

namespace {
BulkArgs bulk_args(uw_ctx* ctx, const uw_bulk_shape& s, const uw_bulk_io& io) {
    BulkArgs a;
    memset(&a, 0, sizeof(a));
    a.f = uw_make_field(ctx->params);
    a.tab = ctx->d_tables;
    a.tasks = io.tasks;
    a.qhead = io.qhead;
    a.max_res = s.max_res;
    a.entries = s.entries;
    a.Lm = s.max_res + 1;
    a.wpc = s.wpc;
    a.nx_tab = s.nx_tab;
    a.ztab = io.ztab;
    a.zstride = io.zstride;
    a.packed = io.packed;
    a.dens = reinterpret_cast<float*>(io.scratch + s.dens_off);
    a.rows = reinterpret_cast<unsigned long long*>(io.scratch + s.rows_off);
    a.cases = reinterpret_cast<uint32_t*>(io.scratch + s.cases_off);
    a.below = reinterpret_cast<uint32_t*>(io.scratch + s.below_off);
    a.act = reinterpret_cast<uint32_t*>(io.scratch + s.act_off);
    a.rec = reinterpret_cast<unsigned long long*>(io.scratch + s.rec_off);
    a.info = io.out->info;
    a.totals = io.out->totals;
    a.verts = io.out->verts;
    a.inds = io.out->inds;
    a.tape = io.tape;
    a.eo = uw_make_emit_opts(io.out);
    a.vcap = io.out->vert_cap;
    a.icap = io.out->ind_cap;
    return a;
}
}  // namespace

int uw_bulk_launch_front(uw_ctx* ctx, const uw_bulk_shape& s, const uw_bulk_io& io, hipStream_t st) {
    if (!s.on) return UW_OK;
    const BulkArgs a = bulk_args(ctx, s, io);
    const BulkDensityKernel kd = pick_bulk_density(ctx->params.octaves);
    if (int rc = uw_prof_start(ctx, st, UW_STAGE_BULK_FRONT)) return rc;
    hipLaunchKernelGGL(kd, dim3(uint32_t(s.grid_density)), dim3(kBlock), s.lds_density, st, a);
    if (s.count_planes == 4)
        hipLaunchKernelGGL(k_bulk_count<4>, dim3(uint32_t(s.grid_count)), dim3(kBlock), 0, st, a);
    else
        hipLaunchKernelGGL(k_bulk_count<8>, dim3(uint32_t(s.grid_count)), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL(k_bulk_scan, dim3(uint32_t(s.grid_scan)), dim3(kBlock), 0, st, a);
    if (int rc = uw_prof_end(ctx, st)) return rc;
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}

int uw_bulk_launch_back(uw_ctx* ctx, const uw_bulk_shape& s, const uw_bulk_io& io, hipStream_t st) {
    if (!s.on) return UW_OK;
    const BulkArgs a = bulk_args(ctx, s, io);
    if (int rc = uw_prof_start(ctx, st, UW_STAGE_BULK_EMIT)) return rc;
    switch (s.emit_planes) {
        case 1: hipLaunchKernelGGL(k_bulk_emit<1>, dim3(uint32_t(s.grid_emit)), dim3(kEmitBlock), 0, st, a); break;
        case 2: hipLaunchKernelGGL(k_bulk_emit<2>, dim3(uint32_t(s.grid_emit)), dim3(kEmitBlock), 0, st, a); break;
        default: hipLaunchKernelGGL(k_bulk_emit<4>, dim3(uint32_t(s.grid_emit)), dim3(kEmitBlock), 0, st, a); break;
    }
    if (int rc = uw_prof_end(ctx, st)) return rc;
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}
