// Plain path: each stage of the chunk builder is one kernel launch over every chunk.
//
//   k_noise_grid   density of every sample (Chunk::build_iso, chunk.rs:105-129) + "some sample is
//                  not above iso" flag (the negation of early_blank_check, chunk.rs:131-133)
//   k_classify     cube case of every cell, per-cell vertex/index counts, chunk-wide prefix sums
//                  and the compacted list of active cells (chunk.rs:141-164)
//   k_offsets      exclusive scan of per-chunk sizes into output offsets, capacity check
//   k_emit         edge lerps of owned records and the index buffer (chunk.rs:164-243)
//
// HBM layout (scratch, per chunk c of resolution n, S = (n+1)^3 samples, C = n^3 cells):
//   dens  [c][S] f32      x-major, z fastest (chunk.rs:351-353)
//   cases [c][C] u8       scan order (x n + y) n + z
//   voff  [c][C] u32      vertex id of the cell's first owned record (chunk-local)
//   act   [c][C] u32      compacted active cells: cell | case << 24
//   ioff  [c][C] u32      index offset of each compacted active cell
#include "uw_mesh.h"
#include "uw_noise.h"
#include "uw_wave.h"

namespace {

constexpr int kThreads = 256;

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------------------------
// Density grid. One workgroup per (chunk, slab of `planes` x-planes): the chunk's per-axis octave
// tables are built in LDS, the slab's densities are evaluated line by line (uw_noise_planes) into
// LDS and then stored to HBM with contiguous, coalesced writes.
template <int OCT>
__global__ void __launch_bounds__(kThreads) k_noise_grid(uw_field f, const uw_tables* __restrict__ tab,
                                                         const int32_t* __restrict__ coords, int32_t n,
                                                         float step, int32_t planes, int32_t slabs,
                                                         const unsigned char* __restrict__ ztab,
                                                         float* __restrict__ dens, uint32_t* __restrict__ below) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int32_t L = n + 1, O = f.octaves;
    const int32_t chunk = blockIdx.x / slabs, slab = blockIdx.x % slabs;
    const int32_t x0 = slab * planes;
    const int32_t nx = min(planes, L - x0);
    float* G3 = reinterpret_cast<float*>(smem);                    // 256 x 3 floats
    uint8_t* P = reinterpret_cast<uint8_t*>(G3 + 3 * 256);         // 256 B
    uw_axis* AX = reinterpret_cast<uw_axis*>(P + 256);             // [O][L] x2
    uw_axis* AY = AX + O * L;
    float* SLAB = reinterpret_cast<float*>(AY + O * L);            // [planes][L][L]

    __shared__ uint32_t s_bits;
    __shared__ double s_freq[UW_MAX_OCTAVES];
    const int32_t cx = coords[3 * chunk], cy = coords[3 * chunk + 1];
    if (threadIdx.x == 0) s_bits = 0u;
    if (threadIdx.x < UW_MAX_OCTAVES) s_freq[threadIdx.x] = f.octave_freq[threadIdx.x];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) {
        const uint8_t p = tab->perm[i];
        P[i] = p;
        const float4 g = uw_grad_vec(p & 15u);
        G3[3 * i] = g.x;
        G3[3 * i + 1] = g.y;
        G3[3 * i + 2] = g.z;
    }
    __syncthreads();
    const uw_chunk_tables T = {AX, AY, nullptr, nullptr, nullptr};
    uw_build_xy_tables(f, s_freq, P, cx, cy, L, step, AX, AY, threadIdx.x, blockDim.x);
    const uw_ztab Zt = uw_ztab_of(ztab, chunk, O, L);
    __syncthreads();
    // plane x lands in slot x % planes = x - x0
    const int bits = uw_noise_planes<OCT>(f, T, Zt, P, G3, L, x0, x0 + nx, SLAB, planes, threadIdx.x, blockDim.x);
    if (bits) atomicOr(&s_bits, uint32_t(bits));
    __syncthreads();
    const int32_t LL = L * L;
    const int32_t count = nx * LL;
    float* dst = dens + uint64_t(chunk) * uint64_t(LL) * uint64_t(L) + uint64_t(x0) * uint64_t(LL);
    for (int32_t s = threadIdx.x; s < count; s += blockDim.x) dst[s] = SLAB[s];
    if (threadIdx.x == 0 && s_bits) atomicOr(&below[chunk], s_bits);
}

// ---------------------------------------------------------------------------------------------
// Block-wide exclusive scan of one u32 per thread (kThreads threads, wave64).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total, uint32_t* lds_waves) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t inc = uw_wave_incl_scan(v);
    if (lane == 63) lds_waves[wave] = inc;
    __syncthreads();
    uint32_t wave_base = 0, sum = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
        const uint32_t t = lds_waves[w];
        if (w < wave) wave_base += t;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return wave_base + inc - v;
}

// Cube cases, counts, per-chunk prefix sums and active-cell compaction. One workgroup per chunk.
__global__ void __launch_bounds__(kThreads) k_classify(float iso, const uw_tables* __restrict__ tab, int32_t n,
                                                       const uint32_t* __restrict__ below,
                                                       const float* __restrict__ dens, uint8_t* __restrict__ cases,
                                                       uint32_t* __restrict__ voff, uint32_t* __restrict__ act,
                                                       uint32_t* __restrict__ ioff, uw_chunk_info* __restrict__ info) {
    __shared__ uint32_t lut[256];  // used-edge mask | triangle count << 16
    __shared__ uint32_t waves[kThreads / 64];
    const int32_t chunk = blockIdx.x;
    if (!(below[chunk] & UW_BELOW_NOT_ABOVE)) {
        if (threadIdx.x == 0) {
            uw_chunk_info r = {UW_CHUNK_BLANK | UW_CHUNK_EARLY_SKIP, 0, 0, 0, 0, 0};
            info[chunk] = r;
        }
        return;
    }
    for (int i = threadIdx.x; i < 256; i += blockDim.x) lut[i] = uint32_t(tab->used[i]) | (uint32_t(tab->tri[i] & 15u) << 16);
    __syncthreads();
    const int32_t L = n + 1, C = n * n * n, nn = n * n;
    const uw_divider dn = uw_make_divider(uint32_t(n)), dnn = uw_make_divider(uint32_t(nn));
    const float* D = dens + uint64_t(chunk) * uint64_t(L) * uint64_t(L) * uint64_t(L);
    const uint64_t cbase = uint64_t(chunk) * uint64_t(C);
    uint32_t carry_v = 0, carry_i = 0, carry_a = 0;
    for (int32_t r0 = 0; r0 < C; r0 += kThreads) {
        const int32_t cell = r0 + threadIdx.x;
        uint32_t packed = 0, cs = 0;
        if (cell < C) {
            int32_t x, y, z;
        uw_cell_xyz(uint32_t(cell), dn, dnn, &x, &y, &z);
            float r[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const uint32_t o = uw_corner_ofs(i);
                r[i] = D[(x + int32_t(o & 1u)) * L * L + (y + int32_t((o >> 1) & 1u)) * L + z + int32_t((o >> 2) & 1u)];
            }
            cs = uw_case_of(r, iso);
            const uint32_t w = lut[cs];
            const uint32_t nv = __popc(w & 0xFFFu & uw_owned_mask(x, y, z));
            const uint32_t nt = w >> 16;
            const uint32_t a = (cs != 0u && cs != 255u) ? 1u : 0u;
            packed = nv | (nt << 12) | (a << 23);
            cases[cbase + cell] = uint8_t(cs);
        }
        uint32_t tot;
        const uint32_t pre = block_exclusive_scan(packed, &tot, waves);
        if (cell < C) {
            voff[cbase + cell] = carry_v + (pre & 0xFFFu);
            if (packed >> 23) {
                const uint32_t slot = carry_a + (pre >> 23);
                act[cbase + slot] = uint32_t(cell) | (cs << 24);
                ioff[cbase + slot] = carry_i + 3u * ((pre >> 12) & 0x7FFu);
            }
        }
        carry_v += tot & 0xFFFu;
        carry_i += 3u * ((tot >> 12) & 0x7FFu);
        carry_a += tot >> 23;
    }
    if (threadIdx.x == 0) {
        uw_chunk_info r = {carry_i == 0 ? uint32_t(UW_CHUNK_BLANK) : 0u, carry_v, carry_i, carry_a, 0, 0};
        info[chunk] = r;
    }
}

// Exclusive scan of per-chunk sizes into output offsets. One workgroup of 1024 threads.
__global__ void __launch_bounds__(1024) k_offsets(uw_chunk_info* __restrict__ info, int32_t nchunks, uint64_t vcap,
                                                  uint64_t icap, uint32_t ind16, uw_totals* __restrict__ totals) {
    __shared__ unsigned long long wv[16], wi[16];
    __shared__ unsigned int wn[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long carry_v = 0, carry_i = 0;
    unsigned int carry_n = 0, ovf = 0;
    for (int32_t c0 = 0; c0 < nchunks; c0 += 1024) {
        const int32_t c = c0 + threadIdx.x;
        uw_chunk_info r = {};
        if (c < nchunks) r = info[c];
        const unsigned long long v = r.n_verts, i = r.n_inds;
        const unsigned int nb = (c < nchunks && r.n_inds > 0) ? 1u : 0u;
        const unsigned long long sv = uw_wave_incl_scan(v), si = uw_wave_incl_scan(i);
        const unsigned int sn = uw_wave_incl_scan(nb);
        if (lane == 63) { wv[wave] = sv; wi[wave] = si; wn[wave] = sn; }
        __syncthreads();
        unsigned long long bv = 0, bi = 0, tv = 0, ti = 0;
        unsigned int tn = 0;
        for (int w = 0; w < 16; ++w) {
            if (w < wave) { bv += wv[w]; bi += wi[w]; }
            tv += wv[w];
            ti += wi[w];
            tn += wn[w];
        }
        __syncthreads();
        if (c < nchunks) {
            const unsigned long long vo = carry_v + bv + sv - v, io = carry_i + bi + si - i;
            r.vert_off = vo;
            r.ind_off = io;
            if (vo + v > vcap || io + i > icap) {
                r.flags |= UW_CHUNK_OVERFLOW;
                ovf = 1;
            }
            if (ind16 && v > 65535u) r.flags |= UW_CHUNK_INDEX_WRAP;
            info[c] = r;
        }
        carry_v += tv;
        carry_i += ti;
        carry_n += tn;
    }
    ovf = __syncthreads_or(ovf);
    if (threadIdx.x == 0) {
        uw_totals t;
        t.n_verts = carry_v;
        t.n_inds = carry_i;
        t.n_nonblank = carry_n;
        t.overflow = ovf ? 1u : 0u;
        *totals = t;
    }
}

// Vertex lerps of owned records and the index buffer. One workgroup per chunk, one thread per
// active cell.
__global__ void __launch_bounds__(kThreads) k_emit(float iso, int32_t extent, const uw_tables* __restrict__ tab,
                                                   const int32_t* __restrict__ coords, int32_t n, float step,
                                                   const float* __restrict__ dens, const uint8_t* __restrict__ cases,
                                                   const uint32_t* __restrict__ voff, const uint32_t* __restrict__ act,
                                                   const uint32_t* __restrict__ ioff,
                                                   const uw_chunk_info* __restrict__ info, float* __restrict__ verts,
                                                   uint32_t* __restrict__ inds, uw_emit_opts eo) {
    const int32_t chunk = blockIdx.x;
    const uw_chunk_info ci = info[chunk];
    if (ci.flags & (UW_CHUNK_BLANK | UW_CHUNK_OVERFLOW)) return;
    const int32_t L = n + 1, C = n * n * n, nn = n * n;
    const uw_divider dn = uw_make_divider(uint32_t(n)), dnn = uw_make_divider(uint32_t(nn));
    const float* D = dens + uint64_t(chunk) * uint64_t(L) * uint64_t(L) * uint64_t(L);
    const uint64_t cbase = uint64_t(chunk) * uint64_t(C);
    const float ox = float(extent * coords[3 * chunk]), oy = float(extent * coords[3 * chunk + 1]),
                oz = float(extent * coords[3 * chunk + 2]);
    __shared__ uint64_t s_tri[256];
    __shared__ __align__(16) uint16_t s_bef[256 * 12];
    __shared__ uint16_t s_used[256];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) {
        s_tri[i] = tab->tri[i];
        s_used[i] = tab->used[i];
    }
    uw_lds_load_before(s_bef, tab, int(threadIdx.x), int(blockDim.x));
    __syncthreads();
    for (uint32_t a = threadIdx.x; a < ci.n_active; a += blockDim.x) {
        const uint32_t ent = act[cbase + a];
        const int32_t cell = int32_t(ent & 0xFFFFFFu);
        const uint32_t cs = ent >> 24;
        int32_t x, y, z;
        uw_cell_xyz(uint32_t(cell), dn, dnn, &x, &y, &z);
        float r[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t o = uw_corner_ofs(i);
            r[i] = D[(x + int32_t(o & 1u)) * L * L + (y + int32_t((o >> 1) & 1u)) * L + z + int32_t((o >> 2) & 1u)];
        }
        const uint32_t owned = uw_owned_mask(x, y, z);
        const uint32_t myv = voff[cbase + cell];
        const uint64_t vb = ci.vert_off + myv;
        const uint64_t ib = ci.ind_off + ioff[cbase + a];
        // ends of the chunk's output ranges: writes stay inside them whatever the ids
        const uint64_t vlim = ci.vert_off + ci.n_verts, ilim = ci.ind_off + ci.n_inds;
        // vertices: one per owned edge of the case, at its rank in first-appearance order
        for (uint32_t vm = uint32_t(s_used[cs]) & owned; vm; vm &= vm - 1u) {
            const uint32_t e = uint32_t(__builtin_ctz(vm));
            const uint32_t rank = __popc(uint32_t(s_bef[cs * 12 + e]) & owned);
            if (vb + rank >= vlim) continue;
            float p[3];
            uw_edge_vertex(e, x, y, z, r, iso, step, ox, oy, oz, p);
            float* dst = verts + eo.vstride * (vb + rank);
            dst[0] = p[0];
            dst[1] = p[1];
            dst[2] = p[2];
            if (eo.vstride == 6) uw_vertex_color(p[2], uw_edge_b(e), float(extent), dst + 3);
        }
        // indices in the row's order; records owned elsewhere take the owner's id
        const uint64_t row = s_tri[cs];
        const int nidx = 3 * int(row & 15u);
        const bool coll = eo.tri_normals || eo.tri_cells;
        float q0[3] = {0.f, 0.f, 0.f}, q1[3] = {0.f, 0.f, 0.f};
        for (int j = 0; j < nidx; ++j) {
            const uint32_t e = uint32_t(row >> (4 + 4 * j)) & 15u;
            uint32_t id;
            if (owned & (1u << e)) {
                id = myv + __popc(uint32_t(s_bef[cs * 12 + e]) & owned);
            } else {
                int32_t qx, qy, qz;
                uint32_t qe;
                uw_edge_owner(x, y, z, e, &qx, &qy, &qz, &qe);
                const int32_t q = (qx * n + qy) * n + qz;
                const uint32_t qc = cases[cbase + q];
                id = voff[cbase + q] + __popc(uint32_t(s_bef[qc * 12 + qe]) & uw_owned_mask(qx, qy, qz));
            }
            if (ib + j >= ilim) continue;
            uw_store_index(inds, eo.ind16, ib + j, id);
            if (coll) {
                float p[3];
                uw_edge_vertex(e, x, y, z, r, iso, step, ox, oy, oz, p);
                const int k = j % 3;
                if (k == 0) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) q0[d] = p[d];
                } else if (k == 1) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) q1[d] = p[d];
                } else {
                    const uint64_t t = (ib + j) / 3;
                    if (eo.tri_normals) {
                        float nm[3];
                        uw_tri_normal(q0, q1, p, nm);
                        eo.tri_normals[3 * t] = nm[0];
                        eo.tri_normals[3 * t + 1] = nm[1];
                        eo.tri_normals[3 * t + 2] = nm[2];
                    }
                    if (eo.tri_cells) eo.tri_cells[t] = uint32_t(cell);
                }
            }
        }
    }
}

// LDS bytes of k_noise_grid.
size_t noise_lds_bytes(int32_t O, int32_t planes, int32_t L) {
    return 256 * 3 * sizeof(float) + 256 + sizeof(uw_axis) * 2 * size_t(O) * size_t(L) +
           sizeof(float) * size_t(planes) * size_t(L) * size_t(L);
}

// x-planes per workgroup: the whole chunk while it fits 16 KB of LDS, else fewer (at least one).
int32_t noise_planes(int32_t L) {
    int32_t p = (16 * 1024) / (4 * L * L);
    return p < 1 ? 1 : (p > L ? L : p);
}

using NoiseKernel = void (*)(uw_field, const uw_tables*, const int32_t*, int32_t, float, int32_t, int32_t,
                             const unsigned char*, float*, uint32_t*);

NoiseKernel pick_noise(int octaves) {
    switch (octaves) {
        case 3: return k_noise_grid<3>;
        case 4: return k_noise_grid<4>;
        case 6: return k_noise_grid<6>;
        case 8: return k_noise_grid<8>;
        default: return k_noise_grid<0>;
    }
}

int ensure_plain_scratch(uw_ctx* ctx, int32_t nchunks, int32_t n) {
    uw_plain_scratch& s = ctx->plain;
    const size_t L = size_t(n) + 1, S = L * L * L, C = size_t(n) * n * n;
    const size_t need_dens = size_t(nchunks) * S, need_cells = size_t(nchunks) * C;
    const size_t have_dens = s.chunks * (size_t(s.n) + 1) * (size_t(s.n) + 1) * (size_t(s.n) + 1);
    const size_t have_cells = s.chunks * size_t(s.n) * s.n * s.n;
    if (s.dens && need_dens <= have_dens && need_cells <= have_cells && size_t(nchunks) <= s.chunks) return UW_OK;
    (void)hipFree(s.dens);
    (void)hipFree(s.below);
    (void)hipFree(s.cases);
    (void)hipFree(s.voff);
    (void)hipFree(s.act);
    (void)hipFree(s.ioff);
    s = uw_plain_scratch();
    UW_HIP(ctx, hipMalloc(&s.dens, need_dens * sizeof(float)));
    UW_HIP(ctx, hipMalloc(&s.below, size_t(nchunks) * sizeof(uint32_t)));
    UW_HIP(ctx, hipMalloc(&s.cases, need_cells));
    UW_HIP(ctx, hipMalloc(&s.voff, need_cells * sizeof(uint32_t)));
    UW_HIP(ctx, hipMalloc(&s.act, need_cells * sizeof(uint32_t)));
    UW_HIP(ctx, hipMalloc(&s.ioff, need_cells * sizeof(uint32_t)));
    s.chunks = size_t(nchunks);
    s.n = n;
    return UW_OK;
}

int launch_noise(uw_ctx* ctx, const uw_field& f, const int32_t* coords, int32_t nchunks, int32_t n, float* dens,
                 uint32_t* below, hipStream_t st) {
    const int32_t L = n + 1;
    const int32_t planes = noise_planes(L);
    const int32_t slabs = (L + planes - 1) / planes;
    const float step = float(f.extent) / float(n);
    const size_t lds = noise_lds_bytes(f.octaves, planes, L);
    const uint64_t blocks = uint64_t(nchunks) * uint64_t(slabs);
    if (blocks > 0x7FFFFFFFull) return uw_fail(ctx, UW_ERR_INVALID_ARG, "too many chunks for one launch");
    NoiseKernel k = pick_noise(f.octaves);
    UW_HIP(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lds)));
    uw_plain_scratch& s = ctx->plain;
    const size_t zbytes = size_t(nchunks) * uw_ztab_bytes(f.octaves, L);
    if (s.ztab_bytes < zbytes) {
        (void)hipFree(s.ztab);
        s.ztab = nullptr;
        s.ztab_bytes = 0;
        UW_HIP(ctx, hipMalloc(&s.ztab, zbytes));
        s.ztab_bytes = zbytes;
    }
    if (int rc = uw_launch_ztables(ctx, f, coords, nullptr, n, nchunks, L, s.ztab, st)) return rc;
    hipLaunchKernelGGL(k_fill_u32, dim3(uint32_t((nchunks + 255) / 256)), dim3(256), 0, st, below, uint64_t(nchunks), 0u);
    if (int rc = uw_prof_start(ctx, st)) return rc;
    hipLaunchKernelGGL(k, dim3(uint32_t(blocks)), dim3(kThreads), lds, st, f, ctx->d_tables, coords, n, step, planes,
                       slabs, s.ztab, dens, below);
    if (int rc = uw_prof_end(ctx, st)) return rc;
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}

}  // namespace

extern "C" {

int uw_sample_density(uw_ctx* ctx, const int32_t* coords, int32_t nchunks, int32_t n, float* dens, uint32_t* below,
                      void* stream) {
    if (!ctx) return UW_ERR_INVALID_ARG;
    if (int rc = uw_check_resolution(ctx, n)) return rc;
    if (nchunks < 0 || (nchunks > 0 && (!coords || !dens))) return uw_fail(ctx, UW_ERR_INVALID_ARG, "bad arguments");
    if (nchunks == 0) return UW_OK;
    UW_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = uw_refuse_capture(ctx, static_cast<hipStream_t>(stream))) return rc;
    const uw_field f = uw_make_field(ctx->params);
    uint32_t* flags = below;
    if (!flags) {
        if (int rc = ensure_plain_scratch(ctx, nchunks, n)) return rc;
        flags = ctx->plain.below;
    }
    return launch_noise(ctx, f, coords, nchunks, n, dens, flags, static_cast<hipStream_t>(stream));
}

int uw_build_chunks_plain(uw_ctx* ctx, const int32_t* coords, int32_t nchunks, int32_t n, const uw_mesh_out* out,
                          uint8_t* cases, void* stream) {
    if (!ctx) return UW_ERR_INVALID_ARG;
    if (int rc = uw_check_resolution(ctx, n)) return rc;
    if (nchunks < 0 || !out || !out->info || !out->totals || (nchunks > 0 && !coords) ||
        (out->vert_cap > 0 && !out->verts) || (out->ind_cap > 0 && !out->inds) || out->vertex_format > 1 ||
        out->index_format > 1)
        return uw_fail(ctx, UW_ERR_INVALID_ARG, "bad arguments");
    UW_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (int rc = uw_refuse_capture(ctx, st)) return rc;
    if (nchunks == 0) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, reinterpret_cast<uint32_t*>(out->totals),
                           uint64_t(sizeof(uw_totals) / 4), 0u);
        UW_HIP(ctx, hipGetLastError());
        return UW_OK;
    }
    if (int rc = ensure_plain_scratch(ctx, nchunks, n)) return rc;
    uw_plain_scratch& s = ctx->plain;
    const uw_field f = uw_make_field(ctx->params);
    if (int rc = launch_noise(ctx, f, coords, nchunks, n, s.dens, s.below, st)) return rc;
    uint8_t* cs = cases ? cases : s.cases;
    if (int rc = uw_prof_start(ctx, st, UW_STAGE_PLAIN_CLASSIFY)) return rc;
    hipLaunchKernelGGL(k_classify, dim3(uint32_t(nchunks)), dim3(kThreads), 0, st, f.iso, ctx->d_tables, n, s.below,
                       s.dens, cs, s.voff, s.act, s.ioff, out->info);
    if (int rc = uw_prof_end(ctx, st)) return rc;
    if (int rc = uw_prof_start(ctx, st, UW_STAGE_PLAIN_OFFSETS)) return rc;
    hipLaunchKernelGGL(k_offsets, dim3(1), dim3(1024), 0, st, out->info, nchunks, out->vert_cap, out->ind_cap,
                       out->index_format == UW_INDEX_U16 ? 1u : 0u, out->totals);
    if (int rc = uw_prof_end(ctx, st)) return rc;
    const float step = float(f.extent) / float(n);
    if (int rc = uw_prof_start(ctx, st, UW_STAGE_PLAIN_EMIT)) return rc;
    hipLaunchKernelGGL(k_emit, dim3(uint32_t(nchunks)), dim3(kThreads), 0, st, f.iso, f.extent, ctx->d_tables, coords,
                       n, step, s.dens, cs, s.voff, s.act, s.ioff, out->info, out->verts, out->inds,
                       uw_make_emit_opts(out));
    if (int rc = uw_prof_end(ctx, st)) return rc;
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}

int uw_chunk_build_host_out(uw_ctx* ctx, int32_t cx, int32_t cy, int32_t cz, int32_t n, const uw_host_out* ho,
                            uw_chunk_info* info) {
    if (!ctx || !info || !ho) return UW_ERR_INVALID_ARG;
    if (int rc = uw_check_resolution(ctx, n)) return rc;
    if ((ho->vert_cap > 0 && !ho->verts) || (ho->ind_cap > 0 && !ho->inds) || ho->vertex_format > 1 ||
        ho->index_format > 1)
        return uw_fail(ctx, UW_ERR_INVALID_ARG, "bad arguments");
    UW_HIP(ctx, hipSetDevice(ctx->device));
    const uint64_t vmax = uw_max_verts(n), imax = uw_max_inds(n);
    if (!ctx->d_coord1) {
        UW_HIP(ctx, hipMalloc(&ctx->d_coord1, 3 * sizeof(int32_t)));
        UW_HIP(ctx, hipMalloc(&ctx->d_info1, sizeof(uw_chunk_info)));
        UW_HIP(ctx, hipMalloc(&ctx->d_totals1, sizeof(uw_totals)));
    }
    if (ctx->verts1_cap < vmax || ctx->inds1_cap < imax) {
        (void)hipFree(ctx->d_verts1);
        (void)hipFree(ctx->d_inds1);
        (void)hipFree(ctx->d_norm1);
        (void)hipFree(ctx->d_cells1);
        ctx->d_verts1 = nullptr;
        ctx->d_inds1 = nullptr;
        ctx->d_norm1 = nullptr;
        ctx->d_cells1 = nullptr;
        ctx->verts1_cap = ctx->inds1_cap = 0;
        UW_HIP(ctx, hipMalloc(&ctx->d_verts1, vmax * 6 * sizeof(float)));
        UW_HIP(ctx, hipMalloc(&ctx->d_inds1, imax * sizeof(uint32_t)));
        ctx->verts1_cap = vmax;
        ctx->inds1_cap = imax;
    }
    const bool want_norm = ho->tri_normals != nullptr, want_cells = ho->tri_cells != nullptr;
    if (want_norm && !ctx->d_norm1) UW_HIP(ctx, hipMalloc(&ctx->d_norm1, ctx->inds1_cap * sizeof(float)));
    if (want_cells && !ctx->d_cells1) UW_HIP(ctx, hipMalloc(&ctx->d_cells1, ctx->inds1_cap / 3 * sizeof(uint32_t)));
    const int32_t c[3] = {cx, cy, cz};
    UW_HIP(ctx, hipMemcpy(ctx->d_coord1, c, sizeof(c), hipMemcpyHostToDevice));
    uw_mesh_out out = {ctx->d_verts1, ctx->d_inds1, ctx->d_info1, ctx->d_totals1, ctx->verts1_cap, ctx->inds1_cap,
                       ho->vertex_format, ho->index_format, want_norm ? ctx->d_norm1 : nullptr,
                       want_cells ? ctx->d_cells1 : nullptr};
    if (int rc = uw_build_chunks_plain(ctx, ctx->d_coord1, 1, n, &out, nullptr, nullptr)) return rc;
    UW_HIP(ctx, hipMemcpy(info, ctx->d_info1, sizeof(uw_chunk_info), hipMemcpyDeviceToHost));
    if (info->n_verts > ho->vert_cap || info->n_inds > ho->ind_cap)
        return uw_fail(ctx, UW_ERR_OVERFLOW, "output capacity too small for this chunk");
    const size_t vfl = ho->vertex_format == UW_VERTEX_XYZ_RGB ? 6 : 3;
    const size_t isz = ho->index_format == UW_INDEX_U16 ? sizeof(uint16_t) : sizeof(uint32_t);
    const size_t ntri = info->n_inds / 3;
    if (info->n_verts)
        UW_HIP(ctx, hipMemcpy(ho->verts, ctx->d_verts1, size_t(info->n_verts) * vfl * sizeof(float),
                              hipMemcpyDeviceToHost));
    if (info->n_inds) UW_HIP(ctx, hipMemcpy(ho->inds, ctx->d_inds1, size_t(info->n_inds) * isz, hipMemcpyDeviceToHost));
    if (ntri && want_norm)
        UW_HIP(ctx, hipMemcpy(ho->tri_normals, ctx->d_norm1, ntri * 3 * sizeof(float), hipMemcpyDeviceToHost));
    if (ntri && want_cells)
        UW_HIP(ctx, hipMemcpy(ho->tri_cells, ctx->d_cells1, ntri * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return UW_OK;
}

int uw_chunk_build_host(uw_ctx* ctx, int32_t cx, int32_t cy, int32_t cz, int32_t n, float* verts, uint64_t vert_cap,
                        uint32_t* inds, uint64_t ind_cap, uw_chunk_info* info) {
    const uw_host_out ho = {verts, vert_cap, inds, ind_cap, UW_VERTEX_XYZ, UW_INDEX_U32, nullptr, nullptr};
    return uw_chunk_build_host_out(ctx, cx, cy, cz, n, &ho, info);
}

}  // extern "C"
