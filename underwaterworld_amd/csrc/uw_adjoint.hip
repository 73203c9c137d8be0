// Adjoint of the chunk builder: gradients of a loss on the emitted vertex positions with respect to
// the density samples, the octave amplitudes and a world-space offset of every sample coordinate.
// The reference has no differentiation; this follows the forward defined by
//   v = s_a + t (s_b - s_a),  t = (iso - rho_a) / (rho_b - rho_a)       (chunk.rs:178-213)
//   rho = bias(p_z) + sum_k A_k n_k(f_k p) / sum_k A_k                    (perlin_util.rs:6-29)
// with the cube cases held fixed (the mesh is piecewise smooth in its inputs):
//   dv/drho_a = (s_b - s_a) (iso - rho_b) / D^2,   dv/drho_b = (s_b - s_a) (rho_a - iso) / D^2
//   drho/dA_j = (n_j - fbm) / sum A
//   drho/ddelta = freq (sum_k A_k f_k grad n_k / sum A + bias'(p_z) e_z)
// Only the forward's vertex tape is kept: the octave noise values and gradients at each vertex's
// two endpoint samples are recomputed here (per-axis lattice data of the chunk in LDS, then eight
// hashed gradients, fades and fade derivatives per octave), so a large volume needs no stored
// density or per-octave fields.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <cmath>

#include "uw_div.h"
#include "uw_mesh.h"
#include "uw_noise.h"
#include "uw_wave.h"

namespace {

constexpr int kMaxL = UW_BATCH_MAX_RESOLUTION + 1;
constexpr int kBlock = 256;

struct BackArgs {
    uw_field f;
    const uw_tables* tab;
    const int32_t* coords;
    const int32_t* res;
    const uw_chunk_info* info;
    const uw_vertex_tape* tape;
    const float* grad_verts;
    float* grad_dens;
    int64_t dens_stride;
    double* grad_amp;
    double* grad_off;
    int32_t nchunks;
    const void* tabs;  // per-chunk tables (BackTables<OMAX>), written by k_backward_tables
    uint32_t* plan;    // piece plan (BackPlan), written by the first workgroup of k_backward_tables
    int32_t grid;      // workgroups of k_backward
    double* part;      // per (workgroup, chunk) parameter sums at slot w + c, [slot][OMAX + 3]
};

// Work split of the backward. A chunk's vertices are cut into pieces of kPiece vertices (one
// endpoint per thread); the launch's pieces are numbered chunk by chunk, and workgroup w of the
// persistent grid takes the contiguous pieces [w per, min(T, (w + 1) per)), per = ceil(T / grid),
// so every workgroup has the same number of pieces whatever the chunks' vertex counts. Workgroup w
// leaves one parameter sum per chunk it touches, at slot w + c (unique: the chunks of later
// workgroups never come before the last chunk of an earlier one); k_backward_sum adds a chunk's
// sums in workgroup order.
constexpr int kPiece = kBlock / 2;
// plan layout (u32): [0] T, [1] per, [2 ..] piece start of chunk c (c = 0 .. nchunks, the last one
// is T), then the first chunk of every workgroup of the grid
__host__ __device__ inline size_t plan_words(int32_t nchunks, int32_t grid) { return 2 + size_t(nchunks) + 1 + size_t(grid); }

// Per-chunk lattice data of the backward: every (axis, octave, sample) entry of the chunk (x hashed)
// and the fade derivatives s'(d) = 30 d^2 (d - 1)^2, each at [k L + i] for the chunk's own L. Built
// once per launch by k_backward_tables (one thread per entry, so the f64 chains of all chunks run
// side by side) and copied into LDS by each workgroup that takes vertices of the chunk.
template <int OMAX>
struct BackTables {
    uw_axis AX[OMAX * kMaxL], AY[OMAX * kMaxL], AZ[OMAX * kMaxL];
    float DF[3][OMAX * kMaxL];
};

__device__ __forceinline__ bool back_skip(const BackArgs& a, int32_t c, int32_t* n) {
    *n = a.res[c];
    const uw_chunk_info ci = a.info[c];
    return *n < 1 || *n > UW_BATCH_MAX_RESOLUTION || (ci.flags & (UW_CHUNK_OVERFLOW | UW_CHUNK_INVALID)) ||
           ci.n_verts == 0;
}

// The piece plan, by one 256-thread workgroup: pieces per chunk (none for a skipped chunk), their
// exclusive scan, and the first chunk of every workgroup of the backward grid.
__device__ void build_plan(const BackArgs& a) {
    __shared__ uint32_t ws[kBlock / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t per_thread = (a.nchunks + kBlock - 1) / kBlock;
    const int32_t c0 = min(a.nchunks, tid * per_thread), c1 = min(a.nchunks, c0 + per_thread);
    uint32_t mine = 0;
    for (int32_t c = c0; c < c1; ++c) {
        int32_t n;
        if (!back_skip(a, c, &n)) mine += (a.info[c].n_verts + kPiece - 1) / kPiece;
    }
    const uint32_t inc = uw_wave_incl_scan(mine);
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    uint32_t base = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        base += w < wave ? ws[w] : 0u;
        total += ws[w];
    }
    base += inc - mine;
    const uint32_t per = total == 0 ? 1u : (total + uint32_t(a.grid) - 1) / uint32_t(a.grid);
    uint32_t* ps = a.plan + 2;
    uint32_t* first = ps + a.nchunks + 1;
    if (tid == 0) {
        a.plan[0] = total;
        a.plan[1] = per;
        ps[a.nchunks] = total;
    }
    for (int32_t c = c0; c < c1; ++c) {
        int32_t n;
        const uint32_t pc = back_skip(a, c, &n) ? 0u : (a.info[c].n_verts + kPiece - 1) / kPiece;
        ps[c] = base;
        // workgroups whose first piece lies in this chunk
        for (uint32_t w = (base + per - 1) / per; w * per < base + pc; ++w) first[w] = uint32_t(c);
        base += pc;
    }
}

// Workgroups of the table launch: 0 builds the plan, [1, 1 + zero_blocks) zero grad_dens (the rows
// of the launch's chunks, dens_stride floats each), the rest build the tables.
template <int OMAX>
__global__ void __launch_bounds__(256) k_backward_tables(BackArgs a, BackTables<OMAX>* __restrict__ out,
                                                         uint32_t zero_blocks) {
    if (blockIdx.x == 0) {
        build_plan(a);
        return;
    }
    if (blockIdx.x <= zero_blocks) {
        const int64_t total = int64_t(a.nchunks) * a.dens_stride;
        const int64_t span = (total + zero_blocks - 1) / zero_blocks;
        const int64_t b = int64_t(blockIdx.x - 1) * span, e = min(total, b + span);
        for (int64_t i = b + threadIdx.x; i < e; i += 256) a.grad_dens[i] = 0.0f;
        return;
    }
    constexpr int kAx = 3 * OMAX * kMaxL;  // axis entries: [axis][octave][sample]
    constexpr int kE = 2 * kAx;            // + fade derivatives
    const int64_t g = int64_t(blockIdx.x - 1 - zero_blocks) * 256 + threadIdx.x;
    const int32_t c = int32_t(g / kE);
    if (c >= a.nchunks) return;
    int32_t n;
    if (back_skip(a, c, &n)) return;
    const int e = int(g - int64_t(c) * kE);
    const int32_t L = n + 1, O = a.f.octaves;
    const float step = float(a.f.extent) / float(n);
    const int32_t cc[3] = {a.coords[3 * c], a.coords[3 * c + 1], a.coords[3 * c + 2]};
    BackTables<OMAX>& t = out[c];
    {
        const int part = e / kAx, r0 = e - part * kAx;
        const int axis = r0 / (OMAX * kMaxL), r = r0 - axis * (OMAX * kMaxL);
        const int k = r / kMaxL, i = r - k * kMaxL;
        if (k >= O || i >= L) return;
        const double p = uw_sample_coord(a.f, axis, cc[axis], i, step);
        if (part == 0) {
            const uw_axis v = uw_axis_of(p, a.f.octave_freq[k]);
            if (axis == 0) t.AX[k * L + i] = uw_hash_x(v, a.tab->perm);
            else if (axis == 1) t.AY[k * L + i] = v;
            else t.AZ[k * L + i] = v;
        } else {
            const double pk = p * a.f.octave_freq[k];
            const double d = pk - floor(pk);
            t.DF[axis][k * L + i] = float(30.0 * d * d * (d - 1.0) * (d - 1.0));
        }
    }
}

template <int OMAX, int OEXACT>
__global__ void __launch_bounds__(kBlock) k_backward(BackArgs a) {
    __shared__ float G3[3 * 256];
    __shared__ uint8_t P[256];
    __shared__ __align__(16) BackTables<OMAX> TB;
    __shared__ double red[kBlock / 64][OMAX + 3];
    uw_axis* AX = TB.AX;
    uw_axis* AY = TB.AY;
    uw_axis* AZ = TB.AZ;
    float(*DF)[OMAX * kMaxL] = TB.DF;

    const int tid = threadIdx.x, wave = tid >> 6;
    const uint32_t T = a.plan[0], per = a.plan[1];
    const uint32_t p0 = blockIdx.x * per, p1 = min(T, p0 + per);
    if (p0 >= p1) return;
    const uint32_t* ps = a.plan + 2;
    const bool want_params = a.grad_amp != nullptr || a.grad_off != nullptr;
    for (int i = tid; i < 256; i += kBlock) {
        const uint8_t p = a.tab->perm[i];
        P[i] = p;
        const float4 g = uw_grad_vec(p & 15u);
        G3[3 * i] = g.x;
        G3[3 * i + 1] = g.y;
        G3[3 * i + 2] = g.z;
    }
    // OEXACT > 0: the launch's octave count, known at compile time (no per-octave branches)
    const int32_t O = OEXACT > 0 ? OEXACT : a.f.octaves;
    const float iso = a.f.iso;
    const double inv_sum = 1.0 / double(a.f.amp_sum);
    const double slope = a.f.bias_kind == UW_BIAS_LINEAR ? double(a.f.bias_a) : 0.0;
    const double fq = a.f.frequency;
    double acc[OMAX + 3];
#pragma unroll
    for (int k = 0; k < OMAX + 3; ++k) acc[k] = 0.0;
    // the workgroup's sums of chunk c (every thread's accumulators) into slot blockIdx.x + c
    auto flush = [&](int32_t c) {
#pragma unroll
        for (int k = 0; k < OMAX + 3; ++k) {
            const double t = uw_wave_sum_f64(acc[k]);
            if ((tid & 63) == 0) red[wave][k] = t;
            acc[k] = 0.0;
        }
        __syncthreads();
        if (tid < OMAX + 3) {
            double t = 0.0;
            for (int w = 0; w < kBlock / 64; ++w) t += red[w][tid];
            a.part[(int64_t(blockIdx.x) + c) * (OMAX + 3) + tid] = t;
        }
        __syncthreads();  // red is written again by the next flush
    };

    // chunk-level values (block-uniform), set when the walk enters a chunk
    int32_t c = __builtin_amdgcn_readfirstlane(int32_t(a.plan[2 + a.nchunks + 1 + blockIdx.x]));
    uint32_t cbeg = 0, cend = 0, nv = 0;
    uint64_t voff = 0;
    int32_t L = 0;
    float step = 0.0f;
    uw_divider dL = uw_make_divider(1u), dLL = dL;
    bool tables = false;  // TB holds chunk c's tables
    for (uint32_t p = p0; p < p1; ++p) {
        if (p == p0 || p >= cend) {  // the first chunk, or the next one with pieces
            if (p != p0) {
                if (want_params) flush(c);
                do {
                    ++c;
                } while (p >= ps[c + 1]);
            }
            cbeg = ps[c];
            cend = ps[c + 1];
            const int32_t n = a.res[c];
            const uw_chunk_info ci = a.info[c];
            nv = ci.n_verts;
            voff = ci.vert_off;
            L = n + 1;
            step = float(a.f.extent) / float(n);
            dL = uw_make_divider(uint32_t(L));
            dLL = uw_make_divider(uint32_t(L * L));
            tables = false;
        }
        if (want_params && !tables) {
            // the chunk's tables: 16-byte loads, all of a thread's issued before its stores
            static_assert(sizeof(BackTables<OMAX>) % 16 == 0, "tables are copied in 16-byte words");
            constexpr int kWords = int(sizeof(BackTables<OMAX>) / 16);
            const uint4* src = reinterpret_cast<const uint4*>(static_cast<const BackTables<OMAX>*>(a.tabs) + c);
            uint4* dst = reinterpret_cast<uint4*>(&TB);
            __syncthreads();  // the previous chunk's tables are read no more
            for (int q0 = 0; q0 < kWords; q0 += 4 * kBlock) {  // four loads in flight per thread
                const int e0 = q0 + tid, e1 = e0 + kBlock, e2 = e1 + kBlock, e3 = e2 + kBlock;
                uint4 v0 = make_uint4(0, 0, 0, 0), v1 = v0, v2 = v0, v3 = v0;
                if (e0 < kWords) v0 = src[e0];
                if (e1 < kWords) v1 = src[e1];
                if (e2 < kWords) v2 = src[e2];
                if (e3 < kWords) v3 = src[e3];
                if (e0 < kWords) dst[e0] = v0;
                if (e1 < kWords) dst[e1] = v1;
                if (e2 < kWords) dst[e2] = v2;
                if (e3 < kWords) dst[e3] = v3;
            }
            __syncthreads();  // (the first pass also covers G3 and P)
            tables = true;
        }
        const uint32_t vi = (p - cbeg) * kPiece + uint32_t(tid >> 1);
        if (vi >= nv) continue;
        const int e = tid & 1;
        const uint64_t v = voff + vi;
        const uw_vertex_tape tv = a.tape[v];
        const int32_t diff = int32_t(tv.b) - int32_t(tv.a);
        const int axis = (diff == 1 || diff == -1) ? 2 : ((diff == L || diff == -L) ? 1 : 0);
        const float dir = diff > 0 ? step : -step;
        const float gv = a.grad_verts[3 * v + axis] * dir;
        const float D = tv.rho_b - tv.rho_a;
        const float inv_d2 = 1.0f / (D * D);
        const float w = e == 0 ? gv * (iso - tv.rho_b) * inv_d2 : gv * (tv.rho_a - iso) * inv_d2;
        const uint32_t s = e == 0 ? tv.a : tv.b;
        if (a.grad_dens) unsafeAtomicAdd(a.grad_dens + int64_t(c) * a.dens_stride + s, w);
        if (!want_params) continue;
        const uint32_t xy = uw_div(s, dL);
        const int32_t x = int32_t(uw_div(s, dLL)), y = int32_t(xy - uint32_t(x * L)), z = int32_t(s - xy * uint32_t(L));
        float nk[OMAX];
        float total = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
        for (int k = 0; k < OMAX; ++k) {
            if (k < O) {
                float g[3];
                nk[k] = uw_octave_grad(AX[k * L + x], AY[k * L + y], AZ[k * L + z], DF[0][k * L + x], DF[1][k * L + y],
                                       DF[2][k * L + z], P, G3, g);
                const float af = a.f.amp[k] * float(a.f.octave_freq[k]);
                total = fmaf(a.f.amp[k], nk[k], total);
                gx = fmaf(af, g[0], gx);
                gy = fmaf(af, g[1], gy);
                gz = fmaf(af, g[2], gz);
            }
        }
        const double we = double(w);
        const double fbm = double(total) * inv_sum;
#pragma unroll
        for (int k = 0; k < OMAX; ++k)
            if (k < O) acc[k] += we * (double(nk[k]) - fbm) * inv_sum;
        acc[OMAX + 0] += we * fq * double(gx) * inv_sum;
        acc[OMAX + 1] += we * fq * double(gy) * inv_sum;
        acc[OMAX + 2] += we * fq * (double(gz) * inv_sum + slope);
    }
    if (want_params) flush(c);
}

// Parameter gradients of each chunk: the sums its workgroups left at slots w + c, added in
// workgroup order; zero for a chunk without pieces. One thread per (chunk, parameter).
template <int OMAX>
__global__ void __launch_bounds__(256) k_backward_sum(BackArgs a) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= int64_t(a.nchunks) * (OMAX + 3)) return;
    const int64_t c = e / (OMAX + 3);
    const int k = int(e - c * (OMAX + 3));
    const uint32_t per = a.plan[1];
    const uint32_t* ps = a.plan + 2;
    const uint32_t pa = ps[c], pb = ps[c + 1];
    double t = 0.0;
    if (pb > pa)
        for (uint32_t w = pa / per; w <= (pb - 1) / per; ++w) t += a.part[(int64_t(w) + c) * (OMAX + 3) + k];
    if (k < OMAX) {
        if (k < a.f.octaves && a.grad_amp) a.grad_amp[c * a.f.octaves + k] = t;
    } else if (a.grad_off) {
        a.grad_off[c * 3 + (k - OMAX)] = t;
    }
}

int ensure_buffer(uw_ctx* ctx, void** buf, size_t* have, size_t bytes) {
    if (bytes <= *have) return UW_OK;
    (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    UW_HIP(ctx, hipMalloc(buf, bytes));
    *have = bytes;
    return UW_OK;
}

template <int OMAX, int OEXACT>
int launch_backward(uw_ctx* ctx, BackArgs a, int32_t nchunks, hipStream_t st) {
    const bool want_params = a.grad_amp != nullptr || a.grad_off != nullptr;
    int per_cu = 0;
    UW_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_backward<OMAX, OEXACT>),
                                                             kBlock, 0));
    if (per_cu < 1) return uw_fail(ctx, UW_ERR_INVALID_ARG, "backward kernel does not fit on a CU");
    a.grid = per_cu * ctx->num_cus;
    a.nchunks = nchunks;
    if (int rc = ensure_buffer(ctx, reinterpret_cast<void**>(&ctx->d_adj_plan), &ctx->adj_plan_bytes,
                               plan_words(nchunks, a.grid) * sizeof(uint32_t)))
        return rc;
    a.plan = ctx->d_adj_plan;
    a.tabs = nullptr;
    a.part = nullptr;
    uint64_t table_blocks = 0;
    if (want_params) {
        if (int rc = ensure_buffer(ctx, &ctx->d_adj_tabs, &ctx->adj_tabs_bytes, size_t(nchunks) * sizeof(BackTables<OMAX>)))
            return rc;
        if (int rc = ensure_buffer(ctx, reinterpret_cast<void**>(&ctx->d_adj_part), &ctx->adj_part_bytes,
                                   (size_t(a.grid) + size_t(nchunks)) * (OMAX + 3) * sizeof(double)))
            return rc;
        a.tabs = ctx->d_adj_tabs;
        a.part = ctx->d_adj_part;
        table_blocks = (uint64_t(nchunks) * (6 * OMAX * kMaxL) + 255) / 256;
    }
    // grad_dens is written, not accumulated into: its rows are zeroed in the same launch, about
    // 64 KiB per workgroup
    const uint64_t zero_blocks =
        a.grad_dens ? (uint64_t(nchunks) * uint64_t(a.dens_stride) * sizeof(float) + 65535) / 65536 : 0;
    if (table_blocks + zero_blocks + 1 > 0x7FFFFFFFull) return uw_fail(ctx, UW_ERR_INVALID_ARG, "too many chunks for one launch");
    // workgroup 0: the piece plan; then the zeroing of grad_dens; then the tables (when parameter
    // gradients are asked for)
    hipLaunchKernelGGL(k_backward_tables<OMAX>, dim3(uint32_t(table_blocks + zero_blocks + 1)), dim3(256), 0, st, a,
                       static_cast<BackTables<OMAX>*>(ctx->d_adj_tabs), uint32_t(zero_blocks));
    hipLaunchKernelGGL((k_backward<OMAX, OEXACT>), dim3(uint32_t(a.grid)), dim3(kBlock), 0, st, a);
    if (want_params) {
        const int64_t work = int64_t(nchunks) * (OMAX + 3);
        hipLaunchKernelGGL(k_backward_sum<OMAX>, dim3(uint32_t((work + 255) / 256)), dim3(256), 0, st, a);
    }
    return UW_OK;
}

}  // namespace

extern "C" {

int uw_ctx_set_field(uw_ctx* ctx, const double* amplitudes, const double* offset) {
    if (!ctx) return UW_ERR_INVALID_ARG;
    uw_params p = ctx->params;
    if (amplitudes) {
        double s = 0.0;
        for (int k = 0; k < p.octaves; ++k) {
            if (!std::isfinite(amplitudes[k])) return uw_fail(ctx, UW_ERR_INVALID_ARG, "amplitudes must be finite");
            p.amplitudes[k] = amplitudes[k];
            s += amplitudes[k];
        }
        if (s == 0.0) return uw_fail(ctx, UW_ERR_INVALID_ARG, "amplitudes must not sum to zero");
    }
    if (offset)
        for (int i = 0; i < 3; ++i) {
            if (!std::isfinite(offset[i])) return uw_fail(ctx, UW_ERR_INVALID_ARG, "offset must be finite");
            p.offset[i] = offset[i];
        }
    ctx->params = p;
    return UW_OK;
}

int uw_adjoint_backward(uw_ctx* ctx, const int32_t* coords, const int32_t* res, int32_t nchunks,
                        const uw_chunk_info* info, const uw_vertex_tape* tape, const float* grad_verts,
                        float* grad_dens, int64_t dens_stride, double* grad_amp, double* grad_off, void* stream) {
    if (!ctx) return UW_ERR_INVALID_ARG;
    if (nchunks < 0 || (nchunks > 0 && (!coords || !res || !info || !tape || !grad_verts)))
        return uw_fail(ctx, UW_ERR_INVALID_ARG, "bad arguments");
    if (grad_dens && dens_stride < 1) return uw_fail(ctx, UW_ERR_INVALID_ARG, "dens_stride must be positive");
    if (nchunks == 0) return UW_OK;
    UW_HIP(ctx, hipSetDevice(ctx->device));
    BackArgs a;
    memset(&a, 0, sizeof(a));
    a.f = uw_make_field(ctx->params);
    a.tab = ctx->d_tables;
    a.coords = coords;
    a.res = res;
    a.info = info;
    a.tape = tape;
    a.grad_verts = grad_verts;
    a.grad_dens = grad_dens;
    a.dens_stride = dens_stride;
    a.grad_amp = grad_amp;
    a.grad_off = grad_off;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (int rc = uw_refuse_capture(ctx, st)) return rc;
    const int O = a.f.octaves;
    if (int rc = uw_prof_start(ctx, st)) return rc;
    int lrc = UW_OK;
    // exact kernels for the octave counts of the shipped presets (4, 6, 8), bounded ones otherwise
    if (O == 4) lrc = launch_backward<4, 4>(ctx, a, nchunks, st);
    else if (O == 6) lrc = launch_backward<6, 6>(ctx, a, nchunks, st);
    else if (O == 8) lrc = launch_backward<8, 8>(ctx, a, nchunks, st);
    else if (O < 4) lrc = launch_backward<4, 0>(ctx, a, nchunks, st);
    else if (O < 8) lrc = launch_backward<8, 0>(ctx, a, nchunks, st);
    else lrc = launch_backward<16, 0>(ctx, a, nchunks, st);
    if (lrc) return lrc;
    if (int rc = uw_prof_end(ctx, st)) return rc;
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}

}  // extern "C"
