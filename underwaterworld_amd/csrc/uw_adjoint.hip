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

#include "uw_mesh.h"
#include "uw_noise.h"

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
    int32_t nsub;   // workgroups per chunk, each taking every nsub-th run of kBlock vertices
    double* part;   // nsub > 1: per (chunk, share) parameter sums, [c][sub][OMAX + 3], added by k_backward_sum
};

// Per-chunk lattice data of the backward: every (axis, octave, sample) entry of the chunk (x hashed)
// and the fade derivatives s'(d) = 30 d^2 (d - 1)^2, each at [k L + i] for the chunk's own L. Built once per launch by k_backward_tables (one thread per entry, so the f64 chains
// of all chunks run side by side) and copied into LDS by each of the chunk's workgroups.
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

template <int OMAX>
__global__ void __launch_bounds__(256) k_backward_tables(BackArgs a, BackTables<OMAX>* __restrict__ out) {
    constexpr int kAx = 3 * OMAX * kMaxL;  // axis entries: [axis][octave][sample]
    constexpr int kE = 2 * kAx;            // + fade derivatives
    const int64_t g = int64_t(blockIdx.x) * 256 + threadIdx.x;
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

template <int OMAX>
__global__ void __launch_bounds__(kBlock) k_backward(BackArgs a) {
    __shared__ float G3[3 * 256];
    __shared__ uint8_t P[256];
    __shared__ __align__(16) BackTables<OMAX> TB;
    __shared__ double red[kBlock / 64][OMAX + 3];
    uw_axis* AX = TB.AX;
    uw_axis* AY = TB.AY;
    uw_axis* AZ = TB.AZ;
    float(*DF)[OMAX * kMaxL] = TB.DF;

    const int tid = threadIdx.x;
    const int32_t c = int32_t(blockIdx.x / uint32_t(a.nsub)), sub = int32_t(blockIdx.x % uint32_t(a.nsub));
    const int32_t n = a.res[c];
    const int32_t O = a.f.octaves;
    const uw_chunk_info ci = a.info[c];
    const bool skip = n < 1 || n > UW_BATCH_MAX_RESOLUTION || (ci.flags & (UW_CHUNK_OVERFLOW | UW_CHUNK_INVALID)) ||
                      ci.n_verts == 0;
    const bool want_params = a.grad_amp != nullptr || a.grad_off != nullptr;
    if (skip) {
        if (a.nsub > 1) {
            if (tid < OMAX + 3 && want_params) a.part[(int64_t(c) * a.nsub + sub) * (OMAX + 3) + tid] = 0.0;
        } else {
            if (tid < O && a.grad_amp) a.grad_amp[int64_t(c) * O + tid] = 0.0;
            if (tid < 3 && a.grad_off) a.grad_off[int64_t(c) * 3 + tid] = 0.0;
        }
        return;
    }
    const int32_t L = n + 1;
    const float step = float(a.f.extent) / float(n);
    if (want_params) {  // the chunk's tables: 16-byte loads, all of a thread's issued before its stores
        static_assert(sizeof(BackTables<OMAX>) % 16 == 0, "tables are copied in 16-byte words");
        constexpr int kWords = int(sizeof(BackTables<OMAX>) / 16);
        const uint4* src = reinterpret_cast<const uint4*>(static_cast<const BackTables<OMAX>*>(a.tabs) + c);
        uint4* dst = reinterpret_cast<uint4*>(&TB);
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
    }
    for (int i = tid; i < 256; i += kBlock) {
        const uint8_t p = a.tab->perm[i];
        P[i] = p;
        const float4 g = uw_grad_vec(p & 15u);
        G3[3 * i] = g.x;
        G3[3 * i + 1] = g.y;
        G3[3 * i + 2] = g.z;
    }
    __syncthreads();

    const float iso = a.f.iso;
    const double inv_sum = 1.0 / double(a.f.amp_sum);
    const double slope = a.f.bias_kind == UW_BIAS_LINEAR ? double(a.f.bias_a) : 0.0;
    double acc[OMAX + 3];
#pragma unroll
    for (int k = 0; k < OMAX + 3; ++k) acc[k] = 0.0;
    float* gd = a.grad_dens ? a.grad_dens + int64_t(c) * a.dens_stride : nullptr;
    const int32_t LL = L * L;
    for (uint32_t i = uint32_t(sub) * kBlock + tid; i < ci.n_verts; i += uint32_t(a.nsub) * kBlock) {
        const uint64_t v = ci.vert_off + i;
        const uw_vertex_tape tv = a.tape[v];
        const int32_t diff = int32_t(tv.b) - int32_t(tv.a);
        const int axis = (diff == 1 || diff == -1) ? 2 : ((diff == L || diff == -L) ? 1 : 0);
        const float dir = diff > 0 ? step : -step;
        const float gv = a.grad_verts[3 * v + axis] * dir;
        const float D = tv.rho_b - tv.rho_a;
        const float inv_d2 = 1.0f / (D * D);
        const float w[2] = {gv * (iso - tv.rho_b) * inv_d2, gv * (tv.rho_a - iso) * inv_d2};
        const uint32_t s[2] = {tv.a, tv.b};
        if (gd) {
            unsafeAtomicAdd(gd + s[0], w[0]);
            unsafeAtomicAdd(gd + s[1], w[1]);
        }
        if (!want_params) continue;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int32_t x = int32_t(s[e]) / LL, y = (int32_t(s[e]) / L) % L, z = int32_t(s[e]) % L;
            float nk[OMAX];
            float total = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
            for (int k = 0; k < OMAX; ++k) {
                if (k < O) {
                    float g[3];
                    nk[k] = uw_octave_grad(AX[k * L + x], AY[k * L + y], AZ[k * L + z], DF[0][k * L + x],
                                           DF[1][k * L + y], DF[2][k * L + z], P, G3, g);
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
    // block reduction of the per-thread accumulators
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
        for (int w = 0; w < kBlock / 64; ++w) t += red[w][tid];
        if (a.nsub > 1) {
            a.part[(int64_t(c) * a.nsub + sub) * (OMAX + 3) + tid] = t;
        } else if (tid < OMAX) {
            if (tid < O && a.grad_amp) a.grad_amp[int64_t(c) * O + tid] = t;
        } else if (a.grad_off) {
            a.grad_off[int64_t(c) * 3 + (tid - OMAX)] = t;
        }
    }
}

// Parameter sums of the nsub workgroups of each chunk, in workgroup order: one thread per (chunk,
// parameter).
template <int OMAX>
__global__ void __launch_bounds__(256) k_backward_sum(const double* __restrict__ part, int32_t nchunks, int32_t nsub,
                                                     int32_t O, double* grad_amp, double* grad_off) {
    const int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (e >= int64_t(nchunks) * (OMAX + 3)) return;
    const int64_t c = e / (OMAX + 3);
    const int k = int(e - c * (OMAX + 3));
    double t = 0.0;
    for (int32_t s = 0; s < nsub; ++s) t += part[(c * nsub + s) * (OMAX + 3) + k];
    if (k < OMAX) {
        if (k < O && grad_amp) grad_amp[c * O + k] = t;
    } else if (grad_off) {
        grad_off[c * 3 + (k - OMAX)] = t;
    }
}

template <int OMAX>
int launch_backward(uw_ctx* ctx, BackArgs a, int32_t nchunks, hipStream_t st) {
    // several workgroups per chunk when the launch has fewer chunks than ~4 per CU: each takes a
    // share of the chunk's vertices, and the shares' parameter sums are added afterwards
    const bool want_params = a.grad_amp != nullptr || a.grad_off != nullptr;
    int32_t nsub = 1;
    while (nsub < 4 && int64_t(nchunks) * nsub * 2 <= int64_t(ctx->num_cus) * 4) nsub *= 2;
    a.nsub = nsub;
    a.part = nullptr;
    if (nsub > 1 && want_params) {
        const size_t bytes = size_t(nchunks) * size_t(nsub) * (OMAX + 3) * sizeof(double);
        if (bytes > ctx->adj_part_bytes) {
            (void)hipFree(ctx->d_adj_part);
            ctx->d_adj_part = nullptr;
            ctx->adj_part_bytes = 0;
            UW_HIP(ctx, hipMalloc(&ctx->d_adj_part, bytes));
            ctx->adj_part_bytes = bytes;
        }
        a.part = ctx->d_adj_part;
    }
    const uint64_t blocks = uint64_t(nchunks) * uint64_t(nsub);
    if (blocks > 0x7FFFFFFFull) return uw_fail(ctx, UW_ERR_INVALID_ARG, "too many chunks for one launch");
    a.nchunks = nchunks;
    a.tabs = nullptr;
    if (want_params) {
        const size_t bytes = size_t(nchunks) * sizeof(BackTables<OMAX>);
        if (bytes > ctx->adj_tabs_bytes) {
            (void)hipFree(ctx->d_adj_tabs);
            ctx->d_adj_tabs = nullptr;
            ctx->adj_tabs_bytes = 0;
            UW_HIP(ctx, hipMalloc(&ctx->d_adj_tabs, bytes));
            ctx->adj_tabs_bytes = bytes;
        }
        a.tabs = ctx->d_adj_tabs;
        const uint64_t entries = uint64_t(nchunks) * (6 * OMAX * kMaxL);
        hipLaunchKernelGGL(k_backward_tables<OMAX>, dim3(uint32_t((entries + 255) / 256)), dim3(256), 0, st, a,
                           static_cast<BackTables<OMAX>*>(ctx->d_adj_tabs));
    }
    hipLaunchKernelGGL(k_backward<OMAX>, dim3(uint32_t(blocks)), dim3(kBlock), 0, st, a);
    if (nsub > 1 && want_params) {
        const int64_t work = int64_t(nchunks) * (OMAX + 3);
        hipLaunchKernelGGL(k_backward_sum<OMAX>, dim3(uint32_t((work + 255) / 256)), dim3(256), 0, st, a.part, nchunks,
                           nsub, a.f.octaves, a.grad_amp, a.grad_off);
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
    const int64_t Lmax = UW_BATCH_MAX_RESOLUTION + 1;
    if (grad_dens && dens_stride < 1) return uw_fail(ctx, UW_ERR_INVALID_ARG, "dens_stride must be positive");
    if (nchunks == 0) return UW_OK;
    UW_HIP(ctx, hipSetDevice(ctx->device));
    (void)Lmax;
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
    if (O <= 4) lrc = launch_backward<4>(ctx, a, nchunks, st);
    else if (O <= 8) lrc = launch_backward<8>(ctx, a, nchunks, st);
    else lrc = launch_backward<16>(ctx, a, nchunks, st);
    if (lrc) return lrc;
    if (int rc = uw_prof_end(ctx, st)) return rc;
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}

}  // extern "C"
