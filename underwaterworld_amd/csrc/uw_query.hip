// Batched density point queries: perlin_util::iso_at at arbitrary noise-space points
// (underwater_world/src/perlin_util.rs:24-29), as the boids call it one point at a time
// (boid.rs:131-143, 324-331). One thread per point; every octave's lattice cell, fraction and
// fade are formed in f64 from the point (as the crate does) and the corner combination runs in f32.
#include <hip/hip_runtime.h>
#include <string.h>

#include "uw_noise.h"

namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) k_density_at(uw_field f, const uw_tables* __restrict__ tab,
                                                       const double* __restrict__ xyz, int64_t npts,
                                                       float* __restrict__ out) {
    __shared__ float G3[3 * 256];
    __shared__ uint8_t P[256];
    for (int i = threadIdx.x; i < 256; i += kBlock) {
        const uint8_t p = tab->perm[i];
        P[i] = p;
        const float4 g = uw_grad_vec(p & 15u);
        G3[3 * i] = g.x;
        G3[3 * i + 1] = g.y;
        G3[3 * i + 2] = g.z;
    }
    __syncthreads();
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < npts; i += int64_t(gridDim.x) * kBlock) {
        // the world-space offset (0 for the reference) shifts the queried point like every sample
        const double px = xyz[3 * i] + f.offset[0] * f.frequency;
        const double py = xyz[3 * i + 1] + f.offset[1] * f.frequency;
        const double pz = xyz[3 * i + 2] + f.offset[2] * f.frequency;
        float total = 0.0f;
        for (int32_t k = 0; k < f.octaves; ++k) {
            const double fk = f.octave_freq[k];
            const uw_axis X = uw_hash_x(uw_axis_of(px, fk), P);
            const uw_axis Y = uw_axis_of(py, fk);
            const uw_axis Z = uw_axis_of(pz, fk);
            total = fmaf(f.amp[k], uw_perlin_octave(X, Y, Z, P, G3), total);
        }
        float adj, post;
        uw_bias_terms(f, pz, &adj, &post);
        out[i] = uw_density_from(total, f.amp_sum, adj, post);
    }
}

}  // namespace

extern "C" int uw_density_at(uw_ctx* ctx, const double* xyz, int64_t npts, float* out, void* stream) {
    if (!ctx) return UW_ERR_INVALID_ARG;
    if (npts < 0 || (npts > 0 && (!xyz || !out))) return uw_fail(ctx, UW_ERR_INVALID_ARG, "bad arguments");
    if (npts == 0) return UW_OK;
    UW_HIP(ctx, hipSetDevice(ctx->device));
    const uw_field f = uw_make_field(ctx->params);
    int64_t blocks = (npts + kBlock - 1) / kBlock;
    const int64_t cap = int64_t(ctx->num_cus) * 16;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_density_at, dim3(uint32_t(blocks)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), f,
                       ctx->d_tables, xyz, npts, out);
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}
