// Per-chunk z tables in HBM (uw_ztab): the z-axis octave entries and the z-bias split of every
// chunk of a launch, written here so that the density loops of the plain and batched kernels can
// read them with scalar loads. One workgroup per chunk.
#include <hip/hip_runtime.h>

#include "uw_noise.h"

namespace {

constexpr int kThreads = 128;

__global__ void __launch_bounds__(kThreads) k_ztables(uw_field f, const int32_t* __restrict__ coords,
                                                      const int32_t* __restrict__ res, int32_t n_uniform,
                                                      int32_t nchunks, int32_t stride, unsigned char* __restrict__ out) {
    const int32_t c = blockIdx.x;
    if (c >= nchunks) return;
    const int32_t n = res ? res[c] : n_uniform;
    if (n < 1 || n + 1 > stride) return;  // flagged invalid by the mesher, which never reads it
    const int32_t O = f.octaves;
    unsigned char* base = out + size_t(c) * uw_ztab_bytes(O, stride);
    uw_axis* az = reinterpret_cast<uw_axis*>(base);
    float* adj = reinterpret_cast<float*>(az + size_t(O) * stride);
    uw_fill_ztab(f, coords[3 * c + 2], n + 1, float(f.extent) / float(n), az, adj, adj + stride, stride, threadIdx.x,
                 kThreads);
}

}  // namespace

int uw_launch_ztables(uw_ctx* ctx, const uw_field& f, const int32_t* coords, const int32_t* res, int32_t n_uniform,
                      int32_t nchunks, int32_t stride, unsigned char* out, hipStream_t st) {
    if (nchunks <= 0) return UW_OK;
    hipLaunchKernelGGL(k_ztables, dim3(uint32_t(nchunks)), dim3(kThreads), 0, st, f, coords, res, n_uniform, nchunks,
                       stride, out);
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}
