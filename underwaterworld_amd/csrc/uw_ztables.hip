// Per-chunk z tables in HBM (uw_ztab): one record per z sample with the z-axis octave entries, the
// z-bias split and the face flags of every chunk of a launch, written here so that the plain
// path's density loop can read them with scalar loads (the batched path writes them in its set-up
// kernel). One thread per record.
#include <hip/hip_runtime.h>

#include "uw_noise.h"

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) k_ztables(uw_field f, const int32_t* __restrict__ coords,
                                                      const int32_t* __restrict__ res, int32_t n_uniform,
                                                      int32_t nchunks, int32_t stride, unsigned char* __restrict__ out) {
    uw_ztab_record(f, coords, res, n_uniform, nchunks, stride, out, uint64_t(blockIdx.x) * kThreads + threadIdx.x);
}

}  // namespace

int uw_launch_ztables(uw_ctx* ctx, const uw_field& f, const int32_t* coords, const int32_t* res, int32_t n_uniform,
                      int32_t nchunks, int32_t stride, unsigned char* out, hipStream_t st) {
    if (nchunks <= 0) return UW_OK;
    const uint64_t work = uint64_t(nchunks) * uint64_t(stride + 1);
    const uint64_t blocks = (work + kThreads - 1) / kThreads;
    if (blocks > 0x7FFFFFFFull) return uw_fail(ctx, UW_ERR_INVALID_ARG, "too many chunks for one launch");
    hipLaunchKernelGGL(k_ztables, dim3(uint32_t(blocks)), dim3(kThreads), 0, st, f, coords, res, n_uniform, nchunks,
                       stride, out);
    UW_HIP(ctx, hipGetLastError());
    return UW_OK;
}
