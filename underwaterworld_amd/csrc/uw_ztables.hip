// Per-chunk z tables in HBM (uw_ztab): one record per z sample with the z-axis octave entries, the
// z-bias split and the face flags of every chunk of a launch, written here so that the density
// loops of the plain and batched kernels can read them with scalar loads. One thread per record.
#include <hip/hip_runtime.h>

#include "uw_noise.h"

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) k_ztables(uw_field f, const int32_t* __restrict__ coords,
                                                      const int32_t* __restrict__ res, int32_t n_uniform,
                                                      int32_t nchunks, int32_t stride, unsigned char* __restrict__ out) {
    const uint64_t g = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
    const uint32_t per = uint32_t(stride) + 1u;  // records per chunk
    if (g >= uint64_t(nchunks) * per) return;
    const int32_t c = int32_t(g / per), i = int32_t(g - uint64_t(c) * per);
    const int32_t n = res ? res[c] : n_uniform;
    if (n < 1 || n + 1 > stride) return;  // flagged invalid by the mesher, which never reads it
    if (i > n + 1) return;                // past the spare record: never read
    uint32_t* rec = reinterpret_cast<uint32_t*>(out + size_t(c) * uw_ztab_bytes(f.octaves, stride)) +
                    size_t(i) * size_t(uw_zrec_words(f.octaves));
    uw_fill_zrec(f, coords[3 * c + 2], n + 1, float(f.extent) / float(n), i, rec);
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
