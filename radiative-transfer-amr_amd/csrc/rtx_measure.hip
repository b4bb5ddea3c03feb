// Measurement helper: a plain FP64 copy for calibrating HBM traffic counters.
//
// Each lane moves one double per iteration with ordinary global loads and
// stores, the access type of the operator kernels' vector loads, so the
// FETCH_SIZE / WRITE_SIZE readings of a copy of known size give the factor
// between those counters and true bytes for them.
#include "rtx_internal.h"

namespace rtx {

__global__ __launch_bounds__(256) void k_copy_f64(const double *__restrict__ src,
                                                  double *__restrict__ dst, int64_t n) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

}  // namespace rtx

using namespace rtx;

extern "C" int rtx_copy_f64(const double *d_src, double *d_dst, int64_t n, void *stream) {
    RTX_CHECK(d_src && d_dst && n >= 0, "rtx_copy_f64: bad argument");
    if (n == 0) return RTX_OK;
    const int64_t want = (n + 255) / 256;
    const int blocks = int(want < 8192 ? want : 8192);
    k_copy_f64<<<blocks, 256, 0, static_cast<hipStream_t>(stream)>>>(d_src, d_dst, n);
    RTX_HIP(hipGetLastError());
    return RTX_OK;
}
