// Stream copy: the achievable-bandwidth reference for the step kernels (which read and
// write the state once, like a copy) and a kernel of known byte count for scaling the
// HBM byte counters.  16-byte vector accesses, four per thread per iteration in flight,
// grid-stride over the buffer.
#include <hip/hip_runtime.h>

#include "../../include/hydro_hip.h"
#include "hydro_runtime.hpp"

namespace {

constexpr int kUnroll = 4;

__global__ void __launch_bounds__(256) stream_copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                          long long n16)
{
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i + (kUnroll - 1) * stride < n16; i += kUnroll * stride) {
        double2 v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) dst[i + u * stride] = v[u];
    }
    for (; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" int hydro_stream_copy(void* dst, const void* src, size_t bytes, void* stream)
{
    hydro::clear_error();
    if (!dst || !src || (bytes % 16) != 0) {
        hydro::set_error("hydro_stream_copy: null pointer or size not a multiple of 16 bytes");
        return HYDRO_ERR_ARG;
    }
    const long long n16 = static_cast<long long>(bytes / 16);
    int dev = hydro::current_device();
    if (dev < 0) return HYDRO_ERR_HIP;
    int cus = 0;
    HYDRO_HIP_OR(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), HYDRO_ERR_HIP);
    const long long want = (n16 + 255) / 256;
    const unsigned blocks = static_cast<unsigned>(want < 16LL * cus ? (want > 0 ? want : 1) : 16LL * cus);
    hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const double2*>(src), static_cast<double2*>(dst), n16);
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    return 0;
}
