// Stream copy: the achievable-bandwidth reference for the step kernels (which read and
// write the state once, like a copy) and a kernel of known byte count for scaling the
// HBM byte counters.  A read-only and a write-only stream bound the two directions on
// their own.
//
// The copy is the fastest read-once / write-once order measured on MI355X: each
// 256-thread block copies one contiguous 4 KiB chunk (one 16-byte load and one store per
// thread) and exits, and the blocks are renumbered so that each XCD walks one contiguous
// eighth of the buffer.  On 12.8 GB buffers it moves 6.2-6.4 TB/s (read + write), against
// 5.7-6.1 with 8 KiB chunks, 5.5-5.8 with 16 KiB and 5.2-5.4 TB/s with 32-256 KiB chunks
// streamed by long-lived blocks, and 6.2 TB/s with 4 KiB chunks in plain block order.
#include <hip/hip_runtime.h>

#include "../../include/hydro_hip.h"
#include "hydro_runtime.hpp"

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));

constexpr int kThreads = 256;
constexpr int kUnroll = 4;
constexpr long long kChunk = 16384;        // 16-byte elements per block of the read and fill streams (256 KiB)
constexpr long long kCopyChunk = kThreads;  // 16-byte elements per block of the copies (4 KiB)

// Blocks are dealt round-robin over the 8 XCDs: give XCD x the x-th contiguous eighth.
__device__ __forceinline__ unsigned long long xcd_block(unsigned b, unsigned n)
{
    const unsigned q = n / 8u, r = n % 8u, x = b % 8u, k = b / 8u;
    return static_cast<unsigned long long>(x) * q + (x < r ? x : r) + k;
}

__global__ void __launch_bounds__(kThreads) stream_copy_kernel(const v2d* __restrict__ src, v2d* __restrict__ dst,
                                                               long long n16)
{
    const long long j = static_cast<long long>(xcd_block(blockIdx.x, gridDim.x)) * kCopyChunk + threadIdx.x;
    if (j < n16) dst[j] = src[j];
}

// The same copy with non-temporal loads and stores (the data is not kept in the caches).
__global__ void __launch_bounds__(kThreads) stream_copy_nt_kernel(const v2d* __restrict__ src, v2d* __restrict__ dst,
                                                                  long long n16)
{
    const long long j = static_cast<long long>(xcd_block(blockIdx.x, gridDim.x)) * kCopyChunk + threadIdx.x;
    if (j < n16) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[j]), &dst[j]);
}

// Read stream: every 64-bit word of the chunk XOR-folded, per block, into one of 256
// slots.  Non-temporal loads and eight 16-byte loads per thread in flight: on MI355X this
// reads at 6.5-7.2 TB/s, against about 6.4 TB/s with ordinary loads.
constexpr int kReadUnroll = 8;
constexpr int kXorSlots = 256;

__global__ void __launch_bounds__(kThreads) stream_read_kernel(const v2d* __restrict__ src, long long n16,
                                                               unsigned long long* __restrict__ slots)
{
    const long long beg = static_cast<long long>(blockIdx.x) * kChunk;
    const long long end = beg + kChunk < n16 ? beg + kChunk : n16;
    unsigned long long x = 0;
    for (long long i = beg + threadIdx.x; i < end; i += kReadUnroll * kThreads) {
        v2d v[kReadUnroll];
#pragma unroll
        for (int u = 0; u < kReadUnroll; ++u) {
            const long long j = i + u * kThreads;
            v[u] = 0.0;
            if (j < end) v[u] = __builtin_nontemporal_load(&src[j]);
        }
#pragma unroll
        for (int u = 0; u < kReadUnroll; ++u)
            x ^= static_cast<unsigned long long>(__double_as_longlong(v[u].x)) ^
                 static_cast<unsigned long long>(__double_as_longlong(v[u].y));
    }
    for (int off = 32; off > 0; off >>= 1) x ^= __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0 && x != 0) atomicXor(&slots[blockIdx.x % kXorSlots], x);
}

// Write stream: a 64-bit pattern to every word of the chunk, four 16-byte stores per
// thread in flight (about 6.1 TB/s on MI355X).
__global__ void __launch_bounds__(kThreads) stream_fill_kernel(v2d* __restrict__ dst, long long n16,
                                                               unsigned long long pattern)
{
    const long long beg = static_cast<long long>(blockIdx.x) * kChunk;
    const long long end = beg + kChunk < n16 ? beg + kChunk : n16;
    v2d v;
    v.x = __longlong_as_double(static_cast<long long>(pattern));
    v.y = v.x;
    for (long long i = beg + threadIdx.x; i < end; i += kUnroll * kThreads) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const long long j = i + u * kThreads;
            if (j < end) dst[j] = v;
        }
    }
}

long long stream_blocks(const char* fn, const void* p, size_t bytes, long long chunk = kChunk)
{
    if (!p || (bytes % 16) != 0) {
        hydro::set_error("%s: null pointer or size not a multiple of 16 bytes", fn);
        return -1;
    }
    const long long blocks = (static_cast<long long>(bytes / 16) + chunk - 1) / chunk;
    if (blocks > 0x7fffffffLL) {
        hydro::set_error("%s: buffer too large", fn);
        return -1;
    }
    return blocks;
}

}  // namespace

extern "C" int hydro_stream_read(const void* src, size_t bytes, unsigned long long* d_xor256, void* stream)
{
    hydro::clear_error();
    const long long blocks = stream_blocks("hydro_stream_read", src, bytes);
    if (blocks < 0) return HYDRO_ERR_ARG;
    if (!d_xor256) {
        hydro::set_error("hydro_stream_read: null slot buffer");
        return HYDRO_ERR_ARG;
    }
    if (blocks == 0) return 0;
    hipLaunchKernelGGL(stream_read_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0,
                       static_cast<hipStream_t>(stream), static_cast<const v2d*>(src),
                       static_cast<long long>(bytes / 16), d_xor256);
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    return 0;
}

extern "C" int hydro_stream_fill(void* dst, size_t bytes, unsigned long long pattern, void* stream)
{
    hydro::clear_error();
    const long long blocks = stream_blocks("hydro_stream_fill", dst, bytes);
    if (blocks < 0) return HYDRO_ERR_ARG;
    if (blocks == 0) return 0;
    hipLaunchKernelGGL(stream_fill_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0,
                       static_cast<hipStream_t>(stream), static_cast<v2d*>(dst), static_cast<long long>(bytes / 16),
                       pattern);
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    return 0;
}

namespace {

int launch_copy(const char* fn, bool nontemporal, void* dst, const void* src, size_t bytes, void* stream)
{
    hydro::clear_error();
    if (!dst) {
        hydro::set_error("%s: null destination", fn);
        return HYDRO_ERR_ARG;
    }
    const long long blocks = stream_blocks(fn, src, bytes, kCopyChunk);
    if (blocks < 0) return HYDRO_ERR_ARG;
    if (blocks == 0) return 0;
    const long long n16 = static_cast<long long>(bytes / 16);
    hipLaunchKernelGGL(nontemporal ? stream_copy_nt_kernel : stream_copy_kernel, dim3(static_cast<unsigned>(blocks)),
                       dim3(kThreads), 0, static_cast<hipStream_t>(stream), static_cast<const v2d*>(src),
                       static_cast<v2d*>(dst), n16);
    HYDRO_HIP_OR(hipGetLastError(), HYDRO_ERR_HIP);
    return 0;
}

}  // namespace

extern "C" int hydro_stream_copy(void* dst, const void* src, size_t bytes, void* stream)
{
    return launch_copy("hydro_stream_copy", false, dst, src, bytes, stream);
}

extern "C" int hydro_stream_copy_nt(void* dst, const void* src, size_t bytes, void* stream)
{
    return launch_copy("hydro_stream_copy_nt", true, dst, src, bytes, stream);
}
