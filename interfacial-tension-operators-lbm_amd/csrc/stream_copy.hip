// A plain streaming copy (lbm_stream_copy): the bandwidth a read-once / write-once kernel reaches on
// this GPU, the attainable ceiling the step kernels' HBM rates are compared with.  No reference
// counterpart.
//
// Each lane moves one 16-byte chunk (a dwordx4 load, a nontemporal dwordx4 store) and each workgroup
// a contiguous 4 KiB block, so the workgroups in flight work on one contiguous window of memory that
// slides through the buffer in dispatch order.  Measured on MI355X, 1 GiB: 6.35 TB/s read + write,
// against 4.1-4.9 TB/s for grid-stride loops (whose waves in flight are spread over the whole buffer)
// and 5.0 TB/s for the runtime's device-to-device copy.  The few bytes past the last whole chunk are
// copied one by one.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "step_kernels.h"

namespace itolbm {

namespace {

constexpr int COPY_THREADS = 256;

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(COPY_THREADS) stream_copy16(u32x4* __restrict__ dst, const u32x4* __restrict__ src,
                                                               long long n16)
{
    const long long i = (long long)blockIdx.x * COPY_THREADS + threadIdx.x;
    if (i < n16) __builtin_nontemporal_store(src[i], &dst[i]);
}

__global__ void copy_tail(unsigned char* __restrict__ dst, const unsigned char* __restrict__ src, int n)
{
    if ((int)threadIdx.x < n) dst[threadIdx.x] = src[threadIdx.x];
}

}  // namespace

hipError_t launch_stream_copy(void* dst, const void* src, size_t bytes, hipStream_t stream)
{
    if (((uintptr_t)dst | (uintptr_t)src) & 15u) return hipErrorInvalidValue;
    const long long n16 = (long long)(bytes / 16);
    const long long nb = (n16 + COPY_THREADS - 1) / COPY_THREADS;
    if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
    if (nb > 0)
        hipLaunchKernelGGL(stream_copy16, dim3((unsigned)nb), dim3(COPY_THREADS), 0, stream, (u32x4*)dst,
                           (const u32x4*)src, n16);
    const int tail = (int)(bytes - 16 * (size_t)n16);
    if (tail > 0)
        hipLaunchKernelGGL(copy_tail, dim3(1), dim3(64), 0, stream, (unsigned char*)dst + 16 * n16,
                           (const unsigned char*)src + 16 * n16, tail);
    return hipGetLastError();
}

}  // namespace itolbm
