// A plain streaming copy (lbm_stream_copy): the bandwidth a read-once / write-once kernel reaches on
// this GPU, the attainable ceiling the step kernels' HBM rates are compared with.  No reference
// counterpart.
//
// Each lane moves 16-byte chunks: UNROLL global loads (dwordx4) are issued before the first store, so
// every wavefront keeps UNROLL x 1 KiB of reads in flight, and the stores are nontemporal (the copy
// reads nothing it writes).  Grid-stride over the chunks; the few bytes past the last whole chunk are
// copied one by one by the first lanes.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "step_kernels.h"

namespace itolbm {

namespace {

constexpr int COPY_THREADS = 256;
constexpr int UNROLL = 4;

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(COPY_THREADS) stream_copy16(u32x4* __restrict__ dst, const u32x4* __restrict__ src,
                                                               long long n16)
{
    const long long stride = (long long)gridDim.x * COPY_THREADS;
    long long i = (long long)blockIdx.x * COPY_THREADS + threadIdx.x;
    for (; i + (UNROLL - 1) * stride < n16; i += UNROLL * stride) {
        u32x4 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) __builtin_nontemporal_store(v[u], &dst[i + u * stride]);
    }
    for (; i < n16; i += stride) __builtin_nontemporal_store(src[i], &dst[i]);
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
    if (n16 > 0) {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        // 8 workgroups of 256 lanes per CU in flight, capped by the work
        long long nb = (n16 + COPY_THREADS - 1) / COPY_THREADS;
        if (nb > 8LL * cus) nb = 8LL * cus;
        hipLaunchKernelGGL(stream_copy16, dim3((unsigned)nb), dim3(COPY_THREADS), 0, stream, (u32x4*)dst,
                           (const u32x4*)src, n16);
    }
    const int tail = (int)(bytes - 16 * (size_t)n16);
    if (tail > 0)
        hipLaunchKernelGGL(copy_tail, dim3(1), dim3(64), 0, stream, (unsigned char*)dst + 16 * n16,
                           (const unsigned char*)src + 16 * n16, tail);
    return hipGetLastError();
}

}  // namespace itolbm
