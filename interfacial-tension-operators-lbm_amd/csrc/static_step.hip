// Device memory of the allocation-free step entry point (lbm_step_case_noalloc, capi.hip).
//
// lbm_step_case keeps each case's launch table (one CaseDev entry) and its tile -> case map in device
// memory that the library allocates on the first call.  The allocation-free variant keeps them in the
// file-scope array below instead: 256 bytes for the table entry, then one int per tile, for at most
// SCRATCH_MAX_TILES tiles (4 MiB + 256 B in all).  Every GPU the library runs on holds its own copy of
// the array until the process ends.  A small kernel, whose arguments carry the table entry by value,
// writes both before the step launches, so the variant copies nothing from host memory and can be
// recorded into a caller's graph.
//
// The array's name is a host-side stand-in, not a device address: host code uses only the address
// hipGetSymbolAddress returns for the current device (scratch_setup), kept per device.  Calls that
// overlap on one GPU (different streams, or different host threads) share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "step_kernels.h"

static __device__ __align__(256) unsigned char scratch[itolbm::SCRATCH_BYTES];

namespace itolbm {

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[MAX_DEVICES] = {};

// tile -> case map of one case (every tile is case 0) and its table entry
__global__ void __launch_bounds__(256) fill_static_table(CaseDev d, CaseDev* __restrict__ tab, int* __restrict__ tiles,
                                                         int ntiles)
{
    const int t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 == 0) *tab = d;
    for (int t = t0; t < ntiles; t += gridDim.x * blockDim.x) tiles[t] = 0;
}

}  // namespace

static_assert(sizeof(CaseDev) <= SCRATCH_TABLE_BYTES, "the table entry fits its slot of the scratch array");

hipError_t scratch_setup(int device, void** out)
{
    *out = nullptr;
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        scratch_ptr[device] = p;
    }
    *out = scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_fill_static_table(void* base, const CaseDev& d, int ntiles, hipStream_t stream)
{
    if (!base || ntiles < 1 || ntiles > SCRATCH_MAX_TILES) return hipErrorInvalidValue;
    const int nb = (ntiles + 255) / 256 < 1024 ? (ntiles + 255) / 256 : 1024;
    hipLaunchKernelGGL(fill_static_table, dim3(nb), dim3(256), 0, stream, d, scratch_table(base),
                       scratch_tiles(base), ntiles);
    return hipGetLastError();
}

}  // namespace itolbm
