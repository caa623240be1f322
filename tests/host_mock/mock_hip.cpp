// Host-only stand-in for the HIP runtime and the kernel launchers, so the C ABI's host logic
// (csrc/capi.hip compiled with --cuda-host-only) runs on a machine without a GPU.
//
// Test infrastructure: built only by tests/host_mock/Makefile into libitolbm_mock.so, loaded only by
// tests/test_host_logic.py.  Device memory is host memory, kernel launches do nothing and are counted,
// and the stream capture state is set by the test (mock_set_capturing), so the tests can check which
// entry points allocate, free, copy or synchronise while a caller's graph capture is active.
#include <hip/hip_runtime_api.h>

#include <cstdlib>
#include <cstring>

#include "step_kernels.h"
#include "diag_kernels.h"

namespace {
struct Counts {
    long long malloc_calls, free_calls, memcpy_calls, memset_calls, sync_calls, launches, captures, graph_launches;
};
Counts counts{};
int capture_forced = 0;     // the test's "a caller is capturing this stream"
int capture_internal = 0;   // between the library's own BeginCapture / EndCapture
int dummy_handle;
}  // namespace

extern "C" {

void mock_reset_counts() { counts = Counts{}; }
void mock_set_capturing(int on) { capture_forced = on; }
void mock_get_counts(long long* out)
{
    out[0] = counts.malloc_calls; out[1] = counts.free_calls; out[2] = counts.memcpy_calls;
    out[3] = counts.memset_calls; out[4] = counts.sync_calls; out[5] = counts.launches;
    out[6] = counts.captures; out[7] = counts.graph_launches;
}

hipError_t hipMalloc(void** p, size_t n)
{
    counts.malloc_calls++;
    *p = std::calloc(n ? n : 1, 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p)
{
    counts.free_calls++;
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind)
{
    counts.memcpy_calls++;
    std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t)
{
    counts.memcpy_calls++;
    std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t)
{
    counts.memset_calls++;
    std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t)
{
    counts.sync_calls++;
    return (capture_forced || capture_internal) ? hipErrorStreamCaptureUnsupported : hipSuccess;
}
hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus* st)
{
    *st = (capture_forced || capture_internal) ? hipStreamCaptureStatusActive : hipStreamCaptureStatusNone;
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int* n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDevice(int* d)
{
    *d = 0;
    return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "mock error"; }
hipError_t hipStreamCreate(hipStream_t* s)
{
    *s = (hipStream_t)&dummy_handle;
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int)
{
    *s = (hipStream_t)&dummy_handle;
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned int) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = (hipEvent_t)&dummy_handle;
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned)
{
    *e = (hipEvent_t)&dummy_handle;
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t)
{
    counts.sync_calls++;
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t)
{
    *ms = 1.0f;
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode)
{
    if (capture_forced || capture_internal) return hipErrorIllegalState;
    counts.captures++;
    capture_internal = 1;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g)
{
    capture_internal = 0;
    *g = (hipGraph_t)&dummy_handle;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t, hipGraphNode_t*, char*, size_t)
{
    *e = (hipGraphExec_t)&dummy_handle;
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t)
{
    counts.graph_launches++;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }

}  // extern "C"

// ---------------------------------------------------------------- kernel launchers (no-ops)
namespace itolbm {

#define MOCK_LAUNCH                                                                                \
    {                                                                                              \
        counts.launches++;                                                                         \
        return hipSuccess;                                                                         \
    }

hipError_t launch_step_f64(int, const StepLaunch&, int, hipStream_t) MOCK_LAUNCH
hipError_t launch_step_geo(int, const GeoLaunch&, hipStream_t) MOCK_LAUNCH
hipError_t launch_step_f64_fused(const StepLaunch&, int, hipStream_t) MOCK_LAUNCH
int step_f64_fused_capacity(bool) { return 2048; }
hipError_t launch_step_lp(int, bool, int, const StepLaunch&, int, hipStream_t) MOCK_LAUNCH
hipError_t launch_f64_rows_to_soa(const double*, bool, double, int, int, long long, long long, double*, hipStream_t) MOCK_LAUNCH
hipError_t launch_f64_convert(const double*, double*, bool, double, int, int, long long, hipStream_t) MOCK_LAUNCH
hipError_t launch_lp_to_f64(int, bool, const void*, int, int, long long, long long, double, double, double*, double*,
                            hipStream_t) MOCK_LAUNCH
hipError_t launch_lp_init_drop(int, bool, void*, int, int, int, double, double, int, int, double, double, double,
                               hipStream_t) MOCK_LAUNCH
hipError_t launch_lp_recentre(int, bool, void*, long long, int, double, double, double, double, hipStream_t) MOCK_LAUNCH
hipError_t launch_drop_pass1(const DropPass1Args&, int, hipStream_t) MOCK_LAUNCH
hipError_t launch_drop_pass2(const DropPass2Args&, int, hipStream_t) MOCK_LAUNCH
hipError_t launch_drop_ordered(const OrderedArgs&, hipStream_t) MOCK_LAUNCH
hipError_t launch_drop_modes(const ModesArgs&, hipStream_t) MOCK_LAUNCH
hipError_t launch_incl_pass(const InclArgs&, int, hipStream_t) MOCK_LAUNCH
hipError_t launch_fill_static_table(void*, const CaseDev&, int, hipStream_t) MOCK_LAUNCH
hipError_t launch_stream_copy(void*, const void*, size_t, hipStream_t) MOCK_LAUNCH

// the static scratch array of one device, as host memory
hipError_t scratch_setup(int, void** out)
{
    static unsigned char mem[SCRATCH_BYTES];
    *out = mem;
    return hipSuccess;
}

}  // namespace itolbm
