// hvc_noalloc.hip -- the fixed per-device buffer of hvc_trace_ra_noalloc.
//
// hvc_trace_ra builds its trace tables (the step clock and the per-step decay
// rows) in device memory it allocates per call.  Callers that cannot allocate
// (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc) use hvc_trace_ra_noalloc instead, whose tables live in `scratch`,
// a file-scope device array of a fixed size that each GPU holds until the
// process ends.  Its device address is looked up once per device when a
// context is created on that device (noalloc_setup); host code uses that
// address, never the array's name.  Every call on one GPU shares the array:
// calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hvc_internal.h"
#include "hvc_neuron.hip.h"

namespace hvc {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static_assert(sizeof(DecayRow) == kNoAllocDecayRowBytes, "decay row size");
static_assert(kNoAllocScratchBytes >= kNoAllocClockBytes + sizeof(DecayRow) * (kNoAllocMaxSteps + 1),
              "scratch holds the clock and the decay rows of kNoAllocMaxSteps steps");

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};
}  // namespace

// Look up this device's copy of `scratch`, once per device; `device` must be
// the current device.
hipError_t noalloc_setup(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

// This device's scratch address (nullptr before noalloc_setup on it).
void* noalloc_scratch(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

// clock[0] = 0, clock[s] = clock[s - 1] + dt: the host's binary64 accumulation
// of the trace clock, one add per step in order (one thread).
__global__ void k_clock_table(double* clock, int n_rows, double dt)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double t = 0.0;
    clock[0] = 0.0;
    for (int s = 1; s < n_rows; ++s) {
        t += dt;
        clock[s] = t;
    }
}

void launch_clock_table(hipStream_t st, double* clock, int n_rows, double dt)
{
    k_clock_table<<<1, 64, 0, st>>>(clock, n_rows, dt);
}

}  // namespace hvc
