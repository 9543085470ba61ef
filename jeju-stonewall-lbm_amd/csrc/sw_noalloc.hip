// Allocation-free step: sw_step with the force row of that step, whose per-workgroup FP64
// force partials live in a fixed device array of this library instead of a caller's ring.
// For callers that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).  Declared in include/stonewall_lbm.h.
#include "sw_step.hpp"
#include "sw_api_util.hpp"

#include <mutex>

namespace sw {

// FP64 (Fx, Fz) partial pairs of one step: one pair per (z, row-group) workgroup, i.e.
// nz * ceil(ny / 4) pairs; 4 MiB holds 262,144 of them (the 640x320x160 lattice needs 12,800,
// the reference's 800x160x321 default 12,840).
constexpr long long NOALLOC_PAIRS = 262144;

// Each GPU the library runs on holds its own copy until the process ends.
static __device__ __align__(256) double scratch[2 * NOALLOC_PAIRS];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static double* g_scratch[MAX_DEVICES];   // device address of `scratch` per device, from hipGetSymbolAddress

// Device address of the current device's `scratch`, looked up once per device with that
// device current.  The host-side name `scratch` is not a device address and is never passed
// to a kernel, a memset or a copy.
static int scratch_address(double** out) {
    int dev = -1;
    SW_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return set_error(SW_ERR_INVALID, "sw_step_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return set_hip_error(e, "sw_step_noalloc: hipGetSymbolAddress(scratch)");
        if (!p) return set_error(SW_ERR_INVALID, "sw_step_noalloc: no device address for the scratch array");
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return SW_OK;
}

}  // namespace sw

using namespace sw;

extern "C" {

long long sw_noalloc_max_pairs(void) { return NOALLOC_PAIRS; }

int sw_step_noalloc(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin,
                    const float* gust, int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs,
                    int x_sponge, double nu_sponge, int ybc, float scale, const float* shift, double* force,
                    void* stream) {
    SW_CHECK_ARG(force, "force must be non-null");
    SW_CHECK_RC(check_dims(nx, ny, nz));
    const long long pairs = (long long)nz * ((ny + ROWS - 1) / ROWS);
    if (pairs > NOALLOC_PAIRS)
        return set_error(SW_ERR_TOO_LARGE, "sw_step_noalloc: nz * ceil(ny/4) exceeds the 262,144 force-partial "
                                           "pairs of its fixed scratch array; use sw_step with a partials ring");
    double* partials = nullptr;
    SW_CHECK_RC(scratch_address(&partials));
    hipStream_t s = (hipStream_t)stream;
    // every (z, row-group) slot of the one-step ring is written by the step before the
    // reduction reads it (same stream), so the scratch needs no clearing between calls
    SW_CHECK_RC(step_entry(f_in, f_out, dtype, links, uin, gust, t, t_dev, nx, ny, nz, nu0, cs, x_sponge, nu_sponge,
                           ybc, scale, shift, partials, 1, s));
    return force_reduce_entry(partials, 1, t, t_dev, 1, ny, nz, force, s);
}

}  // extern "C"
