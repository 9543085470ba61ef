// Host-side helpers shared by the C-ABI entry points: error codes, the last-error
// message, argument checks.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <math.h>
#include "../../include/stonewall_lbm.h"

namespace sw {

constexpr int SW_VERSION = 120;   // 1.2.0 (two-step batch; sw_instance.tile0; table holds per-lattice arguments)
// library error codes (SW_ERR_*) come from stonewall_lbm.h; hipError_t values (< 1000)
// are passed through unchanged

// at most this many cells per lattice: 19*N element offsets are 32-bit
constexpr long long SW_MAX_CELLS = 2147483647LL / 19;

inline char* error_buffer() {
    static thread_local char buf[512] = "no error";
    return buf;
}
inline const char* last_error_message() { return error_buffer(); }
inline int set_error(int code, const char* msg) {
    snprintf(error_buffer(), 512, "%s", msg);
    return code;
}
inline int set_hip_error(hipError_t e, const char* where) {
    snprintf(error_buffer(), 512, "%s: HIP error %d (%s)", where, (int)e, hipGetErrorString(e));
    return (int)e;
}

inline int check_dims(int nx, int ny, int nz) {
    if (nx < 2 || ny < 1 || nz < 2) return set_error(SW_ERR_INVALID, "lattice needs nx >= 2, ny >= 1, nz >= 2");
    if (ny > 65535 || nz > 65535) return set_error(SW_ERR_INVALID, "ny and nz must be at most 65535");
    if ((long long)nx * ny * nz > SW_MAX_CELLS)
        return set_error(SW_ERR_TOO_LARGE, "lattice exceeds 113,025,455 cells (19*N must fit in 31 bits)");
    return SW_OK;
}

inline bool is_pow2_float(float s) {
    if (!(s > 0.0f) || s != s) return false;
    int e;
    float m = frexpf(s, &e);
    return m == 0.5f;
}

inline int check_dtype(int dtype, float scale) {
    if (dtype < 0 || dtype > 3) return set_error(SW_ERR_INVALID, "dtype must be 0 (f32), 1 (f16), 2 (bf16) or 3 (f64)");
    if ((dtype == 1 || dtype == 2) && !(is_pow2_float(scale) && scale >= 0x1p-20f && scale <= 0x1p20f))
        return set_error(SW_ERR_INVALID, "scale of a 16-bit storage form must be a power of two in [2^-20, 2^20]");
    return SW_OK;
}

// The step and force-reduction launches behind sw_step / sw_force_reduce, with their
// argument checks (sw_kernels.hip); other entry points of the library reuse them.
int step_entry(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin, const float* gust,
               int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs, int x_sponge, double nu_sponge,
               int ybc, float scale, const float* shift, double* partials, int ring, hipStream_t s,
               float* acc = nullptr);
int force_reduce_entry(const double* partials, int ring, int t0, const int* t_dev, int nsteps, int ny, int nz,
                       double* force, hipStream_t s);

}  // namespace sw

#define SW_CHECK_ARG(cond, msg) \
    do { if (!(cond)) return ::sw::set_error(SW_ERR_INVALID, (msg)); } while (0)
#define SW_CHECK_RC(expr) \
    do { int _rc = (expr); if (_rc != 0) return _rc; } while (0)
#define SW_HIP(expr) \
    do { hipError_t _e = (expr); if (_e != hipSuccess) return ::sw::set_hip_error(_e, #expr); } while (0)
