// Single-lattice entry points of the two-step kernel (the tile function: sw_step2.hpp).
#include "sw_step2.hpp"
#include "sw_api_util.hpp"

namespace sw {

// Tile of workgroup b: with `banded`, the j-th workgroup of XCD x (b = x + j NXCD) takes tile
// band_start(x) + j (XCD x holds ceil((ntiles - x) / NXCD) workgroups, its band's length).
__device__ __forceinline__ int step2_tile_of(int b, int ntiles, bool banded) {
    if (!banded) return b;
    return band_start(b % NXCD, ntiles) + b / NXCD;
}

template <class S, int TY, bool PF = false, int ACC = 0>
__global__ __launch_bounds__(64 * TY, 1) void k_step2(StepArgs a, int ntx, int ntiles, int banded) {
    __shared__ Step2Lds<TY> L;
    step2_tile<S, TY, false, StepArgs, PF, ACC>(a, a.t + (a.t_dev ? *a.t_dev : 0), ntx, ntiles,
                                                step2_tile_of((int)blockIdx.x, ntiles, banded != 0), L);
}

template <class S, int TY>
__global__ __launch_bounds__(64 * TY, 1) void k_step2_one(StepArgs a, int ntx, int ntiles, int banded) {
    __shared__ Step2Lds<TY> L;
    step2_tile<S, TY, true>(a, a.t + (a.t_dev ? *a.t_dev : 0), ntx, ntiles,
                            step2_tile_of((int)blockIdx.x, ntiles, banded != 0), L);
}

// The same with each tile's sweep cut into nseg z segments, for lattices with fewer tiles than
// compute units: work item i = g * ntiles + tile (segment g of `tile`, numbered through the
// banded order), segment g holding the output planes [1 + g (nz-1) / nseg, 1 + (g+1) (nz-1)
// / nseg); each segment recomputes the first step of the plane below it.  Launched with 13-row
// tiles: with a run-time za (step2_tile) the kernel needs 128 VGPRs, which 13 wavefronts per
// workgroup allow without spilling (at 15 or 16 rows it spilled about 30 and ran 1.46x slower
// per plane).
template <class S, int TY, bool SINGLE, bool PF = false, int ACC = 0>
__global__ __launch_bounds__(64 * TY, 1) void k_step2_seg(StepArgs a, int ntx, int ntiles, int nseg, int banded) {
    __shared__ Step2Lds<TY> L;
    const int item = step2_tile_of((int)blockIdx.x, ntiles * nseg, banded != 0);
    const int g = item / ntiles;
    const int za = 1 + (int)(((long long)g * (a.nz - 1)) / nseg);
    const int zb = 1 + (int)(((long long)(g + 1) * (a.nz - 1)) / nseg);
    step2_tile<S, TY, SINGLE, StepArgs, PF, ACC>(a, a.t + (a.t_dev ? *a.t_dev : 0), ntx, ntiles, item - g * ntiles,
                                                 L, za, zb);
}

constexpr int SEG_ROWS = 13;          // tile rows of the segmented launch
constexpr int PF_ROWS = 12;           // tile rows of the launch with next-plane loads in flight (PF)

// Launch shape of a lattice: tile rows TY (the partials layout: sw_step2_partials_per_step)
// and z segments S.  TY is the setting of sw_set_step2_rows, or by default: 15 or 16 rows
// uncut (the one of fewer rounds of workgroups, each weighted by its rows), unless SEG_ROWS
// rows cut into S = 1 .. 32 segments are estimated 3 % faster; in that case (lattices whose
// tiles leave compute units idle) PF_ROWS rows with the next plane's loads in flight, cut
// into their own best S, where that estimate is lower.  The estimate: rounds of workgroups
// over the compute units (one per unit: the LDS) times a workgroup's sweep (its planes, half
// a plane for the recomputed first step below a segment, about two planes of fill and
// drain) times TY / (TY - 2) rows per output row, times PF_COST for PF_ROWS tiles (the
// 800x160x321 lattice's 13-row and 12-row launches measured against it).  S is the setting
// of sw_set_step2_segments for SEG_ROWS and PF_ROWS tiles (1 for 15 and 16), or the
// estimate's.  The settings never change TY's default, so a lattice keeps its layout.
// (On lattices whose 15-row tiles fill the compute units in one round, e.g. 640x320x160,
// every 12-row shape measured slower than the uncut 15 rows.)
constexpr double PF_COST = 0.8;
static int g_step2_rows = 0;
static int g_step2_banded = 1;
static int g_step2_segments = 0;

static int step2_max_segments(int nz) { return nz - 1 < 32 ? (nz - 1 < 1 ? 1 : nz - 1) : 32; }

static double step2_cost(int nx, int ny, int nz, int cus, int ty, int sg) {
    const long long items = (long long)step2_tiles_for(nx, ny, ty) * sg;
    const double rounds = (double)((items + cus - 1) / cus);
    return rounds * ((double)(nz - 1) / sg + (sg > 1 ? 0.5 : 0.0) + 2.0) * ((double)ty / (ty - 2)) *
           (ty == PF_ROWS ? PF_COST : 1.0);
}

static int step2_auto_segments(int nx, int ny, int nz, int cus, int ty) {
    int best = 1;
    for (int sg = 2; sg <= step2_max_segments(nz); sg++)
        if (step2_cost(nx, ny, nz, cus, ty, sg) < step2_cost(nx, ny, nz, cus, ty, best) - 1e-9) best = sg;
    return best;
}

static int step2_rows(int nx, int ny, int nz, int cus) {
    if (g_step2_rows) return g_step2_rows;
    // between the uncut shapes: the fewer rounds of workgroups, each weighted by its rows
    const int r15 = (step2_tiles_for(nx, ny, 15) + cus - 1) / cus, r16 = (step2_tiles_for(nx, ny, 16) + cus - 1) / cus;
    const int uncut = r15 * 15 < r16 * 16 ? 15 : 16;
    const double cseg = step2_cost(nx, ny, nz, cus, SEG_ROWS, step2_auto_segments(nx, ny, nz, cus, SEG_ROWS));
    if (!(cseg < 0.97 * step2_cost(nx, ny, nz, cus, uncut, 1))) return uncut;
    const double cpf = step2_cost(nx, ny, nz, cus, PF_ROWS, step2_auto_segments(nx, ny, nz, cus, PF_ROWS));
    return cpf < cseg ? PF_ROWS : SEG_ROWS;
}

static int step2_segments(int nx, int ny, int nz, int cus, int ty) {
    if (ty != SEG_ROWS && ty != PF_ROWS) return 1;
    if (g_step2_segments) return g_step2_segments < step2_max_segments(nz) ? g_step2_segments : step2_max_segments(nz);
    return step2_auto_segments(nx, ny, nz, cus, ty);
}

template <class S, bool SINGLE, int ACC = 0>
static void launch_step2(const StepArgs& a, int cus, hipStream_t s) {
    static_assert(!(SINGLE && ACC), "the accumulating form is a two-step sweep");
    const int ntx = (a.nx + 63) / 64;
    const int ty = step2_rows(a.nx, a.ny, a.nz, cus);
    const int tiles = step2_tiles_for(a.nx, a.ny, ty);
    if (ty == PF_ROWS) {
        const int nseg = step2_segments(a.nx, a.ny, a.nz, cus, ty);
        if (nseg > 1) {
            hipLaunchKernelGGL((k_step2_seg<S, PF_ROWS, SINGLE, !SINGLE, ACC>), dim3((unsigned)(tiles * nseg)),
                               dim3(64 * PF_ROWS), 0, s, a, ntx, tiles, nseg, g_step2_banded);
        } else if constexpr (SINGLE) {
            hipLaunchKernelGGL((k_step2_one<S, PF_ROWS>), dim3(tiles), dim3(64 * PF_ROWS), 0, s, a, ntx, tiles,
                               g_step2_banded);
        } else {
            hipLaunchKernelGGL((k_step2<S, PF_ROWS, true, ACC>), dim3(tiles), dim3(64 * PF_ROWS), 0, s, a, ntx,
                               tiles, g_step2_banded);
        }
    } else if (ty == SEG_ROWS) {
        const int nseg = step2_segments(a.nx, a.ny, a.nz, cus, ty);
        hipLaunchKernelGGL((k_step2_seg<S, SEG_ROWS, SINGLE, false, ACC>), dim3((unsigned)(tiles * nseg)),
                           dim3(64 * SEG_ROWS), 0, s, a, ntx, tiles, nseg, g_step2_banded);
    } else if (ty == 15) {
        if constexpr (SINGLE) hipLaunchKernelGGL((k_step2_one<S, 15>), dim3(tiles), dim3(64 * 15), 0, s, a, ntx, tiles, g_step2_banded);
        else hipLaunchKernelGGL((k_step2<S, 15, false, ACC>), dim3(tiles), dim3(64 * 15), 0, s, a, ntx, tiles, g_step2_banded);
    } else {
        if constexpr (SINGLE) hipLaunchKernelGGL((k_step2_one<S, 16>), dim3(tiles), dim3(64 * 16), 0, s, a, ntx, tiles, g_step2_banded);
        else hipLaunchKernelGGL((k_step2<S, 16, false, ACC>), dim3(tiles), dim3(64 * 16), 0, s, a, ntx, tiles, g_step2_banded);
    }
}

static int device_cus(int* cus) {
    int dev = 0;
    SW_HIP(hipGetDevice(&dev));
    SW_HIP(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev));
    return SW_OK;
}

__global__ void k_force_reduce2(const double* __restrict__ partials, int ring, int t0, const int* t_dev, int nz, int P,
                                double* __restrict__ force) {
    const int t = t0 + (t_dev ? *t_dev : 0) + (int)blockIdx.y;
    force_row2(partials, ring, t, (int)blockIdx.x, nz, P, (int)threadIdx.x, force);
}

static int step2_entry(bool single, const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin,
                       const float* gust, int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs,
                       int x_sponge, double nu_sponge, int ybc, float scale, const float* shift, double* partials,
                       int ring, void* stream, float* acc = nullptr, int which = 0) {
    SW_CHECK_ARG(f_in && f_out && links && uin, "f_in, f_out, links and uin must be non-null");
    SW_CHECK_ARG(f_in != f_out, "f_in and f_out must be different buffers (pull scheme)");
    SW_CHECK_RC(check_dims(nx, ny, nz));
    SW_CHECK_RC(check_dtype(dtype, scale));
    SW_CHECK_ARG(dtype != DT_F64, "the two-step kernels take FP32, FP16 or BF16 storage");
    SW_CHECK_ARG((size_t)Q * nx * ny * nz * (dtype == DT_F32 ? 4 : 2) < ((size_t)1 << 32),
                 "the two-step kernels need population arrays below 4 GiB; use sw_step");
    SW_CHECK_ARG(ybc == Y_PERIODIC || ybc == Y_FREESLIP, "ybc must be 0 (periodic) or 1 (free-slip)");
    SW_CHECK_ARG(!partials || ring >= 2, "ring must be at least 2 when partials are given");
    SW_CHECK_ARG(t >= 0, "t must be non-negative");
    SW_CHECK_ARG(which == 0 || (!single && acc && dtype == DT_F32 && (which == 1 || which == 2)),
                 "the accumulating two-step launch takes FP32 storage, a non-null acc and which = 1 or 2");
    int cus = 0;
    SW_CHECK_RC(device_cus(&cus));
    StepArgs a{};
    a.f_in = f_in; a.f_out = f_out; a.links = links; a.uin = uin; a.gust = gust;
    a.partials = partials; a.t_dev = t_dev; a.t = t; a.ring = ring > 0 ? ring : 2;
    a.shift = (dtype == DT_F16 || dtype == DT_BF16) ? shift : nullptr;
    a.nx = nx; a.ny = ny; a.nz = nz; a.nyb = (ny + ROWS - 1) / ROWS;
    a.x_sponge = x_sponge; a.ybc = ybc; a.nu0 = nu0; a.cs = cs; a.nu_sponge = nu_sponge;
    a.scale = scale; a.inv_scale = 1.0f / scale;
    a.acc = acc;
    hipStream_t s = (hipStream_t)stream;
    if (which == 1) {
        launch_step2<float, false, 1>(a, cus, s);
        SW_HIP(hipGetLastError());
        return SW_OK;
    }
    if (which == 2) {
        launch_step2<float, false, 2>(a, cus, s);
        SW_HIP(hipGetLastError());
        return SW_OK;
    }
    switch (dtype) {
        case DT_F32: single ? launch_step2<float, true>(a, cus, s) : launch_step2<float, false>(a, cus, s); break;
        case DT_F16: single ? launch_step2<_Float16, true>(a, cus, s) : launch_step2<_Float16, false>(a, cus, s); break;
        case DT_BF16: single ? launch_step2<bf16_t, true>(a, cus, s) : launch_step2<bf16_t, false>(a, cus, s); break;
    }
    SW_HIP(hipGetLastError());
    return SW_OK;
}

}  // namespace sw

using namespace sw;

extern "C" {

int sw_set_step2_rows(int rows) {
    SW_CHECK_ARG(rows == 0 || rows == PF_ROWS || rows == SEG_ROWS || rows == 15 || rows == 16,
                 "rows must be 0 (automatic), 12, 13, 15 or 16");
    g_step2_rows = rows;
    return SW_OK;
}

int sw_get_step2_rows(void) { return g_step2_rows; }

int sw_set_step2_segments(int segments) {
    SW_CHECK_ARG(segments >= 0 && segments <= 32, "segments must be 0 (automatic) .. 32");
    g_step2_segments = segments;
    return SW_OK;
}

int sw_get_step2_segments(int nx, int ny, int nz) {
    if (check_dims(nx, ny, nz)) return -1;
    int cus = 0;
    if (device_cus(&cus)) return -1;
    return step2_segments(nx, ny, nz, cus, step2_rows(nx, ny, nz, cus));
}

int sw_set_step2_tile_order(int banded) {
    SW_CHECK_ARG(banded == 0 || banded == 1, "banded must be 0 (plain) or 1 (one band of tiles per XCD)");
    g_step2_banded = banded;
    return SW_OK;
}

long long sw_step2_partials_per_step(int nx, int ny, int nz) {
    if (check_dims(nx, ny, nz)) return -1;
    int cus = 0;
    if (device_cus(&cus)) return -1;
    return (long long)nz * step2_tiles_for(nx, ny, step2_rows(nx, ny, nz, cus)) * 2;
}

int sw_step2(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin, const float* gust,
             int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs, int x_sponge, double nu_sponge,
             int ybc, float scale, const float* shift, double* partials, int ring, void* stream) {
    return step2_entry(false, f_in, f_out, dtype, links, uin, gust, t, t_dev, nx, ny, nz, nu0, cs, x_sponge,
                       nu_sponge, ybc, scale, shift, partials, ring, stream);
}

int sw_step2_one(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin, const float* gust,
                 int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs, int x_sponge,
                 double nu_sponge, int ybc, float scale, const float* shift, double* partials, int ring, void* stream) {
    return step2_entry(true, f_in, f_out, dtype, links, uin, gust, t, t_dev, nx, ny, nz, nu0, cs, x_sponge,
                       nu_sponge, ybc, scale, shift, partials, ring, stream);
}

int sw_step2_accumulate(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin,
                        const float* gust, int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs,
                        int x_sponge, double nu_sponge, int ybc, float scale, const float* shift, double* partials,
                        int ring, float* acc, int which, void* stream) {
    return step2_entry(false, f_in, f_out, dtype, links, uin, gust, t, t_dev, nx, ny, nz, nu0, cs, x_sponge,
                       nu_sponge, ybc, scale, shift, partials, ring, stream, acc, which);
}

int sw_force_reduce2(const double* partials, int ring, int t0, const int* t_dev, int nsteps, int nx, int ny, int nz,
                     double* force, void* stream) {
    SW_CHECK_ARG(partials && force, "partials and force must be non-null");
    SW_CHECK_ARG(ring > 0 && nsteps > 0 && nsteps <= ring, "need 0 < nsteps <= ring");
    SW_CHECK_RC(check_dims(nx, ny, nz));
    SW_CHECK_ARG(t0 >= 0, "t0 must be non-negative");
    int cus = 0;
    SW_CHECK_RC(device_cus(&cus));
    const int P = step2_tiles_for(nx, ny, step2_rows(nx, ny, nz, cus));
    hipLaunchKernelGGL(k_force_reduce2, dim3(nz, nsteps), dim3(64), 0, (hipStream_t)stream, partials, ring, t0, t_dev,
                       nz, P, force);
    SW_HIP(hipGetLastError());
    return SW_OK;
}

}  // extern "C"
