/*
 * stonewall_lbm.h -- C ABI of the MI355X (gfx950) D3Q19 stone-wall lattice-Boltzmann step.
 *
 * The reference (Cat-yejun/jeju-stonewall-lbm) drives its kernels through Warp's launch
 * FFI, wp.launch(kernel, dim, inputs=[...]), from stonewall3d/lbm3d.py.  Each entry
 * point below replaces one of those launches; the comment on each names the reference
 * line it replaces.  The Python host surface (stonewall3d.lbm3d in this package) calls
 * them through ctypes at the same places the reference calls wp.launch.
 *
 * Conventions
 *   - Every function returns int: 0 = success, a value below 1000 = the hipError_t
 *     reported by the runtime, 1001+ = a library error (SW_ERR_*).  sw_last_error()
 *     returns a message for the calling thread's last failure.
 *   - All pointers are device pointers unless stated.  The caller owns every buffer;
 *     no entry point on the step path allocates or frees memory.
 *   - `stream` is a hipStream_t (NULL = the null stream).  Launches are asynchronous
 *     and ordered on that stream; nothing synchronizes except sw_prepare_links.
 *   - Populations are stored structure-of-arrays with x contiguous:
 *       f[i*N + (z*ny + y)*nx + x],  N = nx*ny*nz,  i = 0..18 in the reference's
 *     direction order (lbm3d.py:20-24).  sw_to_ref_layout / sw_from_ref_layout convert
 *     to and from the reference's (19, nx, ny, nz) z-contiguous arrays.
 *   - dtype: SW_F32 (float), SW_F16 / SW_BF16 (16-bit storage, FP32 arithmetic; see
 *     below), SW_F64 (double storage and arithmetic).
 *   - 16-bit storage: the weight-normalised population phi_i = f_i / w_i is stored as
 *     (phi_i - sigma) * scale, scale a power of two in [2^-20, 2^20], sigma = sigma_c(z) of
 *     the direction's c_x class c (0, +1, -1) at the cell's height, from a `shift` table
 *     [nz][4] float (device; sw_shift_table) or 1 when `shift` is NULL.  The collision runs
 *     on phi in FP32 (moments from opposite-pair sums and differences; velocity, stress
 *     and relaxation time as the reference; outputs rho feq_i / w_i + (keep * 4.5) qq_i
 *     from terms shared across directions).  FP16 values are rounded to nearest even;
 *     BF16 values are rounded dithered: the float's bits b, plus
 *     (b * 0x9E3779B1 mod 2^32) >> 16, truncated to their top 16.
 *     The same table must be passed to every call on one population state.
 *   - A lattice holds at most 113,025,455 cells.
 */
#ifndef STONEWALL_LBM_H
#define STONEWALL_LBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum sw_dtype { SW_F32 = 0, SW_F16 = 1, SW_BF16 = 2, SW_F64 = 3 };
enum sw_ybc { SW_Y_PERIODIC = 0, SW_Y_FREESLIP = 1 };
enum sw_shift { SW_SHIFT_WEIGHTS = 0, SW_SHIFT_INLET = 1 };
enum sw_flag { SW_FLUID = 0, SW_WALL = 1, SW_GROUND = 2 };           /* lbm3d.py:30 */
enum sw_error {
    SW_OK = 0,
    SW_ERR_INVALID = 1001,
    SW_ERR_GEOMETRY = 1002,
    SW_ERR_TOO_LARGE = 1003,
    SW_ERR_NOT_FOUND = 1004,
    SW_ERR_CAPACITY = 1005
};

/* Library version (major*100 + minor*10 + patch). */
int sw_version(void);

/* Message for the calling thread's last failed call. */
const char* sw_last_error(void);

/* Rows (y lines) per step workgroup; the partials ring holds ceil(ny/rows) pairs per z. */
int sw_rows_per_block(void);

/* uint32 words of a links buffer: nx*ny*nz link masks followed by per-row chunk flags
 * (one bit per 64-cell chunk that touches a solid). */
long long sw_links_words(int nx, int ny, int nz);

/* Doubles per step in the force-partials ring: nz * ceil(ny/rows) * 2. */
long long sw_partials_per_step(int ny, int nz);

/* Builds the per-cell link masks from the solid flags.
 * Replaces the flag set-up at lbm3d.py:288-290 as the kernels consume it: the
 * reference's k_step tests the neighbour flag of every link on every step
 * (lbm3d.py:88-91); here that test is done once.
 *   flag  : (nx, ny, nz) uint8, z contiguous, values FLUID/WALL/GROUND as lbm3d.py:30;
 *           the layer z=0 must be solid, and flag >= 2 may only be reached through
 *           links with cz = +1 (the ground), else SW_ERR_GEOMETRY.
 *   links : sw_links_words(nx, ny, nz) uint32 output: [nz][ny][nx] masks, then chunk flags.
 *   ybc   : SW_Y_PERIODIC (the reference, lbm3d.py:79) or SW_Y_FREESLIP.
 * Synchronizes `stream` (it reports geometry errors). */
int sw_prepare_links(const uint8_t* flag, uint32_t* links, int nx, int ny, int nz, int ybc, void* stream);

/* Writes the equilibrium f = feq(1, u) into f.  Replaces k_init, lbm3d.py:39-54,
 * launched at lbm3d.py:340.  Fluid cells get u = (uin[z](1+noise r1), uin[z] noise r2,
 * uin[z] noise r3) with r in [-1,1) from a PCG hash of (seed, cell); solid cells u = 0. */
int sw_init(void* f, int dtype, const uint32_t* links, const float* uin, int nx, int ny, int nz,
            float noise, unsigned int seed, float scale, const float* shift, void* stream);

/* One fused step f_in -> f_out.  Replaces k_step, lbm3d.py:57-155, launched at
 * lbm3d.py:359-361.
 *   uin        : [nz] inlet profile (lbm3d.py:274-275); gust: per-step inlet factor
 *                table indexed by the step index, so it must hold more entries than
 *                that index (NULL = no gust, the reference's steady inlet).
 *   t, t_dev   : step index = t + *t_dev (t_dev may be NULL); selects the gust entry
 *                and the partials slot t % ring.
 *   nu0, cs, x_sponge, nu_sponge : as lbm3d.py:361 (nu_sponge = 0.05 there).
 *   scale, shift : the 16-bit storage's scale and offset table (ignored for F32 / F64).
 *   partials   : FP64 force-partials ring, ring * sw_partials_per_step(ny, nz) doubles,
 *                or NULL to skip the force.  Reduce with sw_force_reduce.
 * Solid cells of f_out are not written (as the reference, lbm3d.py:65-66). */
int sw_step(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin,
            const float* gust, int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs,
            int x_sponge, double nu_sponge, int ybc, float scale, const float* shift, double* partials,
            int ring, void* stream);

/* sw_step that also adds (rho, ux, uy, uz) of every cell of the new state (f_out) into the
 * accumulator acc[4][N] (float, this library's [c][z][y][x] layout), exactly as sw_macro(f_out,
 * Uacc = acc) after the step would (solid cells add (1, 0, 0, 0)), without reading the
 * populations again.  Replaces the averaging launches k_macro + k_accum that follow k_step
 * on sampled steps, lbm3d.py:363-370. */
int sw_step_accumulate(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin,
                       const float* gust, int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs,
                       int x_sponge, double nu_sponge, int ybc, float scale, const float* shift, double* partials,
                       int ring, float* acc, void* stream);

/* Two fused steps per launch: f_in holds step t, f_out receives step t + 2; the state of
 * step t + 1 stays on chip (no reference counterpart: the reference launches k_step once
 * per step, lbm3d.py:359-361; this is two of those launches with one HBM pass).  Arguments
 * as sw_step; populations are bit-identical to sw_step(t) followed by sw_step(t + 1).
 * FP32, FP16 and BF16 storage with population arrays below 4 GiB.  The gust table needs
 * entries t and t + 1.  partials: ring * sw_step2_partials_per_step(nx, ny, nz) doubles
 * (ring >= 2), slots t % ring and (t + 1) % ring; reduce with sw_force_reduce2.  The force
 * partials are per 64-cell x tile and tile row block, so force rows equal sw_step's to
 * rounding (a different fixed summation order), not bit for bit. */
long long sw_step2_partials_per_step(int nx, int ny, int nz);
int sw_step2(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin, const float* gust,
             int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs, int x_sponge, double nu_sponge,
             int ybc, float scale, const float* shift, double* partials, int ring, void* stream);
/* One step with sw_step2's tiles and force-partials layout: the odd step of a run of
 * sw_step2 pairs (no reference counterpart: lbm3d.py:359-361 launches k_step once per step),
 * so the whole run folds into force rows through one sw_force_reduce2 ring.  Arguments and
 * checks as sw_step2; writes slot t % ring; populations bit-identical to sw_step. */
int sw_step2_one(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin, const float* gust,
                 int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs, int x_sponge,
                 double nu_sponge, int ybc, float scale, const float* shift, double* partials, int ring, void* stream);
/* sw_step2 that also adds (rho, ux, uy, uz) of every cell of the state after the pair's first
 * step (which = 1: step t + 1) or second step (which = 2: step t + 2) into acc[4][N], exactly
 * as sw_macro(..., Uacc = acc) on that state would (solid cells add (1, 0, 0, 0)), without a
 * pass over the populations: the averaging launches k_macro + k_accum of lbm3d.py:363-370 on
 * a sampled step inside a pair.  FP32 storage only (the form run two steps per launch). */
int sw_step2_accumulate(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin,
                        const float* gust, int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs,
                        int x_sponge, double nu_sponge, int ybc, float scale, const float* shift, double* partials,
                        int ring, float* acc, int which, void* stream);
/* force[t][z][0..1] from the sw_step2 partials ring, as sw_force_reduce does for sw_step's. */
int sw_force_reduce2(const double* partials, int ring, int t0, const int* t_dev, int nsteps, int nx, int ny, int nz,
                     double* force, void* stream);
/* Tile rows of the two-step kernels (no reference counterpart): 0 (default) picks per lattice
 * 15 or 16 rows, one workgroup per tile, or, for lattices with fewer tiles than the device has
 * compute units, 12 or 13 rows with each tile's z sweep cut into segments
 * (sw_set_step2_segments; the 12-row kernel keeps the next plane's loads in flight across each
 * plane's barrier); 12, 13, 15 or 16 fixes it for the process.  The rows set the partials layout
 * (sw_step2_partials_per_step, sw_force_reduce2): keep one setting for the life of a partials
 * ring.  Populations do not depend on it; force rows change in rounding only.  The two-step
 * batch uses sw_batch_step2_rows(). */
int sw_set_step2_rows(int rows);
int sw_get_step2_rows(void);
/* Order of the two-step kernels' tiles over the workgroups (no reference counterpart):
 * 1 (default) gives each XCD (each with its own L2) one contiguous band of tiles, so the halo
 * rows and columns a tile reads are read by its neighbours in the same L2; 0 numbers them
 * plainly.  Results do not depend on it. */
int sw_set_step2_tile_order(int banded);
/* z segments per tile of the two-step kernels with 12- or 13-row tiles (no reference
 * counterpart): each tile's sweep is cut into this many workgroups (1 .. 32, at most
 * nz - 1), each recomputing the first step of the plane below its segment, so that a
 * lattice with fewer tiles than compute units fills the device.  0 (default) picks the
 * count per lattice.  Lattices with 15- or 16-row tiles are not cut.  Populations and force
 * rows do not depend on it.  sw_get_step2_segments returns the count a lattice gets (-1 on
 * bad dimensions). */
int sw_set_step2_segments(int segments);
int sw_get_step2_segments(int nx, int ny, int nz);

/* Kernel variants of sw_step (no reference counterpart: the reference has one Warp
 * kernel; these are the tile shapes the autotuner chooses between).  The choice is kept
 * per dtype for the process; every variant gives bit-identical populations.
 *   0  one 64-cell chunk per wavefront iteration, one wavefront per row
 *   1  0 + nontemporal stores
 *   2  1 + nontemporal loads
 *   3  1, each row split over 2 wavefronts
 *   4  1, each row split over 4 wavefronts
 *   5  2, each row split over 2 wavefronts
 *   6  1 with two chunks per iteration
 *   7  1 with the next chunk's loads issued first, rows over 2 wavefronts
 *   8  pairs: a lane owns two adjacent cells, one 32-bit (16-bit storage) or 64-bit
 *      (FP32) access per link, x-1 / x+1 neighbours by DPP lane shifts; NT stores
 *   9  8, rows over 2 wavefronts      10  8, rows over 4 wavefronts
 *   11 9 with cached stores
 *   12 quads (16-bit storage): a lane owns four adjacent cells, one 64-bit access per
 *      link; NT stores                 13  12, rows over 2 wavefronts
 * FP64 storage maps 8..13 to 1, FP32 maps 12..13 to 9.  Variant 9 runs as 8 on short rows
 * (16-bit storage: nx < 512; FP32: nx <= 128), where splitting a row over two wavefronts
 * leaves one idle or unbalanced.  Arrays of 4 GiB or more use 64-bit addressing with
 * variant 1's tile whatever the setting. */
int sw_step_variants(void);
int sw_set_step_variant(int dtype, int variant);
int sw_get_step_variant(int dtype);

/* force[t][z][0..1] (FP64 Fx, Fz per height level) for t = t0 + *t_dev .. + nsteps - 1,
 * summed in a fixed order from the partials ring.  Replaces the FP32 atomic scatter
 * wp.atomic_add(force, t, z, c, ...) at lbm3d.py:99-101 (force array :301). */
int sw_force_reduce(const double* partials, int ring, int t0, const int* t_dev, int nsteps, int ny, int nz,
                    double* force, void* stream);

/* (rho, ux, uy, uz) per cell into U[4][N] and/or Uacc[4][N] += that (either may be NULL).
 * Replaces k_macro + k_accum, lbm3d.py:158-188, launched at lbm3d.py:367-369 and :387.
 * Solid cells give (1, 0, 0, 0). */
int sw_macro(const void* f, int dtype, const uint32_t* links, int nx, int ny, int nz, float scale,
             const float* shift, float* U, float* Uacc, void* stream);

/* Offset table of the 16-bit storage forms (no reference counterpart: the reference
 * stores FP32).  shift: [nz][4] float (device) output, (sigma_0, sigma_+1, sigma_-1, 0) per
 * height, sigma_c = feq_i(1, (U, 0, 0)) / w_i for a direction with c_x = c, evaluated in the
 * operand order of lbm3d.py:35-36 in FP32.
 *   SW_SHIFT_WEIGHTS: U = 0, sigma = 1: the stored value is (f_i - w_i) / w_i.
 *   SW_SHIFT_INLET  : U = uin[z], the inlet velocity at that height.  Where the flow is
 *                     near the inlet profile the stored value is ~10-50x smaller than
 *                     (f_i - w_i) / w_i, and so is its rounding error. */
int sw_shift_table(const float* uin, int nz, int kind, float* shift, void* stream);

/* Max-pooled positive Q-criterion frame of U (from sw_macro) over ds^3 blocks of the crop
 * starting at (x0, 0, 0): out[i][j][k], (qx, qy, qz).  Replaces k_qcoarse + q_at,
 * lbm3d.py:203-238, launched at lbm3d.py:316. */
int sw_qcoarse(const float* U, const uint32_t* links, int nx, int ny, int nz, int x0, int ds, int qx, int qy, int qz,
               float* out, void* stream);

/* Self-test of the step's division shortcut (no counterpart in the reference: the step
 * replaces IEEE division by a Newton-refined hardware reciprocal and an FMA-corrected
 * quotient, which must round exactly like IEEE division).  counts: device buffer of two
 * uint64; after the stream completes, counts[0] = normal divisors over all 2^32 bit
 * patterns whose reciprocal differs from IEEE 1/b, counts[1] = differing quotients among
 * 2^32 hashed samples (|a| in [2^-40, 4), b in [2^-4, 4)).  Both must be 0. */
int sw_selftest_division(unsigned long long* counts, void* stream);

/* Streaming copy dst = src of `bytes` (a multiple of 16; both 16-byte aligned): contiguous
 * 16 KB blocks per workgroup, 16-byte loads and nontemporal 16-byte stores, 32 workgroups
 * per compute unit.  No
 * reference counterpart: it measures the HBM read + write ceiling the step is compared
 * against (bench.py), and it is the known-size kernel that scales the PMC byte counters. */
int sw_stream_copy(const void* src, void* dst, long long bytes, void* stream);

/* Allocation-free step.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  It is sw_step followed by sw_force_reduce of that one step, with
 * the FP64 force partials in a fixed device array of the library (4 MiB per GPU, held until
 * the process ends) instead of a caller's ring; it allocates nothing.  The caller passes
 * inputs and outputs as for sw_step, and `force` is the [T][nz][2] FP64 force record whose
 * row t (+ *t_dev) is written.  Lattices with nz * ceil(ny/4) > sw_noalloc_max_pairs()
 * (262,144) are refused with SW_ERR_TOO_LARGE.  All calls on one GPU share that GPU's array:
 * calls that overlap on one GPU (different streams without ordering) race. */
long long sw_noalloc_max_pairs(void);
int sw_step_noalloc(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin,
                    const float* gust, int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs,
                    int x_sponge, double nu_sponge, int ybc, float scale, const float* shift, double* force,
                    void* stream);

/* ---------------------------------------------------------------- batched form
 * No reference counterpart: the reference runs one lattice per process (run_all.sh).
 * Many independent lattices of one dtype advance together: one persistent launch per
 * step strides over the row-groups of every lattice in a device-resident table.  Each
 * lattice's arithmetic is that of sw_step, so its results are bit-identical alone and
 * in any mix.  The table is rewritten between steps when lattices join or leave. */
typedef struct sw_instance {
    void* f0;                 /* population buffer that holds step t when parity == 0 */
    void* f1;                 /* the other buffer */
    const uint32_t* links;    /* from sw_prepare_links */
    const float* uin;         /* [nz] inlet profile */
    const float* gust;        /* [t_end] inlet factor per step, or NULL */
    double* partials;         /* ring * sw_partials_per_step(ny, nz) doubles, or NULL */
    double* force;            /* [t_end][nz][2] force record, or NULL */
    const float* shift;       /* 16-bit storage offsets [nz][4] (sw_shift_table), or NULL */
    double nu0, cs, nu_sponge;
    float scale;
    int nx, ny, nz;
    int ring, x_sponge, ybc, dtype;
    int t, t_end, parity;     /* next step, end step (exclusive), buffer of step t */
    int item0;                /* first work item; written by sw_batch_upload */
    int tile0;                /* first two-step work item (tile); written by sw_batch_upload */
} sw_instance;

/* Bytes of a device table holding up to max_instances lattices. */
long long sw_batch_table_bytes(int max_instances);
/* Writes n instances (item0 computed here) into the device table; synchronizes stream. */
int sw_batch_upload(const sw_instance* instances, int n, int max_instances, void* table, void* stream);
/* Copies the table's n entries back (t, parity, ...); synchronizes stream. */
int sw_batch_download(const void* table, sw_instance* instances, int n, void* stream);
/* Step t_i + k of every unfinished instance; `blocks` persistent workgroups; variant as
 * sw_set_step_variant; big_offsets = 1 when some population array exceeds 4 GiB. */
int sw_batch_step(const void* table, int dtype, int k, int blocks, int variant, int big_offsets, void* stream);
/* Force rows of steps t_i .. t_i + K - 1 (before sw_batch_advance); needs ring >= K. */
int sw_batch_force_reduce(const void* table, int max_instances, int K, void* stream);
/* t_i += K (clamped to t_end) and the matching parity flip. */
int sw_batch_advance(void* table, int max_instances, int K, void* stream);
/* Captures K batch steps + force reduction + advance as one hipGraph (graph handle out);
 * replay with sw_graph_launch on any stream, free with sw_graph_destroy. */
int sw_batch_graph_create(void* table, int max_instances, int dtype, int K, int blocks, int variant, int big_offsets,
                          void** graph);
/* Two-step form of the batch (no reference counterpart, as above): every instance FP32, FP16
 * or BF16 (one dtype per table) with population arrays below 4 GiB.  Partials ring per
 * instance: ring * sw_batch_partials_per_step(nx, ny, nz) doubles (enough for either form).
 * sw_batch_step2(p): steps t_i + 2p and t_i + 2p + 1 of every instance with two or more left
 * at t_i + 2p, persistent `blocks` workgroups (one per compute unit fits) taking 64 x 10-cell
 * tiles (sw_batch_step2_rows() rows with their two halo rows) from a counter in the table header; the buffers swap once per launch that touches
 * an instance.  sw_batch_force_reduce2: force rows of steps t_i .. t_i + K - 1 (ring >= K).
 * sw_batch_advance2: t_i += K (K even, clamped to t_end) and the parity of the launches.
 * sw_batch_graph_create2: K / 2 pair launches + the tail launch + force reduction + advance
 * as one hipGraph.
 * Populations are bit-identical to single steps; force rows to the single-lattice run with
 * sw_set_step2_rows(sw_batch_step2_rows()) (pairs by sw_step2, an odd last step by sw_step2_one). */
long long sw_batch_partials_per_step(int nx, int ny, int nz);
/* Tile rows of the two-step batch's tiles, halo rows included (12). */
int sw_batch_step2_rows(void);
int sw_batch_step2(void* table, int dtype, int p, int blocks, void* stream);
/* After the K / 2 pair launches of a chunk: the odd last step of every instance with fewer
 * than K (and an odd number of) steps left at the chunk's start. */
int sw_batch_step2_tail(void* table, int dtype, int K, int blocks, void* stream);
int sw_batch_force_reduce2(const void* table, int max_instances, int K, void* stream);
int sw_batch_advance2(void* table, int max_instances, int K, void* stream);
int sw_batch_graph_create2(void* table, int max_instances, int dtype, int K, int blocks, void** graph);
int sw_graph_launch(void* graph, void* stream);
int sw_graph_destroy(void* graph);

/* Layout conversion of ncomp stacked fields between [c][z][y][x] (this library) and
 * [c][x][y][z] (the reference's arrays, lbm3d.py:293-303, used by its checkpoint files,
 * lbm3d.py:342-355).  elem_bytes in {1, 2, 4, 8}; src and dst must not overlap. */
int sw_to_ref_layout(const void* src, void* dst, int elem_bytes, int ncomp, int nx, int ny, int nz, void* stream);
int sw_from_ref_layout(const void* src, void* dst, int elem_bytes, int ncomp, int nx, int ny, int nz, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* STONEWALL_LBM_H */
