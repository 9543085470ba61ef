// Batched form: many independent lattices of uneven size advanced by one launch per step.
//
// A device-resident instance table (header + sw_instance entries, include/stonewall_lbm.h)
// describes every lattice: its two population buffers, link masks, inlet, gust table,
// force-partials ring, force record, physics scalars, its next step t and end step
// t_end, and which buffer holds step t (parity).  The work items of a batch are the
// row-groups (ROWS rows at one z) of all lattices, numbered instance by instance
// (item0 = prefix sum).  k_batch_step is persistent: a fixed grid of workgroups strides
// over the items, so one launch configuration (and one captured hipGraph) serves any
// composition of lattices; the composition lives only in the table, which the host
// rewrites between graph replays when lattices are admitted or retired.
//
// Each lattice's arithmetic and force reduction are exactly those of the single-lattice
// kernel (same step_row, same per-workgroup partials, same fixed-order reduction), so a
// lattice's results are bit-identical alone and in any mix.
//
// Two-step form (sw_batch_step2 / sw_batch_graph_create2): the work items are the 64-cell x
// 12-row tiles of all lattices (tile0 = prefix sum), each launch advances every lattice by
// two steps with the two-step tile function of sw_step2.hpp (one step for a lattice with one
// step left), and the workgroups take tiles from a device counter, so lattices of different
// heights (tiles of different sweep lengths) balance across the compute units.  Force
// partials use the two-step per-tile layout, so a lattice's force rows are bit-identical to
// its single-lattice run with sw_set_step2_rows(sw_batch_step2_rows()) (pairs by sw_step2, an
// odd last step by sw_step2_one); populations are bit-identical to single steps in every case.
// The pair launches keep the next plane's loads in flight across each plane's barrier (the
// tile function's PF form: 12 rows leave it the registers; 51.5 against 49.3 G cell updates/s
// for 16 rows without it on the 64-lattice FP32 mix, where the claimed tiles balance the
// rounds and only the rate per row counts).
#include "sw_step2.hpp"
#include "sw_api_util.hpp"

namespace sw {

constexpr int BATCH2_ROWS = 12;   // tile rows of the two-step batch (BATCH2_ROWS - 2 output rows)

struct TableHeader {
    int n;              // instances in the table
    int total_items;    // sum of nz * ceil(ny / ROWS)
    int max_instances;  // capacity
    int total_tiles;    // sum of the two-step tiles, ceil(nx / 64) * ceil(ny / (BATCH2_ROWS - 2))
    int next_tile;      // two-step work counter: the next unclaimed tile of the running launch
    int finished;       // two-step workgroups of the running launch that found no more tiles
    int pad[10];
};
static_assert(sizeof(TableHeader) == 64, "header is 64 bytes");
static_assert(sizeof(sw_instance) % 8 == 0, "instance entries stay 8-byte aligned");

__device__ __forceinline__ const sw_instance* table_entries(const void* table) {
    return reinterpret_cast<const sw_instance*>(static_cast<const char*>(table) + sizeof(TableHeader));
}

// After the max_instances entries the table holds each instance's step arguments in both
// buffer orientations (args[2i]: f0 -> f1, args[2i + 1]: f1 -> f0), written by sw_batch_upload
// and read by the step kernels through the constant address space: no kernel writes them,
// so the compiler may load a field again where it is needed instead of holding every
// argument of the lattice in registers through the sweep (with the arguments moved to
// registers once per work item the two-step batch kernel spilled 44 VGPRs to scratch).
typedef const __attribute__((address_space(4))) StepArgs CStepArgs;

__device__ __forceinline__ CStepArgs* table_args(const void* table, int max_instances) {
    const char* p = static_cast<const char*>(table) + sizeof(TableHeader) + (size_t)max_instances * sizeof(sw_instance);
    return (CStepArgs*)p;
}

static size_t table_bytes(int max_instances) {
    return sizeof(TableHeader) + (size_t)max_instances * (sizeof(sw_instance) + 2 * sizeof(StepArgs));
}

// the step arguments of instance I with the buffers oriented f_in = odd ? f1 : f0
static StepArgs host_instance_args(const sw_instance& I, bool odd) {
    StepArgs a{};
    a.f_in = odd ? I.f1 : I.f0;
    a.f_out = odd ? I.f0 : I.f1;
    a.links = I.links; a.uin = I.uin; a.gust = I.gust; a.partials = I.partials; a.t_dev = nullptr; a.acc = nullptr;
    a.shift = (I.dtype == DT_F16 || I.dtype == DT_BF16) ? I.shift : nullptr;
    a.t = 0;
    a.ring = I.ring; a.nx = I.nx; a.ny = I.ny; a.nz = I.nz; a.nyb = (I.ny + ROWS - 1) / ROWS;
    a.x_sponge = I.x_sponge; a.ybc = I.ybc; a.nu0 = I.nu0; a.cs = I.cs; a.nu_sponge = I.nu_sponge;
    a.scale = I.scale; a.inv_scale = 1.0f / I.scale;
    return a;
}

// instance owning work item `item` (item0 ascending); all arguments wave-uniform
__device__ __forceinline__ int find_instance(const sw_instance* e, int n, int item) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (e[mid].item0 <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <class S, class TL, bool BUF>
__global__ __launch_bounds__(64 * TL::WPB, TL::OCC) void k_batch_step(const void* __restrict__ table, int k) {
    const TableHeader* h = static_cast<const TableHeader*>(table);
    const sw_instance* e = table_entries(table);
    const int n = __builtin_amdgcn_readfirstlane(h->n), total = __builtin_amdgcn_readfirstlane(h->total_items);
    CStepArgs* args = table_args(table, __builtin_amdgcn_readfirstlane(h->max_instances));
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // a workgroup's items ascend, so after one search its instance only moves forward
    int i = total > 0 ? __builtin_amdgcn_readfirstlane(find_instance(e, n, blockIdx.x)) : 0;
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        while (i + 1 < n && __builtin_amdgcn_readfirstlane(e[i + 1].item0) <= item) i++;
        const sw_instance& I = e[i];
        const int t = __builtin_amdgcn_readfirstlane(I.t) + k;
        if (t >= __builtin_amdgcn_readfirstlane(I.t_end)) continue;   // finished: uniform
        CStepArgs& a = args[2 * i + ((__builtin_amdgcn_readfirstlane(I.parity) + k) & 1)];
        const int local = item - __builtin_amdgcn_readfirstlane(I.item0);
        const int z = local / a.nyb, yb = local - z * a.nyb;
        const int y = yb * ROWS + (wave % ROWS);
        double fx = 0.0, fz = 0.0;
        if (y < a.ny) step_row<S, TL, BUF>(a, y, z, t, lane, wave / ROWS, fx, fz);
        block_force<TL::WPB>(a, fx, fz, z, yb, t);
    }
}

// force rows of steps t .. min(t + K, t_end) - 1 of every instance (before the advance)
__global__ __launch_bounds__(256) void k_batch_force(const void* table, int K) {
    const TableHeader* h = static_cast<const TableHeader*>(table);
    const int i = blockIdx.x, s = blockIdx.y;
    if (i >= h->n) return;
    const sw_instance& I = table_entries(table)[i];
    const int t = I.t + s;
    if (t >= I.t_end || s >= I.ring || !I.partials || !I.force) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nyb = (I.ny + ROWS - 1) / ROWS;
    for (int z = wave; z < I.nz; z += 4) {
        const double* p = I.partials + (((size_t)(t % I.ring) * I.nz + z) * nyb) * 2;
        double sx = 0.0, sz = 0.0;
        for (int q = lane; q < nyb; q += 64) { sx += p[2 * q]; sz += p[2 * q + 1]; }
        sx = wave_sum(sx);
        sz = wave_sum(sz);
        if (lane == 0) {
            I.force[((size_t)t * I.nz + z) * 2 + 0] = sx;
            I.force[((size_t)t * I.nz + z) * 2 + 1] = sz;
        }
    }
}

__global__ void k_batch_advance(void* table, int K) {
    TableHeader* h = static_cast<TableHeader*>(table);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h->n) return;
    sw_instance* e = reinterpret_cast<sw_instance*>(static_cast<char*>(table) + sizeof(TableHeader));
    const int t0 = e[i].t;
    const int t1 = min(t0 + K, e[i].t_end);
    e[i].parity = (e[i].parity + (t1 - t0)) & 1;
    e[i].t = t1;
}

// ---------------------------------------------------------------- two-step form
// first instance whose tiles include `tile` (tile0 ascending)
__device__ __forceinline__ int find_instance_tile(const sw_instance* e, int n, int tile) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (e[mid].tile0 <= tile) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Two-step launches of a chunk of K steps (K even).  Pair launch p (TAIL = false, q = p):
// steps t_i + 2p and t_i + 2p + 1 of every instance with two or more steps left at t_i + 2p.
// Tail launch (TAIL = true, q = K), after the K/2 pair launches: the last step of every
// instance whose remaining steps r = t_end - t_i are fewer than K and odd (its step
// t_i + r - 1, in the buffer orientation of pair launch (r - 1)/2); a separate launch so
// that each kernel holds the registers of one tile sweep only.  The buffers swap once per
// launch that touches an instance.  Persistent workgroups claim tiles from the header's
// counter; the last workgroup to run out resets it for the next launch.  Instances of
// another storage form or with population arrays of 4 GiB or more are skipped (the host
// never puts them in a two-step batch).
template <class S, bool TAIL>
__global__ __launch_bounds__(64 * BATCH2_ROWS, 1) void k_batch_step2(void* table, int q, int dtype) {
    __shared__ Step2Lds<BATCH2_ROWS> L;
    __shared__ int claimed;
    TableHeader* h = static_cast<TableHeader*>(table);
    const sw_instance* e = table_entries(table);
    const int n = __builtin_amdgcn_readfirstlane(h->n), total = __builtin_amdgcn_readfirstlane(h->total_tiles);
    CStepArgs* args = table_args(table, __builtin_amdgcn_readfirstlane(h->max_instances));
    int i = -1;
    for (;;) {
        if (threadIdx.x == 0) claimed = atomicAdd(&h->next_tile, 1);
        __syncthreads();
        const int tile = __builtin_amdgcn_readfirstlane(claimed);
        __syncthreads();                                   // every wavefront has read `claimed`
        if (tile >= total) break;
        // a workgroup's claims ascend, so after one search its instance only moves forward
        if (i < 0) i = __builtin_amdgcn_readfirstlane(find_instance_tile(e, n, tile));
        while (i + 1 < n && __builtin_amdgcn_readfirstlane(e[i + 1].tile0) <= tile) i++;
        const sw_instance& I = e[i];
        if (__builtin_amdgcn_readfirstlane(I.dtype) != dtype) continue;
        const int nx = __builtin_amdgcn_readfirstlane(I.nx), ny = __builtin_amdgcn_readfirstlane(I.ny),
                  nz = __builtin_amdgcn_readfirstlane(I.nz);
        if ((uint64_t)Q * (uint64_t)nx * ny * nz * sizeof(S) >= ((uint64_t)1 << 32)) continue;
        const int t_i = __builtin_amdgcn_readfirstlane(I.t), t_end = __builtin_amdgcn_readfirstlane(I.t_end);
        int p;
        if constexpr (TAIL) {
            const int r = t_end - t_i;
            if (r <= 0 || r >= q || (r & 1) == 0) continue;
            p = (r - 1) / 2;
        } else {
            p = q;
            if (t_end - (t_i + 2 * p) < 2) continue;
        }
        CStepArgs& a = args[2 * i + ((__builtin_amdgcn_readfirstlane(I.parity) + p) & 1)];
        const int ntx = (nx + 63) / 64;
        const int P = step2_tiles_for(nx, ny, BATCH2_ROWS);
        const int local = tile - __builtin_amdgcn_readfirstlane(I.tile0);
        step2_tile<S, BATCH2_ROWS, TAIL, CStepArgs, !TAIL>(a, t_i + 2 * p, ntx, P, local, L);
    }
    if (threadIdx.x == 0) {
        if (atomicAdd(&h->finished, 1) == (int)gridDim.x - 1) {   // every workgroup is past its last claim
            atomicExch(&h->next_tile, 0);
            atomicExch(&h->finished, 0);
        }
    }
}

// force rows of steps t .. min(t + K, t_end) - 1 of every instance from the per-tile partials
__global__ __launch_bounds__(256) void k_batch_force2(const void* table, int K) {
    const TableHeader* h = static_cast<const TableHeader*>(table);
    const int i = blockIdx.x, s = blockIdx.y;
    if (i >= h->n) return;
    const sw_instance& I = table_entries(table)[i];
    const int t = I.t + s;
    if (t >= I.t_end || s >= I.ring || !I.partials || !I.force) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = step2_tiles_for(I.nx, I.ny, BATCH2_ROWS);
    for (int z = wave; z < I.nz; z += 4) force_row2(I.partials, I.ring, t, z, I.nz, P, lane, I.force);
}

// t_i += K (clamped to t_end); the buffers swapped once per pair launch that touched instance i
__global__ void k_batch_advance2(void* table, int K) {
    TableHeader* h = static_cast<TableHeader*>(table);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h->n) return;
    sw_instance* e = reinterpret_cast<sw_instance*>(static_cast<char*>(table) + sizeof(TableHeader));
    const int t0 = e[i].t;
    const int t1 = min(t0 + K, e[i].t_end);
    e[i].parity = (e[i].parity + (t1 - t0 + 1) / 2) & 1;
    e[i].t = t1;
}

// -------------------------------------------------------------------- host side
struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

template <class S, class TL, bool BUF>
static void launch_batch_v(const void* table, int k, int blocks, hipStream_t s) {
    hipLaunchKernelGGL((k_batch_step<S, TL, BUF>), dim3(blocks), dim3(64 * TL::WPB), 0, s, table, k);
}

// variants as sw_set_step_variant; a batch holding a lattice with a population array of
// 4 GiB or more uses 64-bit addressing with nontemporal stores (variant 1's tile)
template <class S>
static void launch_batch(const void* table, int k, int blocks, int variant, bool big, hipStream_t s) {
    if (big) { launch_batch_v<S, TileV1, false>(table, k, blocks, s); return; }
    // pairs: 16- and 32-bit storage (rows of odd length fall back in-kernel)
    if (variant >= 8 && sizeof(S) == 8) variant = 1;
    if ((variant == 12 || variant == 13) && sizeof(S) == 4) variant = 9;
    switch (variant) {
#define SW_CASE(id, TL) case id: launch_batch_v<S, TL, true>(table, k, blocks, s); break;
        SW_TILE_LIST(SW_CASE)
#undef SW_CASE
        default: launch_batch_v<S, TileV1, true>(table, k, blocks, s); break;
    }
}

template <bool TAIL>
static int batch_step2_launch(void* table, int dtype, int q, int blocks, hipStream_t s) {
    const dim3 grid(blocks), block(64 * BATCH2_ROWS);
    switch (dtype) {
        case DT_F32: hipLaunchKernelGGL((k_batch_step2<float, TAIL>), grid, block, 0, s, table, q, dtype); break;
        case DT_F16: hipLaunchKernelGGL((k_batch_step2<_Float16, TAIL>), grid, block, 0, s, table, q, dtype); break;
        case DT_BF16: hipLaunchKernelGGL((k_batch_step2<bf16_t, TAIL>), grid, block, 0, s, table, q, dtype); break;
    }
    return (int)hipGetLastError();
}

static int batch_step_launch(const void* table, int dtype, int k, int blocks, int variant, int big, hipStream_t s) {
    switch (dtype) {
        case DT_F32: launch_batch<float>(table, k, blocks, variant, big, s); break;
        case DT_F64: launch_batch<double>(table, k, blocks, variant, big, s); break;
        case DT_F16: launch_batch<_Float16>(table, k, blocks, variant, big, s); break;
        case DT_BF16: launch_batch<bf16_t>(table, k, blocks, variant, big, s); break;
    }
    return (int)hipGetLastError();
}

}  // namespace sw

using namespace sw;

extern "C" {

long long sw_batch_table_bytes(int max_instances) {
    if (max_instances <= 0) return -1;
    return (long long)table_bytes(max_instances);
}

int sw_batch_upload(const sw_instance* instances, int n, int max_instances, void* table, void* stream) {
    SW_CHECK_ARG(table && (n == 0 || instances), "table and instances must be non-null");
    SW_CHECK_ARG(n >= 0 && n <= max_instances, "need 0 <= n <= max_instances");
    const size_t bytes = sizeof(TableHeader) + (size_t)n * sizeof(sw_instance);
    const size_t arg_bytes = (size_t)n * 2 * sizeof(StepArgs);
    char* host = (char*)malloc(bytes + arg_bytes);
    if (!host) return set_error(SW_ERR_CAPACITY, "sw_batch_upload: host allocation failed");
    StepArgs* args = (StepArgs*)(host + bytes);
    TableHeader h{};
    h.n = n;
    h.max_instances = max_instances;
    sw_instance* e = (sw_instance*)(host + sizeof(TableHeader));
    long long items = 0, tiles = 0;
    for (int i = 0; i < n; i++) {
        e[i] = instances[i];
        const sw_instance& I = e[i];
        int rc = check_dims(I.nx, I.ny, I.nz);
        if (rc) { free(host); return rc; }
        if (!I.f0 || !I.f1 || !I.links || !I.uin || I.f0 == I.f1) {
            free(host);
            return set_error(SW_ERR_INVALID, "sw_batch_upload: instance buffers missing or aliased");
        }
        if (I.dtype < 0 || I.dtype > 3 || check_dtype(I.dtype, I.scale)) {
            free(host);
            return set_error(SW_ERR_INVALID, "sw_batch_upload: bad instance dtype / scale");
        }
        if ((I.partials && I.ring <= 0) || I.t < 0 || I.t > I.t_end || (I.ybc != 0 && I.ybc != 1)) {
            free(host);
            return set_error(SW_ERR_INVALID, "sw_batch_upload: bad ring / t / t_end / ybc");
        }
        e[i].item0 = (int)items;
        items += (long long)I.nz * ((I.ny + ROWS - 1) / ROWS);
        e[i].tile0 = (int)tiles;
        tiles += step2_tiles_for(I.nx, I.ny, BATCH2_ROWS);
        args[2 * i] = host_instance_args(I, false);
        args[2 * i + 1] = host_instance_args(I, true);
    }
    if (items > 0x7FFFFFFFLL || tiles > 0x7FFFFFFFLL) {
        free(host);
        return set_error(SW_ERR_TOO_LARGE, "sw_batch_upload: too many rows");
    }
    h.total_items = (int)items;
    h.total_tiles = (int)tiles;
    memcpy(host, &h, sizeof(h));
    hipStream_t s = (hipStream_t)stream;
    hipError_t err = hipMemcpyAsync(table, host, bytes, hipMemcpyHostToDevice, s);
    if (err == hipSuccess && arg_bytes)
        err = hipMemcpyAsync((char*)table + table_bytes(max_instances) - (size_t)max_instances * 2 * sizeof(StepArgs),
                             args, arg_bytes, hipMemcpyHostToDevice, s);
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    free(host);
    if (err != hipSuccess) return set_hip_error(err, "sw_batch_upload");
    return SW_OK;
}

int sw_batch_download(const void* table, sw_instance* instances, int n, void* stream) {
    SW_CHECK_ARG(table && (n == 0 || instances) && n >= 0, "table and instances must be non-null");
    hipStream_t s = (hipStream_t)stream;
    SW_HIP(hipMemcpyAsync(instances, (const char*)table + sizeof(TableHeader), (size_t)n * sizeof(sw_instance),
                          hipMemcpyDeviceToHost, s));
    SW_HIP(hipStreamSynchronize(s));
    return SW_OK;
}

int sw_batch_step(const void* table, int dtype, int k, int blocks, int variant, int big_offsets, void* stream) {
    SW_CHECK_ARG(table, "table must be non-null");
    SW_CHECK_ARG(dtype >= 0 && dtype <= 3, "dtype must be 0..3");
    SW_CHECK_ARG(blocks > 0 && k >= 0 && variant >= 0 && variant < NUM_TILE_VARIANTS, "need blocks > 0, k >= 0 and a listed variant");
    int rc = batch_step_launch(table, dtype, k, blocks, variant, big_offsets != 0, (hipStream_t)stream);
    if (rc) return set_hip_error((hipError_t)rc, "sw_batch_step");
    return SW_OK;
}

int sw_batch_force_reduce(const void* table, int max_instances, int K, void* stream) {
    SW_CHECK_ARG(table && max_instances > 0 && K > 0 && K <= 65535, "bad table / max_instances / K");
    hipLaunchKernelGGL(k_batch_force, dim3(max_instances, K), dim3(256), 0, (hipStream_t)stream, table, K);
    SW_HIP(hipGetLastError());
    return SW_OK;
}

int sw_batch_advance(void* table, int max_instances, int K, void* stream) {
    SW_CHECK_ARG(table && max_instances > 0 && K > 0, "bad table / max_instances / K");
    hipLaunchKernelGGL(k_batch_advance, dim3((max_instances + 255) / 256), dim3(256), 0, (hipStream_t)stream, table,
                       K);
    SW_HIP(hipGetLastError());
    return SW_OK;
}

int sw_batch_graph_create(void* table, int max_instances, int dtype, int K, int blocks, int variant, int big_offsets,
                          void** graph) {
    SW_CHECK_ARG(table && graph && max_instances > 0 && K > 0 && blocks > 0, "bad arguments");
    SW_CHECK_ARG(dtype >= 0 && dtype <= 3 && variant >= 0 && variant < NUM_TILE_VARIANTS, "bad dtype / variant");
    *graph = nullptr;
    hipStream_t cs = nullptr;
    SW_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    BatchGraph* g = new BatchGraph();
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int rc = 0;
    if (e == hipSuccess) {
        for (int k = 0; k < K && !rc; k++) rc = batch_step_launch(table, dtype, k, blocks, variant, big_offsets != 0, cs);
        if (!rc) {
            hipLaunchKernelGGL(k_batch_force, dim3(max_instances, K), dim3(256), 0, cs, (const void*)table, K);
            hipLaunchKernelGGL(k_batch_advance, dim3((max_instances + 255) / 256), dim3(256), 0, cs, table, K);
            rc = (int)hipGetLastError();
        }
        hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
        if (e == hipSuccess) e = e2;
    }
    if (e == hipSuccess && rc == 0) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess || rc != 0) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return set_hip_error(e != hipSuccess ? e : (hipError_t)rc, "sw_batch_graph_create");
    }
    *graph = g;
    return SW_OK;
}

int sw_batch_step2_rows(void) { return BATCH2_ROWS; }

long long sw_batch_partials_per_step(int nx, int ny, int nz) {
    if (check_dims(nx, ny, nz)) return -1;
    const long long one = (long long)nz * ((ny + ROWS - 1) / ROWS) * 2;
    const long long two = (long long)nz * step2_tiles_for(nx, ny, BATCH2_ROWS) * 2;
    return one > two ? one : two;
}

int sw_batch_step2(void* table, int dtype, int p, int blocks, void* stream) {
    SW_CHECK_ARG(table, "table must be non-null");
    SW_CHECK_ARG(dtype == DT_F32 || dtype == DT_F16 || dtype == DT_BF16, "the two-step batch takes FP32, FP16 or BF16");
    SW_CHECK_ARG(blocks > 0 && p >= 0, "need blocks > 0 and p >= 0");
    int rc = batch_step2_launch<false>(table, dtype, p, blocks, (hipStream_t)stream);
    if (rc) return set_hip_error((hipError_t)rc, "sw_batch_step2");
    return SW_OK;
}

int sw_batch_step2_tail(void* table, int dtype, int K, int blocks, void* stream) {
    SW_CHECK_ARG(table, "table must be non-null");
    SW_CHECK_ARG(dtype == DT_F32 || dtype == DT_F16 || dtype == DT_BF16, "the two-step batch takes FP32, FP16 or BF16");
    SW_CHECK_ARG(blocks > 0 && K > 0 && K % 2 == 0, "need blocks > 0 and an even K");
    int rc = batch_step2_launch<true>(table, dtype, K, blocks, (hipStream_t)stream);
    if (rc) return set_hip_error((hipError_t)rc, "sw_batch_step2_tail");
    return SW_OK;
}

int sw_batch_force_reduce2(const void* table, int max_instances, int K, void* stream) {
    SW_CHECK_ARG(table && max_instances > 0 && K > 0 && K <= 65535, "bad table / max_instances / K");
    hipLaunchKernelGGL(k_batch_force2, dim3(max_instances, K), dim3(256), 0, (hipStream_t)stream, table, K);
    SW_HIP(hipGetLastError());
    return SW_OK;
}

int sw_batch_advance2(void* table, int max_instances, int K, void* stream) {
    SW_CHECK_ARG(table && max_instances > 0 && K > 0 && K % 2 == 0, "bad table / max_instances / K (K even)");
    hipLaunchKernelGGL(k_batch_advance2, dim3((max_instances + 255) / 256), dim3(256), 0, (hipStream_t)stream, table,
                       K);
    SW_HIP(hipGetLastError());
    return SW_OK;
}

int sw_batch_graph_create2(void* table, int max_instances, int dtype, int K, int blocks, void** graph) {
    SW_CHECK_ARG(table && graph && max_instances > 0 && K > 0 && K % 2 == 0 && K <= 65535 && blocks > 0,
                 "bad arguments (K must be even)");
    SW_CHECK_ARG(dtype == DT_F32 || dtype == DT_F16 || dtype == DT_BF16, "the two-step batch takes FP32, FP16 or BF16");
    *graph = nullptr;
    hipStream_t cs = nullptr;
    SW_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    BatchGraph* g = new BatchGraph();
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    int rc = 0;
    if (e == hipSuccess) {
        for (int p = 0; p < K / 2 && !rc; p++) rc = batch_step2_launch<false>(table, dtype, p, blocks, cs);
        if (!rc) rc = batch_step2_launch<true>(table, dtype, K, blocks, cs);
        if (!rc) {
            hipLaunchKernelGGL(k_batch_force2, dim3(max_instances, K), dim3(256), 0, cs, (const void*)table, K);
            hipLaunchKernelGGL(k_batch_advance2, dim3((max_instances + 255) / 256), dim3(256), 0, cs, table, K);
            rc = (int)hipGetLastError();
        }
        hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
        if (e == hipSuccess) e = e2;
    }
    if (e == hipSuccess && rc == 0) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess || rc != 0) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return set_hip_error(e != hipSuccess ? e : (hipError_t)rc, "sw_batch_graph_create2");
    }
    *graph = g;
    return SW_OK;
}

int sw_graph_launch(void* graph, void* stream) {
    SW_CHECK_ARG(graph, "graph must be non-null");
    SW_HIP(hipGraphLaunch(static_cast<BatchGraph*>(graph)->exec, (hipStream_t)stream));
    return SW_OK;
}

int sw_graph_destroy(void* graph) {
    if (!graph) return SW_OK;
    BatchGraph* g = static_cast<BatchGraph*>(graph);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return SW_OK;
}

}  // extern "C"
