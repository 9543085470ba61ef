// Single-lattice kernels and their C-ABI entry points (declared in include/stonewall_lbm.h).
#include "sw_step.hpp"
#include "sw_api_util.hpp"

namespace sw {

// ------------------------------------------------------------------ step
template <class S, class TL, bool BUF, bool ACC = false>
__global__ __launch_bounds__(64 * TL::WPB, TL::OCC) void k_step(StepArgs a) {
    // the wave index is uniform per wavefront; readfirstlane lets the compiler keep every
    // row-derived pointer in scalar registers
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int z = blockIdx.x / a.nyb;
    const int yb = blockIdx.x - z * a.nyb;
    const int y = yb * ROWS + (wave % ROWS);
    const int t = a.t + (a.t_dev ? *a.t_dev : 0);
    double fx = 0.0, fz = 0.0;
    if (y < a.ny) step_row<S, TL, BUF, ACC>(a, y, z, t, lane, wave / ROWS, fx, fz);
    block_force<TL::WPB>(a, fx, fz, z, yb, t);
}

// ------------------------------------------------------------------ link masks
// flag is in the reference layout (nx, ny, nz), z contiguous (lbm3d.py:288-290).
__global__ void k_prepare(const uint8_t* __restrict__ flag, uint32_t* __restrict__ links, int nx, int ny, int nz,
                          int ybc, int* err) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= nx) return;
    auto F = [&](int xx, int yy, int zz) { return flag[((size_t)xx * ny + yy) * nz + zz]; };
    uint32_t m = 0;
    if (F(x, y, z) != 0) {
        m = SOLID_BIT;
    } else if (z == 0) {
        atomicOr(err, 1);                     // fluid in the bottom layer: its pull would leave the grid
    } else if (x > 0) {                       // x == 0 is the inlet and does not pull
        for (int i = 1; i < Q; i++) {
            int xs = x - CX[i], ys = y - CY[i], zs = z - CZ[i];
            if (ybc == Y_PERIODIC) ys = (ys + ny) % ny;
            else if (ys < 0 || ys >= ny) ys = y;
            if (xs >= nx) xs = x;
            if (zs >= nz) zs = z;
            const uint8_t fs = F(xs, ys, zs);
            if (fs != 0) {
                m |= 1u << i;
                if (fs != 1) {
                    if (GROUND_SLOT[i] >= 0) m |= 1u << (GROUND_BIT0 + GROUND_SLOT[i]);
                    else atomicOr(err, 2);    // a non-wall solid reached by a link that is not cz=+1
                }
            }
        }
    }
    links[((size_t)z * ny + y) * nx + x] = m;
}

// chunk flags: one thread per (row, 32-chunk word); bit c set when 64-cell chunk c of
// the row holds a solid cell or a cell with a solid neighbour (the inlet cell excluded)
__global__ void k_chunk_flags(uint32_t* __restrict__ links, int nx, int ny, int nz) {
    const uint32_t cw = chunk_words(nx);
    const size_t gid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (gid >= (size_t)ny * nz * cw) return;
    const uint32_t row = (uint32_t)(gid / cw), word = (uint32_t)(gid % cw);
    const uint32_t* lrow = links + (size_t)row * nx;
    uint32_t bits = 0;
    for (uint32_t c = 0; c < 32; c++) {
        const uint32_t x0 = (word * 32 + c) * 64;
        if (x0 >= (uint32_t)nx) break;
        for (uint32_t x = x0; x < x0 + 64 && x < (uint32_t)nx; x++)
            if (x != 0 && lrow[x] != 0u) { bits |= 1u << c; break; }
    }
    links[(size_t)nx * ny * nz + gid] = bits;
}

// ------------------------------------------------------------------ init (lbm3d.py:39-54)
template <class S>
__global__ void k_init(S* __restrict__ f, const uint32_t* __restrict__ links, const float* __restrict__ uin,
                       int nx, int ny, int nz, float noise, uint32_t seed, float scale, const float* __restrict__ shift) {
    using C = Codec<S>;
    using T = typename C::T;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= nx) return;
    const uint32_t N = (uint32_t)nx * ny * nz;
    const uint32_t self = ((uint32_t)z * ny + y) * nx + x;
    T ux = T(0), uy = T(0), uz = T(0);
    if (!(links[self] & SOLID_BIT)) {
        const uint32_t key = ((uint32_t)x * 7919u + (uint32_t)y) * 7907u + (uint32_t)z;
        const uint32_t h0 = pcg_hash(seed ^ pcg_hash(key));
        const uint32_t h1 = pcg_hash(h0), h2 = pcg_hash(h1);
        const T u0 = (T)uin[z], nz_ = (T)noise;
        ux = u0 * (T(1) + nz_ * (T(2) * (T)u01(h0) - T(1)));
        uy = (u0 * nz_) * (T(2) * (T)u01(h1) - T(1));
        uz = (u0 * nz_) * (T(2) * (T)u01(h2) - T(1));
    }
    const T usq15 = T(1.5) * ((ux * ux + uy * uy) + uz * uz);
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        if constexpr (sizeof(S) == 2) {         // 16-bit forms: phi_i = feq_i / w_i, offset sigma
            const T sg = shift ? (T)shift[z * 4 + xclass(CX[i])] : T(1);
            C::store(f, (size_t)i * N + self, feq_norm<i, T>(ux, uy, uz, usq15), sg, (T)scale);
        } else {
            C::store(f, (size_t)i * N + self, feq<i, T>(T(1), ux, uy, uz, usq15), wgt<T>(i), (T)scale);
        }
    });
}

// ------------------------------------------------------------------ macro + accumulate (lbm3d.py:158-188)
template <class S>
__global__ void k_macro(const S* __restrict__ f, const uint32_t* __restrict__ links, int nx, int ny, int nz,
                        float inv_scale, const float* __restrict__ shift, float* __restrict__ U, float* __restrict__ acc) {
    using C = Codec<S>;
    using T = typename C::T;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= nx) return;
    const uint32_t N = (uint32_t)nx * ny * nz;
    const uint32_t self = ((uint32_t)z * ny + y) * nx + x;
    float u[4];
    if (links[self] & SOLID_BIT) {
        u[0] = 1.0f; u[1] = 0.0f; u[2] = 0.0f; u[3] = 0.0f;
    } else {
        T rho = T(0), mx = T(0), my = T(0), mz = T(0);
        if constexpr (sizeof(S) == 2) {         // phi_i = f_i / w_i: the step's pair moments
            T ph[Q];
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                const T sg = shift ? (T)shift[z * 4 + xclass(CX[i])] : T(1);
                ph[i] = C::load(f, (size_t)i * N + self, sg, (T)inv_scale);
            });
            const Moments<T> m = moments_norm<T, false>(ph);
            rho = m.rho; mx = m.mx; my = m.my; mz = m.mz;
        } else {
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                const T v = C::load(f, (size_t)i * N + self, wgt<T>(i), (T)inv_scale);
                rho = rho + v;
                if constexpr (CX[i] != 0) mx = CX[i] > 0 ? mx + v : mx - v;
                if constexpr (CY[i] != 0) my = CY[i] > 0 ? my + v : my - v;
                if constexpr (CZ[i] != 0) mz = CZ[i] > 0 ? mz + v : mz - v;
            });
        }
        u[0] = (float)rho; u[1] = (float)(mx / rho); u[2] = (float)(my / rho); u[3] = (float)(mz / rho);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (U) U[(size_t)k * N + self] = u[k];
        if (acc) acc[(size_t)k * N + self] = acc[(size_t)k * N + self] + u[k];
    }
}

// ------------------------------------------------------------------ Q-criterion frames (lbm3d.py:203-238)
// U is [4][z][y][x] (rho, ux, uy, uz) as sw_macro writes it; a cell is solid when its
// link mask has SOLID_BIT.  One thread per output voxel (i, j, k) of the (qx, qy, qz)
// frame; Q = -1/2 du_i/dx_j du_j/dx_i by central differences, 0 next to solids, max-pooled
// (positive part) over the ds^3 block at (x0 + i*ds, j*ds, k*ds).  Operand order as q_at.
__global__ void k_qcoarse(const float* __restrict__ U, const uint32_t* __restrict__ links, int nx, int ny, int nz,
                          int x0, int ds, int qx, int qy, int qz, float* __restrict__ out) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= qx * qy * qz) return;
    const int k = gid % qz, j = (gid / qz) % qy, i = gid / (qz * qy);
    const size_t N = (size_t)nx * ny * nz;
    auto C = [&](int x, int y, int z) { return ((size_t)z * ny + y) * nx + x; };
    auto solid = [&](int x, int y, int z) { return (links[C(x, y, z)] & SOLID_BIT) != 0u; };
    const float* u1 = U + N;
    const float* u2 = U + 2 * N;
    const float* u3 = U + 3 * N;
    float best = 0.0f;
    for (int a = 0; a < ds; a++)
        for (int b = 0; b < ds; b++)
            for (int c = 0; c < ds; c++) {
                const int x = x0 + i * ds + a, y = j * ds + b, z = k * ds + c;
                if (!(x >= 1 && x < nx - 1 && z >= 1 && z < nz - 1)) continue;
                const int yp = (y + 1) % ny, ym = (y - 1 + ny) % ny;
                float q = 0.0f;
                if (!(solid(x, y, z) || solid(x + 1, y, z) || solid(x - 1, y, z) || solid(x, yp, z) ||
                      solid(x, ym, z) || solid(x, y, z + 1) || solid(x, y, z - 1))) {
                    const float a11 = 0.5f * (u1[C(x + 1, y, z)] - u1[C(x - 1, y, z)]);
                    const float a12 = 0.5f * (u1[C(x, yp, z)] - u1[C(x, ym, z)]);
                    const float a13 = 0.5f * (u1[C(x, y, z + 1)] - u1[C(x, y, z - 1)]);
                    const float a21 = 0.5f * (u2[C(x + 1, y, z)] - u2[C(x - 1, y, z)]);
                    const float a22 = 0.5f * (u2[C(x, yp, z)] - u2[C(x, ym, z)]);
                    const float a23 = 0.5f * (u2[C(x, y, z + 1)] - u2[C(x, y, z - 1)]);
                    const float a31 = 0.5f * (u3[C(x + 1, y, z)] - u3[C(x - 1, y, z)]);
                    const float a32 = 0.5f * (u3[C(x, yp, z)] - u3[C(x, ym, z)]);
                    const float a33 = 0.5f * (u3[C(x, y, z + 1)] - u3[C(x, y, z - 1)]);
                    q = -0.5f * (((a11 * a11 + a22 * a22) + a33 * a33) + 2.0f * ((a12 * a21 + a13 * a31) + a23 * a32));
                }
                best = q > best ? q : best;
            }
    out[((size_t)i * qy + j) * qz + k] = best;
}

// ------------------------------------------------------------------ storage offsets of the 16-bit forms
// shift[z][c] = sigma_c(z) = feq_i(1, (U, 0, 0)) / w_i for a direction of c_x class c (0, +1,
// -1), U = uin[z] (kind SHIFT_INLET) or 0 (SHIFT_WEIGHTS: sigma = 1); shift[z][3] = 0.  The
// 16-bit forms store (f_i / w_i - sigma) * scale, small wherever the flow is close to the
// inlet profile.  Evaluated as the step evaluates the equilibrium (lbm3d.py:35-36 order).
__global__ void k_shift_table(const float* __restrict__ uin, int nz, int kind, float* __restrict__ shift) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nz) return;
    const float u0 = kind == SHIFT_INLET ? uin[z] : 0.0f;
    const float usq15 = 1.5f * (u0 * u0);
    shift[z * 4 + 0] = feq_norm<3, float>(u0, 0.0f, 0.0f, usq15);     // c = (0, 1, 0): c_x = 0
    shift[z * 4 + 1] = feq_norm<1, float>(u0, 0.0f, 0.0f, usq15);     // c_x = +1
    shift[z * 4 + 2] = feq_norm<2, float>(u0, 0.0f, 0.0f, usq15);     // c_x = -1
    shift[z * 4 + 3] = 0.0f;
}

// ------------------------------------------------------------------ force rows
// force[t][z][c] = sum over yb of partials[t % ring][z][yb][c], in a fixed order:
// lane-strided sums, then a fixed butterfly; lane 0's value is written.
__global__ void k_force_reduce(const double* __restrict__ partials, int ring, int t0, const int* t_dev, int nz,
                               int nyb, double* __restrict__ force) {
    const int s = blockIdx.y;                 // step within the batch
    const int z = blockIdx.x;
    const int lane = threadIdx.x;
    const int t = t0 + (t_dev ? *t_dev : 0) + s;
    const double* p = partials + (((size_t)(t % ring) * nz + z) * nyb) * 2;
    double sx = 0.0, sz = 0.0;
    for (int k = lane; k < nyb; k += 64) { sx += p[2 * k]; sz += p[2 * k + 1]; }
    sx = wave_sum(sx);
    sz = wave_sum(sz);
    if (lane == 0) {
        force[((size_t)t * nz + z) * 2 + 0] = sx;
        force[((size_t)t * nz + z) * 2 + 1] = sz;
    }
}

// ------------------------------------------------------------------ layout transposes
// [c][z][y][x]  <->  [c][x][y][z]; one (c, y) plane per blockIdx.z, 32x32 LDS tiles.
template <class E, bool TO_REF>
__global__ void k_transpose(const E* __restrict__ src, E* __restrict__ dst, int nx, int ny, int nz) {
    __shared__ E tile[32][33];
    const int cy = blockIdx.z;
    const int c = cy / ny, y = cy - c * ny;
    const size_t N = (size_t)nx * ny * nz;
    const int tx = threadIdx.x, ty = threadIdx.y;       // 32 x 8
    const int bx = blockIdx.x * 32, bz = blockIdx.y * 32;
    // internal index (z, y, x) = (z*ny + y)*nx + x ; reference index (x, y, z) = (x*ny + y)*nz + z
    if (TO_REF) {
        for (int r = ty; r < 32; r += 8) {                  // read along x
            const int x = bx + tx, z = bz + r;
            if (x < nx && z < nz) tile[r][tx] = src[c * N + ((size_t)z * ny + y) * nx + x];
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {                  // write along z
            const int z = bz + tx, x = bx + r;
            if (x < nx && z < nz) dst[c * N + ((size_t)x * ny + y) * nz + z] = tile[tx][r];
        }
    } else {
        for (int r = ty; r < 32; r += 8) {                  // read along z
            const int z = bz + tx, x = bx + r;
            if (x < nx && z < nz) tile[tx][r] = src[c * N + ((size_t)x * ny + y) * nz + z];
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {                  // write along x
            const int x = bx + tx, z = bz + r;
            if (x < nx && z < nz) dst[c * N + ((size_t)z * ny + y) * nx + x] = tile[r][tx];
        }
    }
}

// ------------------------------------------------------------------ division self-test
// counts[0]: normal b (with a normal reciprocal) where rcp_rn(b) != IEEE 1/b, over all 2^32
// bit patterns; counts[1]: hashed (a, b) samples, |a| in [2^-40, 4), b in [2^-4, 4), where
// div_rn(a, b, rcp_rn(b)) != IEEE a/b.
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

__global__ void k_zero_counts(unsigned long long* counts) {
    if (threadIdx.x < 2) counts[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void k_selftest_division(unsigned long long* counts) {
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bad_rcp = 0, bad_div = 0;
    for (uint64_t u = tid; u < (1ull << 32); u += stride) {
        const uint32_t bits = (uint32_t)u;
        const uint32_t ex = (bits >> 23) & 0xFFu;
        if (ex != 0 && ex < 253) {
            const float b = __uint_as_float(bits);
            if (__float_as_uint(rcp_rn(b)) != __float_as_uint(1.0f / b)) bad_rcp++;
        }
        const uint32_t h1 = mix32(bits ^ 0x9e3779b9u), h2 = mix32(bits * 2654435761u + 0x7f4a7c15u);
        const float a = __uint_as_float((h1 & 0x807FFFFFu) | ((87u + (h1 >> 24) % 42u) << 23));
        const float b = __uint_as_float((h2 & 0x007FFFFFu) | ((123u + (h2 >> 28) % 6u) << 23));
        if (__float_as_uint(div_rn(a, b, rcp_rn(b))) != __float_as_uint(a / b)) bad_div++;
    }
    if (bad_rcp) atomicAdd(&counts[0], bad_rcp);
    if (bad_div) atomicAdd(&counts[1], bad_div);
}

// ------------------------------------------------------------------ streaming copy (bandwidth ceiling)
// dst = src over n16 16-byte words.  A workgroup copies contiguous 16 KB blocks (four
// 16-byte words per lane, all four loads issued before the nontemporal stores) and strides
// over the blocks; 32 workgroups per compute unit.  The same read + write stream a step
// moves, with no arithmetic and no neighbour access: the HBM ceiling the step is compared
// against.  (Fastest of the copy shapes measured on MI355X: grid-stride or blocked, one to
// eight words in flight per lane, plain or nontemporal loads and stores, 1-32 workgroups
// per compute unit; 5.45-5.53 TB/s at best against 4.6 TB/s for one grid-stride word per lane.)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int COPY_WORDS = 4;
__global__ __launch_bounds__(256) void k_stream_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                     size_t n16) {
    constexpr size_t B = 256 * COPY_WORDS;
    for (size_t b = (size_t)blockIdx.x * B; b < n16; b += (size_t)gridDim.x * B) {
        u32x4 v[COPY_WORDS];
#pragma unroll
        for (int u = 0; u < COPY_WORDS; u++) {
            const size_t k = b + u * 256 + threadIdx.x;
            if (k < n16) v[u] = src[k];
        }
#pragma unroll
        for (int u = 0; u < COPY_WORDS; u++) {
            const size_t k = b + u * 256 + threadIdx.x;
            if (k < n16) __builtin_nontemporal_store(v[u], &dst[k]);
        }
    }
}

}  // namespace sw

using namespace sw;

// Step variants (the autotuner's choices, per dtype): the tile list of sw_step.hpp.
// Lattices whose byte offsets exceed 32 bits use 64-bit addressing with nontemporal stores.
static int g_step_variant[4] = {9, 9, 9, 1};   // measured best on MI355X, 640x320x160 (f32, f16, bf16, f64)

template <class S, class TL, bool BUF, bool ACC = false>
static void launch_step_v(const StepArgs& a, hipStream_t s) {
    dim3 grid(a.nz * a.nyb), blk(64 * TL::WPB);
    hipLaunchKernelGGL((k_step<S, TL, BUF, ACC>), grid, blk, 0, s, a);
}

// raw-buffer addressing needs every byte offset of a population array below 4 GiB
static bool fits_buffer(const StepArgs& a, size_t elem) {
    return (size_t)Q * a.nx * a.ny * a.nz * elem < ((size_t)1 << 32);
}

template <class S>
static void launch_step(const StepArgs& a, int dtype, hipStream_t s) {
    // 64-bit addressing (arrays of 4 GiB and more): plain sweep with nontemporal stores
    if (!fits_buffer(a, sizeof(S))) {
        if (a.acc) launch_step_v<S, TileV1, false, true>(a, s);
        else launch_step_v<S, TileV1, false>(a, s);
        return;
    }
    int v = g_step_variant[dtype];
    if (a.acc) {
        // the accumulate epilogue is built for the pair tiles (16/32-bit storage) and the
        // one-cell tile; every tile gives the same populations
        if (sizeof(S) == 8 || (v != 8 && v != 9)) v = sizeof(S) == 8 ? 1 : 9;
        if (v == 9 && (sizeof(S) == 2 ? a.nx < 512 : a.nx <= 128)) v = 8;
        if (v == 1) launch_step_v<S, TileV1, true, true>(a, s);
        else if (v == 8) launch_step_v<S, TileV8, true, true>(a, s);
        else launch_step_v<S, TileV9, true, true>(a, s);
        return;
    }
    // Variant 9 splits each row's 128-cell pair chunks over two wavefronts; with only a few
    // chunks per row one of them idles or gets the smaller share, and one wavefront per row
    // (variant 8) is faster: measured for 16-bit rows below 512 cells and FP32 rows of one
    // chunk (lattices alone, 20 % void, H = 8 ... 40).
    if (v == 9 && (sizeof(S) == 2 ? a.nx < 512 : a.nx <= 128)) v = 8;
    if (v >= 8 && sizeof(S) == 8) v = 1;    // pairs: 16/32-bit storage (odd nx falls back in-kernel)
    if ((v == 12 || v == 13) && sizeof(S) == 4) v = 9;   // quads: 16-bit storage only
    switch (v) {
#define SW_CASE(id, TL) case id: launch_step_v<S, TL, true>(a, s); break;
        SW_TILE_LIST(SW_CASE)
#undef SW_CASE
        default: launch_step_v<S, TileV1, true>(a, s); break;
    }
}

namespace sw {

int step_entry(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin, const float* gust,
               int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs, int x_sponge, double nu_sponge,
               int ybc, float scale, const float* shift, double* partials, int ring, hipStream_t s, float* acc) {
    SW_CHECK_ARG(f_in && f_out && links && uin, "f_in, f_out, links and uin must be non-null");
    SW_CHECK_ARG(f_in != f_out, "f_in and f_out must be different buffers (pull scheme)");
    SW_CHECK_RC(check_dims(nx, ny, nz));
    SW_CHECK_RC(check_dtype(dtype, scale));
    SW_CHECK_ARG(ybc == Y_PERIODIC || ybc == Y_FREESLIP, "ybc must be 0 (periodic) or 1 (free-slip)");
    SW_CHECK_ARG(!partials || ring > 0, "ring must be positive when partials are given");
    SW_CHECK_ARG(t >= 0, "t must be non-negative");
    StepArgs a{};
    a.f_in = f_in; a.f_out = f_out; a.links = links; a.uin = uin; a.gust = gust;
    a.partials = partials; a.t_dev = t_dev; a.t = t; a.ring = ring > 0 ? ring : 1;
    a.shift = (dtype == DT_F16 || dtype == DT_BF16) ? shift : nullptr;
    a.nx = nx; a.ny = ny; a.nz = nz; a.nyb = (ny + ROWS - 1) / ROWS;
    a.x_sponge = x_sponge; a.ybc = ybc; a.nu0 = nu0; a.cs = cs; a.nu_sponge = nu_sponge;
    a.scale = scale; a.inv_scale = 1.0f / scale;
    a.acc = acc;
    switch (dtype) {
        case DT_F32: launch_step<float>(a, dtype, s); break;
        case DT_F64: launch_step<double>(a, dtype, s); break;
        case DT_F16: launch_step<_Float16>(a, dtype, s); break;
        case DT_BF16: launch_step<bf16_t>(a, dtype, s); break;
    }
    SW_HIP(hipGetLastError());
    return SW_OK;
}

int force_reduce_entry(const double* partials, int ring, int t0, const int* t_dev, int nsteps, int ny, int nz,
                       double* force, hipStream_t s) {
    SW_CHECK_ARG(partials && force, "partials and force must be non-null");
    SW_CHECK_ARG(ring > 0 && nsteps > 0 && nsteps <= ring, "need 0 < nsteps <= ring");
    SW_CHECK_ARG(ny > 0 && nz > 0 && t0 >= 0, "bad ny/nz/t0");
    const int nyb = (ny + ROWS - 1) / ROWS;
    hipLaunchKernelGGL(k_force_reduce, dim3(nz, nsteps), dim3(64), 0, s, partials, ring, t0, t_dev, nz, nyb, force);
    SW_HIP(hipGetLastError());
    return SW_OK;
}

}  // namespace sw

extern "C" {

int sw_version(void) { return SW_VERSION; }

const char* sw_last_error(void) { return last_error_message(); }

int sw_rows_per_block(void) { return ROWS; }

long long sw_links_words(int nx, int ny, int nz) {
    if (check_dims(nx, ny, nz)) return -1;
    return (long long)nx * ny * nz + (long long)ny * nz * chunk_words(nx);
}

long long sw_partials_per_step(int ny, int nz) {
    if (ny <= 0 || nz <= 0) return -1;
    return (long long)nz * ((ny + ROWS - 1) / ROWS) * 2;
}

int sw_prepare_links(const uint8_t* flag, uint32_t* links, int nx, int ny, int nz, int ybc, void* stream) {
    SW_CHECK_ARG(flag && links, "flag and links must be non-null");
    SW_CHECK_RC(check_dims(nx, ny, nz));
    SW_CHECK_ARG(ybc == Y_PERIODIC || ybc == Y_FREESLIP, "ybc must be 0 (periodic) or 1 (free-slip)");
    hipStream_t s = (hipStream_t)stream;
    int* err = nullptr;
    SW_HIP(hipMalloc(&err, sizeof(int)));
    int herr = 0;
    hipError_t e = hipMemsetAsync(err, 0, sizeof(int), s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_prepare, dim3((nx + 255) / 256, ny, nz), dim3(256), 0, s, flag, links, nx, ny, nz, ybc, err);
        const size_t nthreads = (size_t)ny * nz * chunk_words(nx);
        hipLaunchKernelGGL(k_chunk_flags, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, s, links, nx, ny, nz);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(err);
    if (e != hipSuccess) return set_hip_error(e, "sw_prepare_links");
    if (herr & 1) return set_error(SW_ERR_GEOMETRY, "sw_prepare_links: the bottom layer z=0 must be solid (ground)");
    if (herr & 2)
        return set_error(SW_ERR_GEOMETRY,
                         "sw_prepare_links: a non-wall solid (flag>=2) is reached by a link other than cz=+1; "
                         "only the ground layer may carry flag 2");
    return SW_OK;
}

int sw_init(void* f, int dtype, const uint32_t* links, const float* uin, int nx, int ny, int nz, float noise,
            unsigned int seed, float scale, const float* shift, void* stream) {
    SW_CHECK_ARG(f && links && uin, "f, links and uin must be non-null");
    SW_CHECK_RC(check_dims(nx, ny, nz));
    SW_CHECK_RC(check_dtype(dtype, scale));
    hipStream_t s = (hipStream_t)stream;
    dim3 grid((nx + 255) / 256, ny, nz), blk(256);
    switch (dtype) {
        case DT_F32: hipLaunchKernelGGL(k_init<float>, grid, blk, 0, s, (float*)f, links, uin, nx, ny, nz, noise, seed, 1.0f, nullptr); break;
        case DT_F64: hipLaunchKernelGGL(k_init<double>, grid, blk, 0, s, (double*)f, links, uin, nx, ny, nz, noise, seed, 1.0f, nullptr); break;
        case DT_F16: hipLaunchKernelGGL(k_init<_Float16>, grid, blk, 0, s, (_Float16*)f, links, uin, nx, ny, nz, noise, seed, scale, shift); break;
        case DT_BF16: hipLaunchKernelGGL(k_init<bf16_t>, grid, blk, 0, s, (bf16_t*)f, links, uin, nx, ny, nz, noise, seed, scale, shift); break;
    }
    SW_HIP(hipGetLastError());
    return SW_OK;
}

int sw_step(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin, const float* gust,
            int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs, int x_sponge, double nu_sponge,
            int ybc, float scale, const float* shift, double* partials, int ring, void* stream) {
    return step_entry(f_in, f_out, dtype, links, uin, gust, t, t_dev, nx, ny, nz, nu0, cs, x_sponge, nu_sponge, ybc,
                      scale, shift, partials, ring, (hipStream_t)stream);
}

int sw_step_accumulate(const void* f_in, void* f_out, int dtype, const uint32_t* links, const float* uin,
                       const float* gust, int t, const int* t_dev, int nx, int ny, int nz, double nu0, double cs,
                       int x_sponge, double nu_sponge, int ybc, float scale, const float* shift, double* partials,
                       int ring, float* acc, void* stream) {
    SW_CHECK_ARG(acc, "acc must be non-null");
    return step_entry(f_in, f_out, dtype, links, uin, gust, t, t_dev, nx, ny, nz, nu0, cs, x_sponge, nu_sponge, ybc,
                      scale, shift, partials, ring, (hipStream_t)stream, acc);
}

int sw_step_variants(void) { return NUM_TILE_VARIANTS; }

int sw_set_step_variant(int dtype, int variant) {
    SW_CHECK_ARG(dtype >= 0 && dtype <= 3, "dtype must be 0..3");
    SW_CHECK_ARG(variant >= 0 && variant < NUM_TILE_VARIANTS, "unknown step variant");
    g_step_variant[dtype] = variant;
    return SW_OK;
}

int sw_get_step_variant(int dtype) { return (dtype >= 0 && dtype <= 3) ? g_step_variant[dtype] : -1; }

int sw_force_reduce(const double* partials, int ring, int t0, const int* t_dev, int nsteps, int ny, int nz,
                    double* force, void* stream) {
    return force_reduce_entry(partials, ring, t0, t_dev, nsteps, ny, nz, force, (hipStream_t)stream);
}

int sw_macro(const void* f, int dtype, const uint32_t* links, int nx, int ny, int nz, float scale,
             const float* shift, float* U, float* Uacc, void* stream) {
    SW_CHECK_ARG(f && links && (U || Uacc), "f, links and one of U / Uacc must be non-null");
    SW_CHECK_RC(check_dims(nx, ny, nz));
    SW_CHECK_RC(check_dtype(dtype, scale));
    hipStream_t s = (hipStream_t)stream;
    dim3 grid((nx + 255) / 256, ny, nz), blk(256);
    const float inv = 1.0f / scale;
    switch (dtype) {
        case DT_F32: hipLaunchKernelGGL(k_macro<float>, grid, blk, 0, s, (const float*)f, links, nx, ny, nz, 1.0f, nullptr, U, Uacc); break;
        case DT_F64: hipLaunchKernelGGL(k_macro<double>, grid, blk, 0, s, (const double*)f, links, nx, ny, nz, 1.0f, nullptr, U, Uacc); break;
        case DT_F16: hipLaunchKernelGGL(k_macro<_Float16>, grid, blk, 0, s, (const _Float16*)f, links, nx, ny, nz, inv, shift, U, Uacc); break;
        case DT_BF16: hipLaunchKernelGGL(k_macro<bf16_t>, grid, blk, 0, s, (const bf16_t*)f, links, nx, ny, nz, inv, shift, U, Uacc); break;
    }
    SW_HIP(hipGetLastError());
    return SW_OK;
}

int sw_shift_table(const float* uin, int nz, int kind, float* shift, void* stream) {
    SW_CHECK_ARG(shift && (uin || kind == SHIFT_WEIGHTS), "shift (and uin for the inlet kind) must be non-null");
    SW_CHECK_ARG(nz > 0, "nz must be positive");
    SW_CHECK_ARG(kind == SHIFT_WEIGHTS || kind == SHIFT_INLET, "kind must be 0 (weights) or 1 (inlet equilibrium)");
    hipLaunchKernelGGL(k_shift_table, dim3((nz + 63) / 64), dim3(64), 0, (hipStream_t)stream, uin, nz, kind, shift);
    SW_HIP(hipGetLastError());
    return SW_OK;
}

int sw_qcoarse(const float* U, const uint32_t* links, int nx, int ny, int nz, int x0, int ds, int qx, int qy, int qz,
               float* out, void* stream) {
    SW_CHECK_ARG(U && links && out, "U, links and out must be non-null");
    SW_CHECK_RC(check_dims(nx, ny, nz));
    SW_CHECK_ARG(ds >= 1 && qx >= 1 && qy >= 1 && qz >= 1 && x0 >= 0, "need ds, qx, qy, qz >= 1 and x0 >= 0");
    SW_CHECK_ARG(x0 + (long long)qx * ds <= nx && (long long)qy * ds <= ny && (long long)qz * ds <= nz,
                 "the pooled crop must lie inside the lattice");
    const long long n = (long long)qx * qy * qz;
    SW_CHECK_ARG(n < (1LL << 31), "frame too large");
    hipLaunchKernelGGL(k_qcoarse, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, U, links, nx,
                       ny, nz, x0, ds, qx, qy, qz, out);
    SW_HIP(hipGetLastError());
    return SW_OK;
}

int sw_selftest_division(unsigned long long* counts, void* stream) {
    SW_CHECK_ARG(counts, "counts must be a device buffer of two 64-bit counters");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, s, counts);
    hipLaunchKernelGGL(k_selftest_division, dim3(16384), dim3(256), 0, s, counts);
    SW_HIP(hipGetLastError());
    return SW_OK;
}

int sw_stream_copy(const void* src, void* dst, long long bytes, void* stream) {
    SW_CHECK_ARG(src && dst && src != dst, "src and dst must be distinct non-null buffers");
    SW_CHECK_ARG(bytes > 0 && bytes % 16 == 0, "bytes must be a positive multiple of 16");
    SW_CHECK_ARG(((uintptr_t)src | (uintptr_t)dst) % 16 == 0, "src and dst must be 16-byte aligned");
    int dev = 0, cus = 0;
    SW_HIP(hipGetDevice(&dev));
    SW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)(cus * 32)), dim3(256), 0, (hipStream_t)stream,
                       (const u32x4*)src, (u32x4*)dst, (size_t)bytes / 16);
    SW_HIP(hipGetLastError());
    return SW_OK;
}

static int transpose(const void* src, void* dst, int elem_bytes, int ncomp, int nx, int ny, int nz, bool to_ref,
                     void* stream) {
    SW_CHECK_ARG(src && dst && src != dst, "src and dst must be distinct non-null buffers");
    SW_CHECK_RC(check_dims(nx, ny, nz));
    SW_CHECK_ARG(ncomp > 0 && (long long)ncomp * ny <= 65535LL, "ncomp * ny must be at most 65535");
    dim3 grid((nx + 31) / 32, (nz + 31) / 32, ncomp * ny), blk(32, 8);
    hipStream_t s = (hipStream_t)stream;
#define SW_T(E)                                                                                         \
    if (to_ref) hipLaunchKernelGGL((k_transpose<E, true>), grid, blk, 0, s, (const E*)src, (E*)dst, nx, ny, nz); \
    else hipLaunchKernelGGL((k_transpose<E, false>), grid, blk, 0, s, (const E*)src, (E*)dst, nx, ny, nz);
    switch (elem_bytes) {
        case 1: SW_T(uint8_t) break;
        case 2: SW_T(uint16_t) break;
        case 4: SW_T(uint32_t) break;
        case 8: SW_T(uint64_t) break;
        default: return set_error(SW_ERR_INVALID, "elem_bytes must be 1, 2, 4 or 8");
    }
#undef SW_T
    SW_HIP(hipGetLastError());
    return SW_OK;
}

int sw_to_ref_layout(const void* src, void* dst, int elem_bytes, int ncomp, int nx, int ny, int nz, void* stream) {
    return transpose(src, dst, elem_bytes, ncomp, nx, ny, nz, true, stream);
}

int sw_from_ref_layout(const void* src, void* dst, int elem_bytes, int ncomp, int nx, int ny, int nz, void* stream) {
    return transpose(src, dst, elem_bytes, ncomp, nx, ny, nz, false, stream);
}

}  // extern "C"
