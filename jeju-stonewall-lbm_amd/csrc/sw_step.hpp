// The fused D3Q19 step: pull-streaming, half-way bounce-back, outlet / top / side
// boundaries, inlet, momentum-exchange wall force and regularized-BGK collision with a
// Smagorinsky eddy viscosity.  Restates stonewall3d/lbm3d.py:57-155 (k_step).
//
// Work decomposition: a row task is one (y, z) row of one lattice.  A wavefront owns a
// row and sweeps it in chunks of 64 cells (CPL chunks per iteration).  A chunk in which
// no cell is solid or has a solid neighbour (all link masks zero) takes the fast path:
// 19 unconditional coalesced loads per cell from wave-uniform row pointers plus one
// 32-bit lane offset, collision, 19 coalesced stores.  Other chunks (next to the ground
// or a stone) take the general path with per-link bounce-back and the wall force.
//
// Every row's FP32 per-cell wall force (fxw, fzw) is accumulated in FP64 per lane,
// reduced over the wave with a fixed butterfly, then over the workgroup's rows in row
// order, and written as one FP64 pair per workgroup and step into a partials ring; a
// second kernel sums the partials of each z level in a fixed order, so the force rows
// are bit-reproducible (the reference uses FP32 atomics, whose order is not).
#pragma once
#include "sw_common.hpp"

namespace sw {

constexpr int ROWS = 4;            // rows per workgroup (a row-group)

struct StepArgs {
    const void* f_in;
    void* f_out;
    const uint32_t* links;   // [nz][ny][nx]
    const float* uin;        // [nz] inlet profile
    const float* gust;       // [t] gust factor per step, or null (factor 1)
    double* partials;        // [ring][nz][nyb][2], or null
    const int* t_dev;        // optional device step counter added to t
    int t;
    int ring;
    int nx, ny, nz, nyb;
    int x_sponge;
    int ybc;                 // Y_PERIODIC or Y_FREESLIP
    double nu0, cs, nu_sponge;
    float scale, inv_scale;
};

// Collision of one cell from its 19 streamed populations fl[] (lbm3d.py:103-155); the
// results are stored through st(i, value).
template <class T, class Store>
__device__ __forceinline__ void collide(const T (&fl)[Q], uint32_t x, const StepArgs& a, T smag, Store&& st) {
    // ---- moments (:103-127), sequential in i; zero-coefficient terms dropped
    T rho = fl[0], mx, my, mz, pxx, pyy, pzz, pxy, pxz, pyz;
    // first contributions: 0 + v == v for the (never -0) populations
    mx = fl[1]; pxx = fl[1];
    my = fl[3]; pyy = fl[3];
    mz = fl[5]; pzz = fl[5];
    pxy = fl[7]; pxz = fl[11]; pyz = fl[15];
    static_for<1, Q>([&](auto I) {
        constexpr int i = I;
        constexpr int cx = CX[i], cy = CY[i], cz = CZ[i];
        const T v = fl[i];
        rho = rho + v;
        if constexpr (cx != 0 && i != 1) { mx = cx > 0 ? mx + v : mx - v; pxx = pxx + v; }
        if constexpr (cy != 0 && i != 3) { my = cy > 0 ? my + v : my - v; pyy = pyy + v; }
        if constexpr (cz != 0 && i != 5) { mz = cz > 0 ? mz + v : mz - v; pzz = pzz + v; }
        if constexpr (cx * cy != 0 && i != 7) pxy = cx * cy > 0 ? pxy + v : pxy - v;
        if constexpr (cx * cz != 0 && i != 11) pxz = cx * cz > 0 ? pxz + v : pxz - v;
        if constexpr (cy * cz != 0 && i != 15) pyz = cy * cz > 0 ? pyz + v : pyz - v;
    });
    const T ux = mx / rho, uy = my / rho, uz = mz / rho;                        // :128-138
    const T third = div3(rho);
    const T nxx = (pxx - third) - (rho * ux) * ux;
    const T nyy = (pyy - third) - (rho * uy) * uy;
    const T nzz = (pzz - third) - (rho * uz) * uz;
    const T nxy = pxy - (rho * ux) * uy;
    const T nxz = pxz - (rho * ux) * uz;
    const T nyz = pyz - (rho * uy) * uz;

    const T nu0 = (T)a.nu0;                                                    // :140-147
    T nu = nu0;
    if ((int)x > a.x_sponge) {
        const T s = (T)((int)x - a.x_sponge) / (T)(a.nx - a.x_sponge);
        nu = nu0 + ((T)a.nu_sponge * s) * s;
    }
    const T tau0 = T(3) * nu + T(0.5);
    const T qn = sqrt(((nxx * nxx + nyy * nyy) + nzz * nzz) + T(2) * ((nxy * nxy + nxz * nxz) + nyz * nyz));
    const T tau = T(0.5) * (tau0 + sqrt(tau0 * tau0 + (smag * qn) / rho));
    const T keep = T(1) - T(1) / tau;
    const T usq15 = T(1.5) * ((ux * ux + uy * uy) + uz * uz);

    // ---- regularized collision (:149-155)
    // Directions come in opposite pairs (i, i+1), i odd, c_{i+1} = -c_i.  Negation is exact
    // in IEEE arithmetic, so the pair's cu values are +-cu, their (0.5*cu)*cu terms are
    // equal and their stress projections qq are identical: each is computed once, and the
    // results are bit-identical to evaluating lbm3d.py:153-155 separately per direction.
    constexpr T third_c = T(1) / T(3);
    constexpr T A1 = T(1) - third_c;      // (c*c - 1/3) for |c| = 1
    constexpr T A0 = T(0) - third_c;      // (c*c - 1/3) for c = 0
    {
        constexpr T w = wgt<T>(0);
        const T qq = (A0 * nxx + A0 * nyy) + A0 * nzz;
        st(std::integral_constant<int, 0>{}, feq<0, T>(rho, ux, uy, uz, usq15) + ((keep * w) * T(4.5)) * qq);
    }
    static_for<0, 9>([&](auto P) {
        constexpr int i = 2 * P + 1, j = i + 1;
        constexpr int cx = CX[i], cy = CY[i], cz = CZ[i];
        constexpr T w = wgt<T>(i);
        constexpr T ax = cx ? A1 : A0, ay = cy ? A1 : A0, az = cz ? A1 : A0;
        T qq = (ax * nxx + ay * nyy) + az * nzz;
        if constexpr (cx * cy != 0) qq = qq + T(2) * (cx * cy > 0 ? nxy : -nxy);
        if constexpr (cx * cz != 0) qq = qq + T(2) * (cx * cz > 0 ? nxz : -nxz);
        if constexpr (cy * cz != 0) qq = qq + T(2) * (cy * cz > 0 ? nyz : -nyz);
        const T neq = ((keep * w) * T(4.5)) * qq;
        const T cu = T(3) * cdotu<i>(ux, uy, uz);
        const T h = (T(0.5) * cu) * cu;
        const T wr = w * rho;
        st(std::integral_constant<int, i>{}, wr * (((T(1) + cu) + h) - usq15) + neq);
        st(std::integral_constant<int, j>{}, wr * (((T(1) + (-cu)) + h) - usq15) + neq);
    });
}

// Number of 32-bit words of chunk flags per row: bit c of a row's words is set when the
// 64-cell chunk c holds a solid cell or a cell with a solid neighbour (x = 0 excluded).
// The words of all rows follow the N link masks in the links buffer.
__host__ __device__ inline uint32_t chunk_words(uint32_t nx) { return (nx + 64 * 32 - 1) / (64 * 32); }

// Tuning knobs of the step kernel (the autotuner's choices):
//   CPL : 64-cell chunks processed per iteration of a wavefront (1 or 2)
//   PF  : issue the next chunk's loads before the current chunk's collision and stores
//         (its loads then never wait behind the current stores in the in-order vmcnt)
//   NT  : nontemporal (streaming) stores
//   NTL : nontemporal loads
//   WPB : wavefronts per workgroup (4, 8 or 16); each of the ROWS rows of a row-group is
//         split into WPB/ROWS contiguous chunk ranges, one per wavefront
//   OCC : minimum wavefronts per SIMD the compiler must fit (0 = compiler's choice).  A cap
//         of 8 forces 64 VGPRs and spills the 16-bit kernels to scratch (measured 100x
//         slower), so the list below uses 0 throughout.
//   PAIR: 16-bit storage only — a lane owns two adjacent cells, each link is one aligned
//         32-bit load (x-1 / x+1 neighbours by whole-wavefront DPP shifts) and one 32-bit
//         store; other storage forms ignore it
template <int CPL_, bool PF_, bool NT_, bool NTL_ = false, int WPB_ = ROWS, int OCC_ = 0, bool PAIR_ = false>
struct Tile {
    static constexpr bool PAIR = PAIR_;
    static constexpr int OCC = OCC_;
    static constexpr int CPL = CPL_;
    static constexpr bool PF = PF_;
    static constexpr bool NT = NT_;
    static constexpr bool NTL = NTL_;
    static constexpr int WPB = WPB_;
    static constexpr int NSEG = WPB_ / ROWS;
};

// The curated variant list (index = sw_set_step_variant's `variant`).
using TileV0 = Tile<1, false, false, false, 4>;   // plain
using TileV1 = Tile<1, false, true, false, 4>;    // nontemporal stores
using TileV2 = Tile<1, false, true, true, 4>;     // nontemporal stores and loads
using TileV3 = Tile<1, false, true, false, 8>;    // NT stores, rows split over 2 wavefronts
using TileV4 = Tile<1, false, true, false, 16>;   // NT stores, rows split over 4 wavefronts
using TileV5 = Tile<1, false, true, true, 8>;     // NT stores and loads, rows over 2 wavefronts
using TileV6 = Tile<2, false, true, false, 4>;    // 2 chunks per iteration, NT stores
using TileV7 = Tile<1, true, true, false, 8>;     // prefetch, NT stores, rows over 2 wavefronts
using TileV8 = Tile<1, false, true, false, 4, 0, true>;   // 16-bit pairs, NT stores
using TileV9 = Tile<1, false, true, false, 8, 0, true>;   // 16-bit pairs, NT stores, rows over 2 wavefronts
constexpr int NUM_TILE_VARIANTS = 10;
#define SW_TILE_LIST(X) X(0, TileV0) X(1, TileV1) X(2, TileV2) X(3, TileV3) X(4, TileV4) X(5, TileV5) \
                        X(6, TileV6) X(7, TileV7) X(8, TileV8) X(9, TileV9)

// One (y, z) row of one lattice; adds this lane's FP64 wall-force sums to fx_acc/fz_acc.
// BUF selects raw-buffer addressing (all byte offsets of a population array < 4 GiB).
template <class S, class TL, bool BUF>
__device__ __forceinline__ void step_row(const StepArgs& a, int y, int z, int t, int lane, int seg,
                                         double& fx_acc, double& fz_acc) {
    using C = Codec<S>;
    using T = typename C::T;
    constexpr int CPL = TL::CPL;
    constexpr bool PF = TL::PF && TL::CPL == 1;
    constexpr uint32_t ES = sizeof(S);
    const PopMem<S, BUF> mem(a.f_in, a.f_out);
    const uint32_t nx = a.nx, ny = a.ny, nz = a.nz;
    const uint32_t N = nx * ny * nz;
    const size_t NB = (size_t)N * ES;                            // bytes per population
    const uint32_t rowid = (uint32_t)z * ny + y;
    const uint32_t selfrow = rowid * nx;
    const size_t selfb = (size_t)selfrow * ES;
    const T inv_scale = (T)a.inv_scale, scale = (T)a.scale;
    // (18*sqrt(2))*cs*cs evaluated left to right as lbm3d.py:146 does
    const T cs = (T)a.cs;
    const T smag = ((T(18) * (T)sqrt(T(2))) * cs) * cs;

    // wave-uniform byte offset of the source row of every link after the y / z boundary
    // rules, pre-shifted by -cx so that every load of a cell uses the lane offset x*ES
    // (cx = -1 loads use the outlet-clamped lane offset instead: lbm3d.py:82-83)
    size_t srow[Q];
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        int zs = z - CZ[i], ys = y - CY[i], j = i;
        if (a.ybc == Y_PERIODIC) {
            ys = ys < 0 ? ys + (int)ny : (ys >= (int)ny ? ys - (int)ny : ys);      // lbm3d.py:79
        } else {
            if (ys < 0 || ys >= (int)ny) { ys = y; j = MIRY[j]; }                 // free-slip sides
        }
        if (zs >= (int)nz) { zs = z; j = MIRZ[j]; }                               // lbm3d.py:84-86
        const size_t r = (size_t)j * NB + (size_t)(((uint32_t)zs * ny + (uint32_t)ys) * nx) * ES;
        srow[i] = (CX[i] > 0) ? r - ES : r;
    });
    const uint32_t* lrow = a.links + selfrow;
    const uint32_t cw = chunk_words(nx);
    const uint32_t* rflags = a.links + N + rowid * cw;           // wave-uniform chunk flags
    const uint32_t nch_row = (nx + 63) / 64;
    const uint32_t seglen = (nch_row + TL::NSEG - 1) / TL::NSEG;
    const uint32_t cbeg = min(nch_row, (uint32_t)seg * seglen);
    const uint32_t nchunks = min(nch_row, cbeg + seglen);        // end of this wavefront's range
    auto special_chunk = [&](uint32_t c) -> bool { return (rflags[c >> 5] >> (c & 31)) & 1u; };

    auto load_fast = [&](uint32_t x, T (&fl)[Q]) {
        const uint32_t xo = x * ES;
        const uint32_t xpo = min(x + 1, nx - 1) * ES;
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            fl[i] = C::decode(mem.template ld<TL::NTL>(srow[i], (CX[i] < 0) ? xpo : xo), wgt<T>(i), inv_scale);
        });
    };
    auto collide_store = [&](uint32_t x, const T (&fl)[Q]) {
        const uint32_t xo = x * ES;
        collide<T>(fl, x, a, smag, [&](auto I, T v) {
            constexpr int i = I;
            mem.template st<TL::NT>((size_t)i * NB + selfb, xo, C::encode(v, wgt<T>(i), scale));
        });
    };
    // general path: per-link bounce-back and wall force (:77-101)
    auto general_cell = [&](uint32_t x) {
        if (!(x < nx && x != 0)) return;
        const uint32_t mm = lrow[x];
        if (mm & SOLID_BIT) return;                                                  // :65-66
        const uint32_t xo = x * ES;
        const uint32_t xpo = min(x + 1, nx - 1) * ES;
        T fl[Q];
        T fxw = T(0), fzw = T(0);
        bool hit = false;
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr T w = wgt<T>(i);
            if constexpr (i == 0) {
                fl[0] = C::decode(mem.ld(selfb, xo), w, inv_scale);
            } else {
                T val;
                if ((mm >> i) & 1u) {                                                // half-way bounce-back
                    val = C::decode(mem.ld((size_t)OPP[i] * NB + selfb, xo), w, inv_scale);
                    bool wall = true;
                    if constexpr (GROUND_SLOT[i] >= 0) wall = !((mm >> (GROUND_BIT0 + GROUND_SLOT[i])) & 1u);
                    if (wall) {                                                      // momentum exchange
                        const T fr2 = T(2) * (val - w);
                        if constexpr (CX[i] > 0) fxw = fxw - fr2;
                        if constexpr (CX[i] < 0) fxw = fxw + fr2;
                        if constexpr (CZ[i] > 0) fzw = fzw - fr2;
                        if constexpr (CZ[i] < 0) fzw = fzw + fr2;
                        hit = true;
                    }
                } else {
                    val = C::decode(mem.ld(srow[i], (CX[i] < 0) ? xpo : xo), w, inv_scale);
                }
                fl[i] = val;
            }
        });
        if (hit) { fx_acc += (double)fxw; fz_acc += (double)fzw; }
        collide_store(x, fl);
    };

    // one-cell-per-lane sweep of chunks [cbeg, nchunks)
    auto plain_sweep = [&]() {
#pragma clang loop unroll(disable)
        for (uint32_t c0 = cbeg; c0 < nchunks; c0 += CPL) {
            bool special = false;
#pragma unroll
            for (int k = 0; k < CPL; k++)
                if (c0 + k < nchunks) special |= special_chunk(c0 + k);
            if (!special) {
                // ---------------- fast path: no solid cell and no solid neighbour in the chunk(s)
                T fl[CPL][Q];
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    const uint32_t x = (c0 + k) * 64 + lane;
                    if (x < nx && x != 0) load_fast(x, fl[k]);
                }
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    const uint32_t x = (c0 + k) * 64 + lane;
                    if (x < nx && x != 0) collide_store(x, fl[k]);
                }
            } else {
#pragma unroll
                for (int k = 0; k < CPL; k++) general_cell((c0 + k) * 64 + lane);
            }
        }
    };

    bool inlet_done = false;
    if constexpr (TL::PAIR && ES == 2) {
      if (nx & 1u) {
        plain_sweep();                    // pairs need an even row length
      } else {
        // 128-cell chunks; lane L owns cells x0 + 2L and x0 + 2L + 1 (nx is even)
        const uint32_t npc = (nx + 127) / 128;
        const uint32_t plen = (npc + TL::NSEG - 1) / TL::NSEG;
        const uint32_t pbeg = min(npc, (uint32_t)seg * plen);
        const uint32_t pend = min(npc, pbeg + plen);
#pragma clang loop unroll(disable)
        for (uint32_t c = pbeg; c < pend; c++) {
            const uint32_t x0 = c * 128;
            const bool special = special_chunk(2 * c) || (2 * c + 1 < nch_row && special_chunk(2 * c + 1));
            if (special) {
                general_cell(x0 + lane);
                general_cell(x0 + 64 + lane);
                continue;
            }
            const uint32_t xa = x0 + 2 * lane;
            const uint32_t xl = min(xa, nx - 2);             // lanes past the row end read in-row
            const uint32_t lb = xl * 2;
            const uint32_t ledge = (x0 >= 2 ? x0 - 2 : 0) * 2;            // dword holding cell x0-1
            const uint32_t redge = (x0 + 128 < nx ? x0 + 128 : x0) * 2;   // dword holding cell x0+128
            uint32_t raw[Q];
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                const size_t row = (CX[i] > 0) ? srow[i] + ES : srow[i];   // unshifted row start
                const uint32_t cur = mem.ld32(row, lb);
                if constexpr (CX[i] == 0) {
                    raw[i] = cur;
                } else if constexpr (CX[i] > 0) {                        // cells (x-1, x)
                    raw[i] = mid16(cur, wave_shr1(cur, mem.ld32(row, ledge)));
                } else {                                                 // cells (x+1, x+2)
                    uint32_t nxt = wave_shl1(cur, mem.ld32(row, redge));
                    if (xl + 2 >= nx) nxt = cur >> 16;                  // outlet: cell nx-1 pulls itself
                    raw[i] = mid16(nxt, cur);
                }
            });
            uint32_t packed[Q];
            {
                T fl[Q];
                static_for<0, Q>([&](auto I) {
                    fl[I] = C::decode((uint16_t)(raw[I] & 0xFFFFu), wgt<T>(I), inv_scale);
                });
                collide<T>(fl, xa, a, smag, [&](auto I, T v) {
                    packed[I] = (uint32_t)C::encode(v, wgt<T>(I), scale);
                });
            }
            if (x0 == 0) {                                               // inlet cell x = 0 (:67-71)
                T u0 = (T)a.uin[z];
                if (a.gust) u0 = u0 * (T)a.gust[t];
                const T usq15 = T(1.5) * (u0 * u0);
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    const uint32_t e =
                        (uint32_t)C::encode(feq<i, T>(T(1), u0, T(0), T(0), usq15), wgt<T>(i), scale);
                    packed[i] = (lane == 0) ? e : packed[i];
                });
                inlet_done = true;
            }
            {
                T fl[Q];
                static_for<0, Q>([&](auto I) {
                    fl[I] = C::decode((uint16_t)(raw[I] >> 16), wgt<T>(I), inv_scale);
                });
                collide<T>(fl, xa + 1, a, smag, [&](auto I, T v) {
                    packed[I] |= (uint32_t)C::encode(v, wgt<T>(I), scale) << 16;
                });
            }
            if (xa < nx) {
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    mem.template st32<TL::NT>((size_t)i * NB + selfb, xa * 2, packed[i]);
                });
            }
        }
      }
    } else if constexpr (PF) {
        // software-pipelined sweep: chunk c+1's loads are in flight during chunk c's work
        T cur[Q], nxt[Q];
        bool cur_loaded = false;
        if (cbeg < nchunks) {
            const uint32_t x = cbeg * 64 + lane;
            if (!special_chunk(cbeg)) {
                if (x < nx && x != 0) load_fast(x, cur);
                cur_loaded = true;
            }
        }
#pragma clang loop unroll(disable)
        for (uint32_t c = cbeg; c < nchunks; c++) {
            const uint32_t x = c * 64 + lane;
            const bool next_fast = (c + 1 < nchunks) && !special_chunk(c + 1);
            if (cur_loaded) {
                const uint32_t xn = x + 64;
                if (next_fast && xn < nx) load_fast(xn, nxt);
                if (x < nx && x != 0) collide_store(x, cur);
            } else {
                general_cell(x);
                if (next_fast && x + 64 < nx) load_fast(x + 64, nxt);
            }
            if (next_fast) {
                static_for<0, Q>([&](auto I) { cur[I] = nxt[I]; });
            }
            cur_loaded = next_fast;
        }
    } else {
        plain_sweep();
    }
    // ---------------- inlet (:67-71): x = 0 fluid cell gets feq(1, (uin[z] g, 0, 0))
    if (seg == 0 && lane == 0 && !inlet_done && !(lrow[0] & SOLID_BIT)) {
        T u0 = (T)a.uin[z];
        if (a.gust) u0 = u0 * (T)a.gust[t];
        const T usq15 = T(1.5) * (u0 * u0);
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            mem.template st<false>((size_t)i * NB + selfb, 0,
                                   C::encode(feq<i, T>(T(1), u0, T(0), T(0), usq15), wgt<T>(i), scale));
        });
    }
}

// Writes this workgroup's FP64 force pair (sum over its WPB wavefronts in wavefront
// order) to partials[slot][z][yb].
template <int WPB>
__device__ __forceinline__ void block_force(const StepArgs& a, double fx, double fz, int z, int yb, int t) {
    __shared__ double red[WPB][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    fx = wave_sum(fx);
    fz = wave_sum(fz);
    if (lane == 0) { red[wave][0] = fx; red[wave][1] = fz; }
    __syncthreads();
    if (threadIdx.x == 0 && a.partials) {
        double sx = red[0][0], sz = red[0][1];
        for (int r = 1; r < WPB; r++) { sx += red[r][0]; sz += red[r][1]; }
        const size_t slot = (size_t)(t % a.ring);
        double* p = a.partials + ((slot * a.nz + z) * a.nyb + yb) * 2;
        p[0] = sx;
        p[1] = sz;
    }
    __syncthreads();   // red[] is reused when a workgroup processes several row-groups
}

}  // namespace sw
