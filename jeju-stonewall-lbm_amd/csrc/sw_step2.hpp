#pragma once
// Two fused steps per pass (temporal blocking): f(t) -> f(t+1) -> f(t+2) in one launch, the
// intermediate state kept on chip, so each pass reads and writes the populations once for
// two updates.  Restates two applications of stonewall3d/lbm3d.py:57-155 (k_step) with the
// arithmetic of sw_step.hpp, so the populations are bit-identical to two single steps.
//
// Work decomposition.  A workgroup owns a tile of 64 cells in x (aligned to 64: whole
// 64-cell chunks, so every load and store of a row is one aligned segment) and TY rows in
// y, of which rows 1..TY-2 are its output rows and rows 0 and TY-1 halo rows, and sweeps
// the tile through z.  Wavefront w owns row w; lane L the cell x0 + L.  Iteration k:
//   phase 1  every wavefront computes the first step at plane k for its row (f1(k)),
//            keeps the populations that stay in its row and column in registers, writes
//            the ones that cross rows (c_y != 0) to LDS, and issues the loads of f1(k+1);
//   barrier
//   phase 2  the output rows compute the second step at plane k-1 (f2(k-1)) from f1 of
//            planes k-2, k-1, k (registers, LDS, x-neighbours by DPP lane shifts) and store
//            it; the two halo-row wavefronts compute f1(k+1) of the cells just outside the
//            tile in x (x0-1 and x0+64, every row), whose c_x != 0 populations the tile's
//            edge lanes pull in the next iteration, and wavefront 0 writes the force
//            partials of the previous plane.
// One barrier per plane: LDS rings are deep enough (4 slots for c_z = +1, 3 for the rest)
// that a wavefront writing plane k+1 never overwrites a slot another wavefront still reads.
//
// The force of step t (f1) and t+1 (f2) is summed per tile and plane (lane FP64 sums, a
// fixed butterfly per wavefront, wavefronts in order) into partials[slot][z][tile][2]; the
// rows of force come from sw_force_reduce2.  The summation order differs from the single
// step's (per row-group partials), so force rows agree with it to rounding, not bit for bit.
#include "sw_step.hpp"

namespace sw {

// LDS of one workgroup.  Row arrays have 66 columns: column L + 1 holds lane L's cell, column
// 0 the cell x0-1 and column 65 the cell x0+64 (written by the halo computation).
template <int TY>
struct Step2Lds {
    float A[4][2][TY][66];   // c_z = +1, c_y != 0: populations 15, 18 (read two planes later)
    float B[3][6][TY][66];   // c_z = 0,  c_y != 0: populations 3, 4, 7, 8, 9, 10
    float C[3][2][TY][66];   // c_z = -1, c_y != 0: populations 16, 17
    float H[4][6][TY];       // halo cells, in-row populations with c_x != 0:
                             // 1, 11, 13 of x0-1 and 2, 12, 14 of x0+64
    double red[3][TY][4];    // per-wavefront force sums (f1 Fx, Fz; f2 Fx, Fz) of three iterations
};

__host__ __device__ constexpr int s2_gidx(int i) {      // index within the LDS group of a c_y != 0 population
    return i == 15 ? 0 : i == 18 ? 1 : i == 16 ? 0 : i == 17 ? 1 : i == 3 ? 0 : i == 4 ? 1 : i == 7 ? 2
         : i == 8 ? 3 : i == 9 ? 4 : i == 10 ? 5 : -1;
}
__host__ __device__ constexpr int s2_hidx(int i) {      // index of an in-row c_x != 0 population in H
    return i == 1 ? 0 : i == 11 ? 1 : i == 13 ? 2 : i == 2 ? 3 : i == 12 ? 4 : i == 14 ? 5 : -1;
}

template <int TY>
__device__ __forceinline__ float& lds_row(Step2Lds<TY>& L, int j, int plane, int gi, int r, int col) {
    if (CZ[j] > 0) return L.A[plane & 3][gi][r][col];
    if (CZ[j] == 0) return L.B[plane % 3][gi][r][col];
    return L.C[plane % 3][gi][r][col];
}

// One tile of a launch: the two-step sweep (SINGLE = false) or, for the odd step of a run
// of pairs, one step with the same tile decomposition and force-partials layout (SINGLE =
// true: the output rows compute the first step of every plane and store it; the halo rows
// idle).  P: tiles per z plane of the partials layout; L: the workgroup's LDS.
// [za, zb): the z planes whose outputs (and force partials) this call writes, a segment of
// the tile's sweep (za = 1: from the plane above the ground); the two-step sweep then starts
// at plane za - 1, recomputing the first step of the plane below the segment, so a tile cut
// into segments gives the populations and partials of one uncut sweep.  (The defaults, the
// whole sweep, are constants the compiler folds: with a run-time za the FP32 kernel needs
// 35 more VGPRs than it has, so the segmented launch is a separate instantiation.)
// PF: the streamed loads of plane k+1's first step are issued at the end of phase 1 of plane
// k, before the barrier, so they are in flight during the barrier and the second step (19
// more registers held across phase 2: with fewer than 13 rows the workgroup runs at most 3
// wavefronts per SIMD, which leaves 168 VGPRs for them).
// ACC (two-step sweep only): 1 or 2 adds (rho, u) of every cell's state after the pair's
// first or second step into a.acc, as sw_macro's accumulate would on that state
// (lbm3d.py:158-188): the output rows sample the planes [za, zb) they own, the segment with
// za = 1 also the solid ground plane.
template <class S, int TY, bool SINGLE, class A, bool PF = false, int ACC = 0>
__device__ __forceinline__ void step2_tile(A& a, int t0, int ntx, int P, int tile, Step2Lds<TY>& L, int za = 1,
                                           int zb = 1 << 30) {
    using Cd = Codec<S>;
    using T = typename Cd::T;
    using B = typename Cd::B;
    constexpr int OY = TY - 2;
    constexpr bool NORM = sizeof(S) == 2;
    constexpr uint32_t ES = sizeof(S);

    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tx = tile % ntx, ty = tile / ntx;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const uint32_t N = (uint32_t)nx * ny * nz;
    const uint32_t NB = N * ES;                       // bytes per population
    const uint32_t PB = (uint32_t)nx * ny * ES;       // bytes per z plane of a population
    const bool periodic = a.ybc == Y_PERIODIC;
    const __amdgpu_buffer_rsrc_t rin = make_rsrc(a.f_in), rout = make_rsrc(a.f_out);
    const T scale = (T)a.scale;
    const T inv_scale = (T)a.inv_scale;
    const int kscale = (int)((__builtin_bit_cast(uint32_t, a.scale) >> 23) & 0xFFu) - 127;
    const T cs = (T)a.cs;
    const T smag = ((T(18) * (T)sqrt(T(2))) * cs) * cs;
    const uint32_t cw = chunk_words(nx);
    const uint32_t* cflags = a.links + N;
    za = za < 1 ? 1 : za;
    zb = zb > nz ? nz : zb;

    // ---- this wavefront's row
    const int yg = ty * OY - 1 + w;                   // may be -1 or >= ny
    bool need = yg <= ny;                             // rows past the tile's last halo row are idle
    int y;
    if (periodic) {
        y = yg < 0 ? yg + ny : (yg >= ny ? yg - ny : yg);
    } else {
        need = need && yg >= 0 && yg < ny;            // free-slip: no row outside the lattice
        y = yg < 0 ? 0 : (yg >= ny ? ny - 1 : yg);
    }
    const bool out_row = w >= 1 && w <= OY && yg < ny;
    // source rows of c_y = +1 / -1 links after the side rule; mlo / mhi: mirrored (free-slip)
    int ym = y - 1, yp = y + 1;
    bool mlo = false, mhi = false;
    if (periodic) {
        ym = ym < 0 ? ym + ny : ym;
        yp = yp >= ny ? yp - ny : yp;
    } else {
        if (ym < 0) { ym = y; mlo = true; }
        if (yp >= ny) { yp = y; mhi = true; }
    }
    const int x0 = tx * 64;
    const int x = x0 + lane;
    const int xl = x < nx ? x : nx - 1;              // lanes past the row end work on in-row addresses
    const uint32_t xo = (uint32_t)xl * ES;
    const uint32_t xpo = (uint32_t)(xl + 1 < nx ? xl + 1 : nx - 1) * ES;
    const bool outlet = x == nx - 1;

    // offsets of the 16-bit storage: sigma_c(z) of the c_x class c at height z
    auto sig = [&](int z, int c) -> T { return a.shift ? (T)ld_const(a.shift, z * 4 + c) : T(1); };

    // byte offset of link I's source row at plane k (lbm3d.py:77-86 with the side rules)
    auto src_off = [&](auto I, int k, bool lo, bool hi, int yrm, int yr0, int yrp) -> uint32_t {
        constexpr int i = I;
        int j = i;
        if constexpr (CY[i] > 0) { if (lo) j = MIRY[i]; }
        if constexpr (CY[i] < 0) { if (hi) j = MIRY[i]; }
        int zs = k - CZ[i];
        if constexpr (CZ[i] < 0) {
            if (zs >= nz) { zs = k; j = (CY[i] > 0 && lo) || (CY[i] < 0 && hi) ? MIRZ[MIRY[i]] : MIRZ[i]; }
        }
        const int ys = CY[i] > 0 ? yrm : (CY[i] < 0 ? yrp : yr0);
        return (uint32_t)j * NB + (uint32_t)zs * PB + (uint32_t)ys * (uint32_t)nx * ES;
    };
    // source-row offset of link I at plane k for this wavefront's row (side rules applied;
    // the top plane's c_z = -1 links mirror onto their own plane, src_off)
    const uint32_t ybm = (uint32_t)ym * (uint32_t)nx * ES, yb0 = (uint32_t)y * (uint32_t)nx * ES,
                   ybp = (uint32_t)yp * (uint32_t)nx * ES;
    auto row_off = [&](auto I, int k) -> uint32_t {
        constexpr int i = I;
        if constexpr (CZ[i] < 0) {
            if (k + 1 >= nz) return src_off(I, k, mlo, mhi, ym, y, yp);
        }
        uint32_t j = i;
        if constexpr (CY[i] > 0) j = mlo ? (uint32_t)MIRY[i] : (uint32_t)i;
        if constexpr (CY[i] < 0) j = mhi ? (uint32_t)MIRY[i] : (uint32_t)i;
        const uint32_t yb = CY[i] > 0 ? ybm : (CY[i] < 0 ? ybp : yb0);
        return j * NB + (uint32_t)(k - CZ[i]) * PB + yb;
    };
    // decode offset of link I's source at plane k (the source height; the top row's c_z = -1
    // links mirror onto its own height)
    auto dsh = [&](auto I, int k) -> T {
        constexpr int i = I;
        if constexpr (NORM) {
            const int zs = CZ[i] > 0 ? k - 1 : (CZ[i] < 0 ? (k + 1 < nz ? k + 1 : k) : k);
            return sig(zs, xclass(CX[i]));
        } else {
            return wgt<T>(i);
        }
    };
    // value a following step reads back from the storage of v written at height k: identity
    // for FP32; for the 16-bit forms the encode with the offsets of height k, then the decode
    auto stored = [&](auto I, T v, int k) -> T {
        constexpr int i = I;
        if constexpr (NORM) {
            const T s = sig(k, xclass(CX[i]));
            return Cd::decode(Cd::encode(v, (T)neg_scaled((float)s, kscale), scale), s, inv_scale);
        } else {
            return v;
        }
    };
    // (ACC) one cell's sample: the accumulator's four values are loaded at the start of the
    // phase that computes the state (acc_fetch), so that their latency passes during its
    // collision, and prev + (rho, u) of the stored populations sv, or + (1, 0, 0, 0) for a
    // solid cell (sw_macro reads a solid cell so), is stored at its end (acc_sample)
    auto acc_fetch = [&](int k, float (&prev)[4]) {
        const size_t cell = ((size_t)k * ny + y) * nx + x;
#pragma unroll
        for (int c = 0; c < 4; c++) prev[c] = a.acc[c * (size_t)N + cell];
    };
    auto acc_sample = [&](int k, const float (&prev)[4], const T (&sv)[Q], bool live) {
        const size_t cell = ((size_t)k * ny + y) * nx + x;
        float u[4] = {1.0f, 0.0f, 0.0f, 0.0f};
        if (live) macro_values<NORM, T>(sv, u);
#pragma unroll
        for (int c = 0; c < 4; c++) a.acc[c * (size_t)N + cell] = prev[c] + u[c];
    };
    auto chunk_fast = [&](int row, int k, int c) -> bool {
        return !((ld_const(cflags, ((uint32_t)k * ny + row) * cw + (c >> 5)) >> (c & 31)) & 1u);
    };
    auto inlet_u0 = [&](int k, int t) -> T {
        T u0 = (T)ld_const(a.uin, k);
        if (a.gust) u0 = u0 * (T)ld_const(a.gust, t);
        return u0;
    };
    auto wall_force = [&](auto I, T f, T& fxw, T& fzw) {
        constexpr int i = I;
        constexpr T wi = wgt<T>(i);
        const T fr2 = NORM ? (T(2) * wi) * (f - T(1)) : T(2) * (f - wi);
        if constexpr (CX[i] > 0) fxw = fxw - fr2;
        if constexpr (CX[i] < 0) fxw = fxw + fr2;
        if constexpr (CZ[i] > 0) fzw = fzw - fr2;
        if constexpr (CZ[i] < 0) fzw = fzw + fr2;
    };
    auto is_wall = [&](auto I, uint32_t mm) -> bool {
        constexpr int i = I;
        bool wall = (mm >> i) & 1u;
        if constexpr (GROUND_SLOT[i] >= 0) wall = wall && !((mm >> (GROUND_BIT0 + GROUND_SLOT[i])) & 1u);
        return wall;
    };

    // ---- first step, fast path loads of plane k for this wavefront's row
    auto f1_issue = [&](int k, B (&raw)[Q]) {
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            const uint32_t so = row_off(I, k) - (CX[i] > 0 ? ES : 0u);
            raw[i] = buf_load<B>(rin, CX[i] < 0 ? xpo : xo, so);
        });
    };
    auto f1_decode = [&](int k, const B (&raw)[Q], T (&fl)[Q]) {
        static_for<0, Q>([&](auto I) { fl[I] = Cd::decode(raw[I], dsh(I, k), inv_scale); });
    };
    auto f1_load = [&](int k, T (&fl)[Q]) {
        B raw[Q];
        f1_issue(k, raw);
        f1_decode(k, raw, fl);
    };
    // ---- first step near solids: the streamed values are loaded as on the fast path; on
    // links whose source is solid (mask bit set) the cell's own opposite population replaces
    // them (half-way bounce-back, lbm3d.py:87-98), loaded only where a lane of the wavefront
    // needs it, all issued before the first is used; then the wall force (lbm3d.py:89-101).
    // self: byte offset of the cell in a population array; k: its plane.
    auto bounce_fixups = [&](int k, uint32_t self, uint32_t mm, T (&fl)[Q], T& fxw, T& fzw) {
        B own[Q];
        static_for<1, Q>([&](auto I) {
            constexpr int i = I;
            if (__builtin_amdgcn_ballot_w64((mm >> i) & 1u) != 0) own[i] = buf_load<B>(rin, self, (uint32_t)OPP[i] * NB);
        });
        static_for<1, Q>([&](auto I) {
            constexpr int i = I;
            if ((mm >> i) & 1u) {
                T off;
                if constexpr (NORM) off = sig(k, xclass(-CX[i]));
                else off = wgt<T>(OPP[i]);
                fl[i] = Cd::decode(own[i], off, inv_scale);
            }
            if (is_wall(I, mm)) wall_force(I, fl[i], fxw, fzw);
        });
    };

    // ---- halo cells (x0-1, x0+64) of every row at plane kh: f1 into H and the row arrays'
    // edge columns.  Lane l of wavefront 0 (TY-1) takes row l/2 (TY/2 + l/2), side l & 1:
    // the streamed values of plane kh, the bounce-back, the collision and the values the
    // tile's edge lanes pull.  It runs in phase 2, where the halo-row wavefronts have no second
    // step to compute.  (Issuing its loads in phase 1 instead holds 19 more registers across
    // the barrier and measured 5 % slower.)
    const bool halo_wave = (w == 0 || w == TY - 1);
    // (the per-lane set-up is recomputed on each call rather than held in registers across
    // the sweep, where the row wavefronts need them)
    auto halo_f1 = [&](int kh) {
        const int hr = ((w == 0) ? 0 : TY / 2) + (lane >> 1);
        const int hside = lane & 1;
        const int hx = hside ? x0 + 64 : x0 - 1;
        const int hyg = ty * OY - 1 + hr;
        if (!(hr < ((w == 0) ? TY / 2 : TY) && (hside ? hx < nx : tx > 0) && hyg <= ny)) return;
        int hy;
        if (periodic) {
            hy = hyg < 0 ? hyg + ny : (hyg >= ny ? hyg - ny : hyg);
        } else {
            if (hyg < 0 || hyg >= ny) return;
            hy = hyg;
        }
        int hym = hy - 1, hyp = hy + 1;
        bool hlo = false, hhi = false;
        if (periodic) {
            hym = hym < 0 ? hym + ny : hym;
            hyp = hyp >= ny ? hyp - ny : hyp;
        } else {
            if (hym < 0) { hym = hy; hlo = true; }
            if (hyp >= ny) { hyp = hy; hhi = true; }
        }
        const uint32_t hxo = (uint32_t)hx * ES, hxm = (uint32_t)(hx > 0 ? hx - 1 : 0) * ES,
                       hxp = (uint32_t)(hx + 1 < nx ? hx + 1 : nx - 1) * ES;
        // the streamed values first, then the cell's link mask: the mask's wait then covers
        // them all (a mask loaded first had its own wait before the other loads were issued)
        B hraw[Q];
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            hraw[i] = buf_load<B>(rin, src_off(I, kh, hlo, hhi, hym, hy, hyp) + (CX[i] < 0 ? hxp : (CX[i] > 0 ? hxm : hxo)),
                                  0u);
        });
        const uint32_t hmask = a.links[((uint32_t)kh * ny + hy) * nx + hx];
        // a solid cell's values are never read (its neighbours bounce back): it writes nothing
        const bool hlive = !(hmask & SOLID_BIT);
        T fl[Q];
        static_for<0, Q>([&](auto I) { fl[I] = Cd::decode(hraw[I], dsh(I, kh), inv_scale); });
        T fxw = T(0), fzw = T(0);
        bounce_fixups(kh, (((uint32_t)kh * ny + hy) * nx + hx) * ES, hmask, fl, fxw, fzw);
        collide<NORM, T>(fl, (uint32_t)hx, a, smag, [&](auto I, T val) {
            constexpr int i = I;
            if constexpr (CX[i] != 0) {
                if (hlive && (CX[i] > 0) == (hside == 0)) {   // x0-1 feeds c_x = +1, x0+64 feeds c_x = -1
                    const T p = stored(I, val, kh);
                    if constexpr (CY[i] == 0) L.H[kh & 3][s2_hidx(i)][hr] = (float)p;
                    else L.B[kh % 3][s2_gidx(i)][hr][hside ? 65 : 0] = (float)p;
                }
            }
        });
    };

    // one wavefront's FP64 sum of its lanes' per-cell FP32 force into red[slot][w][c0..c0+1]
    // (a fixed butterfly; skipped when no lane has a wall link)
    auto wave_force = [&](float fx, float fz, int slot, int c0) {
        double sx = 0.0, sz = 0.0;
        if (__builtin_amdgcn_ballot_w64(fx != 0.0f || fz != 0.0f) != 0) {
            sx = wave_sum((double)fx);
            sz = wave_sum((double)fz);
        }
        if (lane == 0) { L.red[slot][w][c0] = sx; L.red[slot][w][c0 + 1] = sz; }
    };
    // ---- force partials of one plane: wavefront sums in order
    const bool force = a.partials != nullptr;
    // iteration kf's sums sit in red[kf % 3]: phase 1 writes them at iteration kf, wavefront 0
    // reads them in phase 2 of iteration kf + 1, and they are rewritten only at iteration
    // kf + 3, after a barrier that wavefront 0 reaches once it has read them
    auto flush = [&](int slot_plane) {                   // plane whose f1 force sits in red[slot_plane % 3]
        if (!force || lane != 0) return;
        const int s = slot_plane % 3;
        double f1x = 0.0, f1z = 0.0, f2x = 0.0, f2z = 0.0;
        for (int r = 1; r <= OY; r++) {
            f1x += L.red[s][r][0]; f1z += L.red[s][r][1];
            f2x += L.red[s][r][2]; f2z += L.red[s][r][3];
        }
        const int k1 = slot_plane, k2 = slot_plane - 1;
        if (k1 >= za && k1 < zb) {
            double* p = a.partials + (((size_t)(t0 % a.ring) * nz + k1) * P + tile) * 2;
            p[0] = f1x; p[1] = f1z;
        }
        if (k2 >= za && k2 < zb) {
            double* p = a.partials + (((size_t)((t0 + 1) % a.ring) * nz + k2) * P + tile) * 2;
            p[0] = f2x; p[1] = f2z;
        }
    };

    // encode constants of the 16-bit storage at height k (the weights for FP32)
    auto encode_shift = [&](int k, T (&en)[3]) {
        static_for<0, 3>([&](auto C) { en[C] = (T)neg_scaled((float)sig(k, C), kscale); });
    };

    static_assert(!(SINGLE && ACC), "the accumulating form is a two-step sweep");
    if constexpr (SINGLE) {
        // ---- one step: the output rows compute f1 of every plane and store it; the force of
        // step t0 goes to slot t0 % ring in the per-tile layout, summed as the pair's first step
        auto flush1 = [&](int kf) {
            if (!force || lane != 0 || kf < za || kf >= zb) return;
            const int s = kf % 3;
            double fx = 0.0, fz = 0.0;
            for (int r = 1; r <= OY; r++) { fx += L.red[s][r][0]; fz += L.red[s][r][1]; }
            double* p = a.partials + (((size_t)(t0 % a.ring) * nz + kf) * P + tile) * 2;
            p[0] = fx; p[1] = fz;
        };
        if (force && w == 0 && lane == 0 && za == 1) {
            double* p = a.partials + (((size_t)(t0 % a.ring) * nz + 0) * P + tile) * 2;
            p[0] = 0.0; p[1] = 0.0;
        }
#pragma clang loop unroll(disable)
        for (int k = za; k < zb; k++) {
            if (out_row) {
                T fl[Q];
                T fxw = T(0), fzw = T(0);
                bool live = true;
                const bool fast = chunk_fast(y, k, tx);
                uint32_t mm = 0;
                if (!fast) mm = a.links[((uint32_t)k * ny + y) * nx + xl];
                f1_load(k, fl);
                if (!fast) {
                    live = !(mm & SOLID_BIT);
                    bounce_fixups(k, (((uint32_t)k * ny + y) * nx + xl) * ES, mm, fl, fxw, fzw);
                }
                T en[3];
                encode_shift(k, en);
                auto esh = [&](auto I) -> T {
                    constexpr int i = I;
                    if constexpr (NORM) return en[xclass(CX[i])];
                    else return wgt<T>(i);
                };
                const bool st_ok = x < nx && live;
                const uint32_t rowb = ((uint32_t)k * ny + (uint32_t)y) * (uint32_t)nx * ES;
                if (x == 0) {                                                // inlet (lbm3d.py:67-71)
                    const T u0 = inlet_u0(k, t0);
                    const T usq15 = T(1.5) * (u0 * u0);
                    static_for<0, Q>([&](auto I) {
                        constexpr int i = I;
                        if (st_ok) buf_store<B, 2>(Cd::encode(inlet_value<NORM, i, T>(u0, usq15), esh(I), scale), rout,
                                                   xo, (uint32_t)i * NB + rowb);
                    });
                } else {
                    collide<NORM, T>(fl, (uint32_t)xl, a, smag, [&](auto I, T val) {
                        constexpr int i = I;
                        if (st_ok) buf_store<B, 2>(Cd::encode(val, esh(I), scale), rout, xo, (uint32_t)i * NB + rowb);
                    });
                }
                if (force) wave_force(st_ok && x != 0 ? (float)fxw : 0.0f, st_ok && x != 0 ? (float)fzw : 0.0f, k % 3, 0);
            } else if (force && lane == 0) {
                L.red[k % 3][w][0] = 0.0;
                L.red[k % 3][w][1] = 0.0;
            }
            __syncthreads();
            if (w == 0 && k > za) flush1(k - 1);
        }
        __syncthreads();
        if (w == 0) flush1(zb - 1);
        return;
    }

    // ---- prologue: the ground plane's partials are zero; halo of the first plane swept
    // (ks: plane 1, or the plane below the segment, whose first step the second step needs)
    const int ks = za > 1 ? za - 1 : 1;
    const int ke = zb >= nz ? nz : zb;                     // last iteration (plane ke's first step)
    if (force && w == 0 && lane == 0 && za == 1) {
        for (int s = 0; s < 2; s++) {
            double* p = a.partials + (((size_t)((t0 + s) % a.ring) * nz + 0) * P + tile) * 2;
            p[0] = 0.0; p[1] = 0.0;
        }
    }
    if constexpr (ACC != 0) {
        if (out_row && za == 1 && x < nx) solid_accumulate(a.acc, (size_t)N, (size_t)y * nx + x);   // ground plane
    }
    if (halo_wave && ks < nz) halo_f1(ks);
    B pre[Q];                                                // PF: streamed loads of the next plane
    if constexpr (PF) {
        if (need && ks < nz) f1_issue(ks, pre);
    }

    T h0[3] = {}, hp1[3] = {}, hp2[3] = {}, hm1[3] = {};   // own-column f1 populations kept in registers
    T c6 = T(0), c12 = T(0), c13 = T(0);                     // f1(k) of populations 6, 12, 13
    uint32_t m_prev = 0;                                     // link mask of this lane's cell at plane k-1

#pragma clang loop unroll(disable)
    for (int k = ks; k <= ke; k++) {
        // ================= phase 1: f1 at plane k
        T pk[Q];                                  // what the second step reads of f1(k)
        uint32_t m_cur = 0;
        if (need && k < nz) {
            const bool take1 = ACC == 1 && out_row && x < nx && k >= za && k < zb;
            float prev1[4];
            if constexpr (ACC == 1) {
                if (take1) acc_fetch(k, prev1);
            }
            T fl[Q];
            T fxw = T(0), fzw = T(0);
            bool live = true;
            const bool fast = chunk_fast(y, k, tx);
            if (!fast) m_cur = a.links[((uint32_t)k * ny + y) * nx + xl];
            if constexpr (PF) f1_decode(k, pre, fl);
            else f1_load(k, fl);
            if (!fast) {
                live = !(m_cur & SOLID_BIT);
                bounce_fixups(k, (((uint32_t)k * ny + y) * nx + xl) * ES, m_cur, fl, fxw, fzw);
            }
            T v[Q];
            collide<NORM, T>(fl, (uint32_t)xl, a, smag, [&](auto I, T val) { v[I] = val; });
            if (x == 0) {                                                    // inlet (lbm3d.py:67-71)
                const T u0 = inlet_u0(k, t0);
                const T usq15 = T(1.5) * (u0 * u0);
                static_for<0, Q>([&](auto I) { v[I] = inlet_value<NORM, I, T>(u0, usq15); });
            }
            static_for<0, Q>([&](auto I) { pk[I] = stored(I, v[I], k); });
            if constexpr (ACC == 1) {
                if (take1) acc_sample(k, prev1, pk, live);
            }
            if (force) wave_force(out_row && x < nx && x != 0 && live ? (float)fxw : 0.0f,
                                  out_row && x < nx && x != 0 && live ? (float)fzw : 0.0f, k % 3, 0);
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                if constexpr (CY[i] != 0) lds_row(L, i, k, s2_gidx(i), w, lane + 1) = (float)pk[i];
            });
        } else {
            static_for<0, Q>([&](auto I) { pk[I] = T(0); });
            if (force && lane == 0) { L.red[k % 3][w][0] = 0.0; L.red[k % 3][w][1] = 0.0; }
        }
        c6 = pk[6]; c12 = pk[12]; c13 = pk[13];
        if constexpr (PF) {
            if (need && k + 1 < nz && k + 1 <= ke) f1_issue(k + 1, pre);
        }
        __syncthreads();

        // ================= phase 2: f2 at plane kk = k-1 (output rows); halo f1 at k+1 (halo rows)
        const int kk = k - 1;
        float fbx = 0.0f, fbz = 0.0f;
        if (out_row && kk >= za) {
            const bool top = kk == nz - 1;
            float prev2[4];
            if constexpr (ACC == 2) {
                if (x < nx) acc_fetch(kk, prev2);
            }
            T g[Q];
            auto shr = [&](T v, T edge) { return __uint_as_float(wave_shr1(__float_as_uint(v), __float_as_uint(edge))); };
            auto shl = [&](T v, T edge) {
                const T s = __uint_as_float(wave_shl1(__float_as_uint(v), __float_as_uint(edge)));
                return outlet ? v : s;                                       // outlet: the cell pulls itself
            };
            g[0] = h0[0];
            g[1] = shr(h0[1], L.H[kk & 3][0][w]);
            g[2] = shl(h0[2], L.H[kk & 3][3][w]);
            g[5] = hp2[0];
            g[11] = shr(hp2[1], L.H[(kk - 1) & 3][1][w]);
            g[14] = shl(hp2[2], L.H[(kk - 1) & 3][5][w]);
            if (!top) {
                g[6] = c6;
                g[12] = shl(c12, L.H[k & 3][4][w]);
                g[13] = shr(c13, L.H[k & 3][2][w]);
            } else {                                                         // top: mirror onto plane kk
                g[6] = hp1[0];
                g[12] = shl(hp1[2], L.H[kk & 3][5][w]);
                g[13] = shr(hp1[1], L.H[kk & 3][1][w]);
            }
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                if constexpr (CY[i] != 0) {
                    const bool ym_ = CY[i] > 0 ? mlo : mhi;                 // free-slip mirror at this row
                    const int r = ym_ ? w : w - CY[i];
                    const int col = (CX[i] < 0 && outlet) ? lane + 1 : lane + 1 - CX[i];
                    if constexpr (CZ[i] < 0) {
                        if (top) {                                          // onto plane kk, population MIRZ[j]
                            constexpr int jz = MIRZ[i], jyz = MIRZ[MIRY[i]];
                            g[i] = lds_row(L, jz, kk, ym_ ? s2_gidx(jyz) : s2_gidx(jz), r, col);
                        } else {
                            g[i] = lds_row(L, i, k, ym_ ? s2_gidx(MIRY[i]) : s2_gidx(i), r, col);
                        }
                    } else {
                        g[i] = lds_row(L, i, kk - CZ[i], ym_ ? s2_gidx(MIRY[i]) : s2_gidx(i), r, col);
                    }
                }
            });
            // the cell's own f1 populations at plane kk (bounce-back sources)
            auto own = [&](auto J) -> T {
                constexpr int j = J;
                if constexpr (j == 0 || j == 1 || j == 2) return h0[j];
                else if constexpr (j == 5) return hp1[0];
                else if constexpr (j == 11) return hp1[1];
                else if constexpr (j == 14) return hp1[2];
                else if constexpr (j == 6) return hm1[0];
                else if constexpr (j == 12) return hm1[1];
                else if constexpr (j == 13) return hm1[2];
                else return lds_row(L, j, kk, s2_gidx(j), w, lane + 1);
            };
            bool live = true;
            T fxw = T(0), fzw = T(0);
            if (!chunk_fast(y, kk, tx)) {
                const uint32_t mm = m_prev;
                live = !(mm & SOLID_BIT);
                static_for<1, Q>([&](auto I) {
                    constexpr int i = I;
                    if ((mm >> i) & 1u) g[i] = own(std::integral_constant<int, OPP[i]>{});
                    if (is_wall(I, mm)) wall_force(I, g[i], fxw, fzw);
                });
            }
            T en[3];
            encode_shift(kk, en);
            auto esh = [&](auto I) -> T {
                constexpr int i = I;
                if constexpr (NORM) return en[xclass(CX[i])];
                else return wgt<T>(i);
            };
            const bool st_ok = x < nx && live;
            const uint32_t rowb = ((uint32_t)kk * ny + (uint32_t)y) * (uint32_t)nx * ES;
            T sv[ACC == 2 ? Q : 1];                                          // (ACC 2) the stored state
            if (x == 0) {                                                    // inlet
                const T u0 = inlet_u0(kk, t0 + 1);
                const T usq15 = T(1.5) * (u0 * u0);
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    const T val = inlet_value<NORM, i, T>(u0, usq15);
                    if (st_ok) buf_store<B, 2>(Cd::encode(val, esh(I), scale), rout, xo, (uint32_t)i * NB + rowb);
                    if constexpr (ACC == 2) sv[i] = stored(I, val, kk);
                });
            } else {
                collide<NORM, T>(g, (uint32_t)xl, a, smag, [&](auto I, T val) {
                    constexpr int i = I;
                    if (st_ok) buf_store<B, 2>(Cd::encode(val, esh(I), scale), rout, xo, (uint32_t)i * NB + rowb);
                    if constexpr (ACC == 2) sv[i] = stored(I, val, kk);
                });
                if (st_ok) { fbx = (float)fxw; fbz = (float)fzw; }
            }
            if constexpr (ACC == 2) {
                if (x < nx) acc_sample(kk, prev2, sv, live);
            }
        }
        if (halo_wave && k + 1 < nz && k < ke) halo_f1(k + 1);
        if (w == 0 && k > ks) flush(k - 1);
        // per-wavefront force sums of this iteration
        if (force) wave_force(fbx, fbz, k % 3, 2);
        // rotate the registers: plane k becomes k-1
        hp2[0] = hp1[0]; hp2[1] = hp1[1]; hp2[2] = hp1[2];
        hp1[0] = pk[5]; hp1[1] = pk[11]; hp1[2] = pk[14];
        h0[0] = pk[0]; h0[1] = pk[1]; h0[2] = pk[2];
        hm1[0] = c6; hm1[1] = c12; hm1[2] = c13;
        m_prev = m_cur;
    }
    __syncthreads();
    if (w == 0) flush(ke);
}


// Workgroups are handed to the XCDs round-robin (workgroup b to XCD b mod NXCD), each XCD
// with its own L2.  The two-step launches give XCD x one contiguous band of the tile numbers,
// [band_start(x), band_start(x + 1)), so the halo rows and columns a tile reads were mostly
// just read by its neighbours in the same L2.
constexpr int NXCD = 8;
__host__ __device__ constexpr int band_start(int x, int ntiles) {
    return x * (ntiles / NXCD) + (x < ntiles % NXCD ? x : ntiles % NXCD);
}

// Tile rows of a two-step launch: TY (15 or 16) rows per workgroup, TY - 2 of them output rows.
__host__ __device__ constexpr int step2_tiles_for(int nx, int ny, int ty) {
    return ((nx + 63) / 64) * ((ny + ty - 3) / (ty - 2));
}

// force[t][z] = sum over the P tiles of partials[t % ring][z][tile]: lane-strided FP64 sums
// and a fixed butterfly, one wavefront per (t, z)
__device__ __forceinline__ void force_row2(const double* partials, int ring, int t, int z, int nz, int P, int lane,
                                           double* force) {
    const double* p = partials + (((size_t)(t % ring) * nz + z) * P) * 2;
    double sx = 0.0, sz = 0.0;
    for (int k = lane; k < P; k += 64) { sx += p[2 * k]; sz += p[2 * k + 1]; }
    sx = wave_sum(sx);
    sz = wave_sum(sz);
    if (lane == 0) {
        force[((size_t)t * nz + z) * 2 + 0] = sx;
        force[((size_t)t * nz + z) * 2 + 1] = sz;
    }
}

}  // namespace sw
