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
#include <type_traits>

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
    const float* shift;      // 16-bit storage: [nz][4] offsets sigma (sw_shift_table), or null (sigma = 1)
    float* acc;              // [4][N] accumulator += (rho, ux, uy, uz) of the step's output (ACC kernels)
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
//   NORM = false: fl are the populations f_i; the reference's FP32 operation order.
//   NORM = true : fl are the weight-normalised populations phi_i = f_i / w_i of the 16-bit
//                 storage forms, and so are the results.  The moments come from the sums
//                 a_i = phi_i + phi_opp(i) and differences d_i = phi_i - phi_opp(i) of the
//                 nine opposite pairs (moments_norm); the stress, viscosity and relaxation are
//                 those of NORM = false; phi_out = rho feq_i / w_i + (keep * 4.5) qq_i, formed
//                 from terms shared across directions (see the NORM branch below).
template <class T>
struct Moments { T rho, mx, my, mz, pxx, pyy, pzz, pxy, pxz, pyz; };

// Moments of nine opposite pairs (i, i+1), i odd, of weight-normalised populations:
//   s = (a1 + a3) + a5, s' = ((((a7 + a9) + a11) + a13) + a15) + a17,
//   rho = fma(w0, phi0, fma(w1, s, w2 s')),
//   m_x = fma(w1, d1, w2 (((d7 + d9) + d11) + d13)),  m_y = fma(w1, d3, w2 (((d7 - d9) + d15) + d17)),
//   m_z = fma(w1, d5, w2 (((d11 - d13) + d15) - d17)),
//   Pi_xx = fma(w1, a1, w2 (((a7 + a9) + a11) + a13)),  Pi_yy = fma(w1, a3, w2 (((a7 + a9) + a15) + a17)),
//   Pi_zz = fma(w1, a5, w2 (((a11 + a13) + a15) + a17)),
//   Pi_xy = w2 (a7 - a9),  Pi_xz = w2 (a11 - a13),  Pi_yz = w2 (a15 - a17)
// (w0, w1, w2 = 1/3, 1/18, 1/36; the pair leaders 1, 3, 5, 7, 9, 11, 13, 15, 17 have
// c = +x, +y, +z, (1,1,0), (1,-1,0), (1,0,1), (1,0,-1), (0,1,1), (0,1,-1)).
template <class T, bool STRESS>
__device__ __forceinline__ Moments<T> moments_norm(const T (&fl)[Q]) {
    constexpr T w0 = wgt<T>(0), w1 = wgt<T>(1), w2 = wgt<T>(7);
    T ap[Q], dp[Q];
    static_for<0, 9>([&](auto P) {
        constexpr int i = 2 * P + 1;
        ap[i] = fl[i] + fl[i + 1];
        dp[i] = fl[i] - fl[i + 1];
    });
    Moments<T> m;
    const T s1 = (ap[1] + ap[3]) + ap[5];
    const T s2 = ((((ap[7] + ap[9]) + ap[11]) + ap[13]) + ap[15]) + ap[17];
    m.rho = fma(w0, fl[0], fma(w1, s1, w2 * s2));
    m.mx = fma(w1, dp[1], w2 * (((dp[7] + dp[9]) + dp[11]) + dp[13]));
    m.my = fma(w1, dp[3], w2 * (((dp[7] - dp[9]) + dp[15]) + dp[17]));
    m.mz = fma(w1, dp[5], w2 * (((dp[11] - dp[13]) + dp[15]) - dp[17]));
    if constexpr (STRESS) {
        m.pxx = fma(w1, ap[1], w2 * (((ap[7] + ap[9]) + ap[11]) + ap[13]));
        m.pyy = fma(w1, ap[3], w2 * (((ap[7] + ap[9]) + ap[15]) + ap[17]));
        m.pzz = fma(w1, ap[5], w2 * (((ap[11] + ap[13]) + ap[15]) + ap[17]));
        m.pxy = w2 * (ap[7] - ap[9]);
        m.pxz = w2 * (ap[11] - ap[13]);
        m.pyz = w2 * (ap[15] - ap[17]);
    }
    return m;
}

template <bool NORM, class T, class A, class Store>
__device__ __forceinline__ void collide(const T (&fl)[Q], uint32_t x, A& a, T smag, Store&& st) {
    T rho, mx, my, mz, pxx, pyy, pzz, pxy, pxz, pyz;
    if constexpr (NORM) {
        const Moments<T> m = moments_norm<T, true>(fl);
        rho = m.rho; mx = m.mx; my = m.my; mz = m.mz;
        pxx = m.pxx; pyy = m.pyy; pzz = m.pzz; pxy = m.pxy; pxz = m.pxz; pyz = m.pyz;
    } else {
        // ---- moments (:103-127), sequential in i; zero-coefficient terms dropped
        rho = fl[0];
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
    }
    const T yr = rcp_rn(rho);                                                  // :128-138
    const T ux = div_rn(mx, rho, yr), uy = div_rn(my, rho, yr), uz = div_rn(mz, rho, yr);
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
    const T tau = T(0.5) * (tau0 + sqrt(tau0 * tau0 + div_rn(smag * qn, rho, yr)));
    const T keep = T(1) - rcp_rn(tau);
    const T usq15 = T(1.5) * ((ux * ux + uy * uy) + uz * uz);

    // ---- regularized collision (:149-155)
    // Directions come in opposite pairs (i, i+1), i odd, c_{i+1} = -c_i.  Negation is exact
    // in IEEE arithmetic, so the pair's cu values are +-cu, their (0.5*cu)*cu terms are
    // equal and their stress projections qq are identical: each is computed once, and the
    // results are bit-identical to evaluating lbm3d.py:153-155 separately per direction.
    constexpr T third_c = T(1) / T(3);
    constexpr T A1 = T(1) - third_c;      // (c*c - 1/3) for |c| = 1
    constexpr T A0 = T(0) - third_c;      // (c*c - 1/3) for c = 0
    if constexpr (NORM) {
        const T K = keep * T(4.5);        // (keep * 4.5) * qq is the normalised neq
        // Normalised output phi_out_i = rho E_i + K qq_i from terms shared across directions.
        // With the trace term Kt = K (nxx + nyy + nzz) / 3, K qq_i is K n_aa - Kt for an axis
        // direction a, (K n_aa + K n_bb - Kt) +- 2 K n_ab for a diagonal one in the ab plane,
        // and -Kt at rest; rho E_i = rho (1 - 1.5 u.u) + (0.5 rho cu) cu +- rho cu.  Per
        // opposite pair: R = rho (1 - 1.5 u.u) + K qq, B = fma(0.5 rho cu, cu, R),
        // phi_out = fma(+-rho, cu, B).  74 operations for the 19 outputs.
        static_assert(CX[1] == 1 && CY[3] == 1 && CZ[5] == 1, "axis pair leaders");
        static_assert(CX[7] == 1 && CY[7] == 1 && CX[9] == 1 && CY[9] == -1, "xy pair leaders");
        static_assert(CX[11] == 1 && CZ[11] == 1 && CX[13] == 1 && CZ[13] == -1, "xz pair leaders");
        static_assert(CY[15] == 1 && CZ[15] == 1 && CY[17] == 1 && CZ[17] == -1, "yz pair leaders");
        const T u3x = T(3) * ux, u3y = T(3) * uy, u3z = T(3) * uz;
        const T kxx = K * nxx, kyy = K * nyy, kzz = K * nzz;
        const T k2 = T(2) * K;
        const T kxy = k2 * nxy, kxz = k2 * nxz, kyz = k2 * nyz;
        const T sxy = kxx + kyy, sxz = kxx + kzz, syz = kyy + kzz;
        const T kt = (sxy + kzz) * third_c;
        const T r0 = rho * (T(1) - usq15) - kt;
        const T rh = T(0.5) * rho;
        st(std::integral_constant<int, 0>{}, r0);
        auto pair = [&](auto I, T cu, T R) {
            const T B = fma(rh * cu, cu, R);
            st(I, fma(rho, cu, B));
            st(std::integral_constant<int, decltype(I)::value + 1>{}, fma(-rho, cu, B));
        };
        pair(std::integral_constant<int, 1>{}, u3x, r0 + kxx);
        pair(std::integral_constant<int, 3>{}, u3y, r0 + kyy);
        pair(std::integral_constant<int, 5>{}, u3z, r0 + kzz);
        const T gxy = r0 + sxy, gxz = r0 + sxz, gyz = r0 + syz;
        pair(std::integral_constant<int, 7>{}, u3x + u3y, gxy + kxy);
        pair(std::integral_constant<int, 9>{}, u3x - u3y, gxy - kxy);
        pair(std::integral_constant<int, 11>{}, u3x + u3z, gxz + kxz);
        pair(std::integral_constant<int, 13>{}, u3x - u3z, gxz - kxz);
        pair(std::integral_constant<int, 15>{}, u3y + u3z, gyz + kyz);
        pair(std::integral_constant<int, 17>{}, u3y - u3z, gyz - kyz);
        return;
    }
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
        const T cu = T(3) * cdotu<i>(ux, uy, uz);
        const T h = (T(0.5) * cu) * cu;
        const T neq = ((keep * w) * T(4.5)) * qq;
        const T wr = w * rho;
        st(std::integral_constant<int, i>{}, wr * (((T(1) + cu) + h) - usq15) + neq);
        st(std::integral_constant<int, j>{}, wr * (((T(1) + (-cu)) + h) - usq15) + neq);
    });
}

// k_macro (lbm3d.py:158-182) of one cell's stored populations sv[] (decoded as a following
// step or sw_macro reads them), added into the accumulator as k_accum does (:185-188):
// sequential sums in direction order (FP32 / FP64) or the opposite-pair moments of the
// 16-bit forms, u = m / rho by IEEE division, acc[k] = acc[k] + u[k].
template <bool NORM, class T>
__device__ __forceinline__ void macro_values(const T (&sv)[Q], float (&u)[4]) {
    T rho = T(0), mx = T(0), my = T(0), mz = T(0);
    if constexpr (NORM) {
        const Moments<T> m = moments_norm<T, false>(sv);
        rho = m.rho; mx = m.mx; my = m.my; mz = m.mz;
    } else {
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            const T v = sv[i];
            rho = rho + v;
            if constexpr (CX[i] != 0) mx = CX[i] > 0 ? mx + v : mx - v;
            if constexpr (CY[i] != 0) my = CY[i] > 0 ? my + v : my - v;
            if constexpr (CZ[i] != 0) mz = CZ[i] > 0 ? mz + v : mz - v;
        });
    }
    u[0] = (float)rho; u[1] = (float)(mx / rho); u[2] = (float)(my / rho); u[3] = (float)(mz / rho);
}
template <bool NORM, class T>
__device__ __forceinline__ void macro_accumulate(float* acc, size_t N, size_t cell, const T (&sv)[Q]) {
    float u[4];
    macro_values<NORM, T>(sv, u);
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k * N + cell] = acc[k * N + cell] + u[k];
}
// a solid cell's sample: (1, 0, 0, 0) as k_macro gives it
__device__ __forceinline__ void solid_accumulate(float* acc, size_t N, size_t cell) {
    const float u[4] = {1.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k * N + cell] = acc[k * N + cell] + u[k];
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
//   PAIR: 16- and 32-bit storage — a lane owns two adjacent cells, each link is one
//         aligned 32-bit (16-bit forms) or 64-bit (FP32) load, x-1 / x+1 neighbours by
//         whole-wavefront DPP shifts, and one store of the same width; FP64 ignores it
//   QUAD: 16-bit storage — a lane owns four adjacent cells, one aligned 64-bit load and
//         store per link; the host maps it to the pair tile for FP32 and to variant 1
//         for FP64
template <int CPL_, bool PF_, bool NT_, bool NTL_ = false, int WPB_ = ROWS, int OCC_ = 0, bool PAIR_ = false,
          bool QUAD_ = false>
struct Tile {
    static constexpr bool PAIR = PAIR_;
    static constexpr bool QUAD = QUAD_;
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
using TileV8 = Tile<1, false, true, false, 4, 0, true>;   // pairs, NT stores
using TileV9 = Tile<1, false, true, false, 8, 0, true>;   // pairs, NT stores, rows over 2 wavefronts
using TileV10 = Tile<1, false, true, false, 16, 0, true>; // pairs, NT stores, rows over 4 wavefronts
using TileV11 = Tile<1, false, false, false, 8, 0, true>; // pairs, cached stores, rows over 2 wavefronts
using TileV12 = Tile<1, false, true, false, 4, 0, false, true>;  // 16-bit quads, NT stores
using TileV13 = Tile<1, false, true, false, 8, 0, false, true>;  // 16-bit quads, NT stores, rows over 2 wavefronts
constexpr int NUM_TILE_VARIANTS = 14;
#define SW_TILE_LIST(X) X(0, TileV0) X(1, TileV1) X(2, TileV2) X(3, TileV3) X(4, TileV4) X(5, TileV5) \
                        X(6, TileV6) X(7, TileV7) X(8, TileV8) X(9, TileV9) X(10, TileV10) X(11, TileV11) \
                        X(12, TileV12) X(13, TileV13)

// One (y, z) row of one lattice; adds this lane's FP64 wall-force sums to fx_acc/fz_acc.
// BUF selects raw-buffer addressing (all byte offsets of a population array < 4 GiB).
// ACC: also add (rho, u) of every cell's output state into a.acc (k_macro + k_accum on the
// step's result, lbm3d.py:363-370), without re-reading the populations.
// A: StepArgs, or a StepArgs in the constant address space (the batched kernels' per-lattice
// arguments, whose loads the compiler may repeat instead of keeping them in registers).
template <class S, class TL, bool BUF, bool ACC = false, class A>
__device__ __forceinline__ void step_row(A& a, int y, int z, int t, int lane, int seg,
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
    // Byte offsets of rows: 32-bit with raw-buffer addressing (every offset of the array is
    // below 4 GiB), which halves the scalar registers the 19 row offsets take.
    using Off = std::conditional_t<BUF, uint32_t, size_t>;
    const Off selfb = (Off)((size_t)selfrow * ES);
    const Off NBo = (Off)NB;
    T inv_scale = (T)a.inv_scale, scale = (T)a.scale;
    // 16-bit decodes fma(h, inv_scale, sigma): inv_scale in a vector register, sigma scalar.
    // (A vector copy of the encode scale as well costs a VGPR: 73, one wavefront per SIMD less.)
    if constexpr (ES == 2) inv_scale = vgpr_copy(a.inv_scale);
    const int kscale = (int)((__builtin_bit_cast(uint32_t, a.scale) >> 23) & 0xFFu) - 127;
    // (18*sqrt(2))*cs*cs evaluated left to right as lbm3d.py:146 does
    const T cs = (T)a.cs;
    const T smag = ((T(18) * (T)sqrt(T(2))) * cs) * cs;

    // wave-uniform byte offset of the source row of every link after the y / z boundary
    // rules, pre-shifted by -cx so that every load of a cell uses the lane offset x*ES
    // (cx = -1 loads use the outlet-clamped lane offset instead: lbm3d.py:82-83)
    Off srow[Q];
    // 16-bit forms: the stored value of direction i at height z is (phi_i - sigma_c(z)) * scale,
    // phi_i = f_i / w_i, with sigma depending on the direction only through its c_x class
    // (0, +1, -1).  Per row this needs sigma of the three classes at heights z-1, z, z+1
    // (clamped to the lattice; the top row's c_z = -1 links mirror onto its own height) for
    // decodes, and -sigma(z) * scale of the three classes for encodes: twelve wave-uniform
    // values.  Mirror images of a direction share its c_x, so the y / z mirrors need no
    // other offsets.  The FP32 / FP64 codecs ignore offsets: compile-time weights are
    // passed instead.
    constexpr bool NORM = (ES == 2);
    T sg[3][3], en[3];
    if constexpr (NORM) {
        const int zl[3] = {z > 0 ? z - 1 : 0, z, z + 1 < (int)nz ? z + 1 : z};
#pragma unroll
        for (int l = 0; l < 3; l++)
#pragma unroll
            for (int c = 0; c < 3; c++) sg[l][c] = a.shift ? (T)ld_const(a.shift, zl[l] * 4 + c) : T(1);
#pragma unroll
        for (int c = 0; c < 3; c++) en[c] = (T)neg_scaled((float)sg[1][c], kscale);
    }
    auto dsh = [&](auto I) -> T {                     // decode offset of link I's source
        constexpr int i = I;
        if constexpr (NORM) return sg[CZ[i] > 0 ? 0 : (CZ[i] < 0 ? 2 : 1)][xclass(CX[i])];
        else return wgt<T>(i);
    };
    auto esh = [&](auto I) -> T {                     // encode constant of direction I
        constexpr int i = I;
        if constexpr (NORM) return en[xclass(CX[i])];
        else return wgt<T>(i);
    };
    auto bsh = [&](auto I) -> T {                     // decode offset of the own opposite population
        constexpr int i = I;
        if constexpr (NORM) return sg[1][xclass(-CX[i])];
        else return wgt<T>(OPP[i]);
    };
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
        srow[i] = (Off)((CX[i] > 0) ? r - ES : r);
    });
    const uint32_t* lrow = a.links + selfrow;
    const uint32_t cw = chunk_words(nx);
    const uint32_t* rflags = a.links + N + rowid * cw;           // wave-uniform chunk flags
    const uint32_t nch_row = (nx + 63) / 64;
    const uint32_t seglen = (nch_row + TL::NSEG - 1) / TL::NSEG;
    const uint32_t cbeg = min(nch_row, (uint32_t)seg * seglen);
    const uint32_t nchunks = min(nch_row, cbeg + seglen);        // end of this wavefront's range
    auto special_chunk = [&](uint32_t c) -> bool { return (ld_const(rflags, c >> 5) >> (c & 31)) & 1u; };

    auto load_fast = [&](uint32_t x, T (&fl)[Q]) {
        const uint32_t xo = x * ES;
        const uint32_t xpo = min(x + 1, nx - 1) * ES;
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            fl[i] = C::decode(mem.template ld<TL::NTL>(srow[i], (CX[i] < 0) ? xpo : xo), dsh(I), inv_scale);
        });
    };
    auto osh = [&](auto I) -> T {                     // decode offset of the cell's own population I
        constexpr int i = I;
        if constexpr (NORM) return sg[1][xclass(CX[i])];
        else return wgt<T>(i);
    };
    auto collide_store = [&](uint32_t x, const T (&fl)[Q]) {
        const uint32_t xo = x * ES;
        if constexpr (ACC) {
            T sv[Q];                                   // the stored values as sw_macro decodes them
            collide<NORM, T>(fl, x, a, smag, [&](auto I, T v) {
                constexpr int i = I;
                const typename C::B e = C::encode(v, esh(I), scale);
                mem.template st<TL::NT>((Off)i * NBo + selfb, xo, e);
                sv[i] = C::decode(e, osh(I), inv_scale);
            });
            macro_accumulate<NORM, T>(a.acc, N, (size_t)selfrow + x, sv);
        } else {
            collide<NORM, T>(fl, x, a, smag, [&](auto I, T v) {
                constexpr int i = I;
                mem.template st<TL::NT>((Off)i * NBo + selfb, xo, C::encode(v, esh(I), scale));
            });
        }
    };
    // general path: per-link bounce-back and wall force (:77-101).  Each lane picks the
    // address of every link (streamed source, or its own opposite population on a
    // bounce-back link) and all 19 loads are issued before the first is used.  Lanes past
    // the row end, at the inlet or on a solid cell load in-row values and store nothing.
    auto general_cell = [&](uint32_t x) {
        const bool in_row = x < nx && x != 0;
        const uint32_t xs = in_row ? x : 1u;                                         // nx >= 2
        const uint32_t mm = lrow[xs];
        const bool live = in_row && !(mm & SOLID_BIT);                              // :65-66
        const size_t xo = (size_t)xs * ES;
        const size_t xpo = (size_t)min(xs + 1, nx - 1) * ES;
        typename C::B raw[Q];
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            size_t off;
            if constexpr (i == 0) off = selfb + xo;
            else off = ((mm >> i) & 1u) ? (size_t)OPP[i] * NB + selfb + xo : srow[i] + ((CX[i] < 0) ? xpo : xo);
            raw[i] = mem.ld_at(off);
        });
        if (!live) {
            if constexpr (ACC) { if (in_row) solid_accumulate(a.acc, N, (size_t)selfrow + x); }   // solid cell
            return;
        }
        T fl[Q];
        T fxw = T(0), fzw = T(0);
        bool hit = false;
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr T w = wgt<T>(i);
            if constexpr (i == 0) fl[i] = C::decode(raw[i], dsh(I), inv_scale);
            else fl[i] = C::decode(raw[i], ((mm >> i) & 1u) ? bsh(I) : dsh(I), inv_scale);
            if constexpr (i != 0) {
                bool wall = (mm >> i) & 1u;                                          // half-way bounce-back
                if constexpr (GROUND_SLOT[i] >= 0) wall = wall && !((mm >> (GROUND_BIT0 + GROUND_SLOT[i])) & 1u);
                if (wall) {                                                          // momentum exchange
                    // 2 (f_i - w_i); the 16-bit forms hold phi_i = f_i / w_i: (2 w_i) (phi_i - 1)
                    const T fr2 = NORM ? (T(2) * w) * (fl[i] - T(1)) : T(2) * (fl[i] - w);
                    if constexpr (CX[i] > 0) fxw = fxw - fr2;
                    if constexpr (CX[i] < 0) fxw = fxw + fr2;
                    if constexpr (CZ[i] > 0) fzw = fzw - fr2;
                    if constexpr (CZ[i] < 0) fzw = fzw + fr2;
                    hit = true;
                }
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
    static_assert(!(ACC && TL::QUAD), "the accumulate epilogue runs on the one-cell and pair tiles");
    if constexpr (TL::QUAD && ES == 2) {
      if (nx & 3u) {
        plain_sweep();                    // quads need a row length divisible by 4
      } else {
        // 256-cell chunks; lane L owns cells xq = x0 + 4L .. xq + 3.  Each link is one aligned
        // 64-bit load of four 16-bit values (512 B per wavefront instruction); the value
        // that sits in the adjacent lane comes over by DPP and two 16-bit alignbits rebuild
        // the shifted quad; the chunk-edge values come from one broadcast load per shifted
        // link.  A chunk with a special 64-cell sub-chunk runs that sub-chunk on the
        // general path and the others on the one-cell fast path.
        const uint32_t nqc = (nx + 255) / 256;
        const uint32_t qlen = (nqc + TL::NSEG - 1) / TL::NSEG;
        const uint32_t qbeg = min(nqc, (uint32_t)seg * qlen);
        const uint32_t qend = min(nqc, qbeg + qlen);
#pragma clang loop unroll(disable)
        for (uint32_t c = qbeg; c < qend; c++) {
            const uint32_t x0 = c * 256;
            bool special = false;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (4 * c + k < nch_row) special |= special_chunk(4 * c + k);
            if (special) {
#pragma unroll 1
                for (int k = 0; k < 4; k++) {
                    const uint32_t sub = 4 * c + k;
                    if (sub >= nch_row) break;
                    const uint32_t x = sub * 64 + lane;
                    if (special_chunk(sub)) {
                        general_cell(x);
                    } else if (x < nx && x != 0) {
                        T fl[Q];
                        load_fast(x, fl);
                        collide_store(x, fl);
                    }
                }
                continue;
            }
            const uint32_t xq = x0 + 4 * lane;
            const uint32_t xl = min(xq, nx - 4);             // lanes past the row end read in-row
            const uint32_t lb = xl * 2;
            const uint32_t ledge = (x0 >= 2 ? x0 - 2 : 0) * 2;           // dword holding cells x0-2, x0-1
            const uint32_t redge = (x0 + 256 < nx ? x0 + 256 : x0) * 2;   // dword holding cells x0+256, x0+257
            uint32_t rl[Q], rh[Q];             // cells (xq, xq+1) and (xq+2, xq+3), 16 bits each
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                const Off row = (CX[i] > 0) ? srow[i] + ES : srow[i];   // unshifted row start
                const uint64_t cur = mem.ld64(row, lb);
                const uint32_t lo = (uint32_t)cur, hi = (uint32_t)(cur >> 32);
                if constexpr (CX[i] == 0) {
                    rl[i] = lo; rh[i] = hi;
                } else if constexpr (CX[i] > 0) {                        // cells x-1
                    const uint32_t p = wave_shr1(hi, mem.ld32(row, ledge));
                    rl[i] = mid16(lo, p); rh[i] = mid16(hi, lo);
                } else {                                                 // cells x+1
                    uint32_t n = wave_shl1(lo, mem.ld32(row, redge));
                    if (xl + 4 >= nx) n = hi >> 16;                      // outlet: cell nx-1 pulls itself
                    rl[i] = mid16(hi, lo); rh[i] = mid16(n, hi);
                }
            });
            uint32_t ol[Q], oh[Q];
            {
                typename C::Pending pend[Q];
                T fl[Q];
                static_for<0, Q>([&](auto I) { fl[I] = C::decode((uint16_t)(rl[I] & 0xFFFFu), dsh(I), inv_scale); });
                collide<NORM, T>(fl, xq, a, smag, [&](auto I, T v) { pend[I] = C::first(v, esh(I), scale); });
                if (x0 == 0) {                                           // inlet cell x = 0 (:67-71)
                    T u0 = (T)ld_const(a.uin, z);
                    if (a.gust) u0 = u0 * (T)ld_const(a.gust, t);
                    const T usq15 = T(1.5) * (u0 * u0);
                    static_for<0, Q>([&](auto I) {
                        constexpr int i = I;
                        const typename C::Pending e =
                            C::first(inlet_value<NORM, i, T>(u0, usq15), esh(I), scale);
                        pend[i] = (lane == 0) ? e : pend[i];
                    });
                    inlet_done = true;
                }
                static_for<0, Q>([&](auto I) { fl[I] = C::decode((uint16_t)(rl[I] >> 16), dsh(I), inv_scale); });
                collide<NORM, T>(fl, xq + 1, a, smag, [&](auto I, T v) { ol[I] = C::second(pend[I], v, esh(I), scale); });
            }
            {
                typename C::Pending pend[Q];
                T fl[Q];
                static_for<0, Q>([&](auto I) { fl[I] = C::decode((uint16_t)(rh[I] & 0xFFFFu), dsh(I), inv_scale); });
                collide<NORM, T>(fl, xq + 2, a, smag, [&](auto I, T v) { pend[I] = C::first(v, esh(I), scale); });
                static_for<0, Q>([&](auto I) { fl[I] = C::decode((uint16_t)(rh[I] >> 16), dsh(I), inv_scale); });
                collide<NORM, T>(fl, xq + 3, a, smag, [&](auto I, T v) { oh[I] = C::second(pend[I], v, esh(I), scale); });
            }
            if (xq < nx) {
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    mem.template st64<TL::NT>((Off)i * NBo + selfb, xq * 2, (uint64_t)ol[i] | ((uint64_t)oh[i] << 32));
                });
            }
        }
      }
    } else if constexpr (TL::PAIR && ES == 2) {
      if (nx & 1u) {
        plain_sweep();                    // pairs need an even row length
      } else {
        // 128-cell chunks; lane L owns cells xa = x0 + 2L and xa + 1 (nx is even).  Each link
        // is one aligned 32-bit load of two 16-bit values; the value that sits in the
        // adjacent lane comes over by DPP and a 16-bit alignbit rebuilds the shifted pair;
        // the chunk-edge values come from one broadcast load per shifted link, or for the
        // left edge of a c_x = +1 link from the previous chunk (carry, below).  (Issuing the
        // next chunk's loads before this chunk's collisions needs 133 VGPRs, 3 wavefronts
        // per SIMD, and measured 13 % slower.)
        const uint32_t npc = (nx + 127) / 128;
        const uint32_t plen = (npc + TL::NSEG - 1) / TL::NSEG;
        const uint32_t pbeg = min(npc, (uint32_t)seg * plen);
        const uint32_t pend = min(npc, pbeg + plen);
        auto pair_special = [&](uint32_t c) {
            return special_chunk(2 * c) || (2 * c + 1 < nch_row && special_chunk(2 * c + 1));
        };
        // a chunk with a special half: that half on the general path, the other on the
        // one-cell fast path
        auto special_pair_chunk = [&](uint32_t c) {
#pragma unroll 1
            for (uint32_t sub = 2 * c; sub < min(2 * c + 2, nch_row); sub++) {
                const uint32_t x = sub * 64 + lane;
                if (special_chunk(sub)) {
                    general_cell(x);
                } else if (x < nx && x != 0) {
                    T fl[Q];
                    load_fast(x, fl);
                    collide_store(x, fl);
                }
            }
        };
        // The left-edge word of a c_x = +1 link (cells x0-2, x0-1) is lane 63's word of the
        // same link in the previous chunk of this sweep: after a fast pair chunk it is carried
        // over in a scalar register (carry[]) instead of being loaded again, a load whose
        // 64-byte sector has usually left the L2 by then.
        auto issue = [&](uint32_t c, uint32_t (&cur)[Q], uint32_t (&edge)[Q], bool carried,
                         const uint32_t (&carry)[Q]) {
            const uint32_t x0 = c * 128;
            const uint32_t lb = min(x0 + 2 * lane, nx - 2) * 2;          // lanes past the row end read in-row
            const uint32_t ledge = (x0 >= 2 ? x0 - 2 : 0) * 2;            // dword holding cell x0-1
            const uint32_t redge = (x0 + 128 < nx ? x0 + 128 : x0) * 2;   // dword holding cell x0+128
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                const Off row = (CX[i] > 0) ? srow[i] + ES : srow[i];   // unshifted row start
                cur[i] = mem.template ld32<TL::NTL>(row, lb);
                if constexpr (CX[i] < 0) edge[i] = mem.ld32(row, redge);
            });
            if (carried) {
                static_for<0, Q>([&](auto I) {
                    if constexpr (CX[I] > 0) edge[I] = carry[I];
                });
            } else {
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    if constexpr (CX[i] > 0) edge[i] = mem.ld32(srow[i] + ES, ledge);
                });
            }
        };
        auto finish = [&](uint32_t c, const uint32_t (&cur)[Q], const uint32_t (&edge)[Q], uint32_t (&carry)[Q]) {
            const uint32_t x0 = c * 128;
            const uint32_t xa = x0 + 2 * lane;
            const uint32_t xl = min(xa, nx - 2);
            uint32_t raw[Q];
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                if constexpr (CX[i] == 0) {
                    raw[i] = cur[i];
                } else if constexpr (CX[i] > 0) {                        // cells (x-1, x)
                    raw[i] = mid16(cur[i], wave_shr1(cur[i], edge[i]));
                    carry[i] = __builtin_amdgcn_readlane(cur[i], 63);   // cells x0+126, x0+127
                } else {                                                 // cells (x+1, x+2)
                    uint32_t nxt = wave_shl1(cur[i], edge[i]);
                    if (xl + 2 >= nx) nxt = cur[i] >> 16;               // outlet: cell nx-1 pulls itself
                    raw[i] = mid16(nxt, cur[i]);
                }
            });
            typename C::Pending pnd[Q];
            uint32_t out[Q];
            {
                T fl[Q];
                static_for<0, Q>([&](auto I) {
                    fl[I] = C::decode((uint16_t)(raw[I] & 0xFFFFu), dsh(I), inv_scale);
                });
                collide<NORM, T>(fl, xa, a, smag, [&](auto I, T v) { pnd[I] = C::first(v, esh(I), scale); });
            }
            if (x0 == 0) {                                               // inlet cell x = 0 (:67-71)
                T u0 = (T)ld_const(a.uin, z);
                if (a.gust) u0 = u0 * (T)ld_const(a.gust, t);
                const T usq15 = T(1.5) * (u0 * u0);
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    const typename C::Pending e = C::first(inlet_value<NORM, i, T>(u0, usq15), esh(I), scale);
                    pnd[i] = (lane == 0) ? e : pnd[i];
                });
                inlet_done = true;
            }
            {
                T fl[Q];
                static_for<0, Q>([&](auto I) {
                    fl[I] = C::decode((uint16_t)(raw[I] >> 16), dsh(I), inv_scale);
                });
                collide<NORM, T>(fl, xa + 1, a, smag, [&](auto I, T v) { out[I] = C::second(pnd[I], v, esh(I), scale); });
            }
            if (xa < nx) {
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    mem.template st32<TL::NT>((Off)i * NBo + selfb, xa * 2, out[i]);
                });
                if constexpr (ACC) {                     // both cells' stored values, decoded
                    T sa[Q], sb[Q];
                    static_for<0, Q>([&](auto I) {
                        sa[I] = C::decode((uint16_t)(out[I] & 0xFFFFu), osh(I), inv_scale);
                        sb[I] = C::decode((uint16_t)(out[I] >> 16), osh(I), inv_scale);
                    });
                    macro_accumulate<NORM, T>(a.acc, N, (size_t)selfrow + xa, sa);
                    macro_accumulate<NORM, T>(a.acc, N, (size_t)selfrow + xa + 1, sb);
                }
            }
        };
        uint32_t carry[Q] = {};
        bool carried = false;
#pragma clang loop unroll(disable)
        for (uint32_t c = pbeg; c < pend; c++) {
            if (pair_special(c)) {
                special_pair_chunk(c);
                carried = false;
                continue;
            }
            uint32_t cur[Q], edge[Q];
            issue(c, cur, edge, carried, carry);
            finish(c, cur, edge, carry);
            carried = true;
        }
      }
    } else if constexpr (TL::PAIR && ES == 4) {
      if (nx & 1u) {
        plain_sweep();                    // pairs need an even row length
      } else {
        // 128-cell chunks; lane L owns cells xa = x0 + 2L and xa + 1 (nx is even).  Each link
        // is one aligned 64-bit load (512 B per wavefront instruction); the x-1 / x+1
        // neighbours that sit in the adjacent lane move over by DPP, the two chunk-edge
        // values come from one broadcast load per shifted link (the left edge of a c_x = +1
        // link, after a fast pair chunk, from lane 63 of the previous chunk in a scalar
        // register instead), and both cells' outputs go out as one 64-bit store per link.
        const uint32_t npc = (nx + 127) / 128;
        const uint32_t plen = (npc + TL::NSEG - 1) / TL::NSEG;
        const uint32_t pbeg = min(npc, (uint32_t)seg * plen);
        const uint32_t pend = min(npc, pbeg + plen);
        uint32_t carry[Q] = {};
        bool carried = false;
#pragma clang loop unroll(disable)
        for (uint32_t c = pbeg; c < pend; c++) {
            const uint32_t x0 = c * 128;
            const bool special = special_chunk(2 * c) || (2 * c + 1 < nch_row && special_chunk(2 * c + 1));
            if (special) {                   // the special half on the general path, the other one-cell fast
#pragma unroll 1
                for (uint32_t sub = 2 * c; sub < min(2 * c + 2, nch_row); sub++) {
                    const uint32_t x = sub * 64 + lane;
                    if (special_chunk(sub)) {
                        general_cell(x);
                    } else if (x < nx && x != 0) {
                        T fl[Q];
                        load_fast(x, fl);
                        collide_store(x, fl);
                    }
                }
                carried = false;
                continue;
            }
            const uint32_t xa = x0 + 2 * lane;
            const uint32_t xl = min(xa, nx - 2);             // lanes past the row end read in-row
            const uint32_t lb = xl * 4;
            const uint32_t ledge = (x0 >= 1 ? x0 - 1 : 0) * 4;            // cell x0-1 (x0 = 0: the inlet, rewritten)
            const uint32_t redge = (x0 + 128 < nx ? x0 + 128 : x0) * 4;   // cell x0+128
            uint32_t left[Q];
            if (carried) {
                static_for<0, Q>([&](auto I) {
                    if constexpr (CX[I] > 0) left[I] = carry[I];
                });
            } else {
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    if constexpr (CX[i] > 0) left[i] = mem.ld32(srow[i] + ES, ledge);
                });
            }
            T fa[Q], fb[Q];
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                const Off row = (CX[i] > 0) ? srow[i] + ES : srow[i];   // unshifted row start
                const uint64_t cur = mem.template ld64<TL::NTL>(row, lb);
                const uint32_t lo = (uint32_t)cur, hi = (uint32_t)(cur >> 32);
                uint32_t va, vb;
                if constexpr (CX[i] == 0) {
                    va = lo; vb = hi;
                } else if constexpr (CX[i] > 0) {                        // cells (x-1, x)
                    va = wave_shr1(hi, left[i]); vb = lo;
                    carry[i] = __builtin_amdgcn_readlane(hi, 63);       // cell x0+127
                } else {                                                 // cells (x+1, x+2)
                    va = hi;
                    vb = wave_shl1(lo, mem.ld32(row, redge));
                    if (xl + 2 >= nx) vb = hi;                           // outlet: cell nx-1 pulls itself
                }
                fa[i] = C::decode(va, dsh(I), inv_scale);
                fb[i] = C::decode(vb, dsh(I), inv_scale);
            });
            uint32_t oa[Q];
            collide<NORM, T>(fa, xa, a, smag, [&](auto I, T v) { oa[I] = C::encode(v, esh(I), scale); });
            if (x0 == 0) {                                               // inlet cell x = 0 (:67-71)
                T u0 = (T)ld_const(a.uin, z);
                if (a.gust) u0 = u0 * (T)ld_const(a.gust, t);
                const T usq15 = T(1.5) * (u0 * u0);
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    const uint32_t e = C::encode(inlet_value<NORM, i, T>(u0, usq15), esh(I), scale);
                    oa[i] = (lane == 0) ? e : oa[i];
                });
                inlet_done = true;
            }
            uint64_t ob[Q];
            collide<NORM, T>(fb, xa + 1, a, smag, [&](auto I, T v) {
                ob[I] = (uint64_t)oa[I] | ((uint64_t)C::encode(v, esh(I), scale) << 32);
            });
            if (xa < nx) {
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    mem.template st64<TL::NT>((Off)i * NBo + selfb, xa * 4, ob[i]);
                });
                if constexpr (ACC) {
                    T sa[Q], sb[Q];
                    static_for<0, Q>([&](auto I) {
                        sa[I] = C::decode((uint32_t)ob[I], osh(I), inv_scale);
                        sb[I] = C::decode((uint32_t)(ob[I] >> 32), osh(I), inv_scale);
                    });
                    macro_accumulate<NORM, T>(a.acc, N, (size_t)selfrow + xa, sa);
                    macro_accumulate<NORM, T>(a.acc, N, (size_t)selfrow + xa + 1, sb);
                }
            }
            carried = true;
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
        T u0 = (T)ld_const(a.uin, z);
        if (a.gust) u0 = u0 * (T)ld_const(a.gust, t);
        const T usq15 = T(1.5) * (u0 * u0);
        T sv[Q];
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            const typename C::B e = C::encode(inlet_value<NORM, i, T>(u0, usq15), esh(I), scale);
            mem.template st<false>((Off)i * NBo + selfb, 0, e);
            if constexpr (ACC) sv[i] = C::decode(e, osh(I), inv_scale);
        });
        if constexpr (ACC) macro_accumulate<NORM, T>(a.acc, N, (size_t)selfrow, sv);
        (void)sv;
    }
    // a solid inlet-column cell: its sample is (1, 0, 0, 0) (the general path leaves x = 0 out)
    if constexpr (ACC) {
        if (seg == 0 && lane == 0 && (lrow[0] & SOLID_BIT)) solid_accumulate(a.acc, N, (size_t)selfrow);
    }
}

// Writes this workgroup's FP64 force pair (sum over its WPB wavefronts in wavefront
// order) to partials[slot][z][yb].
template <int WPB, class A>
__device__ __forceinline__ void block_force(A& a, double fx, double fz, int z, int yb, int t) {
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
