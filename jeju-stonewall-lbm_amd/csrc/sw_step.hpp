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

constexpr int ROWS = 4;            // rows (wavefronts) per workgroup
constexpr int WG = 64 * ROWS;

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
    const T third = rho / T(3);
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
    constexpr T third_c = T(1) / T(3);
    constexpr T A1 = T(1) - third_c;      // (c*c - 1/3) for |c| = 1
    constexpr T A0 = T(0) - third_c;      // (c*c - 1/3) for c = 0
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        constexpr int cx = CX[i], cy = CY[i], cz = CZ[i];
        constexpr T w = wgt<T>(i);
        constexpr T ax = cx ? A1 : A0, ay = cy ? A1 : A0, az = cz ? A1 : A0;
        T qq = (ax * nxx + ay * nyy) + az * nzz;
        if constexpr (cx * cy != 0) qq = qq + T(2) * (cx * cy > 0 ? nxy : -nxy);
        if constexpr (cx * cz != 0) qq = qq + T(2) * (cx * cz > 0 ? nxz : -nxz);
        if constexpr (cy * cz != 0) qq = qq + T(2) * (cy * cz > 0 ? nyz : -nyz);
        st(I, feq<i, T>(rho, ux, uy, uz, usq15) + ((keep * w) * T(4.5)) * qq);
    });
}

// One (y, z) row of one lattice; adds this lane's FP64 wall-force sums to fx_acc/fz_acc.
// BUF selects raw-buffer addressing (all byte offsets of a population array < 4 GiB).
template <class S, int YBC, int CPL, bool BUF>
__device__ __forceinline__ void step_row(const StepArgs& a, int y, int z, int t, int lane,
                                         double& fx_acc, double& fz_acc) {
    using C = Codec<S>;
    using T = typename C::T;
    using B = typename C::B;
    constexpr uint32_t ES = sizeof(S);
    const PopMem<S, BUF> mem(a.f_in, a.f_out);
    const uint32_t nx = a.nx, ny = a.ny, nz = a.nz;
    const size_t NB = (size_t)nx * ny * nz * ES;                 // bytes per population
    const uint32_t selfrow = ((uint32_t)z * ny + y) * nx;
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
        if constexpr (YBC == Y_PERIODIC) {
            ys = ys < 0 ? ys + (int)ny : (ys >= (int)ny ? ys - (int)ny : ys);      // lbm3d.py:79
        } else {
            if (ys < 0 || ys >= (int)ny) { ys = y; j = MIRY[j]; }                 // free-slip sides
        }
        if (zs >= (int)nz) { zs = z; j = MIRZ[j]; }                               // lbm3d.py:84-86
        const size_t r = (size_t)j * NB + (size_t)(((uint32_t)zs * ny + (uint32_t)ys) * nx) * ES;
        srow[i] = (CX[i] > 0) ? r - ES : r;
    });
    const uint32_t* lrow = a.links + selfrow;

    for (uint32_t x0 = 0; x0 < nx; x0 += 64 * CPL) {
        uint32_t m[CPL];
        bool special = false;
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            const uint32_t x = x0 + 64 * c + lane;
            m[c] = (x < nx) ? lrow[x] : 0u;
            special |= (x < nx) && (x != 0) && (m[c] != 0u);
        }
        if (!__any(special)) {
            // ---------------- fast path: no solid cell and no solid neighbour in the chunk(s)
            T fl[CPL][Q];
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                const uint32_t x = x0 + 64 * c + lane;
                if (x < nx && x != 0) {
                    const uint32_t xo = x * ES;
                    const uint32_t xpo = min(x + 1, nx - 1) * ES;
                    static_for<0, Q>([&](auto I) {
                        constexpr int i = I;
                        fl[c][i] = C::decode(mem.ld(srow[i], (CX[i] < 0) ? xpo : xo), wgt<T>(i), inv_scale);
                    });
                }
            }
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                const uint32_t x = x0 + 64 * c + lane;
                if (x < nx && x != 0) {
                    const uint32_t xo = x * ES;
                    collide<T>(fl[c], x, a, smag, [&](auto I, T v) {
                        constexpr int i = I;
                        mem.st((size_t)i * NB + selfb, xo, C::encode(v, wgt<T>(i), scale));
                    });
                }
            }
        } else {
            // ---------------- general path: per-link bounce-back and wall force (:77-101)
#pragma unroll
            for (int c = 0; c < CPL; c++) {
                const uint32_t x = x0 + 64 * c + lane;
                if (!(x < nx && x != 0)) continue;
                const uint32_t mm = m[c];
                if (mm & SOLID_BIT) continue;                                        // :65-66
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
                        if ((mm >> i) & 1u) {                                        // half-way bounce-back
                            val = C::decode(mem.ld((size_t)OPP[i] * NB + selfb, xo), w, inv_scale);
                            bool wall = true;
                            if constexpr (GROUND_SLOT[i] >= 0)
                                wall = !((mm >> (GROUND_BIT0 + GROUND_SLOT[i])) & 1u);
                            if (wall) {                                              // momentum exchange
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
                collide<T>(fl, x, a, smag, [&](auto I, T v) {
                    constexpr int i = I;
                    mem.st((size_t)i * NB + selfb, xo, C::encode(v, wgt<T>(i), scale));
                });
            }
        }
    }
    // ---------------- inlet (:67-71): x = 0 fluid cell gets feq(1, (uin[z] g, 0, 0))
    if (lane == 0 && !(lrow[0] & SOLID_BIT)) {
        T u0 = (T)a.uin[z];
        if (a.gust) u0 = u0 * (T)a.gust[t];
        const T usq15 = T(1.5) * (u0 * u0);
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            mem.st((size_t)i * NB + selfb, 0, C::encode(feq<i, T>(T(1), u0, T(0), T(0), usq15), wgt<T>(i), scale));
        });
    }
}

// Writes this workgroup's FP64 force pair (sum over its rows) to partials[slot][z][yb].
__device__ __forceinline__ void block_force(const StepArgs& a, double fx, double fz, int z, int yb, int t) {
    __shared__ double red[ROWS][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    fx = wave_sum(fx);
    fz = wave_sum(fz);
    if (lane == 0) { red[wave][0] = fx; red[wave][1] = fz; }
    __syncthreads();
    if (threadIdx.x == 0 && a.partials) {
        double sx = red[0][0], sz = red[0][1];
        for (int r = 1; r < ROWS; r++) { sx += red[r][0]; sz += red[r][1]; }
        const size_t slot = (size_t)(t % a.ring);
        double* p = a.partials + ((slot * a.nz + z) * a.nyb + yb) * 2;
        p[0] = sx;
        p[1] = sz;
    }
}

}  // namespace sw
