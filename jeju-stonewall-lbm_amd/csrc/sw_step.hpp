// The fused D3Q19 step: pull-streaming, half-way bounce-back, outlet / top / side
// boundaries, inlet, momentum-exchange wall force and regularized-BGK collision with a
// Smagorinsky eddy viscosity.  Restates stonewall3d/lbm3d.py:57-155 (k_step).
//
// Work decomposition: a row task is one (y, z) row of one lattice.  A wavefront owns a
// row and sweeps it in 64-cell chunks; a workgroup holds ROWS wavefronts (ROWS rows of
// the same z).  Every row's FP32 per-cell wall force (fxw, fzw) is accumulated in FP64
// per lane, reduced over the wave with a fixed butterfly, then over the workgroup's
// rows in row order, and written as one FP64 pair per workgroup and step into a
// partials ring.  A second kernel sums the partials of each z level in a fixed order,
// so the force rows are bit-reproducible (the reference uses FP32 atomics, whose
// order is not).
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

// One row (y, z) of one lattice.  Returns this lane's FP64 wall-force sums.
template <class S, int YBC>
__device__ __forceinline__ void step_row(const StepArgs& a, int y, int z, int t, int lane,
                                         double& fx_acc, double& fz_acc) {
    using C = Codec<S>;
    using T = typename C::T;
    const S* __restrict__ fin = static_cast<const S*>(a.f_in);
    S* __restrict__ fout = static_cast<S*>(a.f_out);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const uint32_t N = (uint32_t)nx * (uint32_t)ny * (uint32_t)nz;
    const uint32_t selfrow = ((uint32_t)z * ny + y) * nx;
    const T inv_scale = (T)a.inv_scale, scale = (T)a.scale;

    // wave-uniform source row of every link, after the y / z boundary rules
    const S* srcrow[Q];
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        int zs = z - CZ[i], ys = y - CY[i], j = i;
        if constexpr (YBC == Y_PERIODIC) {
            ys = ys < 0 ? ys + ny : (ys >= ny ? ys - ny : ys);                // lbm3d.py:79
        } else {
            if (ys < 0 || ys >= ny) { ys = y; j = MIRY[j]; }                   // free-slip sides
        }
        if (zs >= nz) { zs = z; j = MIRZ[j]; }                                 // lbm3d.py:84-86
        srcrow[i] = fin + (size_t)j * N + ((uint32_t)zs * ny + ys) * nx;
    });

    const T nu0 = (T)a.nu0, cs = (T)a.cs, nu_sponge = (T)a.nu_sponge;
    // (18*sqrt(2))*cs*cs evaluated left to right as lbm3d.py:146 does
    const T smag = ((T(18) * (T)sqrt(T(2))) * cs) * cs;

    for (int x = lane; x < nx; x += 64) {
        const uint32_t self = selfrow + x;
        const uint32_t m = a.links[self];
        if (m & SOLID_BIT) continue;                                           // :65-66
        if (x == 0) {                                                          // :67-71
            T u0 = (T)a.uin[z];
            if (a.gust) u0 = u0 * (T)a.gust[t];
            const T usq15 = T(1.5) * (u0 * u0);
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                C::store(fout, (size_t)i * N + self, feq<i, T>(T(1), u0, T(0), T(0), usq15), wgt<T>(i), scale);
            });
            continue;
        }
        // ---- pull + bounce-back + momentum exchange (:77-101)
        T fl[Q];
        T fxw = T(0), fzw = T(0);
        bool hit = false;
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr T w = wgt<T>(i);
            if constexpr (i == 0) {
                fl[0] = C::load(fin, self, w, inv_scale);
            } else {
                const bool bb = (m >> i) & 1u;
                T val;
                if (bb) {
                    val = C::load(fin, (size_t)OPP[i] * N + self, w, inv_scale);
                    bool wall = true;
                    if constexpr (GROUND_SLOT[i] >= 0) wall = !((m >> (GROUND_BIT0 + GROUND_SLOT[i])) & 1u);
                    if (wall) {
                        const T fr2 = T(2) * (val - w);
                        if constexpr (CX[i] > 0) fxw = fxw - fr2;
                        if constexpr (CX[i] < 0) fxw = fxw + fr2;
                        if constexpr (CZ[i] > 0) fzw = fzw - fr2;
                        if constexpr (CZ[i] < 0) fzw = fzw + fr2;
                        hit = true;
                    }
                } else {
                    int xs;
                    if constexpr (CX[i] < 0) xs = min(x + 1, nx - 1);                // outlet :82-83
                    else xs = x - CX[i];
                    val = C::load(srcrow[i], xs, w, inv_scale);
                }
                fl[i] = val;
            }
        });
        if (hit) { fx_acc += (double)fxw; fz_acc += (double)fzw; }

        // ---- moments (:103-127), sequential in i; zero-coefficient terms dropped
        T rho = T(0), mx = T(0), my = T(0), mz = T(0);
        T pxx = T(0), pyy = T(0), pzz = T(0), pxy = T(0), pxz = T(0), pyz = T(0);
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int cx = CX[i], cy = CY[i], cz = CZ[i];
            const T v = fl[i];
            rho = rho + v;
            if constexpr (cx != 0) { mx = cx > 0 ? mx + v : mx - v; pxx = pxx + v; }
            if constexpr (cy != 0) { my = cy > 0 ? my + v : my - v; pyy = pyy + v; }
            if constexpr (cz != 0) { mz = cz > 0 ? mz + v : mz - v; pzz = pzz + v; }
            if constexpr (cx * cy != 0) pxy = cx * cy > 0 ? pxy + v : pxy - v;
            if constexpr (cx * cz != 0) pxz = cx * cz > 0 ? pxz + v : pxz - v;
            if constexpr (cy * cz != 0) pyz = cy * cz > 0 ? pyz + v : pyz - v;
        });
        const T ux = mx / rho, uy = my / rho, uz = mz / rho;                    // :128-138
        const T third = rho / T(3);
        const T nxx = (pxx - third) - (rho * ux) * ux;
        const T nyy = (pyy - third) - (rho * uy) * uy;
        const T nzz = (pzz - third) - (rho * uz) * uz;
        const T nxy = pxy - (rho * ux) * uy;
        const T nxz = pxz - (rho * ux) * uz;
        const T nyz = pyz - (rho * uy) * uz;

        T nu = nu0;                                                            // :140-147
        if (x > a.x_sponge) {
            const T s = (T)(x - a.x_sponge) / (T)(nx - a.x_sponge);
            nu = nu0 + (nu_sponge * s) * s;
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
            const T fo = feq<i, T>(rho, ux, uy, uz, usq15) + ((keep * w) * T(4.5)) * qq;
            C::store(fout, (size_t)i * N + self, fo, w, scale);
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
