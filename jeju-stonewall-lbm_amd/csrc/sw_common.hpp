// Shared device-side definitions for the D3Q19 stone-wall LBM kernels (gfx950).
//
// Lattice: D3Q19 in the reference's direction order (stonewall3d/lbm3d.py:20-27):
//   C = (0,0,0), (+-1,0,0), (0,+-1,0), (0,0,+-1), (1,1,0),(-1,-1,0),(1,-1,0),(-1,1,0),
//       (1,0,1),(-1,0,-1),(1,0,-1),(-1,0,1), (0,1,1),(0,-1,-1),(0,1,-1),(0,-1,1)
//   W = 1/3, 1/18 x6, 1/36 x12 (rounded to float32 as numpy does, lbm3d.py:25)
//
// Memory layout in HBM (structure of arrays, x contiguous):
//   population i of cell (x,y,z) lives at  f[i*N + (z*ny + y)*nx + x],  N = nx*ny*nz.
//   A wavefront owns one (y,z) row and sweeps it in 64-cell chunks, so each of the
//   19 loads and 19 stores of a chunk is one contiguous 64-element access.
//
// Link masks (one uint32 per cell, built once from the solid flags):
//   bit i (i = 1..18) : the source cell of link i is solid -> half-way bounce-back
//   bits 19..23       : for the five links with cz = +1 (i = 5,11,14,15,18), the solid
//                       source is ground (flag >= 2): bounce back but no wall force
//   bit 31            : the cell itself is solid (not updated)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <utility>

namespace sw {

constexpr int Q = 19;
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
constexpr int MIRZ[Q] = {0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16};
constexpr int MIRY[Q] = {0, 1, 2, 4, 3, 5, 6, 9, 10, 7, 8, 11, 12, 13, 14, 18, 17, 16, 15};

constexpr uint32_t SOLID_BIT = 1u << 31;
constexpr int GROUND_BIT0 = 19;
// index (0..4) of a cz=+1 link among {5,11,14,15,18}, -1 otherwise
constexpr int GROUND_SLOT[Q] = {-1, -1, -1, -1, -1, 0, -1, -1, -1, -1, -1, 1, -1, -1, 2, 3, -1, -1, 4};

enum YBC { Y_PERIODIC = 0, Y_FREESLIP = 1 };

template <class T> __host__ __device__ constexpr T wgt(int i) {
    return i == 0 ? T(1.0 / 3.0) : (i < 7 ? T(1.0 / 18.0) : T(1.0 / 36.0));
}
// float32 weights are the float rounding of the double fractions (numpy float32 cast)
template <> __host__ __device__ constexpr float wgt<float>(int i) {
    return i == 0 ? float(1.0 / 3.0) : (i < 7 ? float(1.0 / 18.0) : float(1.0 / 36.0));
}

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(static_cast<F&&>(f));
    }
}

// ------------------------------------------------------------- storage codecs
enum Dtype { DT_F32 = 0, DT_F16 = 1, DT_BF16 = 2, DT_F64 = 3 };
// Offsets sigma of the 16-bit storage forms (per-height table built by sw_shift_table): the
// normalised equilibrium of the inlet velocity at each height, or 1 (U = 0).
enum ShiftKind { SHIFT_WEIGHTS = 0, SHIFT_INLET = 1 };

struct bf16_t { uint16_t bits; };

// Storage policy: S is the element type in HBM, T the arithmetic type, B the raw bits.
// The 16-bit forms store (v - s) * scale, v the weight-normalised population f_i / w_i and
// s its offset sigma (a per-height value of the direction's c_x class); scale is a power
// of two so the scaling is exact.  decode(b, s, 1/scale) and load/store take the offset s;
// encode / first / second take ns = -s * scale (neg_scaled), which the step keeps in scalar
// registers.
template <class S> struct Codec;

template <> struct Codec<float> {
    using T = float;
    using B = uint32_t;
    static __device__ __forceinline__ float decode(B b, float, float) { return __uint_as_float(b); }
    static __device__ __forceinline__ B encode(float v, float, float) { return __float_as_uint(v); }
    static __device__ __forceinline__ float load(const float* p, size_t i, float, float) { return p[i]; }
    static __device__ __forceinline__ void store(float* p, size_t i, float v, float, float) { p[i] = v; }
};
template <> struct Codec<double> {
    using T = double;
    using B = uint64_t;
    static __device__ __forceinline__ double decode(B b, double, double) { return __builtin_bit_cast(double, b); }
    static __device__ __forceinline__ B encode(double v, double, double) { return __builtin_bit_cast(B, v); }
    static __device__ __forceinline__ double load(const double* p, size_t i, double, double) { return p[i]; }
    static __device__ __forceinline__ void store(double* p, size_t i, double v, double, double) { p[i] = v; }
};
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

template <> struct Codec<_Float16> {
    using T = float;
    using B = uint16_t;
    // Two cells' values packed into one 32-bit word (low half = first cell), encoded exactly
    // as encode(): the first goes to the low half (v_fma_mixlo_f16), the second to the
    // high half (v_fma_mixhi_f16).
    using Pending = half2_t;
    static __device__ __forceinline__ Pending first(float v, float ns, float scale) {
        // the high half is left unset: second() writes it before the word is used (a zero
        // fill here cost one v_pack per direction)
        Pending p;
        p.x = (_Float16)__builtin_fmaf(v, scale, ns);
        return p;
    }
    static __device__ __forceinline__ uint32_t second(Pending p, float v, float ns, float scale) {
        p.y = (_Float16)__builtin_fmaf(v, scale, ns);
        return __builtin_bit_cast(uint32_t, p);
    }
    // w + h*2^-k with one rounding: h*2^-k is exact, so fma == mul-then-add bit for bit
    static __device__ __forceinline__ float decode(B b, float w, float inv_scale) {
        return __builtin_fmaf((float)__builtin_bit_cast(_Float16, b), inv_scale, w);
    }
    // RN(v - w) * 2^k == RN(v*2^k - w*2^k): scaling by a power of two commutes with rounding.
    // ns = -w*2^k (neg_scaled).
    static __device__ __forceinline__ B encode(float v, float ns, float scale) {
        return __builtin_bit_cast(B, (_Float16)__builtin_fmaf(v, scale, ns));
    }
    static __device__ __forceinline__ float load(const _Float16* p, size_t i, float w, float inv_scale) {
        return w + (float)p[i] * inv_scale;
    }
    static __device__ __forceinline__ void store(_Float16* p, size_t i, float v, float w, float scale) {
        p[i] = (_Float16)((v - w) * scale);
    }
};
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// BF16 rounding of the storage forms: dithered (deterministic stochastic) rounding.  The
// float's 32 bits b are truncated to their top 16 after adding a dither r in [0, 2^16) drawn
// from b itself by Knuth's multiplicative hash, r = (b * 0x9E3779B1 mod 2^32) >> 16: the
// magnitude rounds up iff low16(b) + r >= 2^16, so a value a fraction p of the way between
// two BF16 neighbours rounds up for a fraction ~p of the values around it.  Values already
// in BF16 (low16(b) = 0) and +-0 are unchanged.  With round to nearest the BF16 form drifted
// to a +1.6 % mean drag (four seeds, H = 40, 40 % void); with this rounding it stays within
// 1 % of the FP64 run.
__host__ __device__ __forceinline__ uint16_t bf16_dither(float f) {
    const uint32_t b = __builtin_bit_cast(uint32_t, f);
    return (uint16_t)((b + ((b * 0x9E3779B1u) >> 16)) >> 16);
}
template <> struct Codec<bf16_t> {
    using T = float;
    using B = uint16_t;
    // Two cells packed into one 32-bit word (low half = first cell), each rounded as encode().
    using Pending = float;
    static __device__ __forceinline__ Pending first(float v, float ns, float scale) {
        return __builtin_fmaf(v, scale, ns);
    }
    static __device__ __forceinline__ uint32_t second(Pending p, float v, float ns, float scale) {
        return (uint32_t)bf16_dither(p) | ((uint32_t)bf16_dither(__builtin_fmaf(v, scale, ns)) << 16);
    }
    static __device__ __forceinline__ float decode(B b, float w, float inv_scale) {
        return __builtin_fmaf(bf16_to_f32(b), inv_scale, w);
    }
    static __device__ __forceinline__ B encode(float v, float ns, float scale) {
        return bf16_dither(__builtin_fmaf(v, scale, ns));
    }
    static __device__ __forceinline__ float load(const bf16_t* p, size_t i, float w, float inv_scale) {
        return w + bf16_to_f32(p[i].bits) * inv_scale;
    }
    static __device__ __forceinline__ void store(bf16_t* p, size_t i, float v, float w, float scale) {
        p[i].bits = bf16_dither((v - w) * scale);
    }
};

// -s * 2^k by integer operations on the float's bits: exact for a normal s whose scaled
// value stays normal (the storage offsets are near 1 and 2^-20 <= 2^k <= 2^20), and done by
// the scalar ALU when s is wave-uniform, so the result stays in a scalar register.
// s = 0 -> -0.
__host__ __device__ __forceinline__ float neg_scaled(float s, int k) {
    const uint32_t b = __builtin_bit_cast(uint32_t, s);
    const uint32_t r = (b & 0x7fffffffu) ? (b + ((uint32_t)k << 23)) : b;
    return __builtin_bit_cast(float, r ^ 0x80000000u);
}

// ------------------------------------------------------------- memory access
// A load of data that no kernel writes while a step runs (chunk flags, inlet profile, gust and
// storage-offset tables), through the constant address space: at a wave-uniform address the
// compiler then issues a scalar load, and a test of its value waits only for that load
// (lgkmcnt), not for every vector load and store still in flight (vmcnt(0), as a global load
// of the same word needed).
template <class T>
__device__ __forceinline__ T ld_const(const T* p, size_t i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Raw-buffer access: one descriptor per array, a wave-uniform 32-bit byte offset (row
// start, scalar register) plus a per-lane 32-bit byte offset.  Usable while every byte
// offset of the array fits in 32 bits.
constexpr int BUF_FLAGS = 0x00020000;   // raw buffer, 32-bit data format (gfx9 family)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, -1, BUF_FLAGS);
}

template <class B, int AUX = 0>
__device__ __forceinline__ B buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(B) == 2) return (B)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, AUX);
    else if constexpr (sizeof(B) == 4) return (B)__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX);
    else return __builtin_bit_cast(B, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
}
template <class B, int AUX = 0>
__device__ __forceinline__ void buf_store(B v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(B) == 2) __builtin_amdgcn_raw_buffer_store_b16(v, r, voff, soff, AUX);
    else if constexpr (sizeof(B) == 4) __builtin_amdgcn_raw_buffer_store_b32(v, r, voff, soff, AUX);
    else {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, AUX);
    }
}

// Accessor over one population array pair.  BUF = true: raw buffers (offsets < 4 GiB);
// BUF = false: 64-bit pointers for arrays larger than that.
template <class S, bool BUF> struct PopMem;

template <class S> struct PopMem<S, true> {
    using B = typename Codec<S>::B;
    __amdgpu_buffer_rsrc_t rin, rout;
    __device__ __forceinline__ PopMem(const void* fin, void* fout) : rin(make_rsrc(fin)), rout(make_rsrc(fout)) {}
    // 32-bit access (two adjacent 16-bit populations, or one 32-bit edge value)
    template <bool NT = false>
    __device__ __forceinline__ uint32_t ld32(size_t row_bytes, uint32_t lane_bytes) const {
        return buf_load<uint32_t, NT ? 2 : 0>(rin, lane_bytes, (uint32_t)row_bytes);
    }
    template <bool NT = false>
    __device__ __forceinline__ void st32(size_t row_bytes, uint32_t lane_bytes, uint32_t v) const {
        buf_store<uint32_t, NT ? 2 : 0>(v, rout, lane_bytes, (uint32_t)row_bytes);
    }
    // 64-bit access (two adjacent 32-bit populations)
    template <bool NT = false>
    __device__ __forceinline__ uint64_t ld64(size_t row_bytes, uint32_t lane_bytes) const {
        return buf_load<uint64_t, NT ? 2 : 0>(rin, lane_bytes, (uint32_t)row_bytes);
    }
    // per-lane byte offset of any population (all below 4 GiB here)
    __device__ __forceinline__ B ld_at(size_t bytes) const { return buf_load<B>(rin, (uint32_t)bytes, 0u); }
    template <bool NT = false>
    __device__ __forceinline__ void st64(size_t row_bytes, uint32_t lane_bytes, uint64_t v) const {
        buf_store<uint64_t, NT ? 2 : 0>(v, rout, lane_bytes, (uint32_t)row_bytes);
    }
    template <bool NT = false>
    __device__ __forceinline__ B ld(size_t row_bytes, uint32_t lane_bytes) const {
        return buf_load<B, NT ? 2 : 0>(rin, lane_bytes, (uint32_t)row_bytes);
    }
    template <bool NT = false>
    __device__ __forceinline__ void st(size_t row_bytes, uint32_t lane_bytes, B v) const {
        buf_store<B, NT ? 2 : 0>(v, rout, lane_bytes, (uint32_t)row_bytes);
    }
};
template <class S> struct PopMem<S, false> {
    using B = typename Codec<S>::B;
    const char* in;
    char* out;
    __device__ __forceinline__ PopMem(const void* fin, void* fout)
        : in(static_cast<const char*>(fin)), out(static_cast<char*>(fout)) {}
    template <bool NT = false>
    __device__ __forceinline__ uint32_t ld32(size_t row_bytes, uint32_t lane_bytes) const {
        return *reinterpret_cast<const uint32_t*>(in + row_bytes + lane_bytes);
    }
    template <bool NT = false>
    __device__ __forceinline__ void st32(size_t row_bytes, uint32_t lane_bytes, uint32_t v) const {
        uint32_t* p = reinterpret_cast<uint32_t*>(out + row_bytes + lane_bytes);
        if constexpr (NT) __builtin_nontemporal_store(v, p); else *p = v;
    }
    template <bool NT = false>
    __device__ __forceinline__ uint64_t ld64(size_t row_bytes, uint32_t lane_bytes) const {
        return *reinterpret_cast<const uint64_t*>(in + row_bytes + lane_bytes);
    }
    __device__ __forceinline__ B ld_at(size_t bytes) const { return *reinterpret_cast<const B*>(in + bytes); }
    template <bool NT = false>
    __device__ __forceinline__ void st64(size_t row_bytes, uint32_t lane_bytes, uint64_t v) const {
        uint64_t* p = reinterpret_cast<uint64_t*>(out + row_bytes + lane_bytes);
        if constexpr (NT) __builtin_nontemporal_store(v, p); else *p = v;
    }
    template <bool NT = false>
    __device__ __forceinline__ B ld(size_t row_bytes, uint32_t lane_bytes) const {
        const B* p = reinterpret_cast<const B*>(in + row_bytes + lane_bytes);
        if constexpr (NT) return __builtin_nontemporal_load(p); else return *p;
    }
    template <bool NT = false>
    __device__ __forceinline__ void st(size_t row_bytes, uint32_t lane_bytes, B v) const {
        B* p = reinterpret_cast<B*>(out + row_bytes + lane_bytes);
        if constexpr (NT) __builtin_nontemporal_store(v, p); else *p = v;
    }
};

// feq_i of lbm3d.py:33-36 with the operand order kept:
//   (w*rho) * (((1 + cu) + (0.5*cu)*cu) - usq15),  cu = 3*((cx*ux + cy*uy) + cz*uz)
// usq15 = 1.5*((ux*ux + uy*uy) + uz*uz) is shared by all i.  Terms with a zero
// lattice coefficient are dropped: for finite operands x + (+-0) == x in every
// case that reaches a store (the moment sums start at +0 and can never become -0),
// so the result is bit-identical to the full expression.
template <int I, class T>
__device__ __forceinline__ T cdotu(T ux, T uy, T uz) {
    constexpr int cx = CX[I], cy = CY[I], cz = CZ[I];
    T s = T(0);
    bool have = false;
    if constexpr (cx != 0) { s = cx > 0 ? ux : -ux; have = true; }
    if constexpr (cy != 0) {
        T t = cy > 0 ? uy : -uy;
        if (have) s = s + t; else s = t;
        have = true;
    }
    if constexpr (cz != 0) {
        T t = cz > 0 ? uz : -uz;
        if (have) s = s + t; else s = t;
    }
    return s;
}

// c_x class of a direction for the 16-bit storage offsets: 0 (c_x = 0), 1 (+1), 2 (-1)
__host__ __device__ constexpr int xclass(int cx) { return cx == 0 ? 0 : (cx > 0 ? 1 : 2); }

template <int I, class T>
__device__ __forceinline__ T feq(T rho, T ux, T uy, T uz, T usq15);

// feq_i(1, u) / w_i = ((1 + (0.5 cu) cu) - usq15) + cu, cu = 3 c.u: the equilibrium of the
// 16-bit forms, which hold phi_i = f_i / w_i (the terms of lbm3d.py:35-36 without the
// weight, grouped so that a pair of opposite directions shares everything but the sign of cu)
template <int I, class T>
__device__ __forceinline__ T feq_norm(T ux, T uy, T uz, T usq15) {
    if constexpr (I == 0) {
        return T(1) - usq15;
    } else {
        const T cu = T(3) * cdotu<I>(ux, uy, uz);
        return ((T(1) + (T(0.5) * cu) * cu) - usq15) + cu;
    }
}

// value written at an inlet cell (lbm3d.py:67-71): feq_i(1, (u0, 0, 0)), or its normalised
// form for the 16-bit storage (NORM)
template <bool NORM, int I, class T>
__device__ __forceinline__ T inlet_value(T u0, T usq15) {
    if constexpr (NORM) return feq_norm<I, T>(u0, T(0), T(0), usq15);
    else return feq<I, T>(T(1), u0, T(0), T(0), usq15);
}

template <int I, class T>
__device__ __forceinline__ T feq(T rho, T ux, T uy, T uz, T usq15) {
    constexpr T w = wgt<T>(I);
    if constexpr (I == 0) {
        return (w * rho) * (T(1) - usq15);
    } else {
        const T cu = T(3) * cdotu<I>(ux, uy, uz);
        return (w * rho) * (((T(1) + cu) + (T(0.5) * cu) * cu) - usq15);
    }
}

// x / 3 rounded to nearest: x*RN(1/3) corrected by one FMA residual step.  Checked
// exhaustively against IEEE division for all 2^32 float inputs: identical for every finite
// x except -0 (which a density never is).  The double form keeps the plain division.
__device__ __forceinline__ float div3(float x) {
    constexpr float C3 = 1.0f / 3.0f;
    const float q0 = x * C3;
    const float r = __builtin_fmaf(-q0, 3.0f, x);
    return __builtin_fmaf(r, C3, q0);
}
__device__ __forceinline__ double div3(double x) { return x / 3.0; }

// Correctly rounded reciprocal and quotient without the IEEE division sequence
// (v_div_scale / v_div_fmas / v_div_fixup and mode switches).
//   rcp_rn(b): one Newton step on the hardware reciprocal, y = y0 + y0 (1 - b y0) with two
//     FMAs.  Equal to IEEE 1/b for every normal b whose reciprocal is normal (all 2^32 bit
//     patterns checked on gfx950; sw_selftest_division repeats the check).
//   div_rn(a, b, y), y = rcp_rn(b): q0 = a y, r = a - b q0 (exact by FMA), q = q0 + r y.
//     With y = RN(1/b) this is RN(a/b) (Markstein's theorem) whenever nothing underflows;
//     +0 / b gives +0.  The densities, momenta and stress norms divided here are normal or
//     +0 in any finite flow.  The double forms keep the plain division.
__device__ __forceinline__ float rcp_rn(float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    return __builtin_fmaf(__builtin_fmaf(-b, y0, 1.0f), y0, y0);
}
__device__ __forceinline__ float div_rn(float a, float b, float y) {
    const float q0 = a * y;
    return __builtin_fmaf(__builtin_fmaf(-b, q0, a), y, q0);
}
__device__ __forceinline__ double rcp_rn(double b) { return 1.0 / b; }
__device__ __forceinline__ double div_rn(double a, double b, double) { return a / b; }

// PCG hash (Jarzynski & Olano, JCGT 2020) for the optional init noise
__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float u01(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

// ------------------------------------------------------------- lane shifts (whole wavefront)
// dpp wave_shr:1 — lane L receives lane L-1's value, lane 0 receives `edge`
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v, uint32_t edge) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)v, 0x138, 0xF, 0xF, false);
}
// dpp wave_shl:1 — lane L receives lane L+1's value, lane 63 receives `edge`
__device__ __forceinline__ uint32_t wave_shl1(uint32_t v, uint32_t edge) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)v, 0x130, 0xF, 0xF, false);
}
// two 16-bit halves: low = hi half of `lo`, high = lo half of `hi`
__device__ __forceinline__ uint32_t mid16(uint32_t hi, uint32_t lo) {
    return __builtin_amdgcn_alignbit(hi, lo, 16);
}

// A value every lane of the wavefront holds identically, moved to scalar registers word by
// word.  Values loaded from device memory (e.g. an instance table) are not known to be
// uniform to the compiler; without this, buffer descriptors built from them sit in vector
// registers and every buffer access becomes a readfirstlane loop.
template <class T>
__device__ __forceinline__ T wave_uniform(const T& v) {
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (unsigned j = 0; j < sizeof(T) / 4; j++) w[j] = __builtin_amdgcn_readfirstlane(w[j]);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

// A copy of a wave-uniform float held in a vector register.  A gfx9 VALU instruction reads
// at most one scalar register, and the 16-bit decode's multiply-add takes two wave-uniform
// operands (the inverse scale and a per-row offset): with the inverse scale in a vector
// register the offset is read straight from its scalar register instead of being copied to
// a vector register before each decode.
__device__ __forceinline__ float vgpr_copy(float s) {
    float v;
    asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
    return v;
}

// ------------------------------------------------------------- wave reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace sw
