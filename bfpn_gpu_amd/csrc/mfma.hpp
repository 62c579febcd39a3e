// mfma.hpp — split-precision matrix-core helpers for CDNA4 (gfx950).
//
// Split BF16 (the source GEMMs): a product x*y of FP32 values is evaluated
// on the BF16 MFMA pipe as
//     x*y ~= xh*yh + xh*yl + xl*yh,   xh = bf16(x), xl = bf16(x - xh)
// with FP32 accumulation (v_mfma_f32_16x16x32_bf16, three issues per tile).
// xh + xl carries 16 significand bits, so each product is accurate to about
// 2^-16 relative; the dropped xl*yl term is ~2^-18.  Split FP16 (the DST):
// below.
//
// Operand layouts of v_mfma_f32_16x16x32_bf16 (one wave64, q = lane/16):
//   A (16 x 32, row-major):  lane holds A[lane%16][8q .. 8q+7]
//   B (32 x 16):             lane holds B[8q .. 8q+7][lane%16]
//   D (16 x 16):             lane holds D[4q + r][lane%16], r = 0..3
#pragma once

#include <hip/hip_runtime.h>

namespace bfpn {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

struct SplitFrag {
    bf16x8 hi;
    bf16x8 lo;
};

__device__ inline void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = static_cast<__bf16>(x);
    lo = static_cast<__bf16>(x - static_cast<float>(hi));
}

// Splits 8 FP32 values into a hi/lo BF16 fragment (the same bits as
// split_bf16 per value).  hi is converted once, in pairs; its FP32 values are
// read back from the packed words (low half << 16, high half & 0xffff0000).
__device__ inline SplitFrag split8(const float (&v)[8]) {
    typedef float f32x8 __attribute__((ext_vector_type(8)));
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    f32x8 x;
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = v[e];
    SplitFrag f;
    f.hi = __builtin_convertvector(x, bf16x8);
    const u32x4 w = __builtin_bit_cast(u32x4, f.hi);
    f32x8 h;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        h[2 * k] = __uint_as_float(w[k] << 16);
        h[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
    }
    f.lo = __builtin_convertvector(x - h, bf16x8);
    return f;
}

// acc += A*B with both operands split (three MFMA issues).
__device__ inline f32x4 mma_split(const SplitFrag& a, const SplitFrag& b, f32x4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, acc, 0, 0, 0);
    return acc;
}

__device__ inline f32x4 zero4() { return f32x4{0.f, 0.f, 0.f, 0.f}; }

// ---- split FP16 (the DST) ----
// Operators applied thousands of times in a row (the DST of every depth step)
// need products close to FP32: there a split's representation error
// compounds step after step.  A two-term BF16 split (16 significand bits)
// drifts by ~2e-3 over 1000 applications; two FP16 terms carry 22 bits, and
// a per-plane power-of-two scale keeps the state inside the FP16 range.
// Operand layouts of v_mfma_f32_16x16x32_f16 are those of the BF16 form.
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;

// Two FP16 terms, x = hi + lo (+ ~2^-22 |x|): the state (already scaled into
// FP16 range) and the constant bases (split from FP64).
struct HalfFrag2 {
    f16x8 hi, lo;
};

// Splits 8 values v * s (SCALE) or v.  The residual is formed as
// fma(v, s, -hi), one mixed-precision FMA per value (hi read as FP16).
// Unscaled, s must be a 1.0 the compiler cannot see through (a kernel
// argument), or it forms the residual as a conversion plus a subtraction.
template <bool SCALE>
__device__ inline HalfFrag2 split8_h(const float (&v)[8], float s) {
    typedef float f32x8 __attribute__((ext_vector_type(8)));
    f32x8 x, xs;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        x[e] = v[e];
        xs[e] = SCALE ? v[e] * s : v[e];
    }
    HalfFrag2 f;
    f.hi = __builtin_convertvector(xs, f16x8);
    const f32x8 h = __builtin_convertvector(f.hi, f32x8);
    f32x8 r;
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = __builtin_fmaf(x[e], s, -h[e]);
    f.lo = __builtin_convertvector(r, f16x8);
    return f;
}

// acc + A*B with A the state and B a constant: c_lo x_hi, c_hi x_lo, c_hi x_hi
// (three MFMA issues, the small terms first).  The dropped c_lo x_lo is
// ~2^-22 of hi*hi, the size of the two terms' own representation error.  A
// third constant term (one more product, also ~2^-22) lowered the DST's error
// on random 20 x 20 planes only from 2.3e-7 to 1.9e-7 of their maximum, for
// 6 % of the kernel's time.
__device__ inline f32x4 mma_sc(const HalfFrag2& a, const HalfFrag2& b, f32x4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.hi, b.lo, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.lo, b.hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.hi, b.hi, acc, 0, 0, 0);
    return acc;
}
// The same with A the constant and B the state.
__device__ inline f32x4 mma_cs(const HalfFrag2& a, const HalfFrag2& b, f32x4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.lo, b.hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.hi, b.lo, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.hi, b.hi, acc, 0, 0, 0);
    return acc;
}

// Maximum of non-negative v over the wave, as a wave-uniform value: DPP
// maxima within each row of 16 lanes (neighbour pairs, quads, half-row and
// row mirrors), then the four rows' results read into scalar registers.  On
// the bit patterns: for non-negative floats the integer order is the
// numeric order.
__device__ inline float wave_max_nonneg(float v) {
    unsigned u = __float_as_uint(v);
    u = max(u, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(u), 0xB1, 0xF, 0xF, false)));
    u = max(u, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(u), 0x4E, 0xF, 0xF, false)));
    u = max(u, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(u), 0x141, 0xF, 0xF, false)));
    u = max(u, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(u), 0x140, 0xF, 0xF, false)));
    const unsigned r0 = __builtin_amdgcn_readlane(u, 0), r1 = __builtin_amdgcn_readlane(u, 16);
    const unsigned r2 = __builtin_amdgcn_readlane(u, 32), r3 = __builtin_amdgcn_readlane(u, 48);
    return __uint_as_float(max(max(r0, r1), max(r2, r3)));
}

// Orders one wave's LDS writes before its later LDS reads of locations other
// lanes wrote.  The hardware already executes a wave's LDS instructions in
// order; the fence stops the compiler from forwarding or reusing values across
// the exchange (its alias analysis only sees the lane's own stores).
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace bfpn
