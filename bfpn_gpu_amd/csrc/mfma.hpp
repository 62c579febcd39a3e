// mfma.hpp — split-BF16 matrix-core helpers for CDNA4 (gfx950).
//
// A product x*y of FP32 values is evaluated on the BF16 MFMA pipe as
//     x*y ~= xh*yh + xh*yl + xl*yh,   xh = bf16(x), xl = bf16(x - xh)
// with FP32 accumulation (v_mfma_f32_16x16x32_bf16, three issues per tile).
// xh + xl carries 16 significand bits, so each product is accurate to about
// 2^-16 relative; the dropped xl*yl term is ~2^-18.
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

// Three-term split, x = hi + mid + lo exactly for FP32 x (8 + 8 + 8 bits).
// Used where an operator is applied thousands of times in a row (the DST of
// every depth step): there a two-term split's ~2^-17 representation error
// compounds step after step -- for the constant bases, and for the slowly
// varying state as well (a two-term state split with exact bases still
// drifts by 2e-3 over 1000 applications) -- a three-term split's does not.
struct SplitFrag3 {
    bf16x8 hi, mid, lo;
};

__device__ inline SplitFrag3 split8_3(const float (&v)[8]) {
    typedef float f32x8 __attribute__((ext_vector_type(8)));
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    auto widen = [](bf16x8 b) {  // bf16 -> FP32 from the packed words
        const u32x4 w = __builtin_bit_cast(u32x4, b);
        f32x8 h;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            h[2 * k] = __uint_as_float(w[k] << 16);
            h[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
        }
        return h;
    };
    f32x8 x;
#pragma unroll
    for (int e = 0; e < 8; ++e) x[e] = v[e];
    SplitFrag3 f;
    f.hi = __builtin_convertvector(x, bf16x8);
    const f32x8 r1 = x - widen(f.hi);
    f.mid = __builtin_convertvector(r1, bf16x8);
    f.lo = __builtin_convertvector(r1 - widen(f.mid), bf16x8);
    return f;
}

// acc + A*B keeping every product term down to 2^-16 of hi*hi (six MFMA
// issues, the smallest terms first).
__device__ inline f32x4 mma_split3(const SplitFrag3& a, const SplitFrag3& b, f32x4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.mid, b.mid, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.lo, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.lo, b.hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.mid, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.mid, b.hi, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.hi, b.hi, acc, 0, 0, 0);
    return acc;
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
