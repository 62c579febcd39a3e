// buffer.hpp — buffer-resource addressing for the streaming kernels (gfx950).
//
// A buffer resource holds a base address and a byte size in four SGPRs; an
// access adds a uniform scalar offset and a per-lane 32-bit vector offset,
// so no address arithmetic runs on the vector ALUs.  The hardware checks the
// offset against the size: a load past the end returns 0 and a store past
// the end is dropped, which masks lanes without a branch (offset kOutside).
#pragma once

#include <hip/hip_runtime.h>

namespace bfpn {

// 0x00020000: the raw 32-bit data format word of gfx9-family buffer resources.
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes),
                                             0x00020000);
}

// A vector offset past the end of every resource (sizes stay below 2^31).
constexpr unsigned kOutside = 0x80000000u;

template <typename T>
__device__ inline T bload(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ inline float bload<float>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <>
__device__ inline double bload<double>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ inline float4 bload4(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

__device__ inline void bstore(float v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 0);
}
__device__ inline void bstore(double v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}
__device__ inline void bstore4(float4 v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), r, voff, soff, 0);
}

}  // namespace bfpn
