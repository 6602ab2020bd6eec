// mc_spec.h -- storage formats of the alignment spectrum stacks.
//
// The xcf and patch spectrum stacks are read on every alignment iteration.
// They may be stored as FP32 complex (8 bytes), FP16 complex (4 bytes) or
// BF16 complex (4 bytes); arithmetic is FP32 in every case.  A reduced
// format stores X * g_b, with a per-transform factor g_b that maps the
// transform's largest magnitude to about 2^14 (FP16: well inside its range,
// away from subnormals).  For non-negative images the largest magnitude of
// a spectrum is its DC term, and sum_rows |row DC| bounds it in general, so
// the row pass records |row DC| and k_spec_scale turns those into g_b.
// This storage choice has no counterpart in the reference (it keeps cufftComplex
// stacks, CBufferPool.cpp:220-262).
#pragma once
#include <hip/hip_fp16.h>
#include "mc_common.h"

#define MC_SPEC_F32 0
#define MC_SPEC_F16 1
#define MC_SPEC_BF16 2
#define MC_SPEC_TARGET 16384.0f

template <int FMT>
struct Spec;

template <>
struct Spec<MC_SPEC_F32> {
    static constexpr int bytes = 8;
    static __device__ __forceinline__ cfloat ld(const void* p, long long i, float) {
        return reinterpret_cast<const cfloat*>(p)[i];
    }
    static __device__ __forceinline__ void st(void* p, long long i, cfloat v, float) {
        reinterpret_cast<cfloat*>(p)[i] = v;
    }
};

template <>
struct Spec<MC_SPEC_F16> {
    static constexpr int bytes = 4;
    // inv = 1 / g on load, g on store
    static __device__ __forceinline__ cfloat ld(const void* p, long long i, float inv) {
        const __half2 h = reinterpret_cast<const __half2*>(p)[i];
        const float2 f = __half22float2(h);
        return cmk(f.x * inv, f.y * inv);
    }
    static __device__ __forceinline__ void st(void* p, long long i, cfloat v, float g) {
        reinterpret_cast<__half2*>(p)[i] = __floats2half2_rn(v.x * g, v.y * g);
    }
};

// BF16 by round-to-nearest-even of the FP32 bit pattern (NaN kept quiet).
__device__ __forceinline__ unsigned int mc_bf16_bits(float f) {
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (u >> 16) | 0x40u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

template <>
struct Spec<MC_SPEC_BF16> {
    static constexpr int bytes = 4;
    static __device__ __forceinline__ cfloat ld(const void* p, long long i, float inv) {
        const unsigned int w = reinterpret_cast<const unsigned int*>(p)[i];
        const float re = __uint_as_float(w << 16), im = __uint_as_float(w & 0xffff0000u);
        return cmk(re * inv, im * inv);
    }
    static __device__ __forceinline__ void st(void* p, long long i, cfloat v, float g) {
        reinterpret_cast<unsigned int*>(p)[i] = mc_bf16_bits(v.x * g) | (mc_bf16_bits(v.y * g) << 16);
    }
};

// Per-transform factors of a reduced format: g[b] = TARGET / sum_r |dc[b][r]|
// (1 when the sum is 0) and inv[b] = 1 / g[b] (mc_launch_spec_scale).
struct SpecScale {
    const float* g;    // store factors, or null for FP32
    const float* inv;  // load factors, or null for FP32
};
