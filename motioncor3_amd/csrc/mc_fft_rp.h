// mc_fft_rp.h -- register-pass FFT schedules (Stockham with composite radices).
//
// A length-N transform is done in a few passes of radix R_p (N = R_0 R_1 ...).
// Each work item of a pass holds one whole radix-R_p butterfly in registers,
// computed by a composite register DFT (RegDft), so a length like
// 4092 = 31 * 11 * 12 needs three passes and two LDS exchanges instead of one
// LDS round trip per prime factor.  The first pass reads its inputs straight
// from global memory and the last pass writes straight to global memory.
// LDS holds unpadded columns: an element's address is the work item's base
// plus a compile-time offset, so there is no per-element index arithmetic.
//
// Pass p (sub-transform size NS_p = R_0 ... R_{p-1}, M_p = N / R_p items):
//   item j = q NS_p + k reads x[j + t M_p] (t < R_p), multiplies input t by
//   W_{NS_p R_p}^{k t} (forward sign), transforms, and writes output t to
//   q NS_p R_p + k + t NS_p.  For the last pass NS_p R_p = N, so q = 0.
// Twiddles of pass p >= 1 are stored [t-1][k] (lanes with consecutive k read
// consecutive entries) at rp_twoff(N, p).
#pragma once
#include <cmath>
#include <vector>
#include "mc_fft_ct.h"

// ---- pass schedules ----
// Radix of pass p of schedule sc for the lengths with a register-pass
// schedule (0 = none).  Schedule 0 is each length's default order; schedule
// 1, where a length has one, orders the same radices differently (other
// strides, item counts and register footprints per pass, the same
// transform up to FP32 rounding) for mc_autotune to time.
__host__ __device__ constexpr int rp_radix(int n, int p, int sc = 0) {
    int r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    if (sc == 0) {
        if (n == 4092) { r0 = 31; r1 = 11; r2 = 12; }
        else if (n == 4096) { r0 = 16; r1 = 16; r2 = 16; }
        else if (n == 2048) { r0 = 16; r1 = 16; r2 = 8; }
        else if (n == 2880) { r0 = 15; r1 = 12; r2 = 16; }
        else if (n == 408) { r0 = 17; r1 = 24; }
        else if (n == 288) { r0 = 9; r1 = 32; }
        else if (n == 8184) { r0 = 31; r1 = 11; r2 = 24; }
        else if (n == 5760) { r0 = 15; r1 = 24; r2 = 16; }
    } else if (sc == 1) {
        if (n == 4092) { r0 = 31; r1 = 12; r2 = 11; }
        else if (n == 2880) { r0 = 16; r1 = 12; r2 = 15; }
        else if (n == 408) { r0 = 24; r1 = 17; }
        else if (n == 288) { r0 = 32; r1 = 9; }
    }
    return p == 0 ? r0 : p == 1 ? r1 : p == 2 ? r2 : p == 3 ? r3 : 0;
}
__host__ __device__ constexpr int rp_npass(int n, int sc = 0) {
    int c = 0;
    while (c < 4 && rp_radix(n, c, sc) != 0) c++;
    return c;
}
__host__ __device__ constexpr bool rp_has(int n, int sc = 0) {
    if (rp_npass(n, sc) < 2) return false;
    long long prod = 1;
    for (int p = 0; p < rp_npass(n, sc); p++) prod *= rp_radix(n, p, sc);
    return prod == n;
}
__host__ __device__ constexpr int rp_ns(int n, int p, int sc = 0) {
    int ns = 1;
    for (int i = 0; i < p; i++) ns *= rp_radix(n, i, sc);
    return ns;
}
__host__ __device__ constexpr int rp_twoff(int n, int p, int sc = 0) {
    int off = 0;
    for (int i = 1; i < p; i++) off += rp_ns(n, i, sc) * (rp_radix(n, i, sc) - 1);
    return off;
}
__host__ __device__ constexpr int rp_twcount(int n, int sc = 0) { return rp_twoff(n, rp_npass(n, sc), sc); }

// LDS layout of one transform.  Pass 0 stores an item's R_0 outputs at a
// stride of R_0 complex values across lanes; for an even R_0 (the power-of-
// two lengths) that maps a wave's lanes onto one or two LDS banks, and the
// middle passes' stores at NS = 16 collide in groups of 16 lanes.  Those
// lengths get one pad element per 16 (index i -> i + i/16), which makes both
// strides odd in 8-byte units.  Lengths with an odd R_0 keep the unpadded
// layout, so their kernels are unchanged.
__host__ __device__ constexpr bool rp_pad(int n, int sc = 0) {
    return rp_has(n, sc) && rp_radix(n, 0, sc) % 2 == 0 && n % 16 == 0;
}
// LDS elements of one transform (the column / row stride in LDS)
__host__ __device__ constexpr int rp_ld(int n, int sc = 0) { return rp_pad(n, sc) ? n + n / 16 : n; }
// LDS stride between the C transforms of a workgroup: odd in 8-byte units
// when C > 1, so that lanes that step over transforms (the column-fastest
// item order of the column passes) fall on different banks.
__host__ __device__ constexpr int rp_cs(int n, int c, int sc = 0) {
    return c > 1 ? (rp_ld(n, sc) | 1) : rp_ld(n, sc);
}
template <int N, int SC = 0>
static __device__ __forceinline__ int rp_ix(int i) {
    if constexpr (rp_pad(N, SC)) return i + (i >> 4);
    else return i;
}

// Multiply input t (1 <= t < R) of an item with twiddle index k.
template <int N, int P, int SC = 0>
static __device__ __forceinline__ void rp_twiddle(cfloat* v, int k, const cfloat* __restrict__ tw) {
    constexpr int R = rp_radix(N, P, SC), NS = rp_ns(N, P, SC), TWO = rp_twoff(N, P, SC);
    if constexpr (NS > 1) {
#pragma unroll
        for (int t = 1; t < R; t++) v[t] = cmul(v[t], tw[TWO + (t - 1) * NS + k]);
    }
}

// A middle pass over the C columns held in LDS at c * rp_cs(N, C), NT threads;
// item i -> column i / M, butterfly i % M.  Ends with a barrier.
template <int N, int P, int C, int NT, int SC = 0>
static __device__ __forceinline__ void rp_pass_lds(cfloat* lds, const cfloat* __restrict__ tw) {
    constexpr int R = rp_radix(N, P, SC), NS = rp_ns(N, P, SC), M = N / R;
    constexpr int IT = (C * M + NT - 1) / NT, LD = rp_cs(N, C, SC);
    cfloat v[IT][R];
    int dcol[IT], dst[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int i = threadIdx.x + it * NT;
        dst[it] = -1;
        dcol[it] = 0;
        if ((C * M) % NT == 0 || i < C * M) {
            const int c = i / M, j = i - c * M;
            const int k = j % NS, q = j / NS;
            const cfloat* src = lds + c * LD;
#pragma unroll
            for (int t = 0; t < R; t++) v[it][t] = src[rp_ix<N, SC>(j + t * M)];
            rp_twiddle<N, P, SC>(v[it], k, tw);
            RegDft<R>::run(v[it]);
            dcol[it] = c * LD;
            dst[it] = q * NS * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++) {
        if (dst[it] >= 0) {
            cfloat* d = lds + dcol[it];
#pragma unroll
            for (int t = 0; t < R; t++) d[rp_ix<N, SC>(dst[it] + t * NS)] = v[it][t];
        }
    }
    __syncthreads();
}

template <int N, int P, int C, int NT, int SC = 0>
struct RpMid {
    static __device__ __forceinline__ void run(cfloat* lds, const cfloat* tw) {
        if constexpr (P < rp_npass(N, SC) - 1) {
            rp_pass_lds<N, P, C, NT, SC>(lds, tw);
            RpMid<N, P + 1, C, NT, SC>::run(lds, tw);
        }
    }
};

// Passes P .. PEND-1 in LDS.
template <int N, int P, int PEND, int C, int NT, int SC = 0>
struct RpRange {
    static __device__ __forceinline__ void run(cfloat* lds, const cfloat* tw) {
        if constexpr (P < PEND) {
            rp_pass_lds<N, P, C, NT, SC>(lds, tw);
            RpRange<N, P + 1, PEND, C, NT, SC>::run(lds, tw);
        }
    }
};

// Host: the twiddle table of a register-pass schedule (forward sign), or an
// empty vector when n has none.
inline std::vector<cfloat> rp_twiddle_table(int n, int sc = 0) {
    std::vector<cfloat> tw;
    if (!rp_has(n, sc)) return tw;
    tw.resize(rp_twcount(n, sc) > 0 ? rp_twcount(n, sc) : 1);
    for (int pp = 1; pp < rp_npass(n, sc); pp++) {
        const int ns = rp_ns(n, pp, sc), r = rp_radix(n, pp, sc), off = rp_twoff(n, pp, sc);
        for (int t = 1; t < r; t++)
            for (int k = 0; k < ns; k++) {
                const double ang = -2.0 * 3.14159265358979323846 * (double)k * t / ((double)ns * r);
                tw[off + (t - 1) * ns + k] = cfloat{(float)std::cos(ang), (float)std::sin(ang)};
            }
    }
    return tw;
}
