// mc_common.h -- shared device helpers and plan structures for libmc3.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MC_NTHR 512          // threads per FFT workgroup (8 waves of 64)
#define MC_LDS_ELEMS 16384   // complex fp32 elements of LDS per FFT workgroup (128 KiB)
#define MC_MAX_STAGES 16

struct __align__(8) cfloat {
    float x, y;
};

__device__ __forceinline__ cfloat cmk(float x, float y) { cfloat r; r.x = x; r.y = y; return r; }
__device__ __forceinline__ cfloat cadd(cfloat a, cfloat b) { return cmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cfloat csub(cfloat a, cfloat b) { return cmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cfloat cmul(cfloat a, cfloat b) {
    return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ cfloat cconj(cfloat a) { return cmk(a.x, -a.y); }
__device__ __forceinline__ cfloat cscale(cfloat a, float s) { return cmk(a.x * s, a.y * s); }
// multiply by -i
__device__ __forceinline__ cfloat cmuli_neg(cfloat a) { return cmk(a.y, -a.x); }

// One radix stage of a Stockham autosort FFT: radix r, sub-transform size ns
// before the stage, twiddles at tw + tw_off laid out [k][t-1] (k < ns,
// 1 <= t < r), roots of unity of a large prime at tw + root_off.
struct FftStage {
    int r;
    int ns;
    int tw_off;
    int root_off;
    int m;         // butterflies per transform (n / r)
    int nch;       // output chunks per butterfly (large primes)
    float rcp_m;   // 1/m, 1/ns, 1/nch for reciprocal division
    float rcp_ns;
    float rcp_nch;
};

// q = a / d, r = a % d for 0 <= a < 2^22 via a float reciprocal and one
// correction step (integer division is a long instruction sequence).
__device__ __forceinline__ int fdiv(int a, int d, float rcp) {
    int q = (int)((float)a * rcp);
    const int r = a - q * d;
    q += (r >= d) ? 1 : 0;
    q -= (r < 0) ? 1 : 0;
    return q;
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  Map the linear block id so that consecutive tiles (which
// share cache lines of the staging layout) run on the same XCD.
__device__ __forceinline__ int xcd_tile(int bid, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = bid & 7, i = bid >> 3;
    return x * q + (x < r ? x : r) + i;
}

// LDS index with one pad element per 16: strided butterfly writes then
// spread over all banks instead of hitting two of them.
__device__ __forceinline__ int lp(int i) { return i + (i >> 4); }
// LDS elements to allocate for `count` padded complex values.
__host__ __device__ inline int mc_lds_padded(int count) { return count + (count >> 4) + 1; }

// Plan of one complex FFT length, passed to kernels by value.
struct FftPlan1 {
    int n;
    int nst;
    FftStage st[MC_MAX_STAGES];
    const cfloat* tw;  // device table (forward sign)
};

static inline int mc_div_up(long long a, long long b) { return (int)((a + b - 1) / b); }
