// mc_fft_ct.h -- compile-time specialised in-LDS FFT stages.
//
// The runtime-radix stages of mc_fft_dev.h spend most of their instructions
// on index arithmetic (division by run-time sub-transform sizes).  For the
// transform lengths the reference's geometry produces, the radix list, the
// sub-transform size of every stage and the twiddle-table offsets are
// template constants here, so index math folds to shifts and multiply-highs.
// The factorisation and twiddle layout are exactly those of build_plan1 in
// mc_fft.hip (radix 8 while possible, then one 4 or 2, then prime factors in
// ascending order; per stage ns*(r-1) twiddles, plus r roots above 31).
#pragma once
#include "mc_fft_dev.h"

__host__ __device__ constexpr int ct_spf(int n) {
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}
__host__ __device__ constexpr int ct_next(int n) {
    return n % 8 == 0 ? 8 : (n % 4 == 0 ? 4 : (n % 2 == 0 ? 2 : ct_spf(n)));
}
__host__ __device__ constexpr int ct_nst(int n) {
    int c = 0;
    while (n > 1) {
        n /= ct_next(n);
        c++;
    }
    return c;
}
__host__ __device__ constexpr int ct_radix(int n, int s) {
    for (int i = 0; i < s; i++) n /= ct_next(n);
    return ct_next(n);
}
__host__ __device__ constexpr int ct_ns(int n, int s) {
    int ns = 1;
    for (int i = 0; i < s; i++) {
        const int q = ct_next(n);
        ns *= q;
        n /= q;
    }
    return ns;
}
__host__ __device__ constexpr int ct_twoff(int n, int s) {
    int off = 0, ns = 1;
    for (int i = 0; i < s; i++) {
        const int q = ct_next(n);
        off += ns * (q - 1) + (q > 31 ? q : 0);
        ns *= q;
        n /= q;
    }
    return off;
}

// One register-butterfly stage for a transform of N points held at LDS
// index base..base+N-1 (padded), processed by TPT threads (tid < TPT).
template <int N, int S, int TPT>
static __device__ __forceinline__ void ct_stage(cfloat* lds, int base, int tid,
                                                const cfloat* __restrict__ tw) {
    constexpr int R = ct_radix(N, S), NS = ct_ns(N, S), M = N / R;
    constexpr int IT = (M + TPT - 1) / TPT;
    constexpr int TWO = ct_twoff(N, S);
    cfloat v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int j = tid + it * TPT;
        if ((M % TPT) == 0 || j < M) {
            const int k = j % NS;
#pragma unroll
            for (int t = 0; t < R; t++) v[it][t] = lds[lp(base + j + t * M)];
            if (NS > 1) {
                const cfloat* w = tw + TWO + k * (R - 1);
#pragma unroll
                for (int t = 1; t < R; t++) v[it][t] = cmul(v[it][t], w[t - 1]);
            }
            bfly<R>(v[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int j = tid + it * TPT;
        if ((M % TPT) == 0 || j < M) {
            const int k = j % NS, q = j / NS;
            const int d = base + q * NS * R + k;
#pragma unroll
            for (int t = 0; t < R; t++) lds[lp(d + t * NS)] = v[it][t];
        }
    }
    __syncthreads();
}

// Stage of a prime radix above 31: KC outputs per work item by direct sum.
template <int N, int S, int TPT>
static __device__ __forceinline__ void ct_stage_big(cfloat* lds, int base, int tid,
                                                    const cfloat* __restrict__ tw) {
    constexpr int P = ct_radix(N, S), NS = ct_ns(N, S), M = N / P;
    constexpr int KC = MC_BIG_KC, NCH = (P + KC - 1) / KC;
    constexpr int ITEMS = M * NCH, IT = (ITEMS + TPT - 1) / TPT;
    constexpr int TWO = ct_twoff(N, S), ROOT = TWO + NS * (P - 1);
    cfloat acc[IT][KC];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int item = tid + it * TPT;
#pragma unroll
        for (int q = 0; q < KC; q++) acc[it][q] = cmk(0.f, 0.f);
        if (item < ITEMS) {
            const int bf = item / NCH, ch = item % NCH;
            const int k = bf % NS;
            const int q0 = ch * KC;
            int ridx[KC], step[KC];
#pragma unroll
            for (int q = 0; q < KC; q++) {
                ridx[q] = 0;
                step[q] = (q0 + q < P) ? q0 + q : 0;
            }
            for (int t = 0; t < P; t++) {
                cfloat x = lds[lp(base + bf + t * M)];
                if (NS > 1 && t > 0) x = cmul(x, tw[TWO + k * (P - 1) + t - 1]);
#pragma unroll
                for (int q = 0; q < KC; q++) {
                    const cfloat r = tw[ROOT + ridx[q]];
                    acc[it][q].x = fmaf(x.x, r.x, fmaf(-x.y, r.y, acc[it][q].x));
                    acc[it][q].y = fmaf(x.x, r.y, fmaf(x.y, r.x, acc[it][q].y));
                    const int nx = ridx[q] + step[q];
                    ridx[q] = nx >= P ? nx - P : nx;
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int item = tid + it * TPT;
        if (item < ITEMS) {
            const int bf = item / NCH, ch = item % NCH;
            const int k = bf % NS, q = bf / NS;
            const int d = base + q * NS * P + k;
#pragma unroll
            for (int qq = 0; qq < KC; qq++)
                if (ch * KC + qq < P) lds[lp(d + (ch * KC + qq) * NS)] = acc[it][qq];
        }
    }
    __syncthreads();
}

// Largest prime radix above 31 in the factorisation of N (0 if none).
__host__ __device__ constexpr int ct_big_prime(int n) {
    int b = 0;
    for (int s = 0; s < ct_nst(n); s++)
        if (ct_radix(n, s) > 31) b = ct_radix(n, s);
    return b;
}

// Output pairs per work item of the big-prime stage, and its chunk root
// table: after the per-stage twiddles and roots of the whole transform (at
// an even offset, so that rows of it are 16-byte aligned), for every output
// chunk ch and t = 1..(P-1)/2 the roots e^{-2 pi i t k / P} of the chunk's
// KC outputs k = ch*KC+1.. (root 0 where k > (P-1)/2).
#define MC_BIGSYM_KC 4
__host__ __device__ constexpr int ct_rt_off(int n) {
    const int o = ct_twoff(n, ct_nst(n));
    return o + (o & 1);
}
__host__ __device__ constexpr int ct_rt_elems(int n) {
    const int p = ct_big_prime(n), hp = (p - 1) / 2;
    return p ? ((hp + MC_BIGSYM_KC - 1) / MC_BIGSYM_KC) * hp * MC_BIGSYM_KC : 0;
}

// Prime radix above 31, symmetric-pair form: phase 1 forms
// u_t = x_t + x_{P-t}, v_t = x_t - x_{P-t} (twiddled) of every butterfly and,
// after all of the transform's values are in registers, stores them unpadded
// in the transform's own LDS range ([t][j]: u at t*M + j, v after all u, then
// x_0 of every butterfly: M (P-1) + M = N elements); phase 2 gives each work
// item KC output pairs y_k = x0 + sum_t u_t cos(2 pi t k/P)
// - i sum_t v_t sin(2 pi t k/P) and y_{P-k} (the conjugate-sign partner), with
// the KC roots of each t read as two 16-byte rows of the chunk root table
// (lanes of one chunk share them), so the loop carries no index arithmetic.
template <int N, int S, int TPT>
static __device__ __forceinline__ void ct_stage_bigsym(cfloat* lds, int base, int tid,
                                                       const cfloat* __restrict__ tw) {
    constexpr int P = ct_radix(N, S), NS = ct_ns(N, S), M = N / P, HP = (P - 1) / 2;
    constexpr int KC = MC_BIGSYM_KC, NCH = (HP + KC - 1) / KC;
    constexpr int TWO = ct_twoff(N, S), RT = ct_rt_off(N);
    constexpr int IT1 = (M * HP + TPT - 1) / TPT, IX0 = (M + TPT - 1) / TPT;
    static_assert(KC == 4, "two float4 root rows per t");
    static_assert(ct_big_prime(N) == P, "one prime radix above 31 per length");
    // the transform occupies padded indices base..base+N-1, i.e. at least N
    // consecutive elements from lp(base)
    cfloat* row = lds + lp(base);
    cfloat ua[IT1], va[IT1], x0a[IX0];
#pragma unroll
    for (int it = 0; it < IT1; it++) {
        const int i = tid + it * TPT;
        if ((M * HP) % TPT == 0 || i < M * HP) {
            const int t = i / M + 1, j = i - (t - 1) * M;
            const int k = j % NS;
            cfloat a = lds[lp(base + j + t * M)], b = lds[lp(base + j + (P - t) * M)];
            if (NS > 1) {
                a = cmul(a, tw[TWO + k * (P - 1) + t - 1]);
                b = cmul(b, tw[TWO + k * (P - 1) + P - t - 1]);
            }
            ua[it] = cadd(a, b);
            va[it] = csub(a, b);
        }
    }
#pragma unroll
    for (int it = 0; it < IX0; it++) {
        const int j = tid + it * TPT;
        if (M % TPT == 0 || j < M) x0a[it] = lds[lp(base + j)];
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT1; it++) {
        const int i = tid + it * TPT;
        if ((M * HP) % TPT == 0 || i < M * HP) {
            row[i] = ua[it];
            row[M * HP + i] = va[it];
        }
    }
#pragma unroll
    for (int it = 0; it < IX0; it++) {
        const int j = tid + it * TPT;
        if (M % TPT == 0 || j < M) row[2 * M * HP + j] = x0a[it];
    }
    __syncthreads();
    constexpr int ITEMS = M * NCH, IT = (ITEMS + TPT - 1) / TPT;
    cfloat yk[IT][KC], ym[IT][KC], y0[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int item = tid + it * TPT;
        if (item < ITEMS) {
            const int ch = item / M, j = item % M;
            const cfloat x0 = row[2 * M * HP + j];
            float ax[KC], ay[KC], bx[KC], by[KC];
#pragma unroll
            for (int q = 0; q < KC; q++) {
                ax[q] = x0.x;
                ay[q] = x0.y;
                bx[q] = by[q] = 0.f;
            }
            cfloat s0 = x0;  // y_0, kept by the items of chunk 0
            const cfloat* u = row + j;
            const cfloat* v = row + M * HP + j;
            const float4* rt = reinterpret_cast<const float4*>(tw + RT + ch * HP * KC);
#pragma unroll 13
            for (int t = 1; t <= HP; t++) {
                const cfloat ut = u[(t - 1) * M], vt = v[(t - 1) * M];
                s0 = cadd(s0, ut);
                // roots e^{-2 pi i t k / P} = (cos, -sin) of the chunk's KC outputs
                const float4 r01 = rt[2 * (t - 1)], r23 = rt[2 * (t - 1) + 1];
                const cfloat r[KC] = {cmk(r01.x, r01.y), cmk(r01.z, r01.w), cmk(r23.x, r23.y),
                                      cmk(r23.z, r23.w)};
#pragma unroll
                for (int q = 0; q < KC; q++) {
                    ax[q] = fmaf(ut.x, r[q].x, ax[q]);
                    ay[q] = fmaf(ut.y, r[q].x, ay[q]);
                    bx[q] = fmaf(vt.x, -r[q].y, bx[q]);
                    by[q] = fmaf(vt.y, -r[q].y, by[q]);
                }
            }
            y0[it] = s0;
#pragma unroll
            for (int q = 0; q < KC; q++) {
                yk[it][q] = cmk(ax[q] + by[q], ay[q] - bx[q]);
                ym[it][q] = cmk(ax[q] - by[q], ay[q] + bx[q]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int item = tid + it * TPT;
        if (item < ITEMS) {
            const int ch = item / M, j = item % M;
            const int k = j % NS, qd = j / NS;
            const int d = base + qd * NS * P + k;
            if (ch == 0) lds[lp(d)] = y0[it];
#pragma unroll
            for (int q = 0; q < KC; q++) {
                const int kk = ch * KC + q + 1;
                if (kk <= HP) {
                    lds[lp(d + kk * NS)] = yk[it][q];
                    lds[lp(d + (P - kk) * NS)] = ym[it][q];
                }
            }
        }
    }
    __syncthreads();
}

template <int N, int S, int TPT>
struct CtFft {
    static __device__ __forceinline__ void run(cfloat* lds, int base, int tid, const cfloat* tw) {
        if constexpr (S < ct_nst(N)) {
            if constexpr (ct_radix(N, S) > 31) ct_stage_bigsym<N, S, TPT>(lds, base, tid, tw);
            else ct_stage<N, S, TPT>(lds, base, tid, tw);
            CtFft<N, S + 1, TPT>::run(lds, base, tid, tw);
        }
    }
};

// Forward FFT of the group's transform (all threads of the workgroup call).
template <int N, int TPT>
static __device__ __forceinline__ void ct_fft(cfloat* lds, int base, int tid, const cfloat* tw) {
    CtFft<N, 0, TPT>::run(lds, base, tid, tw);
}
