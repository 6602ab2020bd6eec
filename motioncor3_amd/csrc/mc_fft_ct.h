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

// Largest prime radix above 31 in the factorisation of N (0 if none) and
// the LDS scratch (complex elements) the symmetric big-prime stage needs per
// transform: (P-1)/2 pair sums and differences per butterfly, plus P roots.
__host__ __device__ constexpr int ct_big_prime(int n) {
    int b = 0;
    for (int s = 0; s < ct_nst(n); s++)
        if (ct_radix(n, s) > 31) b = ct_radix(n, s);
    return b;
}
__host__ __device__ constexpr int ct_scratch(int n) {
    const int p = ct_big_prime(n);
    return p ? (n / p) * (p - 1) + p : 0;
}
inline int ct_scratch_host(int n) { return ct_scratch(n); }

// Prime radix above 31, symmetric-pair form: phase 1 forms
// u_t = x_t + x_{P-t}, v_t = x_t - x_{P-t} (twiddled) of every butterfly in
// LDS scratch; phase 2 gives each work item KC output pairs
// y_k = x0 + sum_t u_t cos(2 pi t k/P) - i sum_t v_t sin(2 pi t k/P) and
// y_{P-k} (the conjugate-sign partner), with the roots read from LDS.
template <int N, int S, int TPT>
static __device__ __forceinline__ void ct_stage_bigsym(cfloat* lds, int base, int tid,
                                                       const cfloat* __restrict__ tw,
                                                       cfloat* scr) {
    constexpr int P = ct_radix(N, S), NS = ct_ns(N, S), M = N / P, HP = (P - 1) / 2;
    constexpr int KC = 4, NCH = (HP + KC - 1) / KC;
    constexpr int TWO = ct_twoff(N, S), ROOT = TWO + NS * (P - 1);
    // u and v pair sums stored [t][j], so that the lanes of one output chunk
    // (consecutive j) read consecutive entries; roots then are one broadcast
    // per t for all of them
    cfloat* uv = scr;            // [HP][M] u then [HP][M] v
    cfloat* roots = scr + 2 * M * HP;
    for (int i = tid; i < P; i += TPT) roots[i] = tw[ROOT + i];
    for (int i = tid; i < M * HP; i += TPT) {
        const int j = i / HP, t = i % HP + 1;
        const int k = j % NS;
        cfloat a = lds[lp(base + j + t * M)], b = lds[lp(base + j + (P - t) * M)];
        if (NS > 1) {
            a = cmul(a, tw[TWO + k * (P - 1) + t - 1]);
            b = cmul(b, tw[TWO + k * (P - 1) + P - t - 1]);
        }
        uv[(t - 1) * M + j] = cadd(a, b);
        uv[M * HP + (t - 1) * M + j] = csub(a, b);
    }
    __syncthreads();
    constexpr int ITEMS = M * NCH, IT = (ITEMS + TPT - 1) / TPT;
    cfloat yk[IT][KC], ym[IT][KC], y0[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int item = tid + it * TPT;
        if (item < ITEMS) {
            const int ch = item / M, j = item % M;
            const cfloat x0 = lds[lp(base + j)];
            float ax[KC], ay[KC], bx[KC], by[KC];
            int idx[KC];
#pragma unroll
            for (int q = 0; q < KC; q++) {
                ax[q] = x0.x;
                ay[q] = x0.y;
                bx[q] = by[q] = 0.f;
                idx[q] = 0;
            }
            cfloat s0 = x0;
            const cfloat* u = uv + j;
            const cfloat* v = uv + M * HP + j;
            for (int t = 1; t <= HP; t++) {
                const cfloat ut = u[(t - 1) * M], vt = v[(t - 1) * M];
                if (ch == 0) s0 = cadd(s0, ut);
#pragma unroll
                for (int q = 0; q < KC; q++) {
                    const int kk = ch * KC + q + 1;
                    int id = idx[q] + (kk <= HP ? kk : 0);
                    id = id >= P ? id - P : id;
                    idx[q] = id;
                    // roots hold e^{-2 pi i id / P} = (cos, -sin)
                    const cfloat r = roots[id];
                    ax[q] = fmaf(ut.x, r.x, ax[q]);
                    ay[q] = fmaf(ut.y, r.x, ay[q]);
                    bx[q] = fmaf(vt.x, -r.y, bx[q]);
                    by[q] = fmaf(vt.y, -r.y, by[q]);
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
    static __device__ __forceinline__ void run(cfloat* lds, int base, int tid, const cfloat* tw,
                                               cfloat* scr) {
        if constexpr (S < ct_nst(N)) {
            if constexpr (ct_radix(N, S) > 31) ct_stage_bigsym<N, S, TPT>(lds, base, tid, tw, scr);
            else ct_stage<N, S, TPT>(lds, base, tid, tw);
            CtFft<N, S + 1, TPT>::run(lds, base, tid, tw, scr);
        }
    }
};

// Forward FFT of the group's transform (all threads of the workgroup call).
// scr: this group's LDS scratch of ct_scratch(N) elements (unused when N has
// no prime factor above 31).
template <int N, int TPT>
static __device__ __forceinline__ void ct_fft(cfloat* lds, int base, int tid, const cfloat* tw,
                                              cfloat* scr = nullptr) {
    CtFft<N, 0, TPT>::run(lds, base, tid, tw, scr);
}
