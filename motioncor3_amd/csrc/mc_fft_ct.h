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
// ---- compile-time trigonometry (Taylor series in double, |x| <= pi) ----
__host__ __device__ constexpr double rp_pi() { return 3.14159265358979323846; }
__host__ __device__ constexpr double rp_sin_red(double x) {
    double term = x, sum = x;
    for (int n = 1; n < 30; n++) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
__host__ __device__ constexpr double rp_cos_red(double x) {
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 30; n++) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
// cos and sin of 2 pi a / b for integers, argument reduced to [-pi, pi]
__host__ __device__ constexpr double rp_cos2pi(int a, int b) {
    int r = a % b;
    if (r < 0) r += b;
    if (2 * r > b) r -= b;
    return rp_cos_red(2.0 * rp_pi() * r / b);
}
__host__ __device__ constexpr double rp_sin2pi(int a, int b) {
    int r = a % b;
    if (r < 0) r += b;
    if (2 * r > b) r -= b;
    return rp_sin_red(2.0 * rp_pi() * r / b);
}

// ---- composite register DFT: v[k] = sum_t v[t] e^{-2 pi i t k / R} ----
__host__ __device__ constexpr int rp_first(int r) {
    // leading factor of a composite register DFT: 4 or 2 for even sizes
    // (small butterflies), else the smallest odd prime
    return r % 4 == 0 && r != 4 && r != 8 ? 4 : (r % 2 == 0 && r != 2 && r != 4 && r != 8 ? 2 : ct_spf(r));
}
__host__ __device__ constexpr bool rp_leaf(int r) {
    return r == 1 || r == 2 || r == 4 || r == 8 || (ct_spf(r) == r && r <= 31);
}

template <int R, bool LEAF = rp_leaf(R)>
struct RegDft;

template <int R>
struct RegDft<R, true> {
    static __device__ __forceinline__ void run(cfloat* v) {
        if constexpr (R > 1) bfly<R>(v);
    }
};

// cos/sin(2 pi i / R) for a composite register DFT, built at compile time.
struct RpTab {
    float c[64];
    float s[64];
};
template <int R>
__host__ __device__ constexpr RpTab rp_tab() {
    RpTab t{};
    for (int i = 0; i < R && i < 64; i++) {
        t.c[i] = (float)rp_cos2pi(i, R);
        t.s[i] = (float)rp_sin2pi(i, R);
    }
    return t;
}

// R = A * B: A-point DFTs over n1 (stride B), twiddle W_R^{n2 k1},
// B-point DFTs over n2; output k1 + A k2.
template <int R>
struct RegDft<R, false> {
    static_assert(R <= 64, "composite register DFTs up to 64 points");
    static __device__ __forceinline__ void run(cfloat* v) {
        constexpr int A = rp_first(R), B = R / A;
        constexpr RpTab TW = rp_tab<R>();
        cfloat w[R];
#pragma unroll
        for (int n2 = 0; n2 < B; n2++) {
            cfloat s[A];
#pragma unroll
            for (int n1 = 0; n1 < A; n1++) s[n1] = v[n2 + B * n1];
            RegDft<A>::run(s);
#pragma unroll
            for (int k1 = 0; k1 < A; k1++) {
                if (n2 == 0 || k1 == 0) {
                    w[n2 * A + k1] = s[k1];
                } else {
                    const float c = TW.c[(n2 * k1) % R];
                    const float sn = TW.s[(n2 * k1) % R];
                    // multiply by e^{-i theta} = (c, -sn)
                    w[n2 * A + k1] = cmk(s[k1].x * c + s[k1].y * sn, s[k1].y * c - s[k1].x * sn);
                }
            }
        }
#pragma unroll
        for (int k1 = 0; k1 < A; k1++) {
            cfloat s[B];
#pragma unroll
            for (int n2 = 0; n2 < B; n2++) s[n2] = w[n2 * A + k1];
            RegDft<B>::run(s);
#pragma unroll
            for (int k2 = 0; k2 < B; k2++) v[k1 + A * k2] = s[k2];
        }
    }
};


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

// ---- Rader stage for the prime radix above 31 ----
// y_m = sum_t x_t w^{tm} (w = e^{-2 pi i/P}) over the P inputs of a
// butterfly.  With a primitive root g, t = g^q and m = g^-p (q, p < L =
// P - 1) turn the non-zero part into a cyclic convolution:
// y_{g^-p} = x_0 + c_p, c = a (*) b, a_q = x_{g^q}, b_n = w^{g^-n}, and
// y_0 = x_0 + sum_q a_q.  The convolution is done by length-L DFTs,
// L = L1 * L2 (Cooley-Tukey, L2 the largest prime factor): forward
// A[k1 + L1 k2] = DFT_L2 over n2 of (w_L^{n2 k1} DFT_L1 over n1 of
// a[L2 n1 + n2]), product with Bh = DFT_L(b) / L, then the inverse steps in
// reverse order, so c comes out in natural order.  Work per butterfly is
// O(L (L1 + L2)) instead of the direct form's O(P^2).
__host__ __device__ constexpr int ct_powmod(int g, int e, int p) {
    long long r = 1, b = g % p;
    while (e > 0) {
        if (e & 1) r = (r * b) % p;
        b = (b * b) % p;
        e >>= 1;
    }
    return (int)r;
}
__host__ __device__ constexpr int ct_proot(int p) {
    for (int g = 2; g < p; g++) {
        bool ok = true;
        for (int d = 2; d <= p - 1 && ok; d++)
            if ((p - 1) % d == 0 && ct_spf(d) == d && ct_powmod(g, (p - 1) / d, p) == 1) ok = false;
        if (ok) return g;
    }
    return 0;
}
__host__ __device__ constexpr int ct_lpf(int n) {
    int l = 1;
    for (int f = 2; f <= n; f++)
        while (n % f == 0) {
            l = f;
            n /= f;
        }
    return l;
}
// L2 and L1 of the convolution length of prime P (0 when the stage has no
// Rader form: L2 a prime of at most 31 and L1 at most 32 are compiled)
__host__ __device__ constexpr int ct_rd_l2(int p) {
    return (ct_lpf(p - 1) <= 31 && (p - 1) / ct_lpf(p - 1) <= 32) ? ct_lpf(p - 1) : 0;
}
__host__ __device__ constexpr int ct_rd_l1(int p) { return ct_rd_l2(p) ? (p - 1) / ct_rd_l2(p) : 0; }
// Rader table after the chunk root table: Bh[k] (k < L), then w_L^{n2 k1}
// at n2 * L1 + k1
__host__ __device__ constexpr int ct_rd_off(int n) { return ct_rt_off(n) + ct_rt_elems(n); }
__host__ __device__ constexpr int ct_rd_elems(int n) {
    return ct_big_prime(n) && ct_rd_l2(ct_big_prime(n)) ? 2 * (ct_big_prime(n) - 1) : 0;
}
// whether the big-prime stage of N runs in the Rader form
__host__ __device__ constexpr bool ct_rader(int n) {
    const int p = ct_big_prime(n);
    return p && ct_rd_l2(p) && (n / p) * (p - 1) + n / p == n;
}

template <int N, int S, int TPT>
static __device__ __forceinline__ void ct_stage_rader(cfloat* lds, int base, int tid,
                                                      const cfloat* __restrict__ tw) {
    constexpr int P = ct_radix(N, S), NS = ct_ns(N, S), M = N / P, L = P - 1;
    constexpr int L2 = ct_rd_l2(P), L1 = ct_rd_l1(P);
    constexpr int G = ct_proot(P), GI = ct_powmod(G, P - 2, P);
    constexpr int GS = ct_powmod(G, L2, P), GIS = ct_powmod(GI, L2, P);
    constexpr int TWO = ct_twoff(N, S), RD = ct_rd_off(N);
    constexpr int I1 = M * L2, I2 = M * L1;  // items of steps 1 and 3, and of step 2
    constexpr int IT1 = (I1 + TPT - 1) / TPT, IT2 = (I2 + TPT - 1) / TPT;
    static_assert(ct_rader(N), "Rader form not available for this length");
    const cfloat* Bh = tw + RD;
    const cfloat* wl = tw + RD + L;
    // step 1: item (j, n2) gathers a[L2 n1 + n2] = x_{g^(L2 n1 + n2)},
    // transforms over n1 and twiddles by w_L^{n2 k1}; the workspace
    // W[j][k1][n2] (base + j L + k1 L2 + n2) and the x_0 stash (base + M L + j)
    // overwrite the stage's input only after everything is read
    cfloat v1[IT1][L1];
    cfloat x0[IT1];
#pragma unroll
    for (int it = 0; it < IT1; it++) {
        const int i = tid + it * TPT;
        x0[it] = cmk(0.f, 0.f);
        if (I1 % TPT == 0 || i < I1) {
            const int j = i / L2, n2 = i - j * L2, k = j % NS;
            int t = 1;
            for (int e = 0; e < n2; e++) t = (t * G) % P;
            // the gather indices first, then all twiddle loads as one batch
            // (an index chain between them would hold each load to its own
            // round trip)
            int ti[L1];
#pragma unroll
            for (int n1 = 0; n1 < L1; n1++) {
                ti[n1] = t;
                t = (t * GS) % P;
            }
            if (NS > 1) {
#pragma unroll
                for (int n1 = 0; n1 < L1; n1++) v1[it][n1] = tw[TWO + k * (P - 1) + ti[n1] - 1];
            }
#pragma unroll
            for (int n1 = 0; n1 < L1; n1++) {
                const cfloat x = lds[lp(base + j + ti[n1] * M)];
                v1[it][n1] = NS > 1 ? cmul(x, v1[it][n1]) : x;
            }
            if (n2 == 0) x0[it] = lds[lp(base + j)];
            RegDft<L1>::run(v1[it]);
#pragma unroll
            for (int k1 = 1; k1 < L1; k1++) v1[it][k1] = cmul(v1[it][k1], wl[n2 * L1 + k1]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT1; it++) {
        const int i = tid + it * TPT;
        if (I1 % TPT == 0 || i < I1) {
            const int j = i / L2, n2 = i - j * L2;
#pragma unroll
            for (int k1 = 0; k1 < L1; k1++) lds[lp(base + j * L + k1 * L2 + n2)] = v1[it][k1];
            if (n2 == 0) lds[lp(base + M * L + j)] = x0[it];
        }
    }
    __syncthreads();
    // step 2: item (j, k1) transforms over n2, multiplies by Bh, transforms
    // back over k2 and applies the conjugate twiddle in place; (j, 0) also
    // forms y_0 = x_0 + sum_q a_q, written after step 3
    cfloat y0[IT2];
#pragma unroll
    for (int it = 0; it < IT2; it++) {
        const int i = tid + it * TPT;
        y0[it] = cmk(0.f, 0.f);
        if (I2 % TPT == 0 || i < I2) {
            const int j = i / L1, k1 = i - j * L1;
            cfloat v2[L2], bh[L2];
            const int o = base + j * L + k1 * L2;
            // the item's table values as batches of loads ahead of their
            // uses (a load placed at its use, behind a condition, waits out
            // a round trip of its own)
#pragma unroll
            for (int k2 = 0; k2 < L2; k2++) bh[k2] = Bh[k1 + L1 * k2];
#pragma unroll
            for (int n2 = 0; n2 < L2; n2++) v2[n2] = lds[lp(o + n2)];
            RegDft<L2>::run(v2);
            if (k1 == 0) y0[it] = cadd(lds[lp(base + M * L + j)], v2[0]);
#pragma unroll
            for (int k2 = 0; k2 < L2; k2++) v2[k2] = cconj(cmul(v2[k2], bh[k2]));
            cfloat wv[L2];
#pragma unroll
            for (int n2 = 1; n2 < L2; n2++) wv[n2] = wl[n2 * L1 + k1];
            RegDft<L2>::run(v2);
#pragma unroll
            for (int n2 = 0; n2 < L2; n2++) {
                cfloat z = cconj(v2[n2]);
                if (n2 > 0 && k1 > 0) z = cmul(z, cconj(wv[n2]));
                lds[lp(o + n2)] = z;
            }
        }
    }
    __syncthreads();
    // step 3: item (j, n2) transforms back over k1: c[L2 n1 + n2], which is
    // output g^-(L2 n1 + n2), plus x_0
    cfloat v3[IT1][L1];
#pragma unroll
    for (int it = 0; it < IT1; it++) {
        const int i = tid + it * TPT;
        x0[it] = cmk(0.f, 0.f);
        if (I1 % TPT == 0 || i < I1) {
            const int j = i / L2, n2 = i - j * L2;
#pragma unroll
            for (int k1 = 0; k1 < L1; k1++) v3[it][k1] = cconj(lds[lp(base + j * L + k1 * L2 + n2)]);
            x0[it] = lds[lp(base + M * L + j)];
            RegDft<L1>::run(v3[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT1; it++) {
        const int i = tid + it * TPT;
        if (I1 % TPT == 0 || i < I1) {
            const int j = i / L2, n2 = i - j * L2;
            const int q = j / NS, k = j - q * NS;
            const int d = base + q * NS * P + k;
            int m = 1;
            for (int e = 0; e < n2; e++) m = (m * GI) % P;
#pragma unroll
            for (int n1 = 0; n1 < L1; n1++) {
                lds[lp(d + m * NS)] = cadd(x0[it], cconj(v3[it][n1]));
                m = (m * GIS) % P;
            }
        }
    }
#pragma unroll
    for (int it = 0; it < IT2; it++) {
        const int i = tid + it * TPT;
        if (I2 % TPT == 0 || i < I2) {
            const int j = i / L1, k1 = i - j * L1;
            if (k1 == 0) {
                const int q = j / NS, k = j - q * NS;
                lds[lp(base + q * NS * P + k)] = y0[it];
            }
        }
    }
    __syncthreads();
}

template <int N, int S, int TPT>
struct CtFft {
    static __device__ __forceinline__ void run(cfloat* lds, int base, int tid, const cfloat* tw) {
        if constexpr (S < ct_nst(N)) {
            if constexpr (ct_radix(N, S) > 31 && ct_rader(N)) ct_stage_rader<N, S, TPT>(lds, base, tid, tw);
            else if constexpr (ct_radix(N, S) > 31) ct_stage_bigsym<N, S, TPT>(lds, base, tid, tw);
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
