// mc_fft_dev.h -- in-LDS mixed-radix FFT device code (Stockham autosort).
//
// A workgroup of MC_NTHR threads transforms `batch` contiguous sequences of
// length n held in LDS (batch*n <= MC_LDS_ELEMS).  Each stage reads every
// butterfly's inputs into registers, synchronises, then writes the outputs
// back to the same LDS buffer, so one buffer suffices.  Radices 2, 4, 8 and
// the odd primes up to 31 are register butterflies; larger primes (41, 73,
// 131, ... which the reference's un-padded sizes produce, e.g. 1441 =
// 11*131) use a chunked direct DFT.  Included by each translation unit that
// runs LDS FFTs; every unit uploads its own copy of the prime table.
#pragma once
#include "mc_common.h"

// cos/sin(2 pi j / P) for the register-butterfly primes, [cos..., sin...].
#define MC_PRIMES_TABLE_LEN 320
static __constant__ float g_pcs[MC_PRIMES_TABLE_LEN];

__host__ __device__ constexpr int mc_prime_off(int P) {
    // 3,5,7,11,13,17,19,23,29,31 -> cumulative 2*P offsets
    return P == 3 ? 0 : P == 5 ? 6 : P == 7 ? 16 : P == 11 ? 30 : P == 13 ? 52 :
           P == 17 ? 78 : P == 19 ? 112 : P == 23 ? 150 : P == 29 ? 196 : P == 31 ? 254 : -1;
}

static int mc_upload_prime_table() {
    static const int primes[] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31};
    float h[MC_PRIMES_TABLE_LEN] = {0};
    for (int P : primes) {
        const int off = mc_prime_off(P);
        for (int j = 0; j < P; j++) {
            const double a = 2.0 * 3.14159265358979323846 * j / P;
            h[off + j] = (float)cos(a);
            h[off + P + j] = (float)sin(a);
        }
    }
    return hipMemcpyToSymbol(HIP_SYMBOL(g_pcs), h, sizeof(h)) == hipSuccess ? 0 : -1;
}

// y[k] = sum_t x[t] e^{-2 pi i t k / P}, symmetric-pair form.
template <int P>
static __device__ __forceinline__ void bfly_prime(cfloat* v) {
    constexpr int H = (P - 1) / 2;
    constexpr int OFF = mc_prime_off(P);
    cfloat u[H], w[H];
    const cfloat x0 = v[0];
    cfloat y0 = x0;
#pragma unroll
    for (int t = 1; t <= H; t++) {
        u[t - 1] = cadd(v[t], v[P - t]);
        w[t - 1] = csub(v[t], v[P - t]);
        y0 = cadd(y0, u[t - 1]);
    }
    v[0] = y0;
#pragma unroll
    for (int k = 1; k <= H; k++) {
        float ax = x0.x, ay = x0.y, bx = 0.f, by = 0.f;
#pragma unroll
        for (int t = 1; t <= H; t++) {
            const int idx = (t * k) % P;
            const float c = g_pcs[OFF + idx];
            const float s = g_pcs[OFF + P + idx];
            ax = fmaf(u[t - 1].x, c, ax);
            ay = fmaf(u[t - 1].y, c, ay);
            bx = fmaf(w[t - 1].x, s, bx);
            by = fmaf(w[t - 1].y, s, by);
        }
        v[k] = cmk(ax + by, ay - bx);
        v[P - k] = cmk(ax - by, ay + bx);
    }
}

static __device__ __forceinline__ void dft4(cfloat& a0, cfloat& a1, cfloat& a2, cfloat& a3) {
    const cfloat t0 = cadd(a0, a2), t1 = csub(a0, a2);
    const cfloat t2 = cadd(a1, a3), t3 = cmuli_neg(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <int R>
static __device__ __forceinline__ void bfly(cfloat* v);

template <>
__device__ __forceinline__ void bfly<2>(cfloat* v) {
    const cfloat a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void bfly<4>(cfloat* v) {
    dft4(v[0], v[1], v[2], v[3]);
}
template <>
__device__ __forceinline__ void bfly<8>(cfloat* v) {
    cfloat e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cfloat o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    const float r = 0.70710678118654752f;
    o1 = cmk((o1.x + o1.y) * r, (o1.y - o1.x) * r);
    o2 = cmuli_neg(o2);
    o3 = cmk((o3.y - o3.x) * r, (-o3.x - o3.y) * r);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}
#define MC_PRIME_BFLY(P) \
    template <> __device__ __forceinline__ void bfly<P>(cfloat* v) { bfly_prime<P>(v); }
MC_PRIME_BFLY(3)
MC_PRIME_BFLY(5)
MC_PRIME_BFLY(7)
MC_PRIME_BFLY(11)
MC_PRIME_BFLY(13)
MC_PRIME_BFLY(17)
MC_PRIME_BFLY(19)
MC_PRIME_BFLY(23)
MC_PRIME_BFLY(29)
MC_PRIME_BFLY(31)
#undef MC_PRIME_BFLY

// One Stockham stage with register butterflies.
template <int R>
static __device__ __noinline__ void fft_stage(cfloat* lds, int n, int batch, const FftStage st,
                                 const cfloat* __restrict__ tw) {
    constexpr int MAXIT = (MC_LDS_ELEMS + R * MC_NTHR - 1) / (R * MC_NTHR);
    const int m = st.m;
    const int total = batch * m;
    const int ns = st.ns;
    const cfloat* twp = tw + st.tw_off;
    cfloat v[MAXIT][R];
    int dst[MAXIT];
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        const int j = threadIdx.x + it * MC_NTHR;
        dst[it] = -1;
        if (j < total) {
            const int b = fdiv(j, m, st.rcp_m);
            const int jj = j - b * m;
            const int q = fdiv(jj, ns, st.rcp_ns);
            const int k = jj - q * ns;
            const int src = b * n + jj;
#pragma unroll
            for (int t = 0; t < R; t++) v[it][t] = lds[lp(src + t * m)];
            if (ns > 1) {
                const cfloat* w = twp + k * (R - 1);
#pragma unroll
                for (int t = 1; t < R; t++) v[it][t] = cmul(v[it][t], w[t - 1]);
            }
            bfly<R>(v[it]);
            dst[it] = b * n + q * ns * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
        if (dst[it] >= 0) {
#pragma unroll
            for (int t = 0; t < R; t++) lds[lp(dst[it] + t * ns)] = v[it][t];
        }
    }
    __syncthreads();
}

// Large prime stage: each work item computes KC outputs of one butterfly by
// direct summation over its p inputs.
#define MC_BIG_KC 8
#define MC_BIG_MAXIT 6
static __device__ __noinline__ void fft_stage_big(cfloat* lds, int n, int batch, const FftStage st,
                                     const cfloat* __restrict__ tw) {
    const int p = st.r;
    const int m = st.m;
    const int nch = st.nch;
    const int total = batch * m * nch;
    const int ns = st.ns;
    const cfloat* twp = tw + st.tw_off;
    const cfloat* roots = tw + st.root_off;
    cfloat acc[MC_BIG_MAXIT][MC_BIG_KC];
    int dst[MC_BIG_MAXIT], q0s[MC_BIG_MAXIT];
#pragma unroll
    for (int it = 0; it < MC_BIG_MAXIT; it++) {
        const int item = threadIdx.x + it * MC_NTHR;
        dst[it] = -1;
        q0s[it] = 0;
#pragma unroll
        for (int q = 0; q < MC_BIG_KC; q++) acc[it][q] = cmk(0.f, 0.f);
        if (item < total) {
            const int bf = fdiv(item, nch, st.rcp_nch);
            const int ch = item - bf * nch;
            const int b = fdiv(bf, m, st.rcp_m);
            const int jj = bf - b * m;
            const int qd = fdiv(jj, ns, st.rcp_ns);
            const int k = jj - qd * ns;
            const int q0 = ch * MC_BIG_KC;
            const int src = b * n + jj;
            int ridx[MC_BIG_KC], step[MC_BIG_KC];
#pragma unroll
            for (int q = 0; q < MC_BIG_KC; q++) {
                ridx[q] = 0;
                step[q] = (q0 + q < p) ? q0 + q : 0;
            }
            for (int t = 0; t < p; t++) {
                cfloat x = lds[lp(src + t * m)];
                if (ns > 1 && t > 0) x = cmul(x, twp[k * (p - 1) + t - 1]);
#pragma unroll
                for (int q = 0; q < MC_BIG_KC; q++) {
                    const cfloat r = roots[ridx[q]];
                    acc[it][q].x = fmaf(x.x, r.x, fmaf(-x.y, r.y, acc[it][q].x));
                    acc[it][q].y = fmaf(x.x, r.y, fmaf(x.y, r.x, acc[it][q].y));
                    int nx = ridx[q] + step[q];
                    ridx[q] = nx >= p ? nx - p : nx;
                }
            }
            dst[it] = b * n + qd * ns * p + k;
            q0s[it] = q0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MC_BIG_MAXIT; it++) {
        if (dst[it] >= 0) {
#pragma unroll
            for (int q = 0; q < MC_BIG_KC; q++) {
                const int qq = q0s[it] + q;
                if (qq < p) lds[lp(dst[it] + qq * ns)] = acc[it][q];
            }
        }
    }
    __syncthreads();
}

// Forward FFT of `batch` contiguous sequences of length P.n in LDS.
static __device__ void fft_lds_fwd(cfloat* lds, int batch, const FftPlan1& P) {
    for (int s = 0; s < P.nst; s++) {
        const FftStage st = P.st[s];
        switch (st.r) {
            case 2: fft_stage<2>(lds, P.n, batch, st, P.tw); break;
            case 3: fft_stage<3>(lds, P.n, batch, st, P.tw); break;
            case 4: fft_stage<4>(lds, P.n, batch, st, P.tw); break;
            case 5: fft_stage<5>(lds, P.n, batch, st, P.tw); break;
            case 7: fft_stage<7>(lds, P.n, batch, st, P.tw); break;
            case 8: fft_stage<8>(lds, P.n, batch, st, P.tw); break;
            case 11: fft_stage<11>(lds, P.n, batch, st, P.tw); break;
            case 13: fft_stage<13>(lds, P.n, batch, st, P.tw); break;
            case 17: fft_stage<17>(lds, P.n, batch, st, P.tw); break;
            case 19: fft_stage<19>(lds, P.n, batch, st, P.tw); break;
            case 23: fft_stage<23>(lds, P.n, batch, st, P.tw); break;
            case 29: fft_stage<29>(lds, P.n, batch, st, P.tw); break;
            case 31: fft_stage<31>(lds, P.n, batch, st, P.tw); break;
            default: fft_stage_big(lds, P.n, batch, st, P.tw); break;
        }
    }
}
