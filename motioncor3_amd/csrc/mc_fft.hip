// mc_fft.hip -- batched 2-D real FFT passes for gfx950.
//
// Replaces cufftPlan2d/cufftExecR2C/cufftExecC2R of Util/CCufft2D.cpp:23-113
// (one exec per frame there) with batched two-pass transforms: a row pass
// (real <-> half-complex along x, LDS-resident rows) and a column pass
// (complex along y, C adjacent columns per workgroup).  The staging layout
// between the passes is column-blocked so both passes move contiguous runs.
#include <cmath>
#include <cstdio>
#include <type_traits>
#include <vector>

#include "mc_fft.h"
#include "mc_fft_ct.h"
#include "mc_fft_rp.h"

namespace {

struct RowR2CArgs {
    FftPlan1 plan;
    int H, m, C, R;
    const float* in;
    long long pitch, in_bstride;
    int in_mod;            // transform b reads image b % in_mod (0: b)
    const int* win;        // (x0, y0) per b / in_mod, or null
    cfloat* T;
    long long t_bstride;
    const cfloat* post_tw;
    float scale;
    float* dc;             // |row DC| per (transform, row) of the output, or null
};

struct RowC2RArgs {
    FftPlan1 plan;
    int H, m, C, R;
    const cfloat* T;
    long long t_bstride;
    float* out;
    long long pitch, out_bstride;
    const cfloat* post_tw;
    int clamp;
};

struct ColFwdArgs {
    FftPlan1 plan;
    int H, ncols, C;
    const cfloat* T;
    long long t_bstride;
    cfloat* S;
    long long s_bstride;
    cfloat* X;
    int xcx, xcy;
    long long x_bstride;
    const float* s_g;      // store factors of a reduced S format (per transform)
    const float* x_g;      // store factors of a reduced X format
};

struct ColInvArgs {
    FftPlan1 plan;
    int H, ncols, C;
    const cfloat* S;
    long long s_bstride;
    int src_cx, src_cy;
    const cfloat* phA;
    const cfloat* phB;
    cfloat* T;
    long long t_bstride;
    const float* s_inv;    // load factors of a reduced S format (per transform)
};

// Util/GFourierResize2D.cu:15-35 row map: output row yo of a spectrum with
// `outy` rows taken from a spectrum with `iny` rows, -1 if unmapped.
__device__ __forceinline__ int resize_row(int yo, int outy, int iny) {
    int y = yo;
    if (y > outy / 2) {
        y -= outy;
        if (y <= -iny / 2) return -1;
        return y + iny;
    }
    return (y > iny / 2) ? -1 : y;
}

#define MC_IO_UNROLL 8

__global__ __launch_bounds__(MC_NTHR) void k_rows_r2c(RowR2CArgs a) {
    extern __shared__ cfloat lds[];
    const int nrb = (a.H + a.R - 1) / a.R;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int b = t / nrb;
    const int row0 = (t - b * nrb) * a.R;
    const int nrows = min(a.R, a.H - row0);
    const int img = a.in_mod > 0 ? b % a.in_mod : b;
    const int wi = a.in_mod > 0 ? b / a.in_mod : b;
    int x0 = 0, y0 = 0;
    if (a.win) {
        x0 = a.win[2 * wi];
        y0 = a.win[2 * wi + 1];
    }
    const float* base = a.in + (long long)img * a.in_bstride + (long long)(y0 + row0) * a.pitch + x0;
    const float rm = 1.0f / a.m;
    const int total = nrows * a.m;
    const bool vec = ((x0 | (int)(a.pitch & 1) | (int)(a.in_bstride & 1)) & 1) == 0 &&
                     (((uintptr_t)a.in) & 7) == 0;
    for (int e0 = threadIdx.x; e0 < total; e0 += MC_NTHR * MC_IO_UNROLL) {
        float2 v[MC_IO_UNROLL];
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * MC_NTHR;
            if (e < total) {
                const int r = fdiv(e, a.m, rm);
                const float* src = base + (long long)r * a.pitch + 2 * (e - r * a.m);
                v[u] = vec ? *reinterpret_cast<const float2*>(src) : make_float2(src[0], src[1]);
            }
        }
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * MC_NTHR;
            if (e < total) lds[lp(e)] = cmk(v[u].x, v[u].y);
        }
    }
    __syncthreads();
    fft_lds_fwd(lds, nrows, a.plan);
    const int kc = a.m + 1;
    const int nb = (kc + a.C - 1) / a.C;
    cfloat* tb = a.T + (long long)b * a.t_bstride + (long long)row0 * a.C;
    const int tot = nb * nrows * a.C;
    const int lc = __ffs(a.C) - 1;
    const float rnr = 1.0f / nrows;
    const long long kbs = (long long)a.H * a.C;
#pragma unroll 4
    for (int e = threadIdx.x; e < tot; e += MC_NTHR) {
        const int c = e & (a.C - 1);
        const int rc = e >> lc;
        const int kb = fdiv(rc, nrows, rnr);
        const int r = rc - kb * nrows;
        const int k = kb * a.C + c;
        if (k > a.m) continue;
        const cfloat zk = lds[lp(r * a.m + (k == a.m ? 0 : k))];
        const cfloat zm = lds[lp(r * a.m + (k == 0 ? 0 : a.m - k))];
        const cfloat E = cmk(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        const cfloat D = cmk(zk.x - zm.x, zk.y + zm.y);
        const cfloat O = cmk(0.5f * D.y, -0.5f * D.x);
        const cfloat X = cadd(E, cmul(a.post_tw[k], O));
        tb[kb * kbs + r * a.C + c] = cscale(X, a.scale);
    }
}

__global__ __launch_bounds__(MC_NTHR) void k_rows_c2r(RowC2RArgs a) {
    extern __shared__ cfloat lds[];
    const int nrb = (a.H + a.R - 1) / a.R;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int b = t / nrb;
    const int row0 = (t - b * nrb) * a.R;
    const int nrows = min(a.R, a.H - row0);
    const int kc = a.m + 1;
    const int nb = (kc + a.C - 1) / a.C;
    const cfloat* tb = a.T + (long long)b * a.t_bstride + (long long)row0 * a.C;
    const int tot = nb * nrows * a.C;
    const int lc = __ffs(a.C) - 1;
    const float rnr = 1.0f / nrows;
    const long long kbs = (long long)a.H * a.C;
    for (int e0 = threadIdx.x; e0 < tot; e0 += MC_NTHR * MC_IO_UNROLL) {
        cfloat v[MC_IO_UNROLL];
        int dst[MC_IO_UNROLL];
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * MC_NTHR;
            dst[u] = -1;
            if (e < tot) {
                const int c = e & (a.C - 1);
                const int rc = e >> lc;
                const int kb = fdiv(rc, nrows, rnr);
                const int r = rc - kb * nrows;
                const int k = kb * a.C + c;
                if (k <= a.m) {
                    v[u] = tb[kb * kbs + r * a.C + c];
                    // X_m (Nyquist) goes to the slots after the rows (nyq[r])
                    dst[u] = (k < a.m) ? r * a.m + k : nrows * a.m + r;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++)
            if (dst[u] >= 0) lds[lp(dst[u])] = v[u];
    }
    __syncthreads();
    // Hermitian pre-processing into Z'[k] = (X_k + X*_{m-k}) +
    // i e^{+2 pi i k/W} (X_k - X*_{m-k}); imag parts of X_0, X_m dropped.
    const int half = a.m / 2 + 1;
    const float rhalf = 1.0f / half;
    for (int e = threadIdx.x; e < nrows * half; e += MC_NTHR) {
        const int r = fdiv(e, half, rhalf);
        const int k = e - r * half;
        const int row = r * a.m;
        if (k == 0) {
            const float x0 = lds[lp(row)].x, xm = lds[lp(nrows * a.m + r)].x;
            lds[lp(row)] = cconj(cmk(x0 + xm, x0 - xm));
            continue;
        }
        const int k2 = a.m - k;
        if (k2 < k) continue;
        const cfloat xa = lds[lp(row + k)], xb = lds[lp(row + k2)];
        const cfloat twk = cconj(a.post_tw[k]);
        const cfloat s1 = cmk(xa.x + xb.x, xa.y - xb.y);
        const cfloat d1 = cmk(xa.x - xb.x, xa.y + xb.y);
        const cfloat t1 = cmul(twk, d1);
        const cfloat zk = cmk(s1.x - t1.y, s1.y + t1.x);
        lds[lp(row + k)] = cconj(zk);
        if (k2 != k) {
            const cfloat twk2 = cconj(a.post_tw[k2]);
            const cfloat s2 = cmk(xb.x + xa.x, xb.y - xa.y);
            const cfloat d2 = cmk(xb.x - xa.x, xb.y + xa.y);
            const cfloat t2 = cmul(twk2, d2);
            lds[lp(row + k2)] = cconj(cmk(s2.x - t2.y, s2.y + t2.x));
        }
    }
    __syncthreads();
    fft_lds_fwd(lds, nrows, a.plan);
    float* ob = a.out + (long long)b * a.out_bstride + (long long)row0 * a.pitch;
    const float rm = 1.0f / a.m;
    const bool vec = ((a.pitch | a.out_bstride) & 1) == 0 && (((uintptr_t)a.out) & 7) == 0;
    const int total = nrows * a.m;
#pragma unroll 4
    for (int e = threadIdx.x; e < total; e += MC_NTHR) {
        const int r = fdiv(e, a.m, rm);
        const int j = e - r * a.m;
        const cfloat z = lds[lp(e)];
        float v0 = z.x, v1 = -z.y;
        if (a.clamp) {
            v0 = fmaxf(v0, 0.f);
            v1 = fmaxf(v1, 0.f);
        }
        float* dst = ob + (long long)r * a.pitch + 2 * j;
        if (vec) {
            *reinterpret_cast<float2*>(dst) = make_float2(v0, v1);
        } else {
            dst[0] = v0;
            dst[1] = v1;
        }
    }
}

__global__ __launch_bounds__(MC_NTHR) void k_cols_fwd(ColFwdArgs a) {
    extern __shared__ cfloat lds[];
    const int ncb = (a.ncols + a.C - 1) / a.C;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int b = t / ncb;
    const int kb = t - b * ncb;
    const int c0 = kb * a.C;
    const int nc = min(a.C, a.ncols - c0);
    const cfloat* src = a.T + (long long)b * a.t_bstride + (long long)kb * a.H * a.C;
    const int lc = __ffs(a.C) - 1;
    const int total = a.H * a.C;
    for (int e0 = threadIdx.x; e0 < total; e0 += MC_NTHR * MC_IO_UNROLL) {
        cfloat v[MC_IO_UNROLL];
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * MC_NTHR;
            if (e < total && (e & (a.C - 1)) < nc) v[u] = src[e];
        }
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * MC_NTHR;
            const int c = e & (a.C - 1);
            if (e < total && c < nc) lds[lp(c * a.H + (e >> lc))] = v[u];
        }
    }
    __syncthreads();
    fft_lds_fwd(lds, nc, a.plan);
    if (a.S) {
        cfloat* dst = a.S + (long long)b * a.s_bstride + (long long)c0 * a.H;
#pragma unroll 4
        for (int e = threadIdx.x; e < nc * a.H; e += MC_NTHR) dst[e] = lds[lp(e)];
    }
    if (a.X) {
        cfloat* xd = a.X + (long long)b * a.x_bstride;
        const float rx = 1.0f / a.xcy;
#pragma unroll 4
        for (int e = threadIdx.x; e < nc * a.xcy; e += MC_NTHR) {
            const int c = fdiv(e, a.xcy, rx);
            const int yo = e - c * a.xcy;
            const int kx = c0 + c;
            if (kx >= a.xcx) continue;
            const int yi = resize_row(yo, a.xcy, a.H);
            if (yi < 0) continue;
            xd[(long long)kx * a.xcy + yo] = lds[lp(c * a.H + yi)];
        }
    }
}

__global__ __launch_bounds__(MC_NTHR) void k_cols_inv(ColInvArgs a) {
    extern __shared__ cfloat lds[];
    const int ncb = (a.ncols + a.C - 1) / a.C;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int b = t / ncb;
    const int kb = t - b * ncb;
    const int c0 = kb * a.C;
    const int nc = min(a.C, a.ncols - c0);
    const bool same = (a.src_cx == a.ncols && a.src_cy == a.H);
    const cfloat* sb = a.S + (long long)b * a.s_bstride;
    const float rH = 1.0f / a.H;
    const int total = nc * a.H;
    for (int e0 = threadIdx.x; e0 < total; e0 += MC_NTHR * MC_IO_UNROLL) {
        cfloat v[MC_IO_UNROLL];
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * MC_NTHR;
            if (e < total) {
                const int c = fdiv(e, a.H, rH);
                const int y = e - c * a.H;
                const int kx = c0 + c;
                if (same) {
                    v[u] = sb[(long long)kx * a.H + y];
                } else {
                    const int yi = resize_row(y, a.H, a.src_cy);
                    v[u] = (kx < a.src_cx && yi >= 0) ? sb[(long long)kx * a.src_cy + yi]
                                                      : cmk(0.f, 0.f);
                }
                if (a.phA)
                    v[u] = cmul(v[u], cmul(a.phA[(long long)b * a.ncols + kx],
                                           a.phB[(long long)b * a.H + y]));
            }
        }
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * MC_NTHR;
            if (e < total) lds[lp(e)] = cconj(v[u]);
        }
    }
    __syncthreads();
    fft_lds_fwd(lds, nc, a.plan);
    cfloat* dst = a.T + (long long)b * a.t_bstride + (long long)kb * a.H * a.C;
    const int lc = __ffs(a.C) - 1;
#pragma unroll 4
    for (int e = threadIdx.x; e < a.H * a.C; e += MC_NTHR) {
        const int c = e & (a.C - 1);
        if (c < nc) dst[e] = cconj(lds[lp(c * a.H + (e >> lc))]);
    }
}

// ---------------------------------------------------------------------------
// Specialised passes: the row length M (or column length H), rows per
// workgroup RR and column block C are template constants; each transform is
// owned by MC_NTHR/RR (or MC_NTHR/C) threads.
template <int M, int RR, int C>
__global__ __launch_bounds__(MC_NTHR) void k_rows_r2c_ct(RowR2CArgs a) {
    constexpr int TPT = MC_NTHR / RR;
    constexpr int LIT = (M + TPT - 1) / TPT;
    constexpr int NB = (M + 1 + C - 1) / C;
    constexpr int TOT = NB * RR * C;
    extern __shared__ cfloat lds[];
    const int nrb = (a.H + RR - 1) / RR;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int b = t / nrb;
    const int row0 = (t - b * nrb) * RR;
    const int nrows = min(RR, a.H - row0);
    const int g = threadIdx.x / TPT, tid = threadIdx.x % TPT;
    const int img = a.in_mod > 0 ? b % a.in_mod : b;
    const int wi = a.in_mod > 0 ? b / a.in_mod : b;
    int x0 = 0, y0 = 0;
    if (a.win) {
        x0 = a.win[2 * wi];
        y0 = a.win[2 * wi + 1];
    }
    const bool vec = ((x0 | (int)(a.pitch & 1) | (int)(a.in_bstride & 1)) & 1) == 0 &&
                     (((uintptr_t)a.in) & 7) == 0;
    const bool live = g < nrows;
    const float* src = a.in + (long long)img * a.in_bstride +
                       (long long)(y0 + row0 + (live ? g : 0)) * a.pitch + x0;
    float2 v[LIT];
#pragma unroll
    for (int it = 0; it < LIT; it++) {
        const int j = tid + it * TPT;
        v[it] = make_float2(0.f, 0.f);
        if (live && (M % TPT == 0 || j < M))
            v[it] = vec ? *reinterpret_cast<const float2*>(src + 2 * j)
                        : make_float2(src[2 * j], src[2 * j + 1]);
    }
#pragma unroll
    for (int it = 0; it < LIT; it++) {
        const int j = tid + it * TPT;
        if (M % TPT == 0 || j < M) lds[lp(g * M + j)] = cmk(v[it].x, v[it].y);
    }
    __syncthreads();
    ct_fft<M, TPT>(lds, g * M, tid, a.plan.tw);
    cfloat* tb = a.T + (long long)b * a.t_bstride + (long long)row0 * C;
    const long long kbs = (long long)a.H * C;
#pragma unroll 4
    for (int e = threadIdx.x; e < TOT; e += MC_NTHR) {
        const int c = e % C, r = (e / C) % RR, kb = e / (C * RR);
        const int k = kb * C + c;
        if (k > M || r >= nrows) continue;
        const cfloat zk = lds[lp(r * M + (k == M ? 0 : k))];
        const cfloat zm = lds[lp(r * M + (k == 0 ? 0 : M - k))];
        const cfloat E = cmk(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        const cfloat D = cmk(zk.x - zm.x, zk.y + zm.y);
        const cfloat O = cmk(0.5f * D.y, -0.5f * D.x);
        const cfloat X = cadd(E, cmul(a.post_tw[k], O));
        tb[kb * kbs + r * C + c] = cscale(X, a.scale);
    }
}

template <int M, int RR, int C>
__global__ __launch_bounds__(MC_NTHR) void k_rows_c2r_ct(RowC2RArgs a) {
    constexpr int TPT = MC_NTHR / RR;
    constexpr int NB = (M + 1 + C - 1) / C;
    constexpr int TOT = NB * RR * C;
    constexpr int HALF = M / 2 + 1;
    constexpr int SIT = (M + TPT - 1) / TPT;
    extern __shared__ cfloat lds[];
    const int nrb = (a.H + RR - 1) / RR;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int b = t / nrb;
    const int row0 = (t - b * nrb) * RR;
    const int nrows = min(RR, a.H - row0);
    const int g = threadIdx.x / TPT, tid = threadIdx.x % TPT;
    const cfloat* tb = a.T + (long long)b * a.t_bstride + (long long)row0 * C;
    const long long kbs = (long long)a.H * C;
    for (int e0 = threadIdx.x; e0 < TOT; e0 += MC_NTHR * MC_IO_UNROLL) {
        cfloat v[MC_IO_UNROLL];
        int dst[MC_IO_UNROLL];
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * MC_NTHR;
            dst[u] = -1;
            if (e < TOT) {
                const int c = e % C, r = (e / C) % RR, kb = e / (C * RR);
                const int k = kb * C + c;
                if (k <= M && r < nrows) {
                    v[u] = tb[kb * kbs + r * C + c];
                    dst[u] = (k < M) ? r * M + k : RR * M + r;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++)
            if (dst[u] >= 0) lds[lp(dst[u])] = v[u];
    }
    __syncthreads();
    const int row = g * M;
    for (int k = tid; k < HALF; k += TPT) {
        if (k == 0) {
            const float x0 = lds[lp(row)].x, xm = lds[lp(RR * M + g)].x;
            lds[lp(row)] = cconj(cmk(x0 + xm, x0 - xm));
            continue;
        }
        const int k2 = M - k;
        if (k2 < k) continue;
        const cfloat xa = lds[lp(row + k)], xb = lds[lp(row + k2)];
        const cfloat t1 = cmul(cconj(a.post_tw[k]), cmk(xa.x - xb.x, xa.y + xb.y));
        lds[lp(row + k)] = cconj(cmk(xa.x + xb.x - t1.y, xa.y - xb.y + t1.x));
        if (k2 != k) {
            const cfloat t2 = cmul(cconj(a.post_tw[k2]), cmk(xb.x - xa.x, xb.y + xa.y));
            lds[lp(row + k2)] = cconj(cmk(xb.x + xa.x - t2.y, xb.y - xa.y + t2.x));
        }
    }
    __syncthreads();
    ct_fft<M, TPT>(lds, g * M, tid, a.plan.tw);
    if (g >= nrows) return;
    float* ob = a.out + (long long)b * a.out_bstride + (long long)(row0 + g) * a.pitch;
    // one straight-line copy per store form, the clamp as a select (as in
    // k_rows_c2r_rp)
    const bool clamp = a.clamp != 0;
    auto store = [&](auto vec) {
#pragma unroll
        for (int it = 0; it < SIT; it++) {
            const int j = tid + it * TPT;
            if (M % TPT != 0 && j >= M) continue;
            const cfloat z = lds[lp(row + j)];
            float v0 = z.x, v1 = -z.y;
            v0 = clamp ? fmaxf(v0, 0.f) : v0;
            v1 = clamp ? fmaxf(v1, 0.f) : v1;
            if constexpr (decltype(vec)::value) {
                *reinterpret_cast<float2*>(ob + 2 * j) = make_float2(v0, v1);
            } else {
                ob[2 * j] = v0;
                ob[2 * j + 1] = v1;
            }
        }
    };
    if (((a.pitch | a.out_bstride) & 1) == 0 && (((uintptr_t)a.out) & 7) == 0) store(std::true_type{});
    else store(std::false_type{});
}

template <int H, int C, int NT>
__global__ __launch_bounds__(NT) void k_cols_fwd_ct(ColFwdArgs a) {
    constexpr int TPT = NT / C;
    extern __shared__ cfloat lds[];
    const int ncb = (a.ncols + C - 1) / C;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int b = t / ncb;
    const int kb = t - b * ncb;
    const int c0 = kb * C;
    const int nc = min(C, a.ncols - c0);
    const int g = threadIdx.x / TPT, tid = threadIdx.x % TPT;
    const cfloat* src = a.T + (long long)b * a.t_bstride + (long long)kb * H * C;
    for (int e0 = threadIdx.x; e0 < H * C; e0 += NT * MC_IO_UNROLL) {
        cfloat v[MC_IO_UNROLL];
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * NT;
            if (e < H * C && (e % C) < nc) v[u] = src[e];
        }
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * NT;
            if (e < H * C && (e % C) < nc) lds[lp((e % C) * H + e / C)] = v[u];
        }
    }
    __syncthreads();
    static_assert(ct_big_prime(H) == 0, "column lengths with a prime factor above 31 use the generic kernels");
    ct_fft<H, TPT>(lds, g * H, tid, a.plan.tw);
    if (a.S) {
        cfloat* dst = a.S + (long long)b * a.s_bstride + (long long)c0 * H;
#pragma unroll 4
        for (int e = threadIdx.x; e < nc * H; e += NT) dst[e] = lds[lp(e)];
    }
    if (a.X) {
        cfloat* xd = a.X + (long long)b * a.x_bstride;
        const float rx = 1.0f / a.xcy;
#pragma unroll 4
        for (int e = threadIdx.x; e < nc * a.xcy; e += NT) {
            const int c = fdiv(e, a.xcy, rx);
            const int yo = e - c * a.xcy;
            const int kx = c0 + c;
            if (kx >= a.xcx) continue;
            const int yi = resize_row(yo, a.xcy, H);
            if (yi < 0) continue;
            xd[(long long)kx * a.xcy + yo] = lds[lp(c * H + yi)];
        }
    }
}

template <int H, int C, int NT>
__global__ __launch_bounds__(NT) void k_cols_inv_ct(ColInvArgs a) {
    constexpr int TPT = NT / C;
    extern __shared__ cfloat lds[];
    const int ncb = (a.ncols + C - 1) / C;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int b = t / ncb;
    const int kb = t - b * ncb;
    const int c0 = kb * C;
    const int nc = min(C, a.ncols - c0);
    const int g = threadIdx.x / TPT, tid = threadIdx.x % TPT;
    const bool same = (a.src_cx == a.ncols && a.src_cy == H);
    const cfloat* sb = a.S + (long long)b * a.s_bstride;
    const int total = nc * H;
    for (int e0 = threadIdx.x; e0 < total; e0 += NT * MC_IO_UNROLL) {
        cfloat v[MC_IO_UNROLL];
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * NT;
            if (e < total) {
                const int c = e / H, y = e % H;
                const int kx = c0 + c;
                if (same) {
                    v[u] = sb[(long long)kx * H + y];
                } else {
                    const int yi = resize_row(y, H, a.src_cy);
                    v[u] = (kx < a.src_cx && yi >= 0) ? sb[(long long)kx * a.src_cy + yi]
                                                      : cmk(0.f, 0.f);
                }
                if (a.phA)
                    v[u] = cmul(v[u], cmul(a.phA[(long long)b * a.ncols + kx],
                                           a.phB[(long long)b * H + y]));
            }
        }
#pragma unroll
        for (int u = 0; u < MC_IO_UNROLL; u++) {
            const int e = e0 + u * NT;
            if (e < total) lds[lp(e)] = cconj(v[u]);
        }
    }
    __syncthreads();
    static_assert(ct_big_prime(H) == 0, "column lengths with a prime factor above 31 use the generic kernels");
    ct_fft<H, TPT>(lds, g * H, tid, a.plan.tw);
    cfloat* dst = a.T + (long long)b * a.t_bstride + (long long)kb * H * C;
#pragma unroll 4
    for (int e = threadIdx.x; e < H * C; e += NT) {
        const int c = e % C;
        if (c < nc) dst[e] = cconj(lds[lp(c * H + e / C)]);
    }
}

// (row length M, rows per workgroup, column block) of the reference
// geometries: frames 5760x4092, 4096^2, 1024^2; xcf 2882x2048, 2048^2;
// patches 576x408; the 128^2 feature map.
#ifndef MC_ROW_SPECS
#define MC_ROW_SPECS(X) \
    X(2880, 2, MC_TILE_CT) X(2048, 2, MC_TILE_CT) X(512, 2, 8) X(1441, 4, MC_TILE_CT) X(1024, 2, 8) X(288, 2, 16) X(64, 2, 16)
#endif
#ifndef MC_COL_SPECS
#define MC_COL_SPECS(X) \
    X(4092, 2, 384) X(4096, 2, 512) X(1024, 8, 512) X(2048, 4, 512) X(408, 16, 512) X(128, 16, 512)
#endif

// ---------------------------------------------------------------------------
// Column passes with a register-pass schedule (mc_fft_rp.h), C columns per
// workgroup held unpadded in LDS (C*H*8 bytes).  Forward: staging T -> S
// (+ the xcf crop X); inverse: S (phase factors, conjugated) -> T.  Same
// results as the stage-by-stage kernels up to FP32 rounding order.

// Output row of a cropped (xcy-row) spectrum that input row y maps to, or -1
// (the inverse of resize_row for xcy < H).
__device__ __forceinline__ int crop_row(int y, int xcy, int H) {
    if (y <= xcy / 2) return y;
    const int d = H - xcy;
    return (y > xcy / 2 + d) ? y - d : -1;
}

template <int H, int CW, int NT, int FS, int FX, int SC>
__global__ __launch_bounds__(NT) void k_cols_fwd_rp(ColFwdArgs a, const cfloat* __restrict__ ptw) {
    constexpr int C = CW;
    constexpr int NP = rp_npass(H, SC);
    constexpr int R0 = rp_radix(H, 0, SC), M0 = H / R0;
    constexpr int RL = rp_radix(H, NP - 1, SC), ML = H / RL;
    constexpr int IT0 = (C * M0 + NT - 1) / NT, ITL = (C * ML + NT - 1) / NT;
    extern __shared__ cfloat lds[];
    const int CT = a.C;  // chunk width of the staging layout (a multiple of CW)
    const int ncb = (a.ncols + C - 1) / C;
    const int tt = xcd_tile(blockIdx.x, gridDim.x);
    const int b = tt / ncb, cb = tt - b * ncb, c0 = cb * C;
    const int kb = c0 / CT, coff = c0 - kb * CT;
    const int nc = min(C, a.ncols - c0);
    const cfloat* src = a.T + (long long)b * a.t_bstride + (long long)kb * H * CT + coff;
    // pass 0: staging [y][c] -> registers -> LDS (items: column fastest)
#pragma unroll
    for (int it = 0; it < IT0; it++) {
        const int i = threadIdx.x + it * NT;
        if ((C * M0) % NT == 0 || i < C * M0) {
            const int c = i % C, j = i / C;
            cfloat v[R0];
            const bool ok = c < nc;
#pragma unroll
            for (int t = 0; t < R0; t++) v[t] = ok ? src[(j + t * M0) * CT + c] : cmk(0.f, 0.f);
            RegDft<R0>::run(v);
            cfloat* d = lds + c * rp_cs(H, C, SC);
#pragma unroll
            for (int t = 0; t < R0; t++) d[rp_ix<H, SC>(j * R0 + t)] = v[t];
        }
    }
    __syncthreads();
    RpMid<H, 1, C, NT, SC>::run(lds, ptw);
    // last pass: LDS -> registers -> S (column-major) and the crop X; the
    // store factors of a reduced format are loaded before any store (a load
    // behind a store could only be waited for together with it)
    const float gs = FS == MC_SPEC_F32 ? 1.f : a.s_g[b];
    const float gx = FX == MC_SPEC_F32 ? 1.f : a.x_g[b];
#pragma unroll
    for (int it = 0; it < ITL; it++) {
        const int i = threadIdx.x + it * NT;
        if ((C * ML) % NT == 0 || i < C * ML) {
            const int c = i / ML, k = i - c * ML;
            if (c < nc) {
                cfloat v[RL];
                const cfloat* sl = lds + c * rp_cs(H, C, SC);
#pragma unroll
                for (int t = 0; t < RL; t++) v[t] = sl[rp_ix<H, SC>(k + t * ML)];
                rp_twiddle<H, NP - 1, SC>(v, k, ptw);
                RegDft<RL>::run(v);
                const int kx = c0 + c;
                if (a.S) {
                    const long long d = (long long)b * a.s_bstride + (long long)kx * H + k;
#pragma unroll
                    for (int t = 0; t < RL; t++) Spec<FS>::st(a.S, d + t * ML, v[t], gs);
                }
                if (a.X && kx < a.xcx) {
                    const long long xd = (long long)b * a.x_bstride + (long long)kx * a.xcy;
#pragma unroll
                    for (int t = 0; t < RL; t++) {
                        const int yo = crop_row(k + t * ML, a.xcy, H);
                        if (yo >= 0) Spec<FX>::st(a.X, xd + yo, v[t], gx);
                    }
                }
            }
        }
    }
}

template <int H, int CW, int NT, int FS, int SC>
__global__ __launch_bounds__(NT) void k_cols_inv_rp(ColInvArgs a, const cfloat* __restrict__ ptw) {
    constexpr int C = CW;
    constexpr int NP = rp_npass(H, SC);
    constexpr int R0 = rp_radix(H, 0, SC), M0 = H / R0;
    constexpr int RL = rp_radix(H, NP - 1, SC), ML = H / RL;
    constexpr int IT0 = (C * M0 + NT - 1) / NT, ITL = (C * ML + NT - 1) / NT;
    extern __shared__ cfloat lds[];
    const int CT = a.C;  // chunk width of the staging layout (a multiple of CW)
    const int ncb = (a.ncols + C - 1) / C;
    const int tt = xcd_tile(blockIdx.x, gridDim.x);
    const int b = tt / ncb, cb = tt - b * ncb, c0 = cb * C;
    const int kb = c0 / CT, coff = c0 - kb * CT;
    const int nc = min(C, a.ncols - c0);
    // pass 0: spectrum columns (phase factors, conjugated) -> LDS (items: row fastest)
#pragma unroll
    for (int it = 0; it < IT0; it++) {
        const int i = threadIdx.x + it * NT;
        if ((C * M0) % NT == 0 || i < C * M0) {
            const int c = i / M0, j = i - c * M0;
            const int kx = c0 + c;
            cfloat v[R0];
            if (c < nc) {
                const float si = FS == MC_SPEC_F32 ? 1.f : a.s_inv[b];
                const long long sbo = (long long)b * a.s_bstride;
                if (a.src_cx == a.ncols && a.src_cy == H) {
                    const long long sc = sbo + (long long)kx * H + j;
#pragma unroll
                    for (int t = 0; t < R0; t++) v[t] = Spec<FS>::ld(a.S, sc + t * M0, si);
                } else {
                    // Fourier crop or pad from a src_cx x src_cy spectrum
#pragma unroll
                    for (int t = 0; t < R0; t++) {
                        const int yi = resize_row(j + t * M0, H, a.src_cy);
                        v[t] = (kx < a.src_cx && yi >= 0)
                                   ? Spec<FS>::ld(a.S, sbo + (long long)kx * a.src_cy + yi, si)
                                   : cmk(0.f, 0.f);
                    }
                }
                if (a.phA) {
                    const cfloat pa = a.phA[(long long)b * a.ncols + kx];
                    const cfloat* pb = a.phB + (long long)b * H + j;
#pragma unroll
                    for (int t = 0; t < R0; t++) v[t] = cmul(v[t], cmul(pa, pb[t * M0]));
                }
#pragma unroll
                for (int t = 0; t < R0; t++) v[t] = cconj(v[t]);
            } else {
#pragma unroll
                for (int t = 0; t < R0; t++) v[t] = cmk(0.f, 0.f);
            }
            RegDft<R0>::run(v);
            cfloat* d = lds + c * rp_cs(H, C, SC);
#pragma unroll
            for (int t = 0; t < R0; t++) d[rp_ix<H, SC>(j * R0 + t)] = v[t];
        }
    }
    __syncthreads();
    RpMid<H, 1, C, NT, SC>::run(lds, ptw);
    // last pass: LDS -> registers -> staging [y][c] (items: column fastest)
    cfloat* dst = a.T + (long long)b * a.t_bstride + (long long)kb * H * CT + coff;
#pragma unroll
    for (int it = 0; it < ITL; it++) {
        const int i = threadIdx.x + it * NT;
        if ((C * ML) % NT == 0 || i < C * ML) {
            const int c = i % C, k = i / C;
            if (c < nc) {
                cfloat v[RL];
                const cfloat* sl = lds + c * rp_cs(H, C, SC);
#pragma unroll
                for (int t = 0; t < RL; t++) v[t] = sl[rp_ix<H, SC>(k + t * ML)];
                rp_twiddle<H, NP - 1, SC>(v, k, ptw);
                RegDft<RL>::run(v);
#pragma unroll
                for (int t = 0; t < RL; t++) dst[(k + t * ML) * CT + c] = cconj(v[t]);
            }
        }
    }
}

#ifndef MC_RP_COL_SPECS
// (height, columns per workgroup, threads, radix schedule).  The first entry
// of a height is its default tile; later entries are variants that
// mc_autotune may choose (schedule 1: another radix order, rp_radix).
#define MC_RP_COL_SPECS(X) X(4092, 2, 512, 0) X(4096, 2, 256, 0) X(2048, 4, 512, 0) X(408, 16, 384, 0) \
    X(8184, 1, 264, 0) X(4092, 2, 264, 0) X(4092, 2, 384, 0) X(4092, 2, 448, 0) X(4092, 1, 192, 0) \
    X(4092, 1, 256, 0) X(4092, 4, 528, 0) X(4092, 4, 1024, 0) X(4096, 2, 384, 0) X(4096, 2, 512, 0) \
    X(4096, 4, 512, 0) X(2048, 2, 256, 0) X(408, 8, 384, 0) X(4092, 2, 512, 1) X(408, 16, 384, 1)
#endif

// ---------------------------------------------------------------------------
// Row passes with a register-pass schedule for the complex row length M,
// RR rows per workgroup held unpadded in LDS.  R2C: real rows -> registers
// (pass 0) -> LDS passes -> post-processing -> staging T.  C2R: staging T ->
// LDS -> pre-processing -> LDS passes -> last pass -> real rows.

// The 288-point patch rows (16 rows, 37 KB of LDS per workgroup) are
// compiled for six waves per SIMD (80 VGPRs, a few spilled): 1.07 ms per
// launch at eight waves (44 VGPRs spilled), 0.80 at seven, 0.70 at six.
#ifndef MC_R2C288_WPE
#define MC_R2C288_WPE 6
#endif
template <int M, int RR, int C, int NT, int SC>
__global__ __launch_bounds__(NT, (M == 288 ? MC_R2C288_WPE : 1)) void k_rows_r2c_rp(RowR2CArgs a, const cfloat* __restrict__ ptw) {
    constexpr int NP = rp_npass(M, SC), R0 = rp_radix(M, 0, SC), M0 = M / R0;
    constexpr int IT0 = (RR * M0 + NT - 1) / NT;
    constexpr int NB = (M + 1 + C - 1) / C;
    constexpr int TOT = NB * RR * C;
    extern __shared__ cfloat lds[];
    const int nrb = (a.H + RR - 1) / RR;
    const int tt = xcd_tile(blockIdx.x, gridDim.x);
    const int b = tt / nrb;
    const int row0 = (tt - b * nrb) * RR;
    const int nrows = min(RR, a.H - row0);
    const int img = a.in_mod > 0 ? b % a.in_mod : b;
    const int wi = a.in_mod > 0 ? b / a.in_mod : b;
    int x0 = 0, y0 = 0;
    if (a.win) {
        x0 = a.win[2 * wi];
        y0 = a.win[2 * wi + 1];
    }
    const bool vec = ((x0 | (int)(a.pitch & 1) | (int)(a.in_bstride & 1)) & 1) == 0 &&
                     (((uintptr_t)a.in) & 7) == 0;
    const float* base = a.in + (long long)img * a.in_bstride + (long long)(y0 + row0) * a.pitch + x0;
#pragma unroll
    for (int it = 0; it < IT0; it++) {
        const int i = threadIdx.x + it * NT;
        if ((RR * M0) % NT == 0 || i < RR * M0) {
            const int r = i / M0, j = i - r * M0;
            cfloat v[R0];
            if (r < nrows) {
                const float* src = base + (long long)r * a.pitch + 2 * j;
#pragma unroll
                for (int t = 0; t < R0; t++) {
                    const float2 f = vec ? *reinterpret_cast<const float2*>(src + 2 * t * M0)
                                         : make_float2(src[2 * t * M0], src[2 * t * M0 + 1]);
                    v[t] = cmk(f.x, f.y);
                }
            } else {
#pragma unroll
                for (int t = 0; t < R0; t++) v[t] = cmk(0.f, 0.f);
            }
            RegDft<R0>::run(v);
            cfloat* d = lds + r * rp_cs(M, RR, SC);
#pragma unroll
            for (int t = 0; t < R0; t++) d[rp_ix<M, SC>(j * R0 + t)] = v[t];
        }
    }
    __syncthreads();
    RpRange<M, 1, NP, RR, NT, SC>::run(lds, ptw);
    cfloat* tb = a.T + (long long)b * a.t_bstride + (long long)row0 * C;
    const long long kbs = (long long)a.H * C;
    // All post-twiddles of the thread's items are loaded before its first
    // store: a load issued after a (conditional) store can only be waited
    // for together with it (one memory counter), which would hold every
    // store to a full round trip.
    constexpr int ITP = (TOT + NT - 1) / NT;
    cfloat twa[ITP];
#pragma unroll
    for (int u = 0; u < ITP; u++) {
        const int e = threadIdx.x + u * NT;
        twa[u] = a.post_tw[min((e / (C * RR)) * C + e % C, M)];
    }
#pragma unroll
    for (int u = 0; u < ITP; u++) {
        const int e = threadIdx.x + u * NT;
        if (TOT % NT != 0 && e >= TOT) break;
        const int c = e % C, r = (e / C) % RR, kb = e / (C * RR);
        const int k = kb * C + c;
        const cfloat tw = twa[u];
        if (k > M || r >= nrows) continue;
        const cfloat zk = lds[r * rp_cs(M, RR, SC) + rp_ix<M, SC>(k == M ? 0 : k)];
        const cfloat zm = lds[r * rp_cs(M, RR, SC) + rp_ix<M, SC>(k == 0 ? 0 : M - k)];
        const cfloat E = cmk(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        const cfloat D = cmk(zk.x - zm.x, zk.y + zm.y);
        const cfloat O = cmk(0.5f * D.y, -0.5f * D.x);
        const cfloat X = cadd(E, cmul(tw, O));
        tb[kb * kbs + r * C + c] = cscale(X, a.scale);
        if (a.dc && k == 0) a.dc[(long long)b * a.H + row0 + r] = fabsf(X.x * a.scale);
    }
}

// The half-length packing of the real inverse: from X[k] and its Hermitian
// partner X[M-k] of the complex row (length M + 1 in staging),
// Z[k] = conj(X[k] + conj X[M-k] - i e^{+2 pi i k/(2M)} (X[k] - conj X[M-k]))
// (w = e^{-2 pi i k/(2M)}, the post-twiddle table entry).
__device__ __forceinline__ cfloat c2r_pre(cfloat p, cfloat q, cfloat w) {
    const cfloat t = cmul(cconj(w), cmk(p.x - q.x, p.y + q.y));
    return cconj(cmk(p.x + q.x - t.y, p.y - q.y + t.x));
}

// C2R rows: staging T -> pre-processing (one work item per Hermitian pair
// (k, M-k), k <= M/2; Z[0] also takes the Nyquist value X[M]) -> LDS in
// natural order -> passes 0 .. NP-2 in LDS -> last pass -> real rows.  The
// items run in the staging layout's order (C inputs of one row, the next
// row's, then the next C), so the RR rows' C consecutive inputs are read by
// consecutive lanes as one contiguous run.  Each item holds two inputs and
// one twiddle: the kernel needs no more registers than the forward rows.
// Z[M-k] = c2r_pre(X[M-k], X[k], e^{-2 pi i (M-k)/(2M)}), and that twiddle is
// -conj(w_k); the float table holds that symmetry exactly for 0 < k < M/2
// (only k = 0, M/2 and M break it, and they are never paired), so one load
// serves both.
template <int M, int RR, int C, int NT, int SC>
__global__ __launch_bounds__(NT) void k_rows_c2r_rp(RowC2RArgs a, const cfloat* __restrict__ ptw) {
    constexpr int NP = rp_npass(M, SC);
    constexpr int RL = rp_radix(M, NP - 1, SC), ML = M / RL;
    constexpr int NK = M / 2 + 1;                       // pre-processing items per row
    constexpr int NIP = ((NK + C - 1) / C) * C * RR;    // in staging order, padded
    constexpr int ITP = (NIP + NT - 1) / NT;
    constexpr int ITL = (RR * ML + NT - 1) / NT;
    constexpr int LD = rp_cs(M, RR, SC);
    extern __shared__ cfloat lds[];
    const int nrb = (a.H + RR - 1) / RR;
    const int tt = xcd_tile(blockIdx.x, gridDim.x);
    const int b = tt / nrb;
    const int row0 = (tt - b * nrb) * RR;
    const int nrows = min(RR, a.H - row0);
    const cfloat* tb = a.T + (long long)b * a.t_bstride + (long long)row0 * C;
    const long long kbs = (long long)a.H * C;
#pragma unroll
    for (int it = 0; it < ITP; it++) {
        const int e = threadIdx.x + it * NT;
        const int c = e % C, r = (e / C) % RR, k = (e / (C * RR)) * C + c;
        if ((NIP % NT == 0 || e < NIP) && k < NK && r < nrows) {
            const cfloat* xr = tb + r * C;
            const int m = k == 0 ? M : M - k;
            const cfloat p = xr[(long long)(k / C) * kbs + k % C];
            const cfloat q = xr[(long long)(m / C) * kbs + m % C];
            cfloat* d = lds + r * LD;
            if (k == 0) {
                d[rp_ix<M, SC>(0)] = cconj(cmk(p.x + q.x, p.x - q.x));
            } else {
                const cfloat w = a.post_tw[k];
                d[rp_ix<M, SC>(k)] = c2r_pre(p, q, w);
                if (m != k) d[rp_ix<M, SC>(m)] = c2r_pre(q, p, cmk(-w.x, w.y));
            }
        }
    }
    __syncthreads();
    RpRange<M, 0, NP - 1, RR, NT, SC>::run(lds, ptw);
    // Last pass, in one straight-line copy per store form: a run-time choice
    // between paired and single stores (and the clamp as a branch) inside
    // the unrolled store loop keeps about 40 more registers live.
    const bool clamp = a.clamp != 0;
    auto last = [&](auto vec) {
#pragma unroll
        for (int it = 0; it < ITL; it++) {
            const int i = threadIdx.x + it * NT;
            if ((RR * ML) % NT == 0 || i < RR * ML) {
                const int r = i / ML, k = i - r * ML;
                if (r < nrows) {
                    cfloat v[RL];
                    const cfloat* sl = lds + r * rp_cs(M, RR, SC);
#pragma unroll
                    for (int t = 0; t < RL; t++) v[t] = sl[rp_ix<M, SC>(k + t * ML)];
                    rp_twiddle<M, NP - 1, SC>(v, k, ptw);
                    RegDft<RL>::run(v);
                    float* ob = a.out + (long long)b * a.out_bstride + (long long)(row0 + r) * a.pitch;
#pragma unroll
                    for (int t = 0; t < RL; t++) {
                        const int y = k + t * ML;
                        float v0 = v[t].x, v1 = -v[t].y;
                        v0 = clamp ? fmaxf(v0, 0.f) : v0;
                        v1 = clamp ? fmaxf(v1, 0.f) : v1;
                        if constexpr (decltype(vec)::value) {
                            *reinterpret_cast<float2*>(ob + 2 * y) = make_float2(v0, v1);
                        } else {
                            ob[2 * y] = v0;
                            ob[2 * y + 1] = v1;
                        }
                    }
                }
            }
        }
    };
    if (((a.pitch | a.out_bstride) & 1) == 0 && (((uintptr_t)a.out) & 7) == 0) last(std::true_type{});
    else last(std::false_type{});
}

#ifndef MC_RP_ROW_SPECS
// (row length, rows per workgroup, staging chunk width, threads, radix
// schedule); the first entry of a length is its default, later ones are
// autotuner variants.
#define MC_RP_ROW_SPECS(X) X(2880, 2, MC_TILE_CT, 512, 0) X(2048, 2, MC_TILE_CT, 512, 0) \
    X(288, 16, 16, 512, 0) X(5760, 1, MC_TILE_CT, 512, 0) X(2880, 1, MC_TILE_CT, 256, 0) \
    X(2880, 2, MC_TILE_CT, 384, 0) X(2880, 1, MC_TILE_CT, 192, 0) X(288, 8, 16, 256, 0) \
    X(2880, 4, MC_TILE_CT, 1024, 0) X(2880, 4, MC_TILE_CT, 768, 0) X(2880, 2, MC_TILE_CT, 512, 1) \
    X(288, 16, 16, 512, 1)
#endif

// Twiddle table of a plan's register-pass rows (cols = 0) or columns for
// radix schedule sc (null when the plan has none).
static const cfloat* rp_tw(const Fft2D& p, int cols, int sc) {
    return cols ? (sc ? p.col_ptw1 : p.col_ptw) : (sc ? p.row_ptw1 : p.row_ptw);
}


// Launch a register-pass column kernel for the given storage formats.
template <int H, int CW, int NT, int SC>
static hipError_t launch_cols_fwd_rp(const ColFwdArgs& a, const cfloat* ptw, int batch, int s_fmt,
                                     int x_fmt, hipStream_t s) {
    const dim3 grid(mc_div_up(a.ncols, CW) * batch), blk(NT);
    const size_t lds = (size_t)CW * rp_cs(H, CW, SC) * sizeof(cfloat);
    const int key = s_fmt * 3 + x_fmt;
    switch (key) {
        case 0: hipLaunchKernelGGL((k_cols_fwd_rp<H, CW, NT, 0, 0, SC>), grid, blk, lds, s, a, ptw); break;
        case 1: hipLaunchKernelGGL((k_cols_fwd_rp<H, CW, NT, 0, 1, SC>), grid, blk, lds, s, a, ptw); break;
        case 2: hipLaunchKernelGGL((k_cols_fwd_rp<H, CW, NT, 0, 2, SC>), grid, blk, lds, s, a, ptw); break;
        case 3: hipLaunchKernelGGL((k_cols_fwd_rp<H, CW, NT, 1, 0, SC>), grid, blk, lds, s, a, ptw); break;
        case 4: hipLaunchKernelGGL((k_cols_fwd_rp<H, CW, NT, 1, 1, SC>), grid, blk, lds, s, a, ptw); break;
        case 6: hipLaunchKernelGGL((k_cols_fwd_rp<H, CW, NT, 2, 0, SC>), grid, blk, lds, s, a, ptw); break;
        case 8: hipLaunchKernelGGL((k_cols_fwd_rp<H, CW, NT, 2, 2, SC>), grid, blk, lds, s, a, ptw); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template <int H, int CW, int NT, int SC>
static hipError_t launch_cols_inv_rp(const ColInvArgs& a, const cfloat* ptw, int batch, int s_fmt,
                                     hipStream_t s) {
    const dim3 grid(mc_div_up(a.ncols, CW) * batch), blk(NT);
    const size_t lds = (size_t)CW * rp_cs(H, CW, SC) * sizeof(cfloat);
    switch (s_fmt) {
        case 0: hipLaunchKernelGGL((k_cols_inv_rp<H, CW, NT, 0, SC>), grid, blk, lds, s, a, ptw); break;
        case 1: hipLaunchKernelGGL((k_cols_inv_rp<H, CW, NT, 1, SC>), grid, blk, lds, s, a, ptw); break;
        case 2: hipLaunchKernelGGL((k_cols_inv_rp<H, CW, NT, 2, SC>), grid, blk, lds, s, a, ptw); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// g[b] = TARGET / sum_r |dc[b][r]| and inv[b] = 1 / g[b], summed in a fixed
// order (deterministic).  One workgroup per transform.
__global__ __launch_bounds__(256) void k_spec_scale(const float* __restrict__ dc, int rows,
                                                    float* __restrict__ g, float* __restrict__ inv) {
    __shared__ float part[256];
    const float* d = dc + (long long)blockIdx.x * rows;
    float acc = 0.f;
    for (int r = threadIdx.x; r < rows; r += 256) acc += d[r];
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float tot = part[0];
        const float gg = tot > 0.f ? MC_SPEC_TARGET / tot : 1.f;
        g[blockIdx.x] = gg;
        inv[blockIdx.x] = 1.f / gg;
    }
}


static bool g_attr_fail = false;
template <class F>
static void set_lds_attr(F* f) {
    const int bytes = (mc_lds_padded(MC_LDS_ELEMS) + 64) * (int)sizeof(cfloat);
    if (hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) !=
        hipSuccess)
        g_attr_fail = true;
}

static void set_ct_attrs() {
#define X(M, RR, C) set_lds_attr(k_rows_r2c_ct<M, RR, C>); set_lds_attr(k_rows_c2r_ct<M, RR, C>);
    MC_ROW_SPECS(X)
#undef X
#define X(H, C, NT) set_lds_attr(k_cols_fwd_ct<H, C, NT>); set_lds_attr(k_cols_inv_ct<H, C, NT>);
    MC_COL_SPECS(X)
#undef X
#define X(H, C, NT, SC)                                                                            \
    set_lds_attr(k_cols_fwd_rp<H, C, NT, 0, 0, SC>); set_lds_attr(k_cols_fwd_rp<H, C, NT, 0, 1, SC>); \
    set_lds_attr(k_cols_fwd_rp<H, C, NT, 0, 2, SC>); set_lds_attr(k_cols_fwd_rp<H, C, NT, 1, 0, SC>); \
    set_lds_attr(k_cols_fwd_rp<H, C, NT, 2, 0, SC>); set_lds_attr(k_cols_fwd_rp<H, C, NT, 1, 1, SC>); \
    set_lds_attr(k_cols_fwd_rp<H, C, NT, 2, 2, SC>); set_lds_attr(k_cols_inv_rp<H, C, NT, 0, SC>);    \
    set_lds_attr(k_cols_inv_rp<H, C, NT, 1, SC>); set_lds_attr(k_cols_inv_rp<H, C, NT, 2, SC>);
    MC_RP_COL_SPECS(X)
#undef X
#define X(M, RR, C, NT, SC) \
    set_lds_attr(k_rows_r2c_rp<M, RR, C, NT, SC>); set_lds_attr(k_rows_c2r_rp<M, RR, C, NT, SC>);
    MC_RP_ROW_SPECS(X)
#undef X
}

// Same factorisation as the compile-time stages (mc_fft_ct.h).
void factor_device(int n, std::vector<int>& r) {
    r.clear();
    for (int s = 0; s < ct_nst(n); s++) r.push_back(ct_radix(n, s));
}

int build_plan1(FftPlan1* P, cfloat** dev_tw, int n) {
    std::vector<int> rad;
    factor_device(n, rad);
    if ((int)rad.size() > MC_MAX_STAGES) return -1;
    std::vector<cfloat> h;
    P->n = n;
    P->nst = (int)rad.size();
    int ns = 1;
    for (size_t s = 0; s < rad.size(); s++) {
        const int r = rad[s];
        FftStage& st = P->st[s];
        st.r = r;
        st.ns = ns;
        st.m = n / r;
        st.nch = (r + MC_BIG_KC - 1) / MC_BIG_KC;
        st.rcp_m = 1.0f / st.m;
        st.rcp_ns = 1.0f / ns;
        st.rcp_nch = 1.0f / st.nch;
        st.tw_off = (int)h.size();
        for (int k = 0; k < ns; k++)
            for (int t = 1; t < r; t++) {
                const double a = -2.0 * M_PI * (double)t * k / ((double)ns * r);
                cfloat c;
                c.x = (float)std::cos(a);
                c.y = (float)std::sin(a);
                h.push_back(c);
            }
        st.root_off = (int)h.size();
        if (r > 31) {
            for (int q = 0; q < r; q++) {
                const double a = -2.0 * M_PI * q / r;
                cfloat c;
                c.x = (float)std::cos(a);
                c.y = (float)std::sin(a);
                h.push_back(c);
            }
        }
        ns *= r;
    }
    // chunk root table of the big-prime stage (ct_stage_bigsym), copied from
    // that stage's roots so both forms use the same values
    if (const int bp = ct_big_prime(n)) {
        int root = -1;
        for (int s = 0; s < P->nst; s++)
            if (P->st[s].r == bp) root = P->st[s].root_off;
        if (root < 0 || (int)h.size() != ct_twoff(n, ct_nst(n))) return -1;
        if (h.size() & 1) h.push_back(cfloat{0.f, 0.f});
        const int hp = (bp - 1) / 2, nch = (hp + MC_BIGSYM_KC - 1) / MC_BIGSYM_KC;
        const std::vector<cfloat> roots(h.begin() + root, h.begin() + root + bp);
        for (int ch = 0; ch < nch; ch++)
            for (int t = 1; t <= hp; t++)
                for (int q = 0; q < MC_BIGSYM_KC; q++) {
                    const int k = ch * MC_BIGSYM_KC + q + 1;
                    h.push_back(roots[k <= hp ? (t * k) % bp : 0]);
                }
        if ((int)h.size() != ct_rt_off(n) + ct_rt_elems(n)) return -1;
        // Rader table of the same stage (ct_stage_rader): Bh = DFT_L(b) / L
        // with b_n = e^{-2 pi i g^-n / P}, then e^{-2 pi i n2 k1 / L}
        if (ct_rd_elems(n)) {
            const int L = bp - 1, L1 = ct_rd_l1(bp), L2 = ct_rd_l2(bp);
            const int gi = ct_powmod(ct_proot(bp), bp - 2, bp);
            std::vector<double> br(L), bi(L);
            for (int k = 0, e = 1; k < L; k++, e = (int)((long long)e * gi % bp)) {
                br[k] = std::cos(-2.0 * M_PI * e / bp);
                bi[k] = std::sin(-2.0 * M_PI * e / bp);
            }
            for (int k = 0; k < L; k++) {
                double sr = 0.0, si = 0.0;
                for (int q = 0; q < L; q++) {
                    const double a = -2.0 * M_PI * (double)((long long)q * k % L) / L;
                    sr += br[q] * std::cos(a) - bi[q] * std::sin(a);
                    si += br[q] * std::sin(a) + bi[q] * std::cos(a);
                }
                h.push_back(cfloat{(float)(sr / L), (float)(si / L)});
            }
            for (int n2 = 0; n2 < L2; n2++)
                for (int k1 = 0; k1 < L1; k1++) {
                    const double a = -2.0 * M_PI * (double)(n2 * k1) / L;
                    h.push_back(cfloat{(float)std::cos(a), (float)std::sin(a)});
                }
            if ((int)h.size() != ct_rd_off(n) + ct_rd_elems(n)) return -1;
        }
    }
    if (h.empty()) h.push_back(cfloat{1.f, 0.f});
    if (hipMalloc(dev_tw, h.size() * sizeof(cfloat)) != hipSuccess) return -2;
    if (hipMemcpy(*dev_tw, h.data(), h.size() * sizeof(cfloat), hipMemcpyHostToDevice) != hipSuccess)
        return -2;
    P->tw = *dev_tw;
    return 0;
}

bool g_attr_done = false;

}  // namespace

hipError_t mc_launch_spec_scale(const float* dc, int rows, int batch, float* g, float* inv,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_spec_scale, dim3(batch), dim3(256), 0, s, dc, rows, g, inv);
    return hipGetLastError();
}

int mc_fft_init_constants() {
    if (mc_upload_prime_table() != 0) return -1;
    if (!g_attr_done) {
        set_lds_attr(k_rows_r2c);
        set_lds_attr(k_rows_c2r);
        set_lds_attr(k_cols_fwd);
        set_lds_attr(k_cols_inv);
        set_ct_attrs();
        if (g_attr_fail) return -1;
        g_attr_done = true;
    }
    return 0;
}

int mc_fft_supported(int n) { return n >= 2 && n <= MC_LDS_ELEMS; }

int mc_fft2d_create(Fft2D* p, int W, int H) {
    if (W < 4 || (W % 2) || H < 2) return -1;
    const int m = W / 2;
    if (m + 1 > MC_LDS_ELEMS || H > MC_LDS_ELEMS) return -1;
    p->W = W;
    p->H = H;
    p->m = m;
    p->cx = m + 1;
    // half the LDS per workgroup, so that two workgroups share a CU
    const int budget = MC_LDS_ELEMS / 2;
    int C = MC_TILE_CMAX;
    while (C > 1 && (long long)C * H > budget) C /= 2;
    // Heights with a register-pass column schedule read a few columns of a
    // wider staging chunk, so the row passes move whole 128-byte lines.
    if (rp_has(H) && H >= 1024) C = MC_TILE_CT;
    p->C = C;
    int R = 1;
    const int want = std::max((MC_TILE_RC + C - 1) / C, 2);
    const int per_row = m + 1;
    // a prime factor above 31 without a Rader form: two rows per workgroup
    // (more waves per CU for the direct big-prime stage); with the Rader
    // form (the 1441-point xcf rows) four, so that a row pass reads 128
    // contiguous bytes of each staging chunk
    const int rmax = (ct_big_prime(m) && !ct_rader(m)) ? 2 : want;
    while (R * 2 <= rmax && (R * 2) * per_row <= budget) R *= 2;
    if (R * per_row > MC_LDS_ELEMS) return -1;
    p->R = R;
    if (build_plan1(&p->row, &p->row_tw, m) != 0) return -2;
    if (build_plan1(&p->col, &p->col_tw, H) != 0) return -2;
    std::vector<cfloat> pt(m + 1);
    for (int k = 0; k <= m; k++) {
        const double a = -2.0 * M_PI * k / W;
        pt[k].x = (float)std::cos(a);
        pt[k].y = (float)std::sin(a);
    }
    if (hipMalloc(&p->post_tw, pt.size() * sizeof(cfloat)) != hipSuccess) return -2;
    if (hipMemcpy(p->post_tw, pt.data(), pt.size() * sizeof(cfloat), hipMemcpyHostToDevice) !=
        hipSuccess)
        return -2;
    for (int dim = 0; dim < 2; dim++)
        for (int sc = 0; sc < 2; sc++) {
            const std::vector<cfloat> tw = rp_twiddle_table(dim ? H : m, sc);
            if (tw.empty()) continue;
            cfloat** dstp = dim ? (sc ? &p->col_ptw1 : &p->col_ptw) : (sc ? &p->row_ptw1 : &p->row_ptw);
            if (hipMalloc(dstp, tw.size() * sizeof(cfloat)) != hipSuccess) return -2;
            if (hipMemcpy(*dstp, tw.data(), tw.size() * sizeof(cfloat), hipMemcpyHostToDevice) !=
                hipSuccess)
                return -2;
        }
    const int nb = (m + 1 + C - 1) / C;
    p->t_elems = (long long)nb * H * C;
    p->s_elems = (long long)(m + 1) * H;
    return 0;
}

void mc_fft2d_destroy(Fft2D* p) {
    if (p->row_tw) (void)hipFree(p->row_tw);
    if (p->col_tw) (void)hipFree(p->col_tw);
    if (p->post_tw) (void)hipFree(p->post_tw);
    for (cfloat* t : {p->col_ptw, p->row_ptw, p->col_ptw1, p->row_ptw1})
        if (t) (void)hipFree(t);
    p->row_tw = p->col_tw = p->post_tw = p->col_ptw = p->row_ptw = p->col_ptw1 = p->row_ptw1 = nullptr;
}

bool mc_fft_rp_rows(const Fft2D& p) {
#define X(MM, RRR, CCC, NT, SC) \
    if (p.m == MM && p.C == CCC && p.row_ptw) return true;
    MC_RP_ROW_SPECS(X)
#undef X
    return false;
}

bool mc_fft_rp_cols(const Fft2D& p) {
#define X(HH, CC, NT, SC) \
    if (p.H == HH && p.C % CC == 0 && p.col_ptw) return true;
    MC_RP_COL_SPECS(X)
#undef X
    return false;
}

// "a.b.c" radix order of schedule sc of length n
static std::string rp_order(int n, int sc) {
    std::string o;
    for (int p = 0; p < rp_npass(n, sc); p++) o += (p ? "." : "") + std::to_string(rp_radix(n, p, sc));
    return o;
}

int mc_fft_variants(const Fft2D& p, int cols, std::vector<std::string>* names, std::vector<int>* scheds) {
    int n = 0;
    char nm[128];
    if (cols) {
#define X(HH, CC, NT, SC)                                                                      \
    if (p.H == HH && p.C % CC == 0 && rp_tw(p, 1, SC)) {                                       \
        snprintf(nm, sizeof(nm), "cols<%d> %d col/wg %d thr radix %s", HH, CC, NT,            \
                 rp_order(HH, SC).c_str());                                                    \
        if (names) names->push_back(nm);                                                       \
        if (scheds) scheds->push_back(SC);                                                     \
        n++;                                                                                   \
    }
        MC_RP_COL_SPECS(X)
#undef X
    } else {
#define X(MM, RRR, CCC, NT, SC)                                                                \
    if (p.m == MM && p.C == CCC && rp_tw(p, 0, SC)) {                                          \
        snprintf(nm, sizeof(nm), "rows<%d> %d row/wg %d thr radix %s", MM, RRR, NT,           \
                 rp_order(MM, SC).c_str());                                                    \
        if (names) names->push_back(nm);                                                       \
        if (scheds) scheds->push_back(SC);                                                     \
        n++;                                                                                   \
    }
        MC_RP_ROW_SPECS(X)
#undef X
    }
    return n;
}

hipError_t mc_rows_r2c(const Fft2D& p, const float* in, long long pitch, long long in_bstride,
                       const int* win_xy, cfloat* T, long long t_bstride, int batch, float scale,
                       hipStream_t s) {
    return mc_rows_r2c_mod(p, in, pitch, in_bstride, 0, win_xy, T, t_bstride, batch, scale, s);
}

hipError_t mc_rows_r2c_mod(const Fft2D& p, const float* in, long long pitch, long long in_bstride,
                           int in_mod, const int* win_xy, cfloat* T, long long t_bstride,
                           int batch, float scale, hipStream_t s, float* dc) {
    RowR2CArgs a;
    a.plan = p.row;
    a.H = p.H;
    a.m = p.m;
    a.C = p.C;
    a.R = p.R;
    a.in = in;
    a.pitch = pitch;
    a.in_bstride = in_bstride;
    a.in_mod = in_mod;
    a.win = win_xy;
    a.T = T;
    a.t_bstride = t_bstride;
    a.post_tw = p.post_tw;
    a.scale = scale;
    a.dc = dc;
    dim3 grid(mc_div_up(p.H, p.R) * batch);
    const size_t lds = (size_t)mc_lds_padded(p.R * p.m) * sizeof(cfloat);
    int v_ = 0;
#define X(MM, RRR, CCC, NT, SC)                                                             \
    if (p.m == MM && p.C == CCC && rp_tw(p, 0, SC) && v_++ == p.row_var) {                 \
        hipLaunchKernelGGL((k_rows_r2c_rp<MM, RRR, CCC, NT, SC>),                          \
                           dim3(mc_div_up(p.H, RRR) * batch), dim3(NT),                    \
                           (size_t)RRR * rp_cs(MM, RRR, SC) * sizeof(cfloat), s, a, rp_tw(p, 0, SC)); \
        return hipGetLastError();                                                          \
    }
    MC_RP_ROW_SPECS(X)
#undef X
    if (dc) return hipErrorInvalidValue;  // row DCs only from the register-pass rows
#define X(MM, RRR, CCC)                                                             \
    if (p.m == MM && p.R == RRR && p.C == CCC) {                                    \
        const size_t ldsct =                                                        \
            (size_t)mc_lds_padded(RRR * (MM + 1)) * sizeof(cfloat); \
        hipLaunchKernelGGL((k_rows_r2c_ct<MM, RRR, CCC>), grid, dim3(MC_NTHR), ldsct, s, a); \
        return hipGetLastError();                                                   \
    }
    MC_ROW_SPECS(X)
#undef X
    hipLaunchKernelGGL(k_rows_r2c, grid, dim3(MC_NTHR), lds, s, a);
    return hipGetLastError();
}

hipError_t mc_rows_c2r(const Fft2D& p, const cfloat* T, long long t_bstride, float* out,
                       long long pitch, long long out_bstride, int batch, int clamp,
                       hipStream_t s) {
    RowC2RArgs a;
    a.plan = p.row;
    a.H = p.H;
    a.m = p.m;
    a.C = p.C;
    a.R = p.R;
    a.T = T;
    a.t_bstride = t_bstride;
    a.out = out;
    a.pitch = pitch;
    a.out_bstride = out_bstride;
    a.post_tw = p.post_tw;
    a.clamp = clamp;
    dim3 grid(mc_div_up(p.H, p.R) * batch);
    const size_t lds = (size_t)mc_lds_padded(p.R * (p.m + 1)) * sizeof(cfloat);
    int v_ = 0;
#define X(MM, RRR, CCC, NT, SC)                                                             \
    if (p.m == MM && p.C == CCC && rp_tw(p, 0, SC) && v_++ == p.row_var_inv) {                 \
        hipLaunchKernelGGL((k_rows_c2r_rp<MM, RRR, CCC, NT, SC>),                          \
                           dim3(mc_div_up(p.H, RRR) * batch), dim3(NT),                    \
                           (size_t)RRR * (rp_cs(MM, RRR, SC) + 1) * sizeof(cfloat), s, a,  \
                           rp_tw(p, 0, SC));                                               \
        return hipGetLastError();                                                          \
    }
    MC_RP_ROW_SPECS(X)
#undef X
#define X(MM, RRR, CCC)                                                             \
    if (p.m == MM && p.R == RRR && p.C == CCC) {                                    \
        const size_t ldsct =                                                        \
            (size_t)mc_lds_padded(RRR * (MM + 1)) * sizeof(cfloat); \
        hipLaunchKernelGGL((k_rows_c2r_ct<MM, RRR, CCC>), grid, dim3(MC_NTHR), ldsct, s, a); \
        return hipGetLastError();                                                   \
    }
    MC_ROW_SPECS(X)
#undef X
    hipLaunchKernelGGL(k_rows_c2r, grid, dim3(MC_NTHR), lds, s, a);
    return hipGetLastError();
}

hipError_t mc_cols_fwd(const Fft2D& p, const cfloat* T, long long t_bstride, cfloat* S,
                       long long s_bstride, cfloat* X, int xcx, int xcy, long long x_bstride,
                       int batch, hipStream_t s, int s_fmt, const float* s_g, int x_fmt,
                       const float* x_g) {
    ColFwdArgs a;
    a.plan = p.col;
    a.H = p.H;
    a.ncols = p.cx;
    a.C = p.C;
    a.T = T;
    a.t_bstride = t_bstride;
    a.S = S;
    a.s_bstride = s_bstride;
    a.X = X;
    a.xcx = xcx;
    a.xcy = xcy;
    a.x_bstride = x_bstride;
    a.s_g = s_g;
    a.x_g = x_g;
    dim3 grid(mc_div_up(p.cx, p.C) * batch);
    const size_t lds = (size_t)mc_lds_padded(p.C * p.H) * sizeof(cfloat);
    int v_ = 0;
#define X(HH, CC, NT, SC)                                                          \
    if (p.H == HH && p.C % CC == 0 && rp_tw(p, 1, SC) && v_++ == p.col_var)        \
        return launch_cols_fwd_rp<HH, CC, NT, SC>(a, rp_tw(p, 1, SC), batch, s_fmt, x_fmt, s);
    MC_RP_COL_SPECS(X)
#undef X
    if (s_fmt != MC_SPEC_F32 || x_fmt != MC_SPEC_F32) return hipErrorInvalidValue;
#define X(HH, CC, NT)                                                             \
    if (p.H == HH && p.C == CC) {                                                 \
        hipLaunchKernelGGL((k_cols_fwd_ct<HH, CC, NT>), grid, dim3(NT), lds, s, a);  \
        return hipGetLastError();                                                 \
    }
    MC_COL_SPECS(X)
#undef X
    if ((long long)p.C * p.H > MC_LDS_ELEMS) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_cols_fwd, grid, dim3(MC_NTHR), lds, s, a);
    return hipGetLastError();
}

hipError_t mc_cols_inv(const Fft2D& p, const cfloat* S, long long s_bstride, int src_cx,
                       int src_cy, const PhasorTabs* ph, cfloat* T, long long t_bstride, int batch,
                       hipStream_t s, int s_fmt, const float* s_inv) {
    ColInvArgs a;
    a.plan = p.col;
    a.H = p.H;
    a.ncols = p.cx;
    a.C = p.C;
    a.S = S;
    a.s_bstride = s_bstride;
    a.src_cx = src_cx;
    a.src_cy = src_cy;
    a.phA = ph ? ph->A : nullptr;
    a.phB = ph ? ph->B : nullptr;
    a.T = T;
    a.t_bstride = t_bstride;
    a.s_inv = s_inv;
    dim3 grid(mc_div_up(p.cx, p.C) * batch);
    const size_t lds = (size_t)mc_lds_padded(p.C * p.H) * sizeof(cfloat);
    int v_ = 0;
#define X(HH, CC, NT, SC)                                                          \
    if (p.H == HH && p.C % CC == 0 && rp_tw(p, 1, SC) && v_++ == p.col_var_inv)        \
        return launch_cols_inv_rp<HH, CC, NT, SC>(a, rp_tw(p, 1, SC), batch, s_fmt, s);
    MC_RP_COL_SPECS(X)
#undef X
    if (s_fmt != MC_SPEC_F32) return hipErrorInvalidValue;
#define X(HH, CC, NT)                                                             \
    if (p.H == HH && p.C == CC) {                                                 \
        hipLaunchKernelGGL((k_cols_inv_ct<HH, CC, NT>), grid, dim3(NT), lds, s, a);  \
        return hipGetLastError();                                                 \
    }
    MC_COL_SPECS(X)
#undef X
    if ((long long)p.C * p.H > MC_LDS_ELEMS) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL(k_cols_inv, grid, dim3(MC_NTHR), lds, s, a);
    return hipGetLastError();
}
