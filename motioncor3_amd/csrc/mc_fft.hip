// mc_fft.hip -- batched 2-D real FFT passes for gfx950.
//
// Replaces cufftPlan2d/cufftExecR2C/cufftExecC2R of Util/CCufft2D.cpp:23-113
// (one exec per frame there) with batched two-pass transforms: a row pass
// (real <-> half-complex along x, LDS-resident rows) and a column pass
// (complex along y, C adjacent columns per workgroup).  The staging layout
// between the passes is column-blocked so both passes move contiguous runs.
#include <cmath>
#include <vector>

#include "mc_fft.h"
#include "mc_fft_dev.h"

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

__global__ __launch_bounds__(MC_NTHR) void k_rows_r2c(RowR2CArgs a) {
    extern __shared__ cfloat lds[];
    const int b = blockIdx.y;
    const int row0 = blockIdx.x * a.R;
    const int nrows = min(a.R, a.H - row0);
    const int img = a.in_mod > 0 ? b % a.in_mod : b;
    const int wi = a.in_mod > 0 ? b / a.in_mod : b;
    int x0 = 0, y0 = 0;
    if (a.win) {
        x0 = a.win[2 * wi];
        y0 = a.win[2 * wi + 1];
    }
    const float* base = a.in + (long long)img * a.in_bstride;
    const float rm = 1.0f / a.m;
    if (((x0 | (int)(a.pitch & 1) | (int)(a.in_bstride & 1)) & 1) == 0 &&
        (((uintptr_t)a.in) & 7) == 0) {
        for (int e = threadIdx.x; e < nrows * a.m; e += MC_NTHR) {
            const int r = fdiv(e, a.m, rm);
            const int j = e - r * a.m;
            const float2 v = *reinterpret_cast<const float2*>(
                base + (long long)(y0 + row0 + r) * a.pitch + x0 + 2 * j);
            lds[e] = cmk(v.x, v.y);
        }
    } else {
        for (int e = threadIdx.x; e < nrows * a.m; e += MC_NTHR) {
            const int r = fdiv(e, a.m, rm);
            const int j = e - r * a.m;
            const float* src = base + (long long)(y0 + row0 + r) * a.pitch + x0 + 2 * j;
            lds[e] = cmk(src[0], src[1]);
        }
    }
    __syncthreads();
    fft_lds_fwd(lds, nrows, a.plan);
    const int kc = a.m + 1;
    const int nb = (kc + a.C - 1) / a.C;
    cfloat* tb = a.T + (long long)b * a.t_bstride;
    const int tot = nb * nrows * a.C;
    const int lc = __ffs(a.C) - 1;
    const float rnr = 1.0f / nrows;
    for (int e = threadIdx.x; e < tot; e += MC_NTHR) {
        const int c = e & (a.C - 1);
        const int rc = e >> lc;
        const int kb = fdiv(rc, nrows, rnr);
        const int r = rc - kb * nrows;
        const int k = kb * a.C + c;
        if (k > a.m) continue;
        const cfloat zk = lds[r * a.m + (k == a.m ? 0 : k)];
        const cfloat zm = lds[r * a.m + (k == 0 ? 0 : a.m - k)];
        const cfloat E = cmk(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        const cfloat D = cmk(zk.x - zm.x, zk.y + zm.y);
        const cfloat O = cmk(0.5f * D.y, -0.5f * D.x);
        const cfloat X = cadd(E, cmul(a.post_tw[k], O));
        tb[(long long)kb * a.H * a.C + (long long)(row0 + r) * a.C + c] = cscale(X, a.scale);
    }
}

__global__ __launch_bounds__(MC_NTHR) void k_rows_c2r(RowC2RArgs a) {
    extern __shared__ cfloat lds[];
    const int b = blockIdx.y;
    const int row0 = blockIdx.x * a.R;
    const int nrows = min(a.R, a.H - row0);
    cfloat* nyq = lds + nrows * a.m;
    const int kc = a.m + 1;
    const int nb = (kc + a.C - 1) / a.C;
    const cfloat* tb = a.T + (long long)b * a.t_bstride;
    const int tot = nb * nrows * a.C;
    const int lc = __ffs(a.C) - 1;
    const float rnr = 1.0f / nrows;
    for (int e = threadIdx.x; e < tot; e += MC_NTHR) {
        const int c = e & (a.C - 1);
        const int rc = e >> lc;
        const int kb = fdiv(rc, nrows, rnr);
        const int r = rc - kb * nrows;
        const int k = kb * a.C + c;
        if (k > a.m) continue;
        const cfloat v = tb[(long long)kb * a.H * a.C + (long long)(row0 + r) * a.C + c];
        if (k < a.m) lds[r * a.m + k] = v;
        else nyq[r] = v;
    }
    __syncthreads();
    // Hermitian pre-processing into Z'[k] = (X_k + X*_{m-k}) +
    // i e^{+2 pi i k/W} (X_k - X*_{m-k}); imag parts of X_0, X_m dropped.
    const int half = a.m / 2 + 1;
    const float rhalf = 1.0f / half;
    for (int e = threadIdx.x; e < nrows * half; e += MC_NTHR) {
        const int r = fdiv(e, half, rhalf);
        const int k = e - r * half;
        cfloat* row = lds + r * a.m;
        if (k == 0) {
            const float x0 = row[0].x, xm = nyq[r].x;
            row[0] = cconj(cmk(x0 + xm, x0 - xm));
            continue;
        }
        const int k2 = a.m - k;
        if (k2 < k) continue;
        const cfloat xa = row[k], xb = row[k2];
        const cfloat twk = cconj(a.post_tw[k]);
        const cfloat s1 = cmk(xa.x + xb.x, xa.y - xb.y);
        const cfloat d1 = cmk(xa.x - xb.x, xa.y + xb.y);
        const cfloat t1 = cmul(twk, d1);
        const cfloat zk = cmk(s1.x - t1.y, s1.y + t1.x);
        row[k] = cconj(zk);
        if (k2 != k) {
            const cfloat twk2 = cconj(a.post_tw[k2]);
            const cfloat s2 = cmk(xb.x + xa.x, xb.y - xa.y);
            const cfloat d2 = cmk(xb.x - xa.x, xb.y + xa.y);
            const cfloat t2 = cmul(twk2, d2);
            row[k2] = cconj(cmk(s2.x - t2.y, s2.y + t2.x));
        }
    }
    __syncthreads();
    fft_lds_fwd(lds, nrows, a.plan);
    float* ob = a.out + (long long)b * a.out_bstride;
    const float rm = 1.0f / a.m;
    const bool vec = ((a.pitch | a.out_bstride) & 1) == 0 && (((uintptr_t)a.out) & 7) == 0;
    for (int e = threadIdx.x; e < nrows * a.m; e += MC_NTHR) {
        const int r = fdiv(e, a.m, rm);
        const int j = e - r * a.m;
        const cfloat z = lds[e];
        float v0 = z.x, v1 = -z.y;
        if (a.clamp) {
            v0 = fmaxf(v0, 0.f);
            v1 = fmaxf(v1, 0.f);
        }
        float* dst = ob + (long long)(row0 + r) * a.pitch + 2 * j;
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
    const int b = blockIdx.y;
    const int kb = blockIdx.x;
    const int c0 = kb * a.C;
    const int nc = min(a.C, a.ncols - c0);
    const cfloat* src = a.T + (long long)b * a.t_bstride + (long long)kb * a.H * a.C;
    const int lc = __ffs(a.C) - 1;
    for (int e = threadIdx.x; e < a.H * a.C; e += MC_NTHR) {
        const int y = e >> lc;
        const int c = e & (a.C - 1);
        if (c < nc) lds[c * a.H + y] = src[e];
    }
    __syncthreads();
    fft_lds_fwd(lds, nc, a.plan);
    if (a.S) {
        cfloat* dst = a.S + (long long)b * a.s_bstride + (long long)c0 * a.H;
        for (int e = threadIdx.x; e < nc * a.H; e += MC_NTHR) dst[e] = lds[e];
    }
    if (a.X) {
        cfloat* xd = a.X + (long long)b * a.x_bstride;
        const float rx = 1.0f / a.xcy;
        for (int e = threadIdx.x; e < nc * a.xcy; e += MC_NTHR) {
            const int c = fdiv(e, a.xcy, rx);
            const int yo = e - c * a.xcy;
            const int kx = c0 + c;
            if (kx >= a.xcx) continue;
            const int yi = resize_row(yo, a.xcy, a.H);
            if (yi < 0) continue;
            xd[(long long)kx * a.xcy + yo] = lds[c * a.H + yi];
        }
    }
}

__global__ __launch_bounds__(MC_NTHR) void k_cols_inv(ColInvArgs a) {
    extern __shared__ cfloat lds[];
    const int b = blockIdx.y;
    const int kb = blockIdx.x;
    const int c0 = kb * a.C;
    const int nc = min(a.C, a.ncols - c0);
    const bool same = (a.src_cx == a.ncols && a.src_cy == a.H);
    const cfloat* sb = a.S + (long long)b * a.s_bstride;
    const float rH = 1.0f / a.H;
    for (int e = threadIdx.x; e < nc * a.H; e += MC_NTHR) {
        const int c = fdiv(e, a.H, rH);
        const int y = e - c * a.H;
        const int kx = c0 + c;
        cfloat v;
        if (same) {
            v = sb[(long long)kx * a.H + y];
        } else {
            const int yi = resize_row(y, a.H, a.src_cy);
            v = (kx < a.src_cx && yi >= 0) ? sb[(long long)kx * a.src_cy + yi] : cmk(0.f, 0.f);
        }
        if (a.phA) v = cmul(v, cmul(a.phA[(long long)b * a.ncols + kx], a.phB[(long long)b * a.H + y]));
        lds[e] = cconj(v);
    }
    __syncthreads();
    fft_lds_fwd(lds, nc, a.plan);
    cfloat* dst = a.T + (long long)b * a.t_bstride + (long long)kb * a.H * a.C;
    const int lc = __ffs(a.C) - 1;
    for (int e = threadIdx.x; e < a.H * a.C; e += MC_NTHR) {
        const int y = e >> lc;
        const int c = e & (a.C - 1);
        if (c < nc) dst[e] = cconj(lds[c * a.H + y]);
    }
}

void factor_device(int n, std::vector<int>& r) {
    r.clear();
    while (n % 8 == 0) { r.push_back(8); n /= 8; }
    if (n % 4 == 0) { r.push_back(4); n /= 4; }
    if (n % 2 == 0) { r.push_back(2); n /= 2; }
    static const int small[] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31};
    for (int p : small)
        while (n % p == 0) { r.push_back(p); n /= p; }
    for (int p = 37; n > 1; p += 2) {
        if ((long long)p * p > n) { r.push_back(n); break; }
        while (n % p == 0) { r.push_back(p); n /= p; }
    }
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
    if (h.empty()) h.push_back(cfloat{1.f, 0.f});
    if (hipMalloc(dev_tw, h.size() * sizeof(cfloat)) != hipSuccess) return -2;
    if (hipMemcpy(*dev_tw, h.data(), h.size() * sizeof(cfloat), hipMemcpyHostToDevice) != hipSuccess)
        return -2;
    P->tw = *dev_tw;
    return 0;
}

bool g_attr_done = false;

}  // namespace

int mc_fft_init_constants() {
    if (mc_upload_prime_table() != 0) return -1;
    if (!g_attr_done) {
        const int bytes = (MC_LDS_ELEMS + 64) * (int)sizeof(cfloat);
        hipFuncSetAttribute((const void*)k_rows_r2c, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        hipFuncSetAttribute((const void*)k_rows_c2r, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        hipFuncSetAttribute((const void*)k_cols_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        hipFuncSetAttribute((const void*)k_cols_inv, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
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
    int C = 16;
    while (C > 1 && (long long)C * H > MC_LDS_ELEMS) C /= 2;
    p->C = C;
    int R = (16 + C - 1) / C;
    if (R < 2) R = 2;
    const int rmax = MC_LDS_ELEMS / (m + 1);
    if (R > rmax) R = rmax;
    if (R < 1) return -1;
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
    hipMemcpy(p->post_tw, pt.data(), pt.size() * sizeof(cfloat), hipMemcpyHostToDevice);
    const int nb = (m + 1 + C - 1) / C;
    p->t_elems = (long long)nb * H * C;
    p->s_elems = (long long)(m + 1) * H;
    return 0;
}

void mc_fft2d_destroy(Fft2D* p) {
    if (p->row_tw) hipFree(p->row_tw);
    if (p->col_tw) hipFree(p->col_tw);
    if (p->post_tw) hipFree(p->post_tw);
    p->row_tw = p->col_tw = p->post_tw = nullptr;
}

hipError_t mc_rows_r2c(const Fft2D& p, const float* in, long long pitch, long long in_bstride,
                       const int* win_xy, cfloat* T, long long t_bstride, int batch, float scale,
                       hipStream_t s) {
    return mc_rows_r2c_mod(p, in, pitch, in_bstride, 0, win_xy, T, t_bstride, batch, scale, s);
}

hipError_t mc_rows_r2c_mod(const Fft2D& p, const float* in, long long pitch, long long in_bstride,
                           int in_mod, const int* win_xy, cfloat* T, long long t_bstride,
                           int batch, float scale, hipStream_t s) {
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
    dim3 grid(mc_div_up(p.H, p.R), batch);
    const size_t lds = (size_t)p.R * p.m * sizeof(cfloat);
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
    dim3 grid(mc_div_up(p.H, p.R), batch);
    const size_t lds = (size_t)p.R * (p.m + 1) * sizeof(cfloat);
    hipLaunchKernelGGL(k_rows_c2r, grid, dim3(MC_NTHR), lds, s, a);
    return hipGetLastError();
}

hipError_t mc_cols_fwd(const Fft2D& p, const cfloat* T, long long t_bstride, cfloat* S,
                       long long s_bstride, cfloat* X, int xcx, int xcy, long long x_bstride,
                       int batch, hipStream_t s) {
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
    dim3 grid(mc_div_up(p.cx, p.C), batch);
    const size_t lds = (size_t)p.C * p.H * sizeof(cfloat);
    hipLaunchKernelGGL(k_cols_fwd, grid, dim3(MC_NTHR), lds, s, a);
    return hipGetLastError();
}

hipError_t mc_cols_inv(const Fft2D& p, const cfloat* S, long long s_bstride, int src_cx,
                       int src_cy, const PhasorTabs* ph, cfloat* T, long long t_bstride, int batch,
                       hipStream_t s) {
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
    dim3 grid(mc_div_up(p.cx, p.C), batch);
    const size_t lds = (size_t)p.C * p.H * sizeof(cfloat);
    hipLaunchKernelGGL(k_cols_inv, grid, dim3(MC_NTHR), lds, s, a);
    return hipGetLastError();
}
