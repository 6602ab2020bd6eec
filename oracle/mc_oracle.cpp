/*
 * mc_oracle.cpp -- FP64 CPU restatement of MotionCor3's frame-alignment path.
 *
 * TEST INFRASTRUCTURE.  This is the checker for the HIP product path, never
 * part of it: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  Every function cites the reference file:line it
 * restates (paths relative to the MotionCor3 source tree).  Arithmetic is
 * FP64 throughout; the reference computes in FP32 on the GPU and host.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   FFT            numpy.fft golden vectors (tests/golden/fft_golden.npz)
 *   lin-eqs        Util_LinEqs::Test KAT and oracle/_ref build of
 *                  LibSrc/Util/Util_LinEqs.cpp
 *   peak           sampled paraboloid => exact vertex; tie/edge rules
 *   phase shift    integer shift of a delta, analytic sub-pixel shift
 *   correlate      frame vs its integer-shifted copy => exact peak
 *   spline/spikes  exact quadratic in sqrt(t) => zero residual, spike KAT
 *   dose weights   closed form of Correct/GWeightFrame.cu:19-30
 *   end to end     recovery of seeded ground-truth drift
 */
#include "mc_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cd;
const double kPi = 3.14159265358979323846;

/* ====================================================================== */
/* FFT: Stockham autosort, any radix (O(r^2) DFT per butterfly).          */
/* Stands in for cuFFT (Util/CCufft2D.cpp:23-146): same unnormalised      */
/* forward (e^-i) / inverse (e^+i) conventions.                           */
/* ====================================================================== */
struct FftPlan {
    int n = 0, sign = -1;
    std::vector<int> rad;
    std::vector<std::vector<cd>> tw;     /* per stage: [k*r + t] */
    std::vector<std::vector<cd>> roots;  /* per stage: [q] = e^{sign 2pi i q/r} */
};

void factorize(int n, std::vector<int>& rad) {
    rad.clear();
    while (n % 4 == 0) { rad.push_back(4); n /= 4; }
    while (n % 2 == 0) { rad.push_back(2); n /= 2; }
    for (int p = 3; n > 1; p += 2) {
        while (n % p == 0) { rad.push_back(p); n /= p; }
        if ((long long)p * p > n && n > 1) { rad.push_back(n); n = 1; }
    }
}

std::mutex g_plan_mu;
std::map<std::pair<int, int>, FftPlan*> g_plans;

const FftPlan& get_plan(int n, int sign) {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto key = std::make_pair(n, sign);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) return *it->second;
    FftPlan* p = new FftPlan;
    p->n = n;
    p->sign = sign;
    factorize(n, p->rad);
    int ns = 1;
    for (int r : p->rad) {
        std::vector<cd> tw((size_t)ns * r), ro(r);
        for (int k = 0; k < ns; k++)
            for (int t = 0; t < r; t++)
                tw[(size_t)k * r + t] =
                    std::polar(1.0, sign * 2.0 * kPi * t * k / ((double)ns * r));
        for (int q = 0; q < r; q++) ro[q] = std::polar(1.0, sign * 2.0 * kPi * q / r);
        p->tw.push_back(tw);
        p->roots.push_back(ro);
        ns *= r;
    }
    g_plans[key] = p;
    return *p;
}

void fft_exec(const FftPlan& p, cd* x, cd* work) {
    const int n = p.n;
    cd* a = x;
    cd* b = work;
    int ns = 1;
    thread_local std::vector<cd> v(256), y(256);
    for (size_t s = 0; s < p.rad.size(); s++) {
        const int r = p.rad[s];
        const int m = n / r;
        if ((int)v.size() < r) { v.resize(r); y.resize(r); }
        const cd* tw = p.tw[s].data();
        const cd* ro = p.roots[s].data();
        for (int j = 0; j < m; j++) {
            const int k = j % ns;
            for (int t = 0; t < r; t++) v[t] = a[j + t * m] * tw[(size_t)k * r + t];
            for (int q = 0; q < r; q++) {
                cd acc = 0.0;
                int e = 0;   /* (t * q) mod r */
                for (int t = 0; t < r; t++) {
                    acc += v[t] * ro[e];
                    e += q;
                    if (e >= r) e -= r;
                }
                y[q] = acc;
            }
            const int idx = (j / ns) * ns * r + k;
            for (int q = 0; q < r; q++) b[idx + q * ns] = y[q];
        }
        std::swap(a, b);
        ns *= r;
    }
    if (a != x) std::copy(a, a + n, x);
}

void fft1(cd* x, int n, int sign) {
    if (n <= 1) return;
    const FftPlan& p = get_plan(n, sign);
    thread_local std::vector<cd> w;
    if ((int)w.size() < n) w.resize(n);
    fft_exec(p, x, w.data());
}

/* 2-D real forward transform into the half spectrum, layout [y][kx]
 * (cufftExecR2C semantics, Util/CCufft2D.cpp:65-72). */
void rfft2(const double* img, int W, int H, cd* cmp) {
    const int cx = W / 2 + 1;
#pragma omp parallel
    {
        std::vector<cd> row(W);
#pragma omp for schedule(static)
        for (int y = 0; y < H; y++) {
            for (int x = 0; x < W; x++) row[x] = img[(size_t)y * W + x];
            fft1(row.data(), W, -1);
            for (int k = 0; k < cx; k++) cmp[(size_t)y * cx + k] = row[k];
        }
        /* columns in blocks of CB adjacent kx, so the strided gathers read
         * whole cache lines; each column's transform is unchanged */
        const int CB = 8;
        std::vector<cd> col((size_t)CB * H);
#pragma omp for schedule(static)
        for (int k0 = 0; k0 < cx; k0 += CB) {
            const int nb = std::min(CB, cx - k0);
            for (int y = 0; y < H; y++)
                for (int j = 0; j < nb; j++) col[(size_t)j * H + y] = cmp[(size_t)y * cx + k0 + j];
            for (int j = 0; j < nb; j++) fft1(&col[(size_t)j * H], H, -1);
            for (int y = 0; y < H; y++)
                for (int j = 0; j < nb; j++) cmp[(size_t)y * cx + k0 + j] = col[(size_t)j * H + y];
        }
    }
}

/* 2-D inverse of a half spectrum (cufftExecC2R semantics,
 * Util/CCufft2D.cpp:104-113): complex inverse along y, then Hermitian
 * completion along x with the imaginary parts of kx=0 and kx=W/2 dropped. */
void irfft2(const cd* cmp_in, int W, int H, double* img) {
    const int cx = W / 2 + 1;
    std::vector<cd> cmp(cmp_in, cmp_in + (size_t)cx * H);
#pragma omp parallel
    {
        const int CB = 8;
        std::vector<cd> col((size_t)CB * H);
#pragma omp for schedule(static)
        for (int k0 = 0; k0 < cx; k0 += CB) {
            const int nb = std::min(CB, cx - k0);
            for (int y = 0; y < H; y++)
                for (int j = 0; j < nb; j++) col[(size_t)j * H + y] = cmp[(size_t)y * cx + k0 + j];
            for (int j = 0; j < nb; j++) fft1(&col[(size_t)j * H], H, +1);
            for (int y = 0; y < H; y++)
                for (int j = 0; j < nb; j++) cmp[(size_t)y * cx + k0 + j] = col[(size_t)j * H + y];
        }
        std::vector<cd> row(W);
#pragma omp for schedule(static)
        for (int y = 0; y < H; y++) {
            const cd* c = &cmp[(size_t)y * cx];
            row[0] = cd(c[0].real(), 0.0);
            for (int k = 1; k < cx - 1; k++) {
                row[k] = c[k];
                row[W - k] = std::conj(c[k]);
            }
            row[cx - 1] = cd(c[cx - 1].real(), 0.0);
            fft1(row.data(), W, +1);
            for (int x = 0; x < W; x++) img[(size_t)y * W + x] = row[x].real();
        }
    }
}

/* ====================================================================== */
/* Spectrum operators                                                      */
/* ====================================================================== */

/* Util/GPhaseShift2D.cu:12-38,47-66: out (+)= in * e^{i(kx*sx' + ky*sy')},
 * sx' = sx*2pi/Nx, sy' = sy*2pi/Ny, ky signed. */
void phase_shift(const cd* in, int cx, int cy, double sx, double sy, bool sum, cd* out) {
    const int nx = (cx - 1) * 2;
    const double fsx = sx * (2.0 * kPi / nx);
    const double fsy = sy * (2.0 * kPi / cy);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < cy; y++) {
        const int ys = (y > cy / 2) ? y - cy : y;
        for (int k = 0; k < cx; k++) {
            const double ph = k * fsx + ys * fsy;
            const cd r = in[(size_t)y * cx + k] * cd(std::cos(ph), std::sin(ph));
            if (sum) out[(size_t)y * cx + k] += r;
            else out[(size_t)y * cx + k] = r;
        }
    }
}

/* Util/GPhaseShift2D.cu:68-79: in-place shift, skipped when |s| < 0.1 px. */
void phase_shift_skip(cd* frm, int cx, int cy, double sx, double sy) {
    if (std::sqrt(sx * sx + sy * sy) < 0.1) return;
    phase_shift(frm, cx, cy, sx, sy, false, frm);
}

/* Util/GFourierResize2D.cu:7-39: crop/pad with signed-row remap. Output
 * elements with no source are left untouched. */
void fourier_resize(const cd* in, int inx, int iny, cd* out, int outx, int outy, bool sum) {
    for (int yo = 0; yo < outy; yo++) {
        int y = yo;
        if (y > outy / 2) {
            y -= outy;
            if (y <= -iny / 2) continue;
            y += iny;
        } else if (y > iny / 2) {
            continue;
        }
        for (int k = 0; k < outx && k < inx; k++) {
            const cd v = in[(size_t)y * inx + k];
            if (sum) out[(size_t)yo * outx + k] += v;
            else out[(size_t)yo * outx + k] = v;
        }
    }
}

/* Align/GCorrelateSum2D.cu:13-140 (mGConv, mGLowpass, mGCenterOrigin, C2R,
 * central window copy).  The window is evaluated exactly from the filtered
 * spectrum without forming the full real map: first the complex inverse
 * along y at the window rows, then the Hermitian sum along x at the window
 * columns.  Same values as C2R-then-crop. */
void correlate_window(const cd* sum, const cd* grp, int cx, int cy, double bft,
                      bool phase_only, int seax, int seay, double* win) {
    const int W = (cx - 1) * 2, H = cy;
    std::vector<cd> X((size_t)cx * cy);
    const double filt0 = -0.25 * bft / ((double)(cx - 1) * cy);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < cy; y++) {
        const int ys = (y > cy / 2) ? y - cy : y;
        for (int k = 0; k < cx; k++) {
            const size_t i = (size_t)y * cx + k;
            const cd g = grp[i];
            const cd s = sum[i] - g;
            cd x = std::conj(g) * s;
            double f = std::exp(filt0 * ((double)k * k + (double)ys * ys));
            if (phase_only) f = f / (std::abs(x) + 1e-20);
            x = (i > 0) ? x * f : cd(0.0, 0.0);
            if (((k + ys) % 2) != 0) x = -x;
            X[i] = x;
        }
    }
    const int x0 = (W - seax) / 2, y0 = (H - seay) / 2;
    std::vector<cd> T((size_t)seay * cx);
#pragma omp parallel
    {
        std::vector<cd> col(H);
#pragma omp for schedule(static)
        for (int k = 0; k < cx; k++) {
            for (int y = 0; y < H; y++) col[y] = X[(size_t)y * cx + k];
            fft1(col.data(), H, +1);
            for (int j = 0; j < seay; j++) T[(size_t)j * cx + k] = col[y0 + j];
        }
    }
    std::vector<cd> tw(W);
    for (int i = 0; i < W; i++) tw[i] = std::polar(1.0, 2.0 * kPi * i / W);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < seay; j++) {
        const cd* t = &T[(size_t)j * cx];
        for (int i = 0; i < seax; i++) {
            const int x = x0 + i;
            double v = t[0].real() + t[cx - 1].real() * ((x % 2) ? -1.0 : 1.0);
            long long idx = 0;
            for (int k = 1; k < cx - 1; k++) {
                idx += x;
                if (idx >= W) idx %= W;
                v += 2.0 * (t[k] * tw[idx]).real();
            }
            win[(size_t)j * seax + i] = v;
        }
    }
}

/* Align/CPeak2D.cpp:39-109: first maximum wins (strict >), the integer peak
 * is moved off the border, then a separable 3-point parabola per axis with
 * |delta| > 1 rejected.  shift = peak - size/2. */
void peak2d(const double* img, int nx, int ny, double* shift) {
    double pk = img[0];
    int ip = 0;
    for (int i = 1; i < nx * ny; i++) {
        if (pk >= img[i]) continue;
        pk = img[i];
        ip = i;
    }
    int px = ip % nx, py = ip / nx;
    if (px == 0) px = 1;
    if (py == 0) py = 1;
    if (px == nx - 1) px -= 1;
    if (py == ny - 1) py -= 1;
    const int ic = py * nx + px;
    const double a = (img[ic + 1] + img[ic - 1]) * 0.5 - img[ic];
    const double b = (img[ic + 1] - img[ic - 1]) * 0.5;
    const double c = (img[ic + nx] + img[ic - nx]) * 0.5 - img[ic];
    const double d = (img[ic + nx] - img[ic - nx]) * 0.5;
    double cxv = -b / (2 * a + 1e-30);
    double cyv = -d / (2 * c + 1e-30);
    if (std::fabs(cxv) > 1) cxv = 0;
    if (std::fabs(cyv) > 1) cyv = 0;
    shift[0] = (px + cxv) - nx / 2;
    shift[1] = (py + cyv) - ny / 2;
}

/* ====================================================================== */
/* Host numerics: Util_LinEqs, CSplineFit1D, CRemoveSpikes1D, CStackShift  */
/* ====================================================================== */

/* LibSrc/Util/Util_LinEqs.cpp:19-148: Gauss-Jordan, full pivot search over
 * un-pivoted rows/columns, pivot row swapped onto the diagonal. */
bool lin_eqs(double* a, double* v, int n) {
    std::vector<int> ci(n), ri(n), piv(n, 0);
    for (int ip = 0; ip < n; ip++) {
        int ix = 0, iy = 0;
        bool found = false;
        double fmax = 0.0;
        for (int y = 0; y < n; y++) {
            if (piv[y] == 1) continue;
            for (int x = 0; x < n; x++) {
                if (piv[x] == 1) continue;
                const double c = std::fabs(a[y * n + x]);
                if (c <= fmax) continue;
                fmax = c;
                ix = x;
                iy = y;
                found = true;
            }
        }
        if (!found) return false;
        piv[ix] += 1;
        ci[ip] = ix;
        ri[ip] = iy;
        if (a[iy * n + ix] == 0.0) return false;
        if (ix != iy) {
            for (int x = 0; x < n; x++) std::swap(a[ix * n + x], a[iy * n + x]);
            std::swap(v[ix], v[iy]);
        }
        const int col = ix;
        const double pinv = 1.0 / (a[col * n + col] + 1e-30);
        for (int x = 0; x < n; x++) a[col * n + x] *= pinv;
        v[col] *= pinv;
        for (int y = 0; y < n; y++) {
            if (y == col) continue;
            const double dum = a[y * n + col];
            for (int x = 0; x < n; x++) a[y * n + x] -= a[col * n + x] * dum;
            v[y] -= v[col] * dum;
        }
    }
    return true;
}

/* Util/CSplineFit1D.cpp:31-115: least squares in the basis
 * {1, t, t^2, (t-k_i)^2 for t >= k_i}, with the 1e-20 curvature term. */
struct SplineFit {
    int nk = 0, dim = 3;
    const double* knots = nullptr;
    void terms(double x, double* t) const {
        t[0] = 1;
        t[1] = x;
        t[2] = x * x;
        for (int i = 0; i < nk; i++) t[3 + i] = 0;
        for (int i = 0; i < nk; i++) {
            if (x < knots[i]) break;
            const double v = x - knots[i];
            t[3 + i] = v * v;
        }
    }
    void cur_terms(double x, double* c) const {
        for (int i = 0; i < dim; i++) c[i] = 0;
        c[1] = 1.0;
        c[2] = 2.0 * x;
        for (int i = 0; i < nk; i++) {
            const double v = x - knots[i];
            if (v > 0) break;
            c[3 + i] = 2.0 * v;
        }
    }
    bool fit(const double* x, const double* y, int n, double* out) const {
        const double reg = 1e-20;
        std::vector<double> m(dim * dim, 0.0), cf(dim, 0.0), t(dim), c(dim);
        for (int i = 0; i < n; i++) {
            terms(x[i], t.data());
            cur_terms(x[i], c.data());
            for (int r = 0; r < dim; r++) {
                for (int q = 0; q < dim; q++) m[r * dim + q] += t[r] * t[q] + reg * c[r] * c[q];
                cf[r] += t[r] * y[i];
            }
        }
        if (!lin_eqs(m.data(), cf.data(), dim)) return false;
        for (int i = 0; i < n; i++) {
            terms(x[i], t.data());
            double f = 0;
            for (int r = 0; r < dim; r++) f += cf[r] * t[r];
            out[i] = f;
        }
        return true;
    }
};

/* Util/CRemoveSpikes1D.cpp:126-148 */
void remove_single(double* raw, int n) {
    std::vector<double> buf(raw, raw + n);
    for (int i = 1; i < n - 1; i++) {
        if (std::fabs(buf[i] - buf[i - 1]) < 5.0) continue;
        if (std::fabs(buf[i] - buf[i + 1]) < 5.0) continue;
        raw[i] = (buf[i - 1] + buf[i + 1]) / 2.0;
    }
    const int e = n - 1;
    if (std::fabs(buf[0] - buf[1]) > 10.0 && std::fabs(buf[0]) > std::fabs(buf[1]))
        raw[0] = 2.0 * buf[1] - buf[2];
    if (std::fabs(buf[e] - buf[e - 1]) > 10.0 && std::fabs(buf[e]) > std::fabs(buf[e - 1]))
        raw[e] = 2.0 * buf[e - 1] - buf[e - 2];
}

/* Util/CRemoveSpikes1D.cpp:107-124 */
bool find_bad(double* raw, const double* fit, int n, double tol) {
    bool has = false;
    for (int i = 0; i < n; i++) {
        if (std::fabs(fit[i] - raw[i]) > tol && std::fabs(raw[i]) > std::fabs(fit[i])) {
            raw[i] = fit[i];
            has = true;
        }
    }
    return has;
}

/* Util/CRemoveSpikes1D.cpp:31-105 (SetDataSize + DoIt).  With quirks on, the
 * y refit uses the original array (:88), as the reference does. */
void remove_spikes(double* sx, double* sy, int n, bool single, bool quirks) {
    for (int i = 0; i < 2; i++) {
        remove_single(sx, n);
        remove_single(sy, n);
    }
    if (single) return;
    std::vector<double> t(n);
    const float ffact = 1.0f / n;
    for (int i = 0; i < n; i++) t[i] = std::sqrt((double)(i * ffact));
    const int kd = 15;
    int nk = (n - 2 * kd) / kd;
    if (nk < 0) nk = 0;
    else if (nk > 6) nk = 6;
    if (nk == 0 && n >= 2 * kd) nk = 1;
    std::vector<double> knots(nk);
    for (int i = 0; i < nk; i++) knots[i] = t[(i + 1) * kd];
    SplineFit sf;
    sf.nk = nk;
    sf.dim = nk + 3;
    sf.knots = knots.data();
    std::vector<double> fx(n), fy(n), buf(n);
    sf.fit(t.data(), sx, n, fx.data());
    sf.fit(t.data(), sy, n, fy.data());
    std::copy(sx, sx + n, buf.begin());
    if (find_bad(buf.data(), fx.data(), n, 3.0)) sf.fit(t.data(), buf.data(), n, fx.data());
    std::copy(sy, sy + n, buf.begin());
    if (find_bad(buf.data(), fy.data(), n, 3.0))
        sf.fit(t.data(), quirks ? sy : buf.data(), n, fy.data());
    for (int i = 0; i < n; i++) {
        double tol = 5.0;
        if (i > 20) tol = 4.0;
        if (i < 50) tol = 3.0;
        if (std::fabs(fx[i] - sx[i]) > tol) sx[i] = fx[i];
        if (std::fabs(fy[i] - sy[i]) > tol) sy[i] = fy[i];
    }
}

/* Align/CStackShift.cpp:180-198 */
void smooth(double* sx, double* sy, int n, double w) {
    if (w >= 1) return;
    std::vector<double> ox(sx, sx + n), oy(sy, sy + n);
    const double w2 = 0.5 * (1.0 - w);
    for (int i = 1; i < n - 1; i++) {
        sx[i] = w * ox[i] + w2 * (ox[i - 1] + ox[i + 1]);
        sy[i] = w * oy[i] + w2 * (oy[i - 1] + oy[i + 1]);
    }
}

/* ====================================================================== */
/* Frame grouping and dose (DataUtil/CFmGroupParam.cpp, CFmIntParam.cpp)  */
/* ====================================================================== */
struct Groups {
    int ng = 0, binz = 1;
    std::vector<int> start, size;
    std::vector<double> center;
};

/* DataUtil/CFmGroupParam.cpp:26-97: groups of at most binz x (smallest
 * integrated frame) raw frames; frame centres from
 * DataUtil/CFmIntParam.cpp:138-146,280-287. */
Groups make_groups(int binz, int n, const std::vector<double>& fmcent, const int* fmsize) {
    Groups g;
    g.binz = binz;
    int minraw = 1 << 30;
    for (int i = 0; i < n; i++) minraw = std::min(minraw, fmsize ? fmsize[i] : 1);
    const int maxraw = binz * minraw;
    int f = 0;
    for (int i = 0; i < n; i++) {
        g.start.push_back(f);
        int cnt = 0, raw = 0;
        while (true) {
            raw += fmsize ? fmsize[f] : 1;
            cnt += 1;
            f += 1;
            if (f >= n) break;
            if (raw >= maxraw) break;
        }
        g.size.push_back(cnt);
        g.ng += 1;
        if (f >= n) break;
    }
    for (int i = 0; i < g.ng; i++) {
        double s = 0;
        for (int j = 0; j < g.size[i]; j++) s += fmcent[g.start[i] + j];
        g.center.push_back(s / g.size[i]);
    }
    return g;
}

/* Align/CInterpolateShift.cpp:41-142 */
void interpolate(const std::vector<double>& gx, const std::vector<double>& gy,
                 const Groups& g, const std::vector<double>& fmcent, bool nearest,
                 std::vector<double>& ox, std::vector<double>& oy) {
    const int n = (int)fmcent.size();
    ox.assign(n, 0.0);
    oy.assign(n, 0.0);
    if (n == g.ng) {
        ox = gx;
        oy = gy;
        return;
    }
    if (nearest) {
        for (int k = 0; k < g.ng; k++)
            for (int i = 0; i < g.size[k]; i++) {
                ox[g.start[k] + i] = gx[k];
                oy[g.start[k] + i] = gy[k];
            }
        return;
    }
    for (int i = 0; i < n; i++) {
        int k;
        if (fmcent[i] <= g.center[0]) k = 1;
        else {
            k = g.ng - 1;
            for (int q = 1; q < g.ng; q++)
                if (fmcent[i] < g.center[q]) { k = q; break; }
        }
        const int f = k - 1;
        const double dg = g.center[k] - g.center[f];
        const double df = fmcent[i] - g.center[f];
        ox[i] = gx[f] + (gx[k] - gx[f]) / dg * df;
        oy[i] = gy[f] + (gy[k] - gy[f]) / dg * df;
    }
}

/* Correct/GWeightFrame.cu:19-30 (mGCalcWeight) and :79-133 (BuildWeight). */
double kv_factor(int kv) {
    if (kv >= 300) return 1.0;
    if (kv >= 200) return 0.002 * (kv - 200) + 0.8;
    if (kv >= 120) return 0.004375 * (kv - 120) + 0.45;
    return 0.45;
}

double dose_weight(int kx, int y, int cx, int cy, double pix, double kvf, double dose) {
    double fx = kx * 0.5 / (cx - 1);
    double fy = y / (double)cy;
    if (fy >= 0.5) fy -= 1.0;
    const double f = std::sqrt(fx * fx + fy * fy) / pix;
    const double crit = (0.24499 * std::pow(f, -1.6649) + 2.8141) * kvf;
    return std::exp(-0.5 * dose / crit);
}

/* ====================================================================== */
/* Path context                                                            */
/* ====================================================================== */
struct Geo {
    int W, H, N;
    int cx, cy;          /* full half spectrum */
    int xcx, xcy;        /* xcf half spectrum */
    int xw, xh;          /* xcf image */
    int pw, ph, pcx;     /* patch image, patch half-spectrum x */
    double xbin[2];
    int ocx, ocy, ow, oh;/* output (FtBin) */
};

/* CBufferPool.cpp:220-262 (xcf and patch sizes, float arithmetic kept). */
Geo make_geo(int W, int H, int N, int px, int py, double ftbin) {
    Geo g;
    g.W = W; g.H = H; g.N = N;
    g.cx = W / 2 + 1; g.cy = H;
    float fbx = W / 2048.0f, fby = H / 2048.0f;
    float fb = std::fmin(fbx, fby);
    if (fb < 1) fb = 1.0f;
    int ixx = (int)(W / fb + 0.5f);
    int ixy = (int)(H / fb + 0.5f);
    g.xcx = ixx / 2 + 1;
    g.xcy = ixy / 2 * 2;
    g.xw = (g.xcx - 1) * 2;
    g.xh = g.xcy;
    g.xbin[0] = (double)((g.cx - 1.0f) / (g.xcx - 1.0f));
    g.xbin[1] = (double)(g.cy * 1.0f / g.xcy);
    g.pw = g.ph = g.pcx = 0;
    if (px > 1 && py > 1) {
        int ipx = g.xw / px, ipy = g.xh / py;
        g.pcx = ipx / 2 + 1;
        g.pw = (g.pcx - 1) * 2;
        g.ph = ipy / 2 * 2;
    }
    /* Align/CAlignBase.cpp:41-47 + Util/GFourierResize2D.cu:41-71 */
    g.ocx = g.cx; g.ocy = g.cy;
    if ((float)ftbin != 1.0f) {
        int nx = (int)(((g.cx - 1) * 2) / (float)ftbin);
        int ny = (int)(g.cy / (float)ftbin);
        nx = nx / 2 * 2;
        ny = ny / 2 * 2;
        g.ocx = nx / 2 + 1;
        g.ocy = ny;
    }
    g.ow = (g.ocx - 1) * 2;
    g.oh = g.ocy;
    return g;
}

struct Stack {   /* a stack of half spectra, reference layout */
    int cx = 0, cy = 0, n = 0;
    std::vector<cd> d;
    cd* frame(int i) { return &d[(size_t)i * cx * cy]; }
    const cd* frame(int i) const { return &d[(size_t)i * cx * cy]; }
    void alloc(int x, int y, int k) { cx = x; cy = y; n = k; d.assign((size_t)x * y * k, cd(0, 0)); }
};

/* One CIterativeAlign run (Align/CIterativeAlign.cpp:32-175 with
 * CAlignedSum Align/CAlignedSum.cpp:25-126 and CAlignStack
 * Align/CAlignStack.cpp:77-215).  Returns frame shifts after both smooths. */
struct IterResult {
    std::vector<double> sx, sy, errors;
    int iters = 0;
};

IterResult iterative_align(const Stack& stk, const Groups& grp, const std::vector<double>& fmcent,
                           double bft, bool phase_only, int max_iter, double tol, bool quirks) {
    const int n = stk.n, cx = stk.cx, cy = stk.cy;
    const int W = (cx - 1) * 2;
    const int seax = std::min(64, W), seay = std::min(64, cy);
    IterResult res;
    std::vector<double> tx(n, 0.0), ty(n, 0.0);
    std::vector<double> totx(grp.ng, 0.0), toty(grp.ng, 0.0);
    std::vector<double> gsx(grp.ng, 0.0), gsy(grp.ng, 0.0);
    double b = bft, maxerr = 0.0;
    const size_t fsz = (size_t)cx * cy;
    std::vector<cd> sum(fsz);
    for (int it = 0; it < max_iter; it++) {
        /* CAlignedSum: S = sum_i shift(F_i, -tmp_i) */
        std::fill(sum.begin(), sum.end(), cd(0, 0));
#pragma omp parallel
        {
            std::vector<cd> part(fsz, cd(0, 0));
#pragma omp for schedule(static)
            for (int i = 0; i < n; i++)
                phase_shift(stk.frame(i), cx, cy, -tx[i], -ty[i], true, part.data());
#pragma omp critical
            for (size_t q = 0; q < fsz; q++) sum[q] += part[q];
        }
        maxerr = 0.0;
        std::vector<double> win((size_t)seax * seay);
        std::vector<cd> gs(fsz);
        for (int g = 0; g < grp.ng; g++) {
            std::fill(gs.begin(), gs.end(), cd(0, 0));
            for (int j = 0; j < grp.size[g]; j++) {
                const int f = grp.start[g] + j;
                phase_shift(stk.frame(f), cx, cy, -tx[f], -ty[f], true, gs.data());
            }
            correlate_window(sum.data(), gs.data(), cx, cy, b, phase_only, seax, seay, win.data());
            double sh[2];
            peak2d(win.data(), seax, seay, sh);
            gsx[g] = sh[0];
            gsy[g] = sh[1];
            const double e = std::sqrt(sh[0] * sh[0] + sh[1] * sh[1]);
            if (maxerr < e) maxerr = e;
        }
        b -= 2;
        if (b < 20) b = 20.0;
        for (int g = 0; g < grp.ng; g++) { totx[g] += gsx[g]; toty[g] += gsy[g]; }
        std::vector<double> ix, iy;
        interpolate(gsx, gsy, grp, fmcent, true, ix, iy);
        for (int i = 0; i < n; i++) { tx[i] += ix[i]; ty[i] += iy[i]; }
        res.errors.push_back(maxerr);
        res.iters += 1;
        if (maxerr < tol && it > 0) break;
    }
    interpolate(totx, toty, grp, fmcent, false, res.sx, res.sy);
    remove_spikes(res.sx.data(), res.sy.data(), n, false, quirks);
    smooth(res.sx.data(), res.sy.data(), n, 0.5);
    double w = 0.3 + 0.1 * grp.binz;
    if (w > 0.9) w = 0.9;
    smooth(res.sx.data(), res.sy.data(), n, w);
    return res;
}

/* GNormByStd2D (Align/GNormByStd2D.cu:7-42), evaluated out of place: the
 * reference kernel reads neighbours of pixels that other blocks of the same
 * launch overwrite, so its result depends on block scheduling. */
void norm_by_var(std::vector<double>& img, int w, int h, int wx, int wy) {
    std::vector<double> src(img);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double m = 0, s = 0;
            int c = 0;
            for (int j = 0; j < wy; j++) {
                int yy = y - wy / 2 + j;
                if (yy < 0 || yy >= h) continue;
                for (int i = 0; i < wx; i++) {
                    int xx = x - wx / 2 + i;
                    if (xx < 0 || xx >= w) continue;
                    const double v = src[(size_t)yy * w + xx];
                    m += v;
                    s += v * v;
                    c++;
                }
            }
            m /= c;
            s = s / c - m * m;
            if (s < 0) img[(size_t)y * w + x] = 0.0;
            else img[(size_t)y * w + x] = src[(size_t)y * w + x] / (s + 1e-20);
        }
}

/* Align/CDetectFeatures.cpp:58-100: the B x B real image made positive
 * (+ 1 - min, Util/GPositivity2D AddVal) and divided by its local 5 x 5
 * variance (Align/GNormByStd2D.cu). */
void feature_prepare(std::vector<double>& img, int B) {
    double mn = 1e300;
    for (double v : img) mn = std::min(mn, v);
    for (double& v : img) v += 1.0 - mn;
    norm_by_var(img, B, B, 5, 5);
}

/* Align/CDetectFeatures.cpp:134-186: mean and sigma of the map, the same
 * over the values inside mean +- 4 sigma, then a pixel is a feature when
 * mu - 4 sigma <= v <= mu - 0.08 sigma of the clipped statistics. */
void feature_threshold(const double* img, int n, unsigned char* feat, double* thr) {
    double mean = 0, m2 = 0;
    for (int i = 0; i < n; i++) { mean += img[i]; m2 += img[i] * img[i]; }
    mean /= n;
    m2 /= n;
    double sd = m2 - mean * mean;
    sd = (sd < 0) ? 0.0 : std::sqrt(sd);
    double lo = mean - 4.0 * sd, hi = mean + 4.0 * sd;
    double dm = 0, ds = 0;
    int cnt = 0;
    for (int i = 0; i < n; i++) {
        const double v = img[i];
        if (v < lo || v > hi) continue;
        dm += v;
        ds += v * v;
        cnt++;
    }
    dm /= (cnt + 1e-30);
    ds = ds / (cnt + 1e-30) - dm * dm;
    ds = (ds < 0) ? 0.0 : std::sqrt(ds);
    lo = dm - 4.0 * ds;
    hi = dm - 0.08 * ds;
    if (thr) { thr[0] = lo; thr[1] = hi; }
    for (int i = 0; i < n; i++) feat[i] = (img[i] < lo || img[i] > hi) ? 0 : 1;
}

/* Align/CDetectFeatures.cpp:58-186: the aligned xcf sum Fourier-cropped to
 * 128 x 128, C2R, made positive, normalised, thresholded. */
void feature_map_from_sum(const cd* sumx, int xcx, int xcy, unsigned char* feat) {
    const int B = 128, bcx = B / 2 + 1;
    std::vector<cd> bin((size_t)bcx * B, cd(0, 0));
    fourier_resize(sumx, xcx, xcy, bin.data(), bcx, B, false);
    std::vector<double> img((size_t)B * B);
    irfft2(bin.data(), B, B, img.data());
    feature_prepare(img, B);
    feature_threshold(img.data(), B * B, feat, nullptr);
}

/* Align/CDetectFeatures.cpp:272-396 and Align/CPatchCenters.cpp:42-105.
 * These are discrete decisions (which feature is nearest, where a window
 * starts), so they are evaluated in FP32 as the reference writes them: in
 * FP64 a tie between two features at equal FP32 distance breaks the other
 * way. */
struct FeatureCenters {
    int B = 128, npx, npy;
    float pat[2];
    int sea[4];
    std::vector<unsigned char> feat, used;
    std::vector<float> cent;   /* 3 per patch */
    int check_range(int start, int size, const int* rng) const {
        if (start < rng[0]) return rng[0];
        if (start + size > rng[1]) return rng[1] - size;
        return start;
    }
    void set_used(float cxv, float cyv) {
        int wx = (int)(pat[0] * 0.8f), wy = (int)(pat[1] * 0.8f);
        int sx = (int)(cxv - wx / 2), sy = (int)(cyv - wy / 2);
        sx = check_range(sx, wx, &sea[0]);
        sy = check_range(sy, wy, &sea[2]);
        for (int y = 0; y < wy; y++)
            for (int x = 0; x < wx; x++) used[(size_t)(sy + y) * B + sx + x] = 1;
    }
    bool check_feature(float cxv, float cyv) const {
        int ix = (int)cxv, iy = (int)cyv;
        if (ix < sea[0] || ix > sea[1]) return false;
        if (iy < sea[2] || iy > sea[3]) return false;
        return feat[(size_t)iy * B + ix] != 0;
    }
    void find_center(float* c) {
        int wx = B / 3, wy = B / 3;
        int sx = (int)(c[0] - wx * 0.5f + 0.5f), sy = (int)(c[1] - wy * 0.5f + 0.5f);
        sx = check_range(sx, wx, &sea[0]);
        sy = check_range(sy, wy, &sea[2]);
        float minr = 1e20f;
        float ac[2] = {-1.0f, 0.0f};
        for (int y = 0; y < wy; y++)
            for (int x = 0; x < wx; x++) {
                size_t j = (size_t)(sy + y) * B + sx + x;
                if (used[j] || !feat[j]) continue;
                float fx = x + sx + 0.5f, fy = y + sy + 0.5f;
                float dx = fx - c[0], dy = fy - c[1];
                float r = sqrtf(dx * dx + dy * dy);
                if (r >= minr) continue;
                minr = r;
                ac[0] = fx;
                ac[1] = fy;
            }
        if (ac[0] > 0) { c[0] = ac[0]; c[1] = ac[1]; }
        set_used(c[0], c[1]);
    }
    void run(const unsigned char* f, int px, int py) {
        npx = px; npy = py;
        feat.assign(f, f + B * B);
        used.assign(B * B, 0);
        pat[0] = B * 1.0f / px;
        pat[1] = B * 1.0f / py;
        sea[0] = (int)(pat[0] * 0.5f - 0.5f);
        if (sea[0] < 4) sea[0] = 4;
        sea[1] = B - sea[0];
        sea[2] = (int)(pat[1] * 0.5f - 0.5f);
        if (sea[2] < 4) sea[2] = 4;
        sea[3] = B - sea[2];
        const int np = px * py;
        cent.assign(3 * np, 0.0f);
        for (int i = 0; i < np; i++) {
            int x = i % px, y = i / px;
            float fx = (x + 0.5f) * pat[0], fy = (y + 0.5f) * pat[1];
            bool has = check_feature(fx, fy);
            cent[3 * i] = fx;
            cent[3 * i + 1] = fy;
            if (has) { cent[3 * i + 2] = 1.0f; set_used(fx, fy); }
            else cent[3 * i + 2] = -1.0f;
        }
        for (int i = 0; i < np; i++) {
            if (cent[3 * i + 2] > 0) continue;
            find_center(&cent[3 * i]);
        }
    }
};

void patch_starts_from_features(const unsigned char* feat, int px, int py, int xw, int xh,
                                int* starts) {
    FeatureCenters fc;
    fc.run(feat, px, py);
    const int psx = xw / px / 2 * 2, psy = xh / py / 2 * 2;
    /* CPatchCenters::mCalcStart (Align/CPatchCenters.cpp:99-105) */
    auto calc_start = [](float c, int ps, int xs) {
        int s = (int)(c - ps / 2);
        if (s < 0) s = 0;
        else if (s + ps > xs) s = xs - ps;
        return s;
    };
    /* CDetectFeatures::GetCenter (Align/CDetectFeatures.cpp:188-195) */
    const float binx = xw * 1.0f / 128, biny = xh * 1.0f / 128;
    for (int i = 0; i < px * py; i++) {
        const float cxv = fc.cent[3 * i] * binx;
        const float cyv = fc.cent[3 * i + 1] * biny;
        starts[2 * i] = calc_start(cxv, psx, xw);
        starts[2 * i + 1] = calc_start(cyv, psy, xh);
    }
}

/* Align/CPatchShifts.cpp:263-286.  shifts: [frame][patch][xy].  With quirks
 * the y component subtracts element k+2 of the reference frame's row and
 * the reference row is modified while it is used; `tail` holds the values a
 * read past the last frame would see (the flag bytes, zero at this point). */
int make_relative(double* s, int nf, int np, bool quirks) {
    auto mean_std = [&](int f, double* ms) {
        const double* p = s + (size_t)f * np * 2;
        double mx = 0, my = 0;
        for (int i = 0; i < np; i++) { mx += p[2 * i]; my += p[2 * i + 1]; }
        mx /= np;
        my /= np;
        double sd = 0;
        for (int i = 0; i < np; i++) {
            double dx = p[2 * i] - mx, dy = p[2 * i + 1] - my;
            sd += dx * dx + dy * dy;
        }
        ms[0] = std::sqrt(mx * mx + my * my);
        ms[1] = std::sqrt(sd / np);
    };
    double fmin = 1e20;
    int ref = 0;
    for (int i = 1; i < nf; i++) {
        double ms[2];
        mean_std(i, ms);
        const double v = ms[0] + ms[1];
        if (v > fmin) continue;
        fmin = v;
        ref = i;
    }
    const size_t total = (size_t)nf * np * 2;
    if (!quirks) {
        std::vector<double> r(s + (size_t)ref * np * 2, s + (size_t)(ref + 1) * np * 2);
        for (int i = 0; i < nf; i++)
            for (int j = 0; j < np; j++) {
                s[((size_t)i * np + j) * 2] -= r[2 * j];
                s[((size_t)i * np + j) * 2 + 1] -= r[2 * j + 1];
            }
        return ref;
    }
    auto at = [&](size_t idx) -> double { return idx < total ? s[idx] : 0.0; };
    const size_t rb = (size_t)ref * np * 2;
    for (int i = 0; i < nf; i++) {
        const size_t pb = (size_t)i * np * 2;
        for (int j = 0; j < np; j++) {
            const size_t k = 2 * j;
            s[pb + k] -= at(rb + k);
            s[pb + k + 1] -= at(rb + k + 2);
        }
    }
    return ref;
}

/* Align/CPatchShifts.cpp:288-355 */
void detect_bads(const double* s, const double* cen, int nf, int np, int W, int H,
                 unsigned char* bad) {
    for (int f = 0; f < nf; f++) {
        const double* p = s + (size_t)f * np * 2;
        for (int i = 0; i < np; i++) {
            const double ds = std::sqrt(p[2 * i] * p[2 * i] + p[2 * i + 1] * p[2 * i + 1]);
            double rms = 0, sw = 0;
            for (int j = 0; j < np; j++) {
                if (j == i) continue;
                const double fx = (cen[2 * j] - cen[2 * i]) / W;
                const double fy = (cen[2 * j + 1] - cen[2 * i + 1]) / H;
                const double w = std::exp(-100.0 * (fx * fx + fy * fy));
                rms += w * (p[2 * j] * p[2 * j] + p[2 * j + 1] * p[2 * j + 1]);
                sw += w;
            }
            double lr = std::sqrt(rms / sw);
            if (lr < 2) lr = 2.0;
            bad[(size_t)f * np + i] = (ds > lr) ? 1 : 0;
        }
    }
}

/* Correct/GCorrectPatchShift.cu:17-66 (mGCorrect3D, upsample 1): weighted
 * patch shift with w = exp(-100 r), r <= 0.6, nearest gather with (int)
 * truncation and 811*|x| mod N reflection.  `wts`, when given, holds every
 * pixel's np weights (0 where r > 0.6, see warp_weights): the same values the
 * loop would compute, so the result is identical. */
void local_warp(const double* img, int W, int H, const double* cen, const double* sh,
                const unsigned char* bad, int np, double* out, const double* wts = nullptr) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            double sx = 0, sy = 0, w = 0;
            const double* wp = wts ? wts + ((size_t)y * W + x) * np : nullptr;
            for (int p = 0; p < np; p++) {
                if (bad && bad[p]) continue;
                double e;
                if (wp) {
                    e = wp[p];
                    if (e == 0.0) continue;
                } else {
                    double ax = (x - cen[2 * p]) / W;
                    double ay = (y - cen[2 * p + 1]) / H;
                    double r = std::sqrt(ax * ax + ay * ay);
                    if (r > 0.6) continue;
                    e = std::exp(-100.0 * r);
                }
                w += e;
                sx += sh[2 * p] * e;
                sy += sh[2 * p + 1] * e;
            }
            if (w > 0) { sx /= w; sy /= w; }
            /* The gather index is the FP32 difference of the reference kernel
             * ((int)((blockIdx.x - fSx) * iUpSample), :52-53): its rounding
             * (6e-5 px near x = 1000) decides which pixel is read. */
            int xs = (int)((float)x - (float)sx), ys = (int)((float)y - (float)sy);
            if (xs < 0 || ys < 0 || xs >= W || ys >= H) {
                xs = xs < 0 ? -xs : xs;
                ys = ys < 0 ? -ys : ys;
                xs = (int)((811LL * xs) % W);
                ys = (int)((811LL * ys) % H);
            }
            out[(size_t)y * W + x] = img[(size_t)ys * W + xs];
        }
}

/* The patch weights exp(-100 r) of every pixel (0 where r > 0.6): they do
 * not depend on the frame, so a movie computes them once. */
std::vector<double> warp_weights(int W, int H, const double* cen, int np) {
    std::vector<double> wts((size_t)W * H * np);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            for (int p = 0; p < np; p++) {
                double ax = (x - cen[2 * p]) / W;
                double ay = (y - cen[2 * p + 1]) / H;
                double r = std::sqrt(ax * ax + ay * ay);
                wts[((size_t)y * W + x) * np + p] = (r > 0.6) ? 0.0 : std::exp(-100.0 * r);
            }
    return wts;
}

}  // namespace

/* ====================================================================== */
/* extern "C" surface                                                      */
/* ====================================================================== */
extern "C" {

void orc_default_params(orc_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->iter = 15;
    p->tol = 0.1;
    p->bft[0] = 500.0;
    p->bft[1] = 100.0;
    p->kv = 300;
    p->group[0] = 1;
    p->group[1] = 4;
    p->fm_ref = -1;
    p->ft_bin = 1.0;
    p->sum_range[0] = 3.0;
    p->sum_range[1] = 25.0;
    p->ref_quirks = 1;
    p->out_stack[1] = 1;
    p->align = 1;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_geometry(int W, int H, int px, int py, int* sizes, double* xbin) {
    Geo g = make_geo(W, H, 1, px, py, 1.0);
    sizes[0] = g.xw;
    sizes[1] = g.xh;
    sizes[2] = g.pw;
    sizes[3] = g.ph;
    xbin[0] = g.xbin[0];
    xbin[1] = g.xbin[1];
}

void orc_fft_c2c(double* data, int n, int howmany, int sign) {
    cd* c = reinterpret_cast<cd*>(data);
    for (int i = 0; i < howmany; i++) fft1(c + (size_t)i * n, n, sign < 0 ? -1 : 1);
}

void orc_rfft2(const double* img, int W, int H, double* cmp) {
    rfft2(img, W, H, reinterpret_cast<cd*>(cmp));
}

void orc_irfft2(const double* cmp, int W, int H, double* img) {
    irfft2(reinterpret_cast<const cd*>(cmp), W, H, img);
}

void orc_phase_shift(const double* in, int cx, int cy, double sx, double sy, int acc, double* out) {
    phase_shift(reinterpret_cast<const cd*>(in), cx, cy, sx, sy, acc != 0, reinterpret_cast<cd*>(out));
}

void orc_fourier_resize(const double* in, int inx, int iny, double* out, int outx, int outy, int acc) {
    fourier_resize(reinterpret_cast<const cd*>(in), inx, iny, reinterpret_cast<cd*>(out), outx, outy,
                   acc != 0);
}

void orc_correlate_window(const double* sum, const double* grp, int cx, int cy, double b, int po,
                          int seax, int seay, double* win) {
    correlate_window(reinterpret_cast<const cd*>(sum), reinterpret_cast<const cd*>(grp), cx, cy, b,
                     po != 0, seax, seay, win);
}

void orc_peak2d(const double* img, int nx, int ny, double* shift) { peak2d(img, nx, ny, shift); }

int orc_lin_eqs(double* coeff, double* vals, int dim) { return lin_eqs(coeff, vals, dim) ? 1 : 0; }

void orc_remove_spikes(double* sx, double* sy, int n, int single, int quirks) {
    remove_spikes(sx, sy, n, single != 0, quirks != 0);
}

void orc_smooth(double* sx, double* sy, int n, double w) { smooth(sx, sy, n, w); }

void orc_dose_weights(int cx, int cy, double pix, int kv, const double* acc, int nf, double* wsum) {
    const double kvf = kv_factor(kv);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < cy; y++)
        for (int k = 0; k < cx; k++) {
            const size_t i = (size_t)y * cx + k;
            if (i == 0) { wsum[i] = 0.0; continue; }
            double s = 0;
            for (int f = 0; f < nf; f++) {
                const double w = dose_weight(k, y, cx, cy, pix, kvf, acc[f]);
                s += w * w;
            }
            wsum[i] = std::sqrt(s);
        }
}

void orc_make_relative(double* s, unsigned char* unused, int nf, int np, int quirks, int* ref) {
    (void)unused;
    int r = make_relative(s, nf, np, quirks != 0);
    if (ref) *ref = r;
}

void orc_detect_bads(const double* s, const double* c, int nf, int np, int W, int H,
                     unsigned char* bad) {
    detect_bads(s, c, nf, np, W, H, bad);
}

void orc_local_warp(const double* img, int W, int H, const double* c, const double* s,
                    const unsigned char* bad, int np, double* out) {
    local_warp(img, W, H, c, s, bad, np, out);
}

void orc_feature_map(const double* sum_cmp, int xw, int xh, unsigned char* feat) {
    feature_map_from_sum(reinterpret_cast<const cd*>(sum_cmp), xw / 2 + 1, xh, feat);
}

void orc_patch_centers(const unsigned char* feat, int px, int py, int xw, int xh, int* starts) {
    patch_starts_from_features(feat, px, py, xw, xh, starts);
}

void orc_norm_by_var(double* img, int w, int h) {
    std::vector<double> v(img, img + (size_t)w * h);
    norm_by_var(v, w, h, 5, 5);
    std::copy(v.begin(), v.end(), img);
}

void orc_feature_threshold(const double* img, int n, unsigned char* feat, double* thr) {
    feature_threshold(img, n, feat, thr);
}

void orc_interpolate(const double* gx, const double* gy, int binz, int n, const int* fmsize,
                     int nearest, double* ox, double* oy, int* ngroups) {
    std::vector<double> fmcent(n);
    int st = 0;
    for (int i = 0; i < n; i++) {
        const int sz = fmsize ? fmsize[i] : 1;
        fmcent[i] = st + 0.5 * (sz - 1.0);
        st += sz;
    }
    const Groups g = make_groups(binz, n, fmcent, fmsize);
    std::vector<double> vx(gx, gx + g.ng), vy(gy, gy + g.ng), rx, ry;
    interpolate(vx, vy, g, fmcent, nearest != 0, rx, ry);
    std::copy(rx.begin(), rx.end(), ox);
    std::copy(ry.begin(), ry.end(), oy);
    if (ngroups) *ngroups = g.ng;
}

/* Per-frequency weights applied to a corrected frame spectrum.
 * -CorrInterp (Util/GCorrLinearInterp.cu:7-24, patch mode only, in
 * Correct/CGenRealStack.cpp:132-138): divide by sinc(pi r)^(1/8), r the
 * frequency radius in cycles per pixel.  -InFmMotion
 * (MotionDecon/GMotionWeight.cu:14-40,62-74,96-110): for the frame's motion
 * m (CInFrameMotion::GetFullMotion, MotionDecon/CInFrameMotion.cpp:79-97),
 * multiply by sinc(pi |m| d / (2 (cx - 1))) with d the distance of (kx, ky)
 * from the line through the origin perpendicular to m (the reference's
 * slope form, kx itself when the slope exceeds 1e5), 1 when d < 0.1 and 0
 * once |m| d / (2 (cx - 1)) > 1; no weight when |m| < 0.1.  DC is untouched. */
double frame_weight(int k, int y, int cx, int cy, double mx, double my, bool motion, bool corr) {
    if (k == 0 && y == 0) return 1.0;
    double w = 1.0;
    if (corr) {
        double fx = k * 0.5 / (cx - 1), fy = (double)y / cy;
        if (fy > 0.5) fy -= 1.0;
        const double r = std::sqrt(fx * fx + fy * fy) * M_PI;
        w /= std::pow(std::sin(r) / r, 0.125);
    }
    const double ms = std::sqrt(mx * mx + my * my);
    if (motion && !(ms < 0.1)) {
        const double slope = -mx / (my + 1e-20);
        double d;
        if (std::fabs(slope) > 1e5) d = k;
        else {
            const int yy = (y > cy / 2) ? y - cy : y;
            d = std::fabs(slope * k - yy) / std::sqrt(slope * slope + 1.0);
        }
        if (!(d < 0.1)) {
            d /= 2.0 * (cx - 1);
            if (ms * d > 1.0) w *= 0.0;
            else w *= std::sin(M_PI * ms * d) / (M_PI * ms * d);
        }
    }
    return w;
}

/* CInFrameMotion::GetFullMotion: the shift to the next frame, the last
 * frame's to the one before (float differences of the stored shifts). */
void frame_motion(const std::vector<double>& gx, const std::vector<double>& gy, int i, double* m) {
    const int n = (int)gx.size();
    int a = i, b = i + 1;
    if (i == n - 1) {
        a = i - 1;
        b = i;
    }
    m[0] = (float)gx[b] - (float)gx[a];
    m[1] = (float)gy[b] - (float)gy[a];
}

void weigh_frame(cd* f, int cx, int cy, const double* m, bool motion, bool corr) {
    if (!motion && !corr) return;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < cy; y++)
        for (int k = 0; k < cx; k++) f[(size_t)y * cx + k] *= frame_weight(k, y, cx, cy, m[0], m[1], motion, corr);
}

void orc_frame_weight(int cx, int cy, double mx, double my, int motion, int corr, double* w) {
    for (int y = 0; y < cy; y++)
        for (int k = 0; k < cx; k++)
            w[(size_t)y * cx + k] = frame_weight(k, y, cx, cy, mx, my, motion != 0, corr != 0);
}

/* Sum buffers of Correct/CCorrectFullShift.cpp:187-300 (mInit, mGenSums,
 * mDoseWeight, mSum) and the output step :146-185 (mUnpadSums). */
struct SumBank {
    int cx, cy, nsum;
    bool dw, dws, split;
    double pix, kvf, r0, r1;
    const double* accdose;
    std::vector<double> wsum;
    std::vector<cd> sums;
    const Geo* geo = nullptr;
    double* stack = nullptr;  // aligned stack, one output-size frame per frame
    /* every loop below is element-wise over the spectrum: the frames are
     * still added in frame order, element by element */
    void add(cd* f, int i) {
        const size_t fs = (size_t)cx * cy;
        int ns = 0;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < fs; q++) sums[q] += f[q];
        ns++;
        if (stack) {
            /* Correct/CCorrectFullShift.cpp:336-356 (mCropFrame): the
             * corrected frame before dose weighting, cropped to the -FtBin
             * size, C2R, unpadded; no positivity clamp */
            std::vector<cd> ob((size_t)geo->ocx * geo->ocy, cd(0, 0));
            fourier_resize(f, cx, cy, ob.data(), geo->ocx, geo->ocy, false);
            irfft2(ob.data(), geo->ow, geo->oh, stack + (size_t)i * geo->ow * geo->oh);
        }
        if (dw) {
#pragma omp parallel for schedule(static)
            for (int y = 0; y < cy; y++)
                for (int k = 0; k < cx; k++) {
                    const size_t q = (size_t)y * cx + k;
                    if (q == 0) continue;
                    f[q] *= dose_weight(k, y, cx, cy, pix, kvf, accdose[i]) / wsum[q];
                }
            cd* s1 = &sums[(size_t)ns * fs];
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < fs; q++) s1[q] += f[q];
            ns++;
            if (dws) {
                if (!(accdose[i] < r0) && !(accdose[i] > r1)) {
                    cd* s2 = &sums[(size_t)ns * fs];
#pragma omp parallel for schedule(static)
                    for (size_t q = 0; q < fs; q++) s2[q] += f[q];
                }
                ns++;
            }
        }
        if (split) {
            cd* sp = &sums[(size_t)(ns + (i % 2)) * fs];
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < fs; q++) sp[q] += f[q];
        }
    }
    void finish(const Geo& g, double* out) {
        std::vector<cd> ob((size_t)g.ocx * g.ocy);
        std::vector<double> oimg((size_t)g.ow * g.oh);
        for (int s = 0; s < nsum; s++) {
            std::fill(ob.begin(), ob.end(), cd(0, 0));
            fourier_resize(&sums[(size_t)s * cx * cy], cx, cy, ob.data(), g.ocx, g.ocy, false);
            irfft2(ob.data(), g.ow, g.oh, oimg.data());
            for (size_t q = 0; q < oimg.size(); q++)
                out[(size_t)s * g.ow * g.oh + q] = std::max(0.0, oimg[q]);
        }
    }
};

/* Align/CZbinStack.cpp:27-70: the aligned stack summed over the frame
 * groups of CFmGroupParam::Setup(zbin) in frame order, in place; returns the
 * number of stack frames. */
int zbin_stack(double* stack, int N, size_t npix, int zbin, const std::vector<double>& fmcent,
               const int* fmsize) {
    if (zbin <= 1) return N;
    const Groups gr = make_groups(zbin, N, fmcent, fmsize);
    for (int k = 0; k < gr.ng; k++) {
        double* dst = stack + (size_t)k * npix;
        const double* src = stack + (size_t)gr.start[k] * npix;
        if (dst != src) std::copy(src, src + npix, dst);
        for (int j = 1; j < gr.size[k]; j++) {
            const double* fj = stack + (size_t)(gr.start[k] + j) * npix;
            for (size_t q = 0; q < npix; q++) dst[q] += fj[q];
        }
    }
    return gr.ng;
}

/* -Align 0: Align/CSimpleSum.cpp:22-135 with MrcUtil/CSumFFTStack.cpp:60-210.
 * The frames are summed as they are (all frames; with -SplitSum also the
 * even-indexed frames, and all minus those), and every sum and every stack
 * frame is cropped to the -FtBin size through a normalised R2C and a C2R,
 * with no clamp at zero. */
int simple_sum(const float* frames, int W, int H, int N, const orc_params* prm, orc_out* out,
               const std::vector<double>& fmcent) {
    const Geo g = make_geo(W, H, N, 0, 0, prm->ft_bin);
    const bool split = prm->split_sum != 0;
    const size_t np = (size_t)W * H, onp = (size_t)g.ow * g.oh;
    std::vector<double> s((split ? 3 : 1) * np, 0.0);
    for (int i = 0; i < N; i++)
        for (size_t q = 0; q < np; q++) {
            const double v = frames[(size_t)i * np + q];
            s[q] += v;
            if (split && i % 2 == 0) s[np + q] += v;
        }
    if (split)
        for (size_t q = 0; q < np; q++) s[2 * np + q] = s[q] - s[np + q];
    std::vector<cd> spec((size_t)g.cx * g.cy), ob((size_t)g.ocx * g.ocy);
    const double sc = 1.0 / ((double)W * H);
    auto crop = [&](const double* img, double* dst) {
        rfft2(img, W, H, spec.data());
        for (auto& v : spec) v *= sc;
        std::fill(ob.begin(), ob.end(), cd(0, 0));
        fourier_resize(spec.data(), g.cx, g.cy, ob.data(), g.ocx, g.ocy, false);
        irfft2(ob.data(), g.ow, g.oh, dst);
    };
    out->out_w = g.ow;
    out->out_h = g.oh;
    out->num_sums = split ? 3 : 1;
    for (int k = 0; k < out->num_sums; k++) crop(&s[k * np], out->sums + k * onp);
    if (out->global_iters) out->global_iters[0] = 0;
    if (out->global_shift) std::fill(out->global_shift, out->global_shift + 2 * N, 0.0);
    out->stack_frames = 0;
    if (prm->out_stack[0] && out->stack) {
        std::vector<double> img(np);
        for (int i = 0; i < N; i++) {
            for (size_t q = 0; q < np; q++) img[q] = frames[(size_t)i * np + q];
            crop(img.data(), out->stack + i * onp);
        }
        out->stack_frames = zbin_stack(out->stack, N, onp, prm->out_stack[1], fmcent, prm->fm_int_size);
    }
    return 0;
}

/* Stage wall times on stderr when ORC_TIMING is set (where the oracle's
 * time goes; no effect on results). */
struct StageClock {
    bool on = std::getenv("ORC_TIMING") != nullptr;
    double t0 = omp_get_wtime();
    void mark(const char* what) {
        if (!on) return;
        const double t = omp_get_wtime();
        std::fprintf(stderr, "oracle %-28s %8.2f s\n", what, t - t0);
        t0 = t;
    }
};

/* Align::CAlignMain::DoIt (Align/CAlignMain.cpp:39-75) for full-frame
 * (Align/CFullAlign.cpp:70-171) and patch (Align/CPatchAlign.cpp:25-92)
 * modes, with the sums of CCorrectFullShift / GCorrectPatchShift. */
int orc_align_movie(const float* frames, int W, int H, int N, const orc_params* prm,
                    const orc_override* ov, orc_out* out) {
    if (W % 2 || W < 8 || H < 8 || N < 2) return -1;
    StageClock clk;
    orc_override none;
    std::memset(&none, 0, sizeof(none));
    if (!ov) ov = &none;
    const bool quirks = prm->ref_quirks != 0;
    const bool patch_mode = prm->patch[0] > 1 && prm->patch[1] > 1;
    const Geo g = make_geo(W, H, N, prm->patch[0], prm->patch[1], prm->ft_bin);
    if (patch_mode && (g.pw > g.xw / 2 || g.ph > g.xh / 2)) return -2;
    const int np = patch_mode ? prm->patch[0] * prm->patch[1] : 0;
    if (patch_mode && ov->replay_global && (!ov->replay_patch || !ov->replay_centers)) return -3;

    /* DataUtil/CFmIntParam.cpp:80-135,280-287 */
    std::vector<double> fmcent(N), accdose(N);
    for (int i = 0; i < N; i++) {
        const int st = prm->fm_int_start ? prm->fm_int_start[i] : i + prm->throw0;
        const int sz = prm->fm_int_size ? prm->fm_int_size[i] : 1;
        fmcent[i] = st + 0.5 * (sz - 1.0);
        /* accumulated dose: FmDose (i+1) without integration (mSetup),
         * the raw-frame dose times the raw frames up to and including this
         * frame with it (mCalcIntFms, mThrowIntFrames) */
        accdose[i] = prm->fm_int_size ? prm->fm_dose * (st + sz) : prm->fm_dose * (i + 1);
    }
    if (prm->align == 0) return simple_sum(frames, W, H, N, prm, out, fmcent);
    SumBank bank;
    bank.cx = g.cx;
    bank.cy = g.cy;
    bank.dw = !(prm->kv < 80 || prm->kv > 300) && prm->pix_size > 0 && prm->fm_dose > 0;
    bank.dws = bank.dw && !(prm->sum_range[0] < 0) && !(prm->sum_range[1] <= 0) &&
               (prm->sum_range[1] - prm->sum_range[0] > 1.0);
    bank.split = prm->split_sum != 0;
    bank.nsum = 1 + (bank.split ? 2 : 0) + (bank.dw ? 1 + (bank.dws ? 1 : 0) : 0);
    bank.pix = prm->pix_size;
    bank.kvf = kv_factor(prm->kv);
    bank.r0 = prm->sum_range[0];
    bank.r1 = prm->sum_range[1];
    bank.accdose = accdose.data();
    bank.geo = &g;
    bank.stack = prm->out_stack[0] ? out->stack : nullptr;
    out->stack_frames = 0;
    auto finish_stack = [&]() {
        if (bank.stack)
            out->stack_frames = zbin_stack(bank.stack, N, (size_t)g.ow * g.oh, prm->out_stack[1], fmcent,
                                           prm->fm_int_size);
    };
    out->out_w = g.ow;
    out->out_h = g.oh;
    out->num_sums = bank.nsum;

    /* Align/CTransformStack.cpp:26-135 + Util/CCufft2D.cpp:65-80 */
    Stack frm;
    frm.alloc(g.cx, g.cy, N);
    {
        std::vector<double> img((size_t)W * H);
        const double sc = 1.0 / ((double)W * H);
        for (int i = 0; i < N; i++) {
            for (size_t q = 0; q < (size_t)W * H; q++) img[q] = frames[(size_t)i * W * H + q];
            rfft2(img.data(), W, H, frm.frame(i));
            cd* f = frm.frame(i);
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < (size_t)g.cx * g.cy; q++) f[q] *= sc;
        }
    }
    clk.mark("forward R2C");
    const int ref_frame = (prm->fm_ref < 0) ? N / 2 : std::min(std::max(prm->fm_ref - 1, 0), N - 1);
    const double tol = prm->tol / std::max(g.xbin[0], g.xbin[1]);

    std::vector<double> gx(N, 0.0), gy(N, 0.0);
    std::vector<double> xreal;
    std::vector<int> starts(2 * np, 0);
    if (ov->replay_global) {
        for (int i = 0; i < N; i++) { gx[i] = ov->replay_global[2 * i]; gy[i] = ov->replay_global[2 * i + 1]; }
        if (out->global_iters) out->global_iters[0] = 0;
    } else {
        /* Align/CGenXcfStack.cpp:22-107 */
        Stack xcf;
        xcf.alloc(g.xcx, g.xcy, N);
        for (int i = 0; i < N; i++)
            fourier_resize(frm.frame(i), g.cx, g.cy, xcf.frame(i), g.xcx, g.xcy, false);
        if (ov->xcf_shift) {
            for (int i = 0; i < N; i++) { gx[i] = ov->xcf_shift[2 * i]; gy[i] = ov->xcf_shift[2 * i + 1]; }
            if (out->global_iters) out->global_iters[0] = 0;
        } else {
            Groups grp0 = make_groups(prm->group[0], N, fmcent, prm->fm_int_size);
            IterResult r = iterative_align(xcf, grp0, fmcent, prm->bft[0], prm->phase_only != 0,
                                           prm->iter, tol, quirks);
            gx = r.sx;
            gy = r.sy;
            clk.mark("global alignment");
            if (out->global_iters) out->global_iters[0] = r.iters;
            if (out->global_errors)
                for (int i = 0; i < r.iters; i++) out->global_errors[i] = r.errors[i];
        }
        if (out->xcf_shift)
            for (int i = 0; i < N; i++) { out->xcf_shift[2 * i] = gx[i]; out->xcf_shift[2 * i + 1] = gy[i]; }
        if (patch_mode) {
            /* Align/CFullAlign.cpp:147-159: features, centres, real xcf */
            std::vector<cd> sum((size_t)g.xcx * g.xcy, cd(0, 0));
            for (int i = 0; i < N; i++)
                phase_shift(xcf.frame(i), g.xcx, g.xcy, -gx[i], -gy[i], true, sum.data());
            std::vector<unsigned char> feat(128 * 128);
            feature_map_from_sum(sum.data(), g.xcx, g.xcy, feat.data());
            if (out->feature_map) std::copy(feat.begin(), feat.end(), out->feature_map);
            if (ov->patch_starts) std::copy(ov->patch_starts, ov->patch_starts + 2 * np, starts.begin());
            else patch_starts_from_features(feat.data(), prm->patch[0], prm->patch[1], g.xw, g.xh,
                                            starts.data());
            if (out->patch_starts) std::copy(starts.begin(), starts.end(), out->patch_starts);
            clk.mark("features");
            xreal.assign((size_t)N * g.xw * g.xh, 0.0);
            for (int i = 0; i < N; i++) {
                phase_shift_skip(xcf.frame(i), g.xcx, g.xcy, -gx[i], -gy[i]);
                irfft2(xcf.frame(i), g.xw, g.xh, &xreal[(size_t)i * g.xw * g.xh]);
            }
        }
        /* Align/CFullAlign.cpp:161-163 and :96 (CStackShift::TruncateDecimal) */
        for (int i = 0; i < N; i++) {
            gx[i] *= (float)g.xbin[0];
            gy[i] *= (float)g.xbin[1];
            gx[i] = 0.01 * (double)(long long)(gx[i] * 100.0);
            gy[i] = 0.01 * (double)(long long)(gy[i] * 100.0);
        }
    }

    if (bank.dw) {
        bank.wsum.resize((size_t)g.cx * g.cy);
        orc_dose_weights(g.cx, g.cy, prm->pix_size, prm->kv, accdose.data(), N, bank.wsum.data());
    }
    bank.sums.assign((size_t)bank.nsum * g.cx * g.cy, cd(0, 0));
    const double sc = 1.0 / ((double)W * H);

    if (!patch_mode) {
        /* Correct/CCorrectFullShift.cpp:229-309 with absolute shifts */
        if (out->global_shift)
            for (int i = 0; i < N; i++) { out->global_shift[2 * i] = gx[i]; out->global_shift[2 * i + 1] = gy[i]; }
        for (int i = 0; i < N; i++) {
            phase_shift_skip(frm.frame(i), g.cx, g.cy, -gx[i], -gy[i]);
            /* mGenSums' mMotionDecon (Correct/CCorrectFullShift.cpp:271-320) */
            double m[2];
            frame_motion(gx, gy, i, m);
            weigh_frame(frm.frame(i), g.cx, g.cy, m, prm->in_fm_motion != 0, false);
            bank.add(frm.frame(i), i);
        }
        bank.finish(g, out->sums);
        finish_stack();
        return 0;
    }

    /* Align/CPatchAlign.cpp:47-65: shifts relative to the reference frame,
     * frame spectra shifted in place (CGenRealStack, no inverse FFT) */
    {
        const double rx = gx[ref_frame], ry = gy[ref_frame];
        for (int i = 0; i < N; i++) { gx[i] -= rx; gy[i] -= ry; }
        for (int i = 0; i < N; i++) {
            phase_shift_skip(frm.frame(i), g.cx, g.cy, -gx[i], -gy[i]);
            /* CGenRealStack::mCorrectBilinear, mMotionDecon (:132-166) */
            double m[2];
            frame_motion(gx, gy, i, m);
            weigh_frame(frm.frame(i), g.cx, g.cy, m, prm->in_fm_motion != 0, prm->corr_interp != 0);
        }
    }
    if (out->global_shift)
        for (int i = 0; i < N; i++) { out->global_shift[2 * i] = gx[i]; out->global_shift[2 * i + 1] = gy[i]; }

    std::vector<double> ps((size_t)N * np * 2, 0.0), pc(2 * np, 0.0);
    std::vector<unsigned char> bad((size_t)N * np, 0);
    if (ov->replay_global) {
        ps.assign(ov->replay_patch, ov->replay_patch + (size_t)N * np * 2);
        pc.assign(ov->replay_centers, ov->replay_centers + 2 * np);
        if (ov->replay_bad) bad.assign(ov->replay_bad, ov->replay_bad + (size_t)N * np);
    } else {
        /* Align/CMeasurePatches.cpp:22-83 with CExtractPatch
         * (Align/CExtractPatch.cpp:88-106) and an unnormalised patch R2C */
        Groups grp1 = make_groups(prm->group[1], N, fmcent, prm->fm_int_size);
        for (int p = 0; p < np; p++) {
            const int sx0 = starts[2 * p], sy0 = starts[2 * p + 1];
            Stack pat;
            pat.alloc(g.pcx, g.ph, N);
#pragma omp parallel
            {
                std::vector<double> win((size_t)g.pw * g.ph);
#pragma omp for schedule(static)
                for (int i = 0; i < N; i++) {
                    const double* src = &xreal[(size_t)i * g.xw * g.xh];
                    for (int y = 0; y < g.ph; y++)
                        for (int x = 0; x < g.pw; x++)
                            win[(size_t)y * g.pw + x] = src[(size_t)(sy0 + y) * g.xw + sx0 + x];
                    rfft2(win.data(), g.pw, g.ph, pat.frame(i));
                }
            }
            IterResult r = iterative_align(pat, grp1, fmcent, prm->bft[1], prm->phase_only != 0,
                                           prm->iter, tol, quirks);
            if (out->patch_iters) out->patch_iters[p] = r.iters;
            const int cxp = sx0 + g.pw / 2, cyp = sy0 + g.ph / 2;
            pc[2 * p] = cxp * (float)g.xbin[0];
            pc[2 * p + 1] = cyp * (float)g.xbin[1];
            for (int i = 0; i < N; i++) {
                ps[((size_t)i * np + p) * 2] = r.sx[i] * (float)g.xbin[0];
                ps[((size_t)i * np + p) * 2 + 1] = r.sy[i] * (float)g.xbin[1];
            }
        }
        clk.mark("patch alignment");
        xreal.clear();
        xreal.shrink_to_fit();
        make_relative(ps.data(), N, np, quirks);
        detect_bads(ps.data(), pc.data(), N, np, W, H, bad.data());
    }
    if (out->patch_shift) std::copy(ps.begin(), ps.end(), out->patch_shift);
    if (out->patch_bad) std::copy(bad.begin(), bad.end(), out->patch_bad);
    if (out->patch_center) std::copy(pc.begin(), pc.end(), out->patch_center);

    /* Correct/GCorrectPatchShift.cu:137-285: C2R, warp, normalised R2C, sums */
    std::vector<double> img((size_t)W * H), wimg((size_t)W * H);
    /* frame-independent weights, kept when they take at most 8 GB */
    std::vector<double> wts;
    if ((double)W * H * np * 8.0 <= 8e9) wts = warp_weights(W, H, pc.data(), np);
    double t_c2r = 0, t_warp = 0, t_r2c = 0;
    for (int i = 0; i < N; i++) {
        cd* f = frm.frame(i);
        double ta = omp_get_wtime();
        irfft2(f, W, H, img.data());
        double tb = omp_get_wtime();
        local_warp(img.data(), W, H, pc.data(), &ps[(size_t)i * np * 2], &bad[(size_t)i * np], np,
                   wimg.data(), wts.empty() ? nullptr : wts.data());
        double tc = omp_get_wtime();
        rfft2(wimg.data(), W, H, f);
        t_c2r += tb - ta;
        t_warp += tc - tb;
        t_r2c += omp_get_wtime() - tc;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < (size_t)g.cx * g.cy; q++) f[q] *= sc;
        /* mGenSums' mMotionDecon with the same full shifts
         * (Correct/GCorrectPatchShift.cu:106): applied a second time */
        double m[2];
        frame_motion(gx, gy, i, m);
        weigh_frame(f, g.cx, g.cy, m, prm->in_fm_motion != 0, false);
        bank.add(f, i);
    }
    if (clk.on) std::fprintf(stderr, "oracle   C2R %.2f s, warp %.2f s, R2C %.2f s\n", t_c2r, t_warp, t_r2c);
    clk.mark("local correction + sums");
    bank.finish(g, out->sums);
    clk.mark("output");
    finish_stack();
    return 0;
}

}  // extern "C"
