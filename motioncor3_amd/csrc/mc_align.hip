// mc_align.hip -- alignment, correction and summation kernels for gfx950.
//
// The reference runs one CIterativeAlign at a time with ~100 launches and a
// device->host round trip per iteration (Align/CAlignStack.cpp:103-116) and
// the peak search on the CPU (Align/CPeak2D.cpp).  Here every job (the
// global alignment, or all patches at once) keeps its trajectory state in
// device memory in FP64; one iteration is six launches over all jobs and
// groups, and a converged job's work items exit on entry.
#include <algorithm>
#include <cmath>

#include "mc_align.h"
#include "mc_fft_rp.h"
#include "mc_spec.h"

#define MC_WARP_MAXP 64   // patches of the local-motion kernels
#define MC_PS_MAXN 4096   // frames of the patch-statistics kernel

#define MC_PI 3.14159265358979323846

// ---------------------------------------------------------------------------
// Phase factors of GPhaseShift2D (Util/GPhaseShift2D.cu:12-66): the factor
// e^{i(kx*sx' + ky*sy')} separates into A[kx] * B[ky]; both are formed in
// FP64 per transform.  skip != 0 applies the |s| < 0.1 px rule of
// Util/GPhaseShift2D.cu:68-79 (unit factors).
__global__ void k_phasors(const double* shifts, double mult, int skip, const unsigned char* done,
                          int per_job, int W, int cx, int H, int hn, cfloat* A, cfloat* B) {
    const int b = blockIdx.y;
    if (done && done[b / per_job]) return;
    double sx = shifts[2 * b] * mult, sy = shifts[2 * b + 1] * mult;
    if (skip && sqrt(sx * sx + sy * sy) < 0.1) sx = sy = 0.0;
    const double fsx = sx * (2.0 * MC_PI / W), fsy = sy * (2.0 * MC_PI / hn);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cx) {
        double sn, cs;
        sincos(e * fsx, &sn, &cs);
        A[(long long)b * cx + e] = cmk((float)cs, (float)sn);
    } else if (e < cx + H) {
        const int y = e - cx;
        const int ys = (y > H / 2) ? y - H : y;
        double sn, cs;
        sincos(ys * fsy, &sn, &cs);
        B[(long long)b * H + y] = cmk((float)cs, (float)sn);
    }
}

// The same phase factors for the running shifts of an iteration batch whose
// jobs differ in size: frame b of the batch belongs to job fjob[b] (its row
// length and A rows from the job table), factors e^{-i(...)} (mult -1), no
// skip rule.  Grid x covers the widest job's cx + H.
__global__ void k_phasors_jobs(const double* shifts, const unsigned char* done, const JobGeo* jobs,
                               const int* fjob, int H, cfloat* A, cfloat* B) {
    const int b = blockIdx.y;
    const int j = fjob[b];
    if (done && done[j]) return;
    const JobGeo& g = jobs[j];
    const int cx = g.cx;
    const double sx = -shifts[2 * b], sy = -shifts[2 * b + 1];
    const double fsx = sx * (2.0 * MC_PI / g.W), fsy = sy * (2.0 * MC_PI / H);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cx) {
        double sn, cs;
        sincos(e * fsx, &sn, &cs);
        A[g.ao + (long long)(b - g.fo) * cx + e] = cmk((float)cs, (float)sn);
    } else if (e < cx + H) {
        const int y = e - cx;
        const int ys = (y > H / 2) ? y - H : y;
        double sn, cs;
        sincos(ys * fsy, &sn, &cs);
        B[(long long)b * H + y] = cmk((float)cs, (float)sn);
    }
}

// GFourierResize2D (Util/GFourierResize2D.cu:7-39) on [kx][y] half spectra:
// output row yo reads input row yo (yo <= outy/2) or yo - outy + iny (the
// negative frequencies), when that row exists in the input; columns beyond
// the input's are not written, nor rows without a source.  Consecutive lanes
// take consecutive rows of one column, so both sides are contiguous.
__global__ __launch_bounds__(256) void k_fourier_resize(const cfloat* __restrict__ in, int inx, int iny,
                                                        long long in_stride, cfloat* __restrict__ out,
                                                        int outx, int outy, long long out_stride,
                                                        int sum) {
    const int yo = blockIdx.x * blockDim.x + threadIdx.x;
    const int kx = blockIdx.y, b = blockIdx.z;
    if (yo >= outy || kx >= inx) return;
    int y = yo;
    if (y > outy / 2) {
        y -= outy;
        if (y <= -iny / 2) return;
        y += iny;
    } else if (y > iny / 2) {
        return;
    }
    const cfloat v = in[(long long)b * in_stride + (long long)kx * iny + y];
    cfloat* q = out + (long long)b * out_stride + (long long)kx * outy + yo;
    if (sum) {
        const cfloat o = *q;
        *q = cmk(o.x + v.x, o.y + v.y);
    } else {
        *q = v;
    }
}

// Byte fill of device memory by a kernel (graph-capturable, unlike a memset
// node).
__global__ __launch_bounds__(256) void k_fill_u8(unsigned char* p, long long n, unsigned char v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t mc_launch_fill_u8(unsigned char* p, long long n, unsigned char v, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t mc_launch_fourier_resize(const cfloat* in, int inx, int iny, long long in_stride, cfloat* out,
                                    int outx, int outy, long long out_stride, int batch, int sum,
                                    hipStream_t s) {
    dim3 grid((outy + 255) / 256, std::min(inx, outx), batch);
    hipLaunchKernelGGL(k_fourier_resize, grid, dim3(256), 0, s, in, inx, iny, in_stride, out, outx, outy,
                       out_stride, sum);
    return hipGetLastError();
}

// Per-frequency weights of a corrected frame spectrum [kx][y], in FP32 as the
// reference computes them.  corr_interp: -CorrInterp (Util/GCorrLinearInterp.cu:7-24),
// divide by sinc(pi r)^(1/8).  motion: -InFmMotion
// (MotionDecon/GMotionWeight.cu:14-40,62-74,96-110), the frame's motion is
// its shift to the next frame (the last frame's to the one before,
// MotionDecon/CInFrameMotion.cpp:79-97) and the weight a sinc of the
// distance to the line through the origin perpendicular to it.  DC is
// untouched.
__global__ __launch_bounds__(256) void k_frame_weights(cfloat* S, long long fs, int N, int cx, int H,
                                                       const double* shifts, int motion, int corr_interp) {
    const int kx = blockIdx.y, f = blockIdx.z;
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= H || (kx == 0 && y == 0)) return;
    cfloat* q = S + (long long)f * fs + (long long)kx * H + y;
    cfloat v = *q;
    if (corr_interp) {
        float fx = kx * 0.5f / (cx - 1);
        float fy = y / (float)H;
        if (fy > 0.5f) fy -= 1.0f;
        fx = sqrtf(fx * fx + fy * fy) * 3.141592654f;
        fx = sinf(fx) / fx;
        fx = powf(fx, 0.125f);
        v.x /= fx;
        v.y /= fx;
    }
    if (motion) {
        int a = f, b = f + 1;
        if (f == N - 1) {
            a = f - 1;
            b = f;
        }
        const float mx = (float)shifts[2 * b] - (float)shifts[2 * a];
        const float my = (float)shifts[2 * b + 1] - (float)shifts[2 * a + 1];
        const float ms = sqrtf(mx * mx + my * my);
        if (!((double)ms < 0.1) && !(ms < 0.1f)) {
            const float slope = -mx / (my + 1e-20f);
            float dist;
            if (fabsf(slope) > 1e5f) {
                dist = (float)kx;
            } else {
                const int yy = (y > H / 2) ? y - H : y;
                dist = fabsf(slope * kx - yy) / sqrtf(slope * slope + 1.0f);
            }
            if (!(dist < 0.1f)) {
                dist = dist / (float)(2 * (cx - 1));
                float w = 0.0f;
                if (!(ms * dist > 1.0f)) {
                    const float a_ = 3.141592654f * ms * dist;
                    w = sinf(a_) / a_;
                }
                v.x *= w;
                v.y *= w;
            }
        }
    }
    *q = v;
}

// CAlignedSum (Align/CAlignedSum.cpp:25-126): S = sum_i shift(F_i). Even and
// odd frames accumulate separately and are added at the end, as the
// reference's two streams do (:99-112).
// With a job table the jobs may differ in frame count and row length
// (grid y covers the widest); without one they are J jobs of N frames.
template <int FF>
__global__ void k_aligned_sum(const cfloat* __restrict__ F, const float* __restrict__ finv,
                              const cfloat* __restrict__ A, const cfloat* __restrict__ B,
                              const unsigned char* done, int N, int cx, int H, cfloat* S,
                              const JobGeo* __restrict__ jobs) {
    const int j = blockIdx.z;
    if (done && done[j]) return;
    long long ao = (long long)j * N * cx, bo = (long long)j * N * H, so = (long long)j * cx * H;
    if (jobs) {
        const JobGeo& g = jobs[j];
        F = g.F;
        finv = g.finv;
        N = g.N;
        cx = g.cx;
        ao = g.ao;
        bo = (long long)g.fo * H;
        so = g.so;
    } else {
        F = reinterpret_cast<const cfloat*>(reinterpret_cast<const char*>(F) +
                                            (long long)j * N * cx * H * Spec<FF>::bytes);
        if (finv) finv += (long long)j * N;
    }
    const long long fs = (long long)cx * H;
    const int kx = blockIdx.y;
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= H || kx >= cx) return;
    const long long q = (long long)kx * H + y;
    cfloat a0 = cmk(0.f, 0.f), a1 = cmk(0.f, 0.f);
    // unrolled so that several frames' loads are in flight at once (the
    // sums still run in frame order)
#pragma unroll 4
    for (int i = 0; i < N; i++) {
        const cfloat f = Spec<FF>::ld(F, i * fs + q, FF == MC_SPEC_F32 ? 1.f : finv[i]);
        const cfloat v = cmul(f, cmul(A[ao + (long long)i * cx + kx], B[bo + (long long)i * H + y]));
        if (i & 1) a1 = cadd(a1, v);
        else a0 = cadd(a0, v);
    }
    S[so + q] = cadd(a0, a1);
}

// GCorrelateSum2D for one column block of one group of one job
// (Align/GCorrelateSum2D.cu:13-73 + the group sum of Align/CAlignStack.cpp:
// 145-178): G = sum_{i in group} shift(F_i), X = conj(G)(S - G) * B-filter
// (optionally / |X|), DC = 0, then the inverse transform along ky is done in
// LDS and only the rows of the sea_y lags around zero are kept.  The
// reference's (-1)^(kx+ky) centring and centre crop of the C2R output are
// the same as reading lags -sea/2 .. sea/2-1 of the uncentred transform.
struct CorrArgs {
    FftPlan1 plan;
    const cfloat* F;
    const cfloat* S;
    const cfloat* A;
    const cfloat* B;
    const int* gstart;
    const int* gsize;
    const unsigned char* done;
    int N, G, cx, H, Cc;
    float filt0;
    int phase_only, seay;
    cfloat* T1;
    const float* finv;   // load factors of a reduced F format (per transform)
    const JobGeo* jobs;  // job table (jobs of different sizes), or null
    int J;
    float nbq;           // -0.25 B: a job's filter constant is nbq / ((cx - 1) H)
};

// One job's view of the correlation arguments: its frames, phasor rows,
// sum, output rows and group table.
struct CorrJob {
    const cfloat* F;
    const float* finv;
    const cfloat* A;
    const cfloat* B;
    const cfloat* S;
    cfloat* T1;
    const int* gstart;
    const int* gsize;
    int N, G, cx;
    float filt0;
};

__device__ __forceinline__ CorrJob corr_job(const CorrArgs& a, int j, int H, int fbytes) {
    CorrJob r;
    if (a.jobs) {
        const JobGeo& g = a.jobs[j];
        r.F = g.F;
        r.finv = g.finv;
        r.A = a.A + g.ao;
        r.B = a.B + (long long)g.fo * H;
        r.S = a.S + g.so;
        r.T1 = a.T1 + g.t1o;
        r.gstart = a.gstart + g.gt;
        r.gsize = a.gsize + g.gt;
        r.N = g.N;
        r.G = g.G;
        r.cx = g.cx;
        r.filt0 = a.nbq / (float)((g.cx - 1) * H);
    } else {
        const long long fs = (long long)a.cx * H;
        r.F = reinterpret_cast<const cfloat*>(reinterpret_cast<const char*>(a.F) +
                                              (long long)j * a.N * fs * fbytes);
        r.finv = a.finv ? a.finv + (long long)j * a.N : nullptr;
        r.A = a.A + (long long)j * a.N * a.cx;
        r.B = a.B + (long long)j * a.N * H;
        r.S = a.S + (long long)j * fs;
        r.T1 = a.T1 + (long long)j * a.G * a.seay * a.cx;
        r.gstart = a.gstart;
        r.gsize = a.gsize;
        r.N = a.N;
        r.G = a.G;
        r.cx = a.cx;
        r.filt0 = a.filt0;
    }
    return r;
}

// Grid: (column blocks of the widest job, groups of the job with most, jobs).
__global__ __launch_bounds__(MC_NTHR) void k_correlate(CorrArgs a) {
    extern __shared__ cfloat lds[];
    const int j = blockIdx.z, g = blockIdx.y;
    if (a.done && a.done[j]) return;
    const int H = a.H;
    const CorrJob r = corr_job(a, j, H, 8);
    const int c0 = blockIdx.x * a.Cc;
    if (g >= r.G || c0 >= r.cx) return;
    const int nc = min(a.Cc, r.cx - c0);
    const int gs = r.gstart[g], gn = r.gsize[g];
    const long long fs = (long long)r.cx * H;
    const float rH = 1.0f / H;
    for (int e = threadIdx.x; e < nc * H; e += MC_NTHR) {
        const int c = fdiv(e, H, rH);
        const int y = e - c * H;
        const int kx = c0 + c;
        const long long q = (long long)kx * H + y;
        cfloat G = cmk(0.f, 0.f);
        for (int t = 0; t < gn; t++) {
            const long long b = gs + t;
            G = cadd(G, cmul(r.F[b * fs + q], cmul(r.A[b * r.cx + kx], r.B[b * H + y])));
        }
        const cfloat sd = csub(r.S[q], G);
        cfloat X = cmk(G.x * sd.x + G.y * sd.y, G.x * sd.y - G.y * sd.x);
        const int ys = (y > H / 2) ? y - H : y;
        float f = expf(r.filt0 * (float)(kx * kx + ys * ys));
        if (a.phase_only) f = f / (sqrtf(X.x * X.x + X.y * X.y) + 1e-20f);
        X = (kx == 0 && y == 0) ? cmk(0.f, 0.f) : cscale(X, f);
        lds[lp(e)] = cconj(X);
    }
    __syncthreads();
    fft_lds_fwd(lds, nc, a.plan);
    cfloat* out = r.T1 + (long long)g * a.seay * r.cx;
    const float rnc = 1.0f / nc;
    for (int e = threadIdx.x; e < a.seay * nc; e += MC_NTHR) {
        const int jy = fdiv(e, nc, rnc);
        const int c = e - jy * nc;
        int lag = jy - a.seay / 2;
        if (lag < 0) lag += H;
        out[(long long)jy * r.cx + c0 + c] = cconj(lds[lp(c * H + lag)]);
    }
}

// GCorrelateSum2D with a register-pass column schedule: the filtered cross
// spectrum of C columns is formed straight into the registers of pass 0, and
// the last pass only evaluates the butterflies that own one of the seay
// kept lags (y in [0, seay/2) and [H - seay/2, H)).
// WPE: waves per SIMD the tile is compiled for (a register budget of 512 /
// WPE VGPRs).  The 2048-point tile is built for five, so that three of its
// 49 KB workgroups share a CU (18 waves), with three columns per workgroup
// (80 VGPRs); batching its loads (BL) spills at five waves.  The 408-point
// tile is built for four (128 VGPRs) with both BL bits: each frame's
// group-sum loads and then the sum's loads go out as one batch each (at five
// waves its registers spill, and every spilled value costs a wait for all
// outstanding memory operations).
template <int H, int C, int NT, int FF, int WPE, int BL>
__global__ __launch_bounds__(NT, WPE) void k_correlate_rp(CorrArgs a, const cfloat* __restrict__ ptw) {
    constexpr int NP = rp_npass(H), R0 = rp_radix(H, 0), M0 = H / R0;
    constexpr int RL = rp_radix(H, NP - 1), ML = H / RL;
    constexpr int IT0 = (C * M0 + NT - 1) / NT, ITL = (C * ML + NT - 1) / NT;
    extern __shared__ cfloat lds[];
    // one-dimensional grid over (job, column block, group), group fastest,
    // dealt to the XCDs in contiguous ranges: the workgroups of one column
    // block run on one XCD one after another, so its columns of the sum S
    // (and the phasor tables) come from that XCD's L2 after the first group
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    int j, r0;
    if (a.jobs) {
        // the job whose tile range holds this tile (tile0 ascending)
        int lo = 0, hi = a.J - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.jobs[mid].tile0 <= tile) lo = mid;
            else hi = mid - 1;
        }
        j = lo;
        r0 = tile - a.jobs[j].tile0;
    } else {
        const int per_job = ((a.cx + C - 1) / C) * a.G;
        j = tile / per_job;
        r0 = tile - j * per_job;
    }
    if (a.done && a.done[j]) return;
    const CorrJob r = corr_job(a, j, H, Spec<FF>::bytes);
    const int cb = r0 / r.G;
    const int g = r0 - cb * r.G;
    const int c0 = cb * C;
    const int nc = min(C, r.cx - c0);
    const int gs = r.gstart[g], gn = r.gsize[g];
    const long long fs = (long long)r.cx * H;
#pragma unroll
    for (int it = 0; it < IT0; it++) {
        const int i = threadIdx.x + it * NT;
        if ((C * M0) % NT == 0 || i < C * M0) {
            const int c = i / M0, jj = i - c * M0;
            const int kx = c0 + c;
            cfloat v[R0];
#pragma unroll
            for (int t = 0; t < R0; t++) v[t] = cmk(0.f, 0.f);
            if (c < nc) {
                for (int f = 0; f < gn; f++) {
                    const long long b = gs + f;
                    const long long fo = b * fs + (long long)kx * H + jj;
                    const float fi = FF == MC_SPEC_F32 ? 1.f : r.finv[b];
                    const cfloat* Bb = r.B + b * H + jj;
                    const cfloat pa = r.A[b * r.cx + kx];
                    if constexpr (BL & 1) {
                        // all loads of the frame first, then the products
                        cfloat fv[R0], bv[R0];
#pragma unroll
                        for (int t = 0; t < R0; t++) {
                            fv[t] = Spec<FF>::ld(r.F, fo + t * M0, fi);
                            bv[t] = Bb[t * M0];
                        }
#pragma unroll
                        for (int t = 0; t < R0; t++) v[t] = cadd(v[t], cmul(fv[t], cmul(pa, bv[t])));
                    } else {
#pragma unroll
                        for (int t = 0; t < R0; t++)
                            v[t] = cadd(v[t], cmul(Spec<FF>::ld(r.F, fo + t * M0, fi), cmul(pa, Bb[t * M0])));
                    }
                }
                // BL bit 1: the sum's values first (one batch of loads), then
                // the filter, with the phase-only branch outside the loop;
                // otherwise each value is loaded where the filter needs it
                // (fewer registers)
                const cfloat* Sc = r.S + (long long)kx * H + jj;
                if constexpr (BL & 2) {
                    cfloat sv[R0];
#pragma unroll
                    for (int t = 0; t < R0; t++) sv[t] = Sc[t * M0];
                    auto filter = [&](auto phase_only) {
#pragma unroll
                        for (int t = 0; t < R0; t++) {
                            const int y = jj + t * M0;
                            const cfloat G = v[t];
                            const cfloat sd = csub(sv[t], G);
                            cfloat X = cmk(G.x * sd.x + G.y * sd.y, G.x * sd.y - G.y * sd.x);
                            const int ys = (y > H / 2) ? y - H : y;
                            float fl = expf(r.filt0 * (float)(kx * kx + ys * ys));
                            if (decltype(phase_only)::value) fl = fl / (sqrtf(X.x * X.x + X.y * X.y) + 1e-20f);
                            X = (kx == 0 && y == 0) ? cmk(0.f, 0.f) : cscale(X, fl);
                            v[t] = cconj(X);
                        }
                    };
                    if (a.phase_only) filter(std::true_type{});
                    else filter(std::false_type{});
                } else {
#pragma unroll
                    for (int t = 0; t < R0; t++) {
                        const int y = jj + t * M0;
                        const cfloat G = v[t];
                        const cfloat sd = csub(Sc[t * M0], G);
                        cfloat X = cmk(G.x * sd.x + G.y * sd.y, G.x * sd.y - G.y * sd.x);
                        const int ys = (y > H / 2) ? y - H : y;
                        float fl = expf(r.filt0 * (float)(kx * kx + ys * ys));
                        if (a.phase_only) fl = fl / (sqrtf(X.x * X.x + X.y * X.y) + 1e-20f);
                        X = (kx == 0 && y == 0) ? cmk(0.f, 0.f) : cscale(X, fl);
                        v[t] = cconj(X);
                    }
                }
            }
            RegDft<R0>::run(v);
            cfloat* d = lds + c * rp_cs(H, C);
#pragma unroll
            for (int t = 0; t < R0; t++) d[rp_ix<H>(jj * R0 + t)] = v[t];
        }
    }
    __syncthreads();
    RpMid<H, 1, C, NT>::run(lds, ptw);
    cfloat* out = r.T1 + (long long)g * a.seay * r.cx;
    const int half = a.seay / 2;
#pragma unroll
    for (int it = 0; it < ITL; it++) {
        const int i = threadIdx.x + it * NT;
        if ((C * ML) % NT == 0 || i < C * ML) {
            const int c = i / ML, k = i - c * ML;
            bool need = false;
#pragma unroll
            for (int t = 0; t < RL; t++) {
                const int y = k + t * ML;
                const int jy = (y < H / 2 ? y : y - H) + half;
                need |= (jy >= 0 && jy < a.seay);
            }
            if (c < nc && need) {
                cfloat v[RL];
                const cfloat* sl = lds + c * rp_cs(H, C);
#pragma unroll
                for (int t = 0; t < RL; t++) v[t] = sl[rp_ix<H>(k + t * ML)];
                rp_twiddle<H, NP - 1>(v, k, ptw);
                RegDft<RL>::run(v);
#pragma unroll
                for (int t = 0; t < RL; t++) {
                    const int y = k + t * ML;
                    const int jy = (y < H / 2 ? y : y - H) + half;
                    if (jy >= 0 && jy < a.seay) out[(long long)jy * r.cx + c0 + c] = cconj(v[t]);
                }
            }
        }
    }
}

#ifndef MC_RP_CORR_SPECS
// (column length, columns per workgroup, threads, waves per SIMD, batched
// loads: bit 0 the group sums', bit 1 the sum's)
#define MC_RP_CORR_SPECS(X) X(2048, 3, 384, 5, 0) X(408, 16, 384, 4, 3)
#endif

// Second half of the windowed C2R: for one lag row, the Hermitian sum over
// kx at sea_x lags, c(l) = Re T0 + (-1)^l Re T_M + 2 sum_{0<k<M} Re(T_k e^{2 pi i k l/W}).
// With a job table, grid y covers the job with most groups and each job
// reads its own row length and twiddles (LDS sized for the widest).
__global__ __launch_bounds__(256) void k_row_dft(const cfloat* T1, const cfloat* twW,
                                                 const unsigned char* done, int G, int cx, int W,
                                                 int seax, int seay, float* win,
                                                 const JobGeo* __restrict__ jobs) {
    extern __shared__ cfloat lds[];
    const int jy = blockIdx.x, g = blockIdx.y, j = blockIdx.z;
    if (done && done[j]) return;
    long long t1o = ((long long)j * G + g) * seay * cx, wo = (long long)j * G + g;
    if (jobs) {
        const JobGeo& jg = jobs[j];
        if (g >= jg.G) return;
        cx = jg.cx;
        W = jg.W;
        twW = jg.twW;
        t1o = jg.t1o + (long long)g * seay * cx;
        wo = jg.go + g;
    }
    cfloat* row = lds;
    cfloat* tw = lds + cx;
    float* red = reinterpret_cast<float*>(tw + W);
    const cfloat* src = T1 + t1o + (long long)jy * cx;
    for (int e = threadIdx.x; e < cx; e += blockDim.x) row[e] = src[e];
    for (int e = threadIdx.x; e < W; e += blockDim.x) tw[e] = twW[e];
    __syncthreads();
    const int jx = threadIdx.x & 63;
    const int part = threadIdx.x >> 6;
    const int M = cx - 1;
    float acc = 0.f;
    if (jx < seax) {
        const int lx = jx - seax / 2;
        int lxm = lx % W;
        if (lxm < 0) lxm += W;
        const int k0 = 1 + part;
        int idx = (int)(((long long)k0 * lxm) % W);
        const int step = (int)((4LL * lxm) % W);
        for (int k = k0; k < M; k += 4) {
            const cfloat t = row[k], w = tw[idx];
            acc = fmaf(t.x, w.x, fmaf(-t.y, w.y, acc));
            idx += step;
            if (idx >= W) idx -= W;
        }
        acc *= 2.f;
        if (part == 0) acc += row[0].x + ((lx & 1) ? -row[M].x : row[M].x);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < 64 && jx < seax) {
        const float v = red[threadIdx.x] + red[threadIdx.x + 64] + red[threadIdx.x + 128] +
                        red[threadIdx.x + 192];
        win[(wo * seay + jy) * seax + jx] = v;
    }
}

// CPeak2D (Align/CPeak2D.cpp:39-109): first maximum wins, integer peak moved
// off the border, separable 3-point parabola in FP64, |delta| > 1 rejected.
__device__ void peak_fit(const float* img, int nx, int ny, double* out) {
    __shared__ float sv[256];
    __shared__ int si[256];
    float best = -INFINITY;
    int bi = 0x7fffffff;
    for (int i = threadIdx.x; i < nx * ny; i += blockDim.x) {
        const float v = img[i];
        if (v > best) {
            best = v;
            bi = i;
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const float v2 = sv[threadIdx.x + o];
            const int i2 = si[threadIdx.x + o];
            if (v2 > sv[threadIdx.x] || (v2 == sv[threadIdx.x] && i2 < si[threadIdx.x])) {
                sv[threadIdx.x] = v2;
                si[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int ip = si[0];
        if (ip == 0x7fffffff) ip = 0;
        int px = ip % nx, py = ip / nx;
        if (px == 0) px = 1;
        if (py == 0) py = 1;
        if (px == nx - 1) px -= 1;
        if (py == ny - 1) py -= 1;
        const int ic = py * nx + px;
        const double c0 = img[ic];
        const double a = ((double)img[ic + 1] + img[ic - 1]) * 0.5 - c0;
        const double b = ((double)img[ic + 1] - img[ic - 1]) * 0.5;
        const double c = ((double)img[ic + nx] + img[ic - nx]) * 0.5 - c0;
        const double d = ((double)img[ic + nx] - img[ic - nx]) * 0.5;
        double dx = -b / (2 * a + 1e-30);
        double dy = -d / (2 * c + 1e-30);
        if (fabs(dx) > 1) dx = 0;
        if (fabs(dy) > 1) dy = 0;
        out[0] = (px + dx) - nx / 2;
        out[1] = (py + dy) - ny / 2;
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_peaks(const float* win, const unsigned char* done, int G,
                                               int nx, int ny, double* gshift,
                                               const JobGeo* __restrict__ jobs) {
    const int g = blockIdx.x, j = blockIdx.y;
    if (done && done[j]) return;
    long long o = (long long)j * G + g;
    if (jobs) {
        if (g >= jobs[j].G) return;
        o = jobs[j].go + g;
    }
    peak_fit(win + o * nx * ny, nx, ny, gshift + o * 2);
}

__global__ __launch_bounds__(256) void k_peaks_f(const float* win, int nx, int ny, double* out) {
    peak_fit(win + (long long)blockIdx.x * nx * ny, nx, ny, out + 2 * blockIdx.x);
}

// Iteration bookkeeping of CIterativeAlign::mAlignStack
// (Align/CIterativeAlign.cpp:120-145) and CAlignStack::mUpdateError
// (Align/CAlignStack.cpp:207-215): err = max |group shift|, total += group
// shifts, running frame shifts += nearest-group interpolation
// (Align/CInterpolateShift.cpp:60-71), stop when err < tol after the first
// iteration.  Accumulation in FP64.
struct UpdArgs {
    double* tmp;
    double* total;
    const double* gshift;
    double* errors;
    int* iters;
    unsigned char* done;
    const int* fgroup;
    int N, G, max_iter, it;
    double tol;
    const JobGeo* jobs;  // per-job frame and group counts, offsets, tolerance (or null)
};

__global__ void k_update(UpdArgs a) {
    const int j = blockIdx.x;
    if (a.done[j] || threadIdx.x != 0) return;
    int N = a.N, G = a.G;
    long long fo = (long long)j * N, go = (long long)j * G;
    const int* fgroup = a.fgroup;
    double tol = a.tol;
    if (a.jobs) {
        const JobGeo& jg = a.jobs[j];
        N = jg.N;
        G = jg.G;
        fo = jg.fo;
        go = jg.go;
        fgroup += jg.ft;
        tol = jg.tol;
    }
    const double* gs = a.gshift + go * 2;
    double err = 0.0;
    for (int g = 0; g < G; g++) {
        const double e = sqrt(gs[2 * g] * gs[2 * g] + gs[2 * g + 1] * gs[2 * g + 1]);
        if (err < e) err = e;
        a.total[(go + g) * 2] += gs[2 * g];
        a.total[(go + g) * 2 + 1] += gs[2 * g + 1];
    }
    for (int i = 0; i < N; i++) {
        const int g = fgroup[i];
        a.tmp[(fo + i) * 2] += gs[2 * g];
        a.tmp[(fo + i) * 2 + 1] += gs[2 * g + 1];
    }
    a.errors[(long long)j * a.max_iter + a.it] = err;
    a.iters[j] = a.it + 1;
    if (err < tol && a.it > 0) a.done[j] = 1;
}

// ---------------------------------------------------------------------------
// FP64 trajectory numerics: Util_LinEqs (LibSrc/Util/Util_LinEqs.cpp:19-148),
// CSplineFit1D (Util/CSplineFit1D.cpp:31-115), CRemoveSpikes1D
// (Util/CRemoveSpikes1D.cpp:31-148), CStackShift::Smooth
// (Align/CStackShift.cpp:180-198) and the linear group->frame interpolation
// (Align/CInterpolateShift.cpp:75-142).  One 64-lane wave works on one
// trajectory (k_finalize, k_despike): every element of an array step is one
// lane's, computed with the reference's operations for that element (a sum
// over frames stays a sequential sum per element); the Gauss-Jordan pivot
// search is a lane reduction that keeps the reference's first-maximum rule.
#define MC_TL 64

// Spline terms of CSplineFit1D for increasing knots: 1, x, x^2, then
// (x - k_i)^2 while x >= k_i (the term loop stops at the first knot above
// x); the curvature terms 0, 1, 2x, then 2 (x - k_i) only when x <= k_0
// (the curvature loop stops at the first knot below x).
__device__ __forceinline__ double spl_term(double x, int r, const double* knots) {
    if (r == 0) return 1.0;
    if (r == 1) return x;
    if (r == 2) return x * x;
    const double k = knots[r - 3];
    if (x < k) return 0.0;
    const double v = x - k;
    return v * v;
}
__device__ __forceinline__ double spl_cur(double x, int r, const double* knots) {
    if (r == 0) return 0.0;
    if (r == 1) return 1.0;
    if (r == 2) return 2.0 * x;
    if (x - knots[0] > 0) return 0.0;
    return 2.0 * (x - knots[r - 3]);
}

// Gauss-Jordan with full pivoting on the n x n system (n <= 9) in LDS;
// false when singular.
__device__ bool w_lin_eqs(double* a, double* v, int n) {
    __shared__ int piv[16];
    __shared__ int sel[2];
    __shared__ double rv[MC_TL];
    __shared__ int re[MC_TL];
    const int l = threadIdx.x;
    if (l < n) piv[l] = 0;
    __syncthreads();
    for (int ip = 0; ip < n; ip++) {
        // first maximum of |a| over unpivoted rows and columns, row-major
        double best = 0.0;
        int be = -1;
        for (int e = l; e < n * n; e += MC_TL) {
            const int y = e / n, x = e - y * n;
            if (piv[y] == 1 || piv[x] == 1) continue;
            const double c = fabs(a[e]);
            if (c > best) {
                best = c;
                be = e;
            }
        }
        rv[l] = best;
        re[l] = be;
        __syncthreads();
        if (l == 0) {
            double bv = 0.0;
            int bi = -1;
            for (int k = 0; k < MC_TL; k++) {
                if (re[k] < 0) continue;
                if (rv[k] > bv || (rv[k] == bv && re[k] < bi)) {
                    bv = rv[k];
                    bi = re[k];
                }
            }
            sel[0] = bi < 0 ? -1 : bi % n;  // column
            sel[1] = bi < 0 ? -1 : bi / n;  // row
            if (bi >= 0) piv[bi % n] += 1;
        }
        __syncthreads();
        const int ix = sel[0], iy = sel[1];
        if (ix < 0) return false;
        if (a[iy * n + ix] == 0.0) return false;
        if (ix != iy) {
            if (l < n) {
                const double t = a[ix * n + l];
                a[ix * n + l] = a[iy * n + l];
                a[iy * n + l] = t;
            }
            if (l == 0) {
                const double t = v[ix];
                v[ix] = v[iy];
                v[iy] = t;
            }
            __syncthreads();
        }
        const int col = ix;
        const double pinv = 1.0 / (a[col * n + col] + 1e-30);
        __syncthreads();
        if (l < n) a[col * n + l] *= pinv;
        if (l == 0) v[col] *= pinv;
        __syncthreads();
        // rows y != col: a[y][x] -= a[col][x] * a[y][col], v[y] -= v[col] * a[y][col]
        double na[2] = {0.0, 0.0};
        double nv = 0.0;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int e = l + k * MC_TL;
            if (e < n * n) {
                const int y = e / n, x = e - y * n;
                if (y != col) na[k] = a[e] - a[col * n + x] * a[y * n + col];
            }
        }
        if (l < n && l != col) nv = v[l] - v[col] * a[l * n + col];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int e = l + k * MC_TL;
            if (e < n * n && e / n != col) a[e] = na[k];
        }
        if (l < n && l != col) v[l] = nv;
        __syncthreads();
    }
    return true;
}

// CSplineFit1D::DoIt: least squares of y on the spline terms of x, then the
// fit evaluated at every x into out (left untouched if the system is
// singular).
__device__ void w_fit(const double* knots, int nk, const double* x, const double* y, int n, double* out) {
    __shared__ double m[81], cf[9];
    const int d = nk + 3, l = threadIdx.x;
    for (int e = l; e < d * d + d; e += MC_TL) {
        double acc = 0.0;
        if (e < d * d) {
            const int r = e / d, q = e - r * d;
            for (int i = 0; i < n; i++) {
                const double xi = x[i];
                acc += spl_term(xi, r, knots) * spl_term(xi, q, knots) +
                       1e-20 * spl_cur(xi, r, knots) * spl_cur(xi, q, knots);
            }
            m[e] = acc;
        } else {
            const int r = e - d * d;
            for (int i = 0; i < n; i++) acc += spl_term(x[i], r, knots) * y[i];
            cf[r] = acc;
        }
    }
    __syncthreads();
    if (!w_lin_eqs(m, cf, d)) {
        __syncthreads();
        return;
    }
    for (int i = l; i < n; i += MC_TL) {
        double f = 0.0;
        for (int r = 0; r < d; r++) f += cf[r] * spl_term(x[i], r, knots);
        out[i] = f;
    }
    __syncthreads();
}

__device__ void w_remove_single(double* raw, double* buf, int n) {
    const int l = threadIdx.x;
    for (int i = l; i < n; i += MC_TL) buf[i] = raw[i];
    __syncthreads();
    for (int i = l + 1; i < n - 1; i += MC_TL) {
        if (fabs(buf[i] - buf[i - 1]) < 5.0) continue;
        if (fabs(buf[i] - buf[i + 1]) < 5.0) continue;
        raw[i] = (buf[i - 1] + buf[i + 1]) / 2.0;
    }
    if (l == 0) {
        const int e = n - 1;
        if (fabs(buf[0] - buf[1]) > 10.0 && fabs(buf[0]) > fabs(buf[1])) raw[0] = 2.0 * buf[1] - buf[2];
        if (fabs(buf[e] - buf[e - 1]) > 10.0 && fabs(buf[e]) > fabs(buf[e - 1]))
            raw[e] = 2.0 * buf[e - 1] - buf[e - 2];
    }
    __syncthreads();
}

__device__ bool w_find_bad(double* raw, const double* fit, int n, double tol) {
    int has = 0;
    for (int i = threadIdx.x; i < n; i += MC_TL) {
        if (fabs(fit[i] - raw[i]) > tol && fabs(raw[i]) > fabs(fit[i])) {
            raw[i] = fit[i];
            has = 1;
        }
    }
    return __syncthreads_or(has) != 0;
}

// scratch: 5*n doubles
__device__ void w_remove_spikes(double* sx, double* sy, int n, int quirks, double* scratch) {
    double* t = scratch;
    double* fx = scratch + n;
    double* fy = scratch + 2 * n;
    double* buf = scratch + 3 * n;
    double* buf2 = scratch + 4 * n;
    const int l = threadIdx.x;
    for (int r = 0; r < 2; r++) {
        w_remove_single(sx, buf, n);
        w_remove_single(sy, buf, n);
    }
    const float ffact = 1.0f / n;
    for (int i = l; i < n; i += MC_TL) t[i] = sqrt((double)(i * ffact));
    int nk = (n - 30) / 15;
    if (nk < 0) nk = 0;
    else if (nk > 6) nk = 6;
    if (nk == 0 && n >= 30) nk = 1;
    __shared__ double knots[6];
    __syncthreads();
    if (l < nk) knots[l] = t[(l + 1) * 15];
    for (int i = l; i < n; i += MC_TL) {
        fx[i] = sx[i];
        fy[i] = sy[i];
    }
    __syncthreads();
    w_fit(knots, nk, t, sx, n, fx);
    w_fit(knots, nk, t, sy, n, fy);
    for (int i = l; i < n; i += MC_TL) buf[i] = sx[i];
    __syncthreads();
    if (w_find_bad(buf, fx, n, 3.0)) w_fit(knots, nk, t, buf, n, fx);
    for (int i = l; i < n; i += MC_TL) buf2[i] = sy[i];
    __syncthreads();
    if (w_find_bad(buf2, fy, n, 3.0)) w_fit(knots, nk, t, quirks ? sy : buf2, n, fy);
    for (int i = l; i < n; i += MC_TL) {
        double tol = 5.0;
        if (i > 20) tol = 4.0;
        if (i < 50) tol = 3.0;
        if (fabs(fx[i] - sx[i]) > tol) sx[i] = fx[i];
        if (fabs(fy[i] - sy[i]) > tol) sy[i] = fy[i];
    }
    __syncthreads();
}

__device__ void w_smooth(double* sx, double* sy, int n, double w, double* buf) {
    if (w >= 1) return;
    const int l = threadIdx.x;
    for (int i = l; i < n; i += MC_TL) {
        buf[i] = sx[i];
        buf[n + i] = sy[i];
    }
    __syncthreads();
    const double w2 = 0.5 * (1.0 - w);
    for (int i = l + 1; i < n - 1; i += MC_TL) {
        sx[i] = w * buf[i] + w2 * (buf[i - 1] + buf[i + 1]);
        sy[i] = w * buf[n + i] + w2 * (buf[n + i - 1] + buf[n + i + 1]);
    }
    __syncthreads();
}

struct FinArgs {
    const double* total;
    double* final_;
    double* scratch;
    const unsigned char* dummy;
    const double* gcent;
    const double* fcent;
    int N, G, binz, quirks;
    long long sstride;
    int use_lds;   // work in LDS (sstride doubles) instead of the global scratch
    const JobGeo* jobs;  // per-job counts, offsets and group bin (or null)
};

// End of CIterativeAlign (Align/CIterativeAlign.cpp:63-85,146-149): linear
// interpolation of the accumulated group shifts, RemoveSpikes, Smooth(0.5),
// then Smooth(0.3 + 0.1 * binZ, capped at 0.9).  One wave per job; its work
// arrays live in LDS when they fit.
__global__ __launch_bounds__(MC_TL) void k_finalize(FinArgs a) {
    extern __shared__ double fsh[];
    const int j = blockIdx.x, l = threadIdx.x;
    int N = a.N, G = a.G, binz = a.binz;
    long long fo = (long long)j * N, go = (long long)j * G, sco = (long long)j * a.sstride;
    const double* gcent = a.gcent;
    const double* fcent = a.fcent;
    if (a.jobs) {
        const JobGeo& jg = a.jobs[j];
        N = jg.N;
        G = jg.G;
        binz = jg.binz;
        fo = jg.fo;
        go = jg.go;
        sco = jg.sco;
        gcent += jg.gt;
        fcent += jg.ft;
    }
    double* sc = a.use_lds ? fsh : a.scratch + sco;
    double* sx = sc;
    double* sy = sc + N;
    double* work = sc + 2 * N;
    const double* tot = a.total + go * 2;
    for (int i = l; i < N; i += MC_TL) {
        if (N == G) {
            sx[i] = tot[2 * i];
            sy[i] = tot[2 * i + 1];
            continue;
        }
        const double fc = fcent[i];
        int k;
        if (fc <= gcent[0]) k = 1;
        else {
            k = G - 1;
            for (int q = 1; q < G; q++)
                if (fc < gcent[q]) { k = q; break; }
        }
        const int f = k - 1;
        const double dg = gcent[k] - gcent[f];
        const double df = fc - gcent[f];
        sx[i] = tot[2 * f] + (tot[2 * k] - tot[2 * f]) / dg * df;
        sy[i] = tot[2 * f + 1] + (tot[2 * k + 1] - tot[2 * f + 1]) / dg * df;
    }
    __syncthreads();
    w_remove_spikes(sx, sy, N, a.quirks, work);
    w_smooth(sx, sy, N, 0.5, work);
    double w = 0.3 + 0.1 * binz;
    if (w > 0.9) w = 0.9;
    w_smooth(sx, sy, N, w, work);
    double* out = a.final_ + fo * 2;
    for (int i = l; i < N; i += MC_TL) {
        out[2 * i] = sx[i];
        out[2 * i + 1] = sy[i];
    }
}

__global__ __launch_bounds__(MC_TL) void k_despike(double* traj, double* scratch, int n, int quirks) {
    double* t = traj + (long long)blockIdx.x * 2 * n;
    w_remove_spikes(t, t + n, n, quirks, scratch + (long long)blockIdx.x * 5 * n);
}

// ---------------------------------------------------------------------------
// Feature map of CDetectFeatures::DoIt (Align/CDetectFeatures.cpp:58-186) on
// the 128x128 real map: +(1 - min), divide by the local 5x5 variance
// (Align/GNormByStd2D.cu:7-42, evaluated out of place), moments, clipped
// statistics, feature = value in [mean - 4 sd, mean - 0.08 sd].
__global__ __launch_bounds__(1024) void k_feature_map(const float* img, unsigned char* feat) {
    const int B = 128, NP = B * B;
    extern __shared__ float fm[];
    float* a = fm;
    float* b = fm + NP;
    __shared__ double rd[3][1024];
    __shared__ float rmin[1024];
    float mn = INFINITY;
    for (int i = threadIdx.x; i < NP; i += 1024) {
        a[i] = img[i];
        mn = fminf(mn, a[i]);
    }
    rmin[threadIdx.x] = mn;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) rmin[threadIdx.x] = fminf(rmin[threadIdx.x], rmin[threadIdx.x + o]);
        __syncthreads();
    }
    const float add = 1.0f - rmin[0];
    for (int i = threadIdx.x; i < NP; i += 1024) a[i] += add;
    __syncthreads();
    double s1 = 0, s2 = 0;
    for (int i = threadIdx.x; i < NP; i += 1024) {
        const int x = i % B, y = i / B;
        float m = 0.f, v = 0.f;
        int c = 0;
        for (int jj = 0; jj < 5; jj++) {
            const int yy = y - 2 + jj;
            if (yy < 0 || yy >= B) continue;
            for (int ii = 0; ii < 5; ii++) {
                const int xx = x - 2 + ii;
                if (xx < 0 || xx >= B) continue;
                const float t = a[yy * B + xx];
                m += t;
                v += t * t;
                c++;
            }
        }
        m /= c;
        v = v / c - m * m;
        const float o = (v < 0) ? 0.f : a[i] / (v + 1e-20f);
        b[i] = o;
        s1 += o;
        s2 += (double)o * o;
    }
    rd[0][threadIdx.x] = s1;
    rd[1][threadIdx.x] = s2;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            rd[0][threadIdx.x] += rd[0][threadIdx.x + o];
            rd[1][threadIdx.x] += rd[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    const double mean = rd[0][0] / NP;
    double sd = rd[1][0] / NP - mean * mean;
    sd = (sd < 0) ? 0.0 : sqrt(sd);
    const double lo = mean - 4.0 * sd, hi = mean + 4.0 * sd;
    __syncthreads();
    double c1 = 0, c2 = 0, c3 = 0;
    for (int i = threadIdx.x; i < NP; i += 1024) {
        const double v = b[i];
        if (v < lo || v > hi) continue;
        c1 += v;
        c2 += v * v;
        c3 += 1.0;
    }
    rd[0][threadIdx.x] = c1;
    rd[1][threadIdx.x] = c2;
    rd[2][threadIdx.x] = c3;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            rd[0][threadIdx.x] += rd[0][threadIdx.x + o];
            rd[1][threadIdx.x] += rd[1][threadIdx.x + o];
            rd[2][threadIdx.x] += rd[2][threadIdx.x + o];
        }
        __syncthreads();
    }
    double dm = rd[0][0] / (rd[2][0] + 1e-30);
    double ds = rd[1][0] / (rd[2][0] + 1e-30) - dm * dm;
    ds = (ds < 0) ? 0.0 : sqrt(ds);
    const double flo = dm - 4.0 * ds, fhi = dm - 0.08 * ds;
    for (int i = threadIdx.x; i < NP; i += 1024) feat[i] = (b[i] < flo || b[i] > fhi) ? 0 : 1;
}

// ---------------------------------------------------------------------------
// CPatchShifts::MakeRelative + DetectBads (Align/CPatchShifts.cpp:263-355) in
// FP64.  With quirks the y component subtracts element k+2 of the reference
// frame's row and that row is modified while it is used (:277-285); a read
// past the last frame sees the (still zero) flag bytes.
__global__ __launch_bounds__(256) void k_patch_stats(double* s, const double* cen, int N, int P,
                                                     int W, int H, int quirks,
                                                     unsigned char* bad, int* ref_out) {
    __shared__ double vfr[MC_PS_MAXN];
    __shared__ double rrow[2 * MC_WARP_MAXP + 1];
    __shared__ int sref;
    // mean + std of every frame's patch shifts (frames >= 1), in parallel
    for (int i = 1 + threadIdx.x; i < N; i += blockDim.x) {
        const double* p = s + (long long)i * P * 2;
        double mx = 0, my = 0;
        for (int k = 0; k < P; k++) { mx += p[2 * k]; my += p[2 * k + 1]; }
        mx /= P;
        my /= P;
        double sd = 0;
        for (int k = 0; k < P; k++) {
            const double dx = p[2 * k] - mx, dy = p[2 * k + 1] - my;
            sd += dx * dx + dy * dy;
        }
        vfr[i] = sqrt(mx * mx + my * my) + sqrt(sd / P);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // the reference's scan: the last frame whose value does not exceed
        // the running minimum
        double fmin = 1e20;
        int ref = 0;
        for (int i = 1; i < N; i++) {
            if (vfr[i] > fmin) continue;
            fmin = vfr[i];
            ref = i;
        }
        sref = ref;
        if (ref_out) *ref_out = ref;
    }
    __syncthreads();
    const int ref = sref;
    const long long rb = (long long)ref * P * 2;
    // original values the in-place subtraction reads: the reference row and
    // (quirk) the x of the frame after it
    for (int k = threadIdx.x; k < 2 * P; k += blockDim.x) rrow[k] = s[rb + k];
    if (threadIdx.x == 0) rrow[2 * P] = (ref + 1 < N) ? s[rb + 2 * P] : 0.0;
    __syncthreads();
    for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
        const int i = e / P, k = e - i * P;
        double* q = s + (long long)i * P * 2 + 2 * k;
        if (quirks) {
            // sequential in-place semantics of the reference loop (x uses
            // element k, y element k+2 of the reference row as modified so
            // far): before the reference frame both read original values;
            // the reference frame's own x become 0 first; after it every x
            // read is 0 and only the last y reads the next frame's x_0
            const double nx = (k + 1 < P) ? rrow[2 * k + 2] : rrow[2 * P];
            if (i < ref) {
                q[0] -= rrow[2 * k];
                q[1] -= nx;
            } else if (i == ref) {
                q[0] = 0.0;
                q[1] -= nx;
            } else if (k == P - 1) {
                q[1] -= rrow[2 * P];
            }
        } else {
            if (i == ref) {
                q[0] = 0.0;
                q[1] = 0.0;
            } else {
                q[0] -= rrow[2 * k];
                q[1] -= rrow[2 * k + 1];
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
        const int f = e / P, i = e - f * P;
        const double* p = s + (long long)f * P * 2;
        const double ds = sqrt(p[2 * i] * p[2 * i] + p[2 * i + 1] * p[2 * i + 1]);
        double rms = 0, sw = 0;
        for (int k = 0; k < P; k++) {
            if (k == i) continue;
            const double fx = (cen[2 * k] - cen[2 * i]) / W;
            const double fy = (cen[2 * k + 1] - cen[2 * i + 1]) / H;
            const double w = exp(-100.0 * (fx * fx + fy * fy));
            rms += w * (p[2 * k] * p[2 * k] + p[2 * k + 1] * p[2 * k + 1]);
            sw += w;
        }
        double lr = sqrt(rms / sw);
        if (lr < 2) lr = 2.0;
        bad[e] = (ds > lr) ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// mGCorrect3D (Correct/GCorrectPatchShift.cu:17-66, upsample 1) with the
// frames inside.  Per pixel the reference forms w_p = exp(-100 r_p) for every
// patch with r_p <= 0.6 (r in units of the frame size), skips bad patches,
// sums w, w*sx, w*sy in patch order in FP32, divides, and gathers the pixel at
// (int)(x - sx), (int)(y - sy) with 811*|.| mod N reflection.  Here the
// weights of a 64x4-pixel tile are formed once for all frames.  A patch whose
// weight is below 1e-14 of the largest weight among a frame's good patches
// (r > r_min + 0.3224) for every pixel of the tile is left out: adding it
// cannot change any of the FP32 sums (it is far below half an ulp), so the
// kept patches, taken in the reference's order, give the same sums.  r_min
// is taken over the patches that are good in every frame (bad_any clear), an
// upper bound on every frame's nearest good patch; with none, every patch
// within r <= 0.6 is kept.  The frame table holds
// {ok*sx, ok*sy, ok} per kept patch (ok = 0 for a bad patch), so
// fw = fma(w, ok, fw) and fsx = fma(w, ok*sx, fsx) reproduce the reference's
// skip exactly; up to 16 kept patches the weights live in registers.
#define MC_WARP_MAXP 64
#ifndef MC_WARP_TX
#define MC_WARP_TX 64
#endif
#ifndef MC_WARP_TY
#define MC_WARP_TY 4
#endif
#define MC_WARP_NT (MC_WARP_TX * MC_WARP_TY)
#ifndef MC_WARP_FCH
#define MC_WARP_FCH 16
#endif
#define MC_WARP_KREG 16
#ifndef MC_WARP_U
#define MC_WARP_U 8
#endif

struct WarpArgs {
    const float* R;
    float* Rp;
    long long pitch, fstride;
    int W, H, N, P;
    const float* cen;
    const float* sh;
    const unsigned char* bad;
    unsigned long long bad_any;  // bit p: patch p is bad in some frame
};

// Source offset (within a frame) of output pixel (x, y) for the weighted
// shift sums of one frame.
__device__ __forceinline__ int warp_src(int W, int H, int pitch, int x, int y, float fw, float fsx,
                                        float fsy) {
    if (fw > 0) {
        fsx = fsx / fw;
        fsy = fsy / fw;
    }
    // keep the gather index finite for non-finite or absurd shifts
    if (!(fabsf(fsx) < 1e6f)) fsx = 0.f;
    if (!(fabsf(fsy) < 1e6f)) fsy = 0.f;
    int xs = (int)((float)x - fsx);
    int ys = (int)((float)y - fsy);
    if (xs < 0 || ys < 0 || xs >= W || ys >= H) {
        xs = xs < 0 ? -xs : xs;
        ys = ys < 0 ? -ys : ys;
        xs = (int)((811LL * xs) % W);
        ys = (int)((811LL * ys) % H);
    }
    return ys * pitch + xs;
}

// Fill the table of frames f0..f0+nf-1 for the K (>= nk) kept patches;
// fbad[fi] (cleared by the caller) becomes 1 when a kept patch is bad in
// frame f0 + fi.
__device__ __forceinline__ void warp_table(const WarpArgs& a, const int* klist, int nk, int K,
                                           int f0, int nf, float4* tab, int* fbad) {
    for (int e = threadIdx.x; e < nf * K; e += MC_WARP_NT) {
        const int fi = e / K, q = e - fi * K;
        float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
        if (q < nk) {
            const int p = klist[q];
            const long long o = (long long)(f0 + fi) * a.P + p;
            if (!a.bad[o]) {
                t.x = a.sh[2 * o];
                t.y = a.sh[2 * o + 1];
                t.z = 1.f;
            } else {
                fbad[fi] = 1;
            }
        }
        tab[e] = t;
    }
}

// A frame without bad kept patches has fw = sum_q w_q, the same FMA chain
// (fma(w, 1, fw) is w + fw) for every such frame: it is formed once.
template <int K>
__device__ __forceinline__ void warp_tile_reg(const WarpArgs& a, const int* klist, int nk,
                                              float4* tab, int* fbad, int x, int y, bool live) {
    float w[K];
#pragma unroll
    for (int q = 0; q < K; q++) {
        w[q] = 0.f;
        if (live && q < nk) {
            const int p = klist[q];
            const float ax = ((float)x - a.cen[2 * p]) / (float)a.W;
            const float ay = ((float)y - a.cen[2 * p + 1]) / (float)a.H;
            const float r = sqrtf(ax * ax + ay * ay);
            w[q] = (r > 0.6f) ? 0.f : expf(-100.0f * r);
        }
    }
    float fw_all = 0.f;
#pragma unroll
    for (int q = 0; q < K; q++) fw_all = fmaf(w[q], q < nk ? 1.f : 0.f, fw_all);
    for (int f0 = 0; f0 < a.N; f0 += MC_WARP_FCH) {
        const int nf = min(MC_WARP_FCH, a.N - f0);
        __syncthreads();
        if (threadIdx.x < MC_WARP_FCH) fbad[threadIdx.x] = 0;
        __syncthreads();
        warp_table(a, klist, nk, K, f0, nf, tab, fbad);
        __syncthreads();
        if (!live) continue;
        // gathers of MC_WARP_U frames are issued before their stores so that
        // each wave keeps several loads in flight, and the next group's
        // source offsets are formed while they are; offsets within a frame
        // are 32-bit (the host checks pitch * H < 2^31)
        const int pitch = (int)a.pitch;
        const int oo = y * pitch + x;
        int so[MC_WARP_U];
        auto offsets = [&](int fb) {
#pragma unroll
            for (int u = 0; u < MC_WARP_U; u++) {
                const int fi = min(fb + u, nf - 1);
                const float4* t = tab + fi * K;
                float fw = fw_all, fsx = 0.f, fsy = 0.f;
                if (fbad[fi]) {
                    fw = 0.f;
#pragma unroll
                    for (int q = 0; q < K; q++) fw = fmaf(w[q], t[q].z, fw);
                }
#pragma unroll
                for (int q = 0; q < K; q++) {
                    const float4 v = t[q];
                    fsx = fmaf(v.x, w[q], fsx);
                    fsy = fmaf(v.y, w[q], fsy);
                }
                so[u] = warp_src(a.W, a.H, pitch, x, y, fw, fsx, fsy);
            }
        };
        // Whole groups load and store without per-frame conditions: a
        // conditional store sits in a block of its own, and the compiler
        // then waits for every outstanding memory operation (the previous
        // stores included) before it.  The last nf % MC_WARP_U frames go one
        // at a time.
        const int nfull = nf - nf % MC_WARP_U;
        // Software-pipelined: group g+1's gathers go out before group g's
        // stores.  Loads and stores share one memory counter, counted in
        // issue order, so a wave then waits only for the loads it is about
        // to store, while the previous group's stores stay in flight.
        if (nfull) {
            offsets(0);
            float v[MC_WARP_U];
#pragma unroll
            for (int u = 0; u < MC_WARP_U; u++) v[u] = (a.R + (long long)(f0 + u) * a.fstride)[so[u]];
            for (int fb = 0; fb + MC_WARP_U < nfull; fb += MC_WARP_U) {
                offsets(fb + MC_WARP_U);
                float vn[MC_WARP_U];
#pragma unroll
                for (int u = 0; u < MC_WARP_U; u++)
                    vn[u] = (a.R + (long long)(f0 + fb + MC_WARP_U + u) * a.fstride)[so[u]];
#pragma unroll
                for (int u = 0; u < MC_WARP_U; u++) (a.Rp + (long long)(f0 + fb + u) * a.fstride)[oo] = v[u];
#pragma unroll
                for (int u = 0; u < MC_WARP_U; u++) v[u] = vn[u];
            }
#pragma unroll
            for (int u = 0; u < MC_WARP_U; u++)
                (a.Rp + (long long)(f0 + nfull - MC_WARP_U + u) * a.fstride)[oo] = v[u];
        }
        for (int fi = nfull; fi < nf; fi++) {
            const float4* t = tab + fi * K;
            float fw = fw_all, fsx = 0.f, fsy = 0.f;
            if (fbad[fi]) {
                fw = 0.f;
#pragma unroll
                for (int q = 0; q < K; q++) fw = fmaf(w[q], t[q].z, fw);
            }
#pragma unroll
            for (int q = 0; q < K; q++) {
                const float4 v = t[q];
                fsx = fmaf(v.x, w[q], fsx);
                fsy = fmaf(v.y, w[q], fsy);
            }
            const long long fo = (long long)(f0 + fi) * a.fstride;
            a.Rp[fo + oo] = a.R[fo + warp_src(a.W, a.H, pitch, x, y, fw, fsx, fsy)];
        }
    }
}

// More than MC_WARP_KREG kept patches: weights in LDS, run-time loop.
__device__ __forceinline__ void warp_tile_lds(const WarpArgs& a, const int* klist, int nk,
                                           float4* tab, int* fbad, float* wk, int x, int y, bool live) {
    for (int q = 0; q < nk; q++) {
        float w = 0.f;
        if (live) {
            const int p = klist[q];
            const float ax = ((float)x - a.cen[2 * p]) / (float)a.W;
            const float ay = ((float)y - a.cen[2 * p + 1]) / (float)a.H;
            const float r = sqrtf(ax * ax + ay * ay);
            w = (r > 0.6f) ? 0.f : expf(-100.0f * r);
        }
        wk[q * MC_WARP_NT + threadIdx.x] = w;
    }
    for (int f0 = 0; f0 < a.N; f0 += MC_WARP_FCH) {
        const int nf = min(MC_WARP_FCH, a.N - f0);
        __syncthreads();
        warp_table(a, klist, nk, nk, f0, nf, tab, fbad);
        __syncthreads();
        if (!live) continue;
        for (int fi = 0; fi < nf; fi++) {
            const float4* t = tab + fi * nk;
            float fw = 0.f, fsx = 0.f, fsy = 0.f;
            for (int q = 0; q < nk; q++) {
                const float4 v = t[q];
                const float wq = wk[q * MC_WARP_NT + threadIdx.x];
                fw = fmaf(wq, v.z, fw);
                fsx = fmaf(v.x, wq, fsx);
                fsy = fmaf(v.y, wq, fsy);
            }
            const long long fo = (long long)(f0 + fi) * a.fstride;
            a.Rp[fo + (long long)y * a.pitch + x] =
                a.R[fo + warp_src(a.W, a.H, (int)a.pitch, x, y, fw, fsx, fsy)];
        }
    }
}

// Dynamic LDS: frame table MC_WARP_FCH x max(K) float4, then (only when the
// patch count exceeds MC_WARP_KREG) weights [kept patch][thread].
__global__ __launch_bounds__(MC_WARP_NT) void k_local_warp(WarpArgs a) {
    extern __shared__ float4 tab[];
    __shared__ unsigned int kmask[2];
    __shared__ int klist[MC_WARP_MAXP];
    __shared__ int nk_s;
    __shared__ int fbad[MC_WARP_FCH];
    const int x = blockIdx.x * MC_WARP_TX + (threadIdx.x % MC_WARP_TX);
    const int y = blockIdx.y * MC_WARP_TY + (threadIdx.x / MC_WARP_TX);
    const bool live = x < a.W && y < a.H;
    if (threadIdx.x == 0) kmask[0] = kmask[1] = 0u;
    __syncthreads();
    // Selection only has to be a superset of the patches that can matter, so
    // it uses squared distances with reciprocal scales and a small margin;
    // the weights themselves are computed exactly as the reference does.
    const float rw = 1.0f / (float)a.W, rh = 1.0f / (float)a.H;
    float r2min = 1e30f;
    for (int p = 0; p < a.P && live; p++) {
        if ((a.bad_any >> p) & 1ull) continue;
        const float ax = ((float)x - a.cen[2 * p]) * rw;
        const float ay = ((float)y - a.cen[2 * p + 1]) * rh;
        r2min = fminf(r2min, ax * ax + ay * ay);
    }
    const float rlim = fminf(sqrtf(r2min) + 0.3225f, 0.6001f);
    const float r2lim = rlim * rlim;
    unsigned int lo = 0u, hi = 0u;
    for (int p = 0; p < a.P && live; p++) {
        const float ax = ((float)x - a.cen[2 * p]) * rw;
        const float ay = ((float)y - a.cen[2 * p + 1]) * rh;
        if (ax * ax + ay * ay <= r2lim) {
            if (p < 32) lo |= 1u << p;
            else hi |= 1u << (p - 32);
        }
    }
    if (lo) atomicOr(&kmask[0], lo);
    if (hi) atomicOr(&kmask[1], hi);
    __syncthreads();
    if (threadIdx.x == 0) {
        int n = 0;
        for (int p = 0; p < a.P; p++) {
            const unsigned int wd = kmask[p >> 5];
            if ((wd >> (p & 31)) & 1u) klist[n++] = p;
        }
        nk_s = n;
    }
    __syncthreads();
    const int nk = nk_s;
    switch (nk) {
#define MC_WK(K) \
    case K: warp_tile_reg<K>(a, klist, nk, tab, fbad, x, y, live); return;
        MC_WK(1) MC_WK(2) MC_WK(3) MC_WK(4) MC_WK(5) MC_WK(6) MC_WK(7) MC_WK(8)
        MC_WK(9) MC_WK(10) MC_WK(11) MC_WK(12) MC_WK(13) MC_WK(14) MC_WK(15) MC_WK(16)
#undef MC_WK
        default: break;
    }
    if (nk == 0) {
        warp_tile_reg<1>(a, klist, 0, tab, fbad, x, y, live);
        return;
    }
        warp_tile_lds(a, klist, nk, tab, fbad, reinterpret_cast<float*>(tab + MC_WARP_FCH * a.P), x, y,
                      live);
}

static size_t warp_lds_bytes(int P) {
    const int kt = P > MC_WARP_KREG ? P : MC_WARP_KREG;
    size_t b = (size_t)MC_WARP_FCH * kt * sizeof(float4);
    if (P > MC_WARP_KREG) b += (size_t)P * MC_WARP_NT * sizeof(float);
    return b;
}

// ---------------------------------------------------------------------------
// CCorrectFullShift::mGenSums (Correct/CCorrectFullShift.cpp:271-300) over
// all frames in one pass per frequency: sum0 += F_i; with dose weighting
// F_i *= w_i / sqrt(sum_j w_j^2) (Correct/GWeightFrame.cu:19-71, DC left
// alone), sum1 += F_i, sum2 += F_i if the accumulated dose is in
// -SumRange; -SplitSum adds F_i to the odd/even sum.  Optional phasors apply
// the global shift of Correct/CCorrectFullShift.cpp:302-309 on load.  The
// weights w_i = exp(-0.5 D_i / Dc(k)) with D_i = dose*(i+1)
// (DataUtil/CFmIntParam.cpp:128-135) are generated as powers of
// q = exp(-0.5 dose / Dc) in FP64.
struct DwArgs {
    const cfloat* F;      // frame spectra, stored in the kernel's format FF
    const float* finv;    // per-frame load factors of a reduced format
    long long fstride;
    int N, cx, H;
    const cfloat* A;
    const cfloat* B;
    int dw;
    double pix, kvf, dose;
    int dws;
    double r0, r1;
    int split;
    cfloat* sums;
    long long sstride;
    const int* cum;  // raw frames counted into frame i's dose (null: i + 1)
};

__device__ __forceinline__ double crit_q(int kx, int y, int cx, int H, double pix, double kvf,
                                         double dose) {
    double fx = kx * 0.5 / (cx - 1);
    double fy = y / (double)H;
    if (fy >= 0.5) fy -= 1.0;
    const double f = sqrt(fx * fx + fy * fy) / pix;
    const double crit = (0.24499 * pow(f, -1.6649) + 2.8141) * kvf;
    return exp(-0.5 * dose / crit);
}

template <int FF>
__global__ __launch_bounds__(256) void k_dw_sums(DwArgs a) {
    const int kx = blockIdx.y;
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= a.H) return;
    const long long q = (long long)kx * a.H + y;
    const bool dc = (kx == 0 && y == 0);
    double qd = 1.0, inv_ws = 1.0;
    if (a.dw && !dc) {
        qd = crit_q(kx, y, a.cx, a.H, a.pix, a.kvf, a.dose);
        double w = 1.0, s = 0.0;
        int prev = 0;
        for (int i = 0; i < a.N; i++) {
            const int c = a.cum ? a.cum[i] : i + 1;
            for (int k = prev; k < c; k++) w *= qd;
            prev = c;
            s += w * w;
        }
        inv_ws = 1.0 / sqrt(s);
    }
    cfloat s0 = cmk(0.f, 0.f), s1 = s0, s2 = s0, so = s0, se = s0;
    double w = 1.0;
    int prev = 0;
    for (int i = 0; i < a.N; i++) {
        cfloat v = Spec<FF>::ld(a.F, (long long)i * a.fstride + q, FF == MC_SPEC_F32 ? 1.f : a.finv[i]);
        if (a.A) v = cmul(v, cmul(a.A[(long long)i * a.cx + kx], a.B[(long long)i * a.H + y]));
        s0 = cadd(s0, v);
        if (a.dw) {
            // w_i = exp(-0.5 D_i / Dc) with D_i = dose x c_i, c_i raw frames
            const int c = a.cum ? a.cum[i] : i + 1;
            for (int k = prev; k < c; k++) w *= qd;
            prev = c;
            if (!dc) v = cscale(v, (float)(w * inv_ws));
            s1 = cadd(s1, v);
            if (a.dws) {
                const double d = a.dose * c;
                if (!(d < a.r0) && !(d > a.r1)) s2 = cadd(s2, v);
            }
        }
        if (a.split) {
            if (i & 1) so = cadd(so, v);
            else se = cadd(se, v);
        }
    }
    int ns = 0;
    a.sums[(long long)(ns++) * a.sstride + q] = s0;
    if (a.dw) {
        a.sums[(long long)(ns++) * a.sstride + q] = s1;
        if (a.dws) a.sums[(long long)(ns++) * a.sstride + q] = s2;
    }
    if (a.split) {
        a.sums[(long long)ns * a.sstride + q] = se;
        a.sums[(long long)(ns + 1) * a.sstride + q] = so;
    }
}

__global__ void k_dose_wsum(int cx, int H, int N, double pix, double kvf, double dose, float* out) {
    const long long fs = (long long)cx * H;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= fs) return;
    const int kx = (int)(q / H);
    const int y = (int)(q - (long long)kx * H);
    if (kx == 0 && y == 0) {
        out[q] = 0.f;
        return;
    }
    const double qd = crit_q(kx, y, cx, H, pix, kvf, dose);
    double w = 1.0, s = 0.0;
    for (int i = 0; i < N; i++) {
        w *= qd;
        s += w * w;
    }
    out[q] = (float)sqrt(s);
}

// ---------------------------------------------------------------------------
// Synthetic movie renderer (benchmark input only): Gaussian blobs on a unit
// background from per-tile blob lists, then Poisson counts by CDF inversion
// of a counter-based uniform.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ void k_synth(float* frames, int W, int H, long long pitch, long long fstride,
                        const float* blobs, const float* disp, int nblobs, const int* toff,
                        const int* tidx, int tile, int tiles_x, float scale,
                        unsigned long long seed, int poisson) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int f = blockIdx.z;
    if (x >= W) return;
    const int t = (y / tile) * tiles_x + (x / tile);
    float v = 1.0f;
    const float* d = disp + (long long)f * nblobs * 2;
    for (int e = toff[t]; e < toff[t + 1]; e++) {
        const int k = tidx[e];
        const float bx = blobs[4 * k] + d[2 * k], by = blobs[4 * k + 1] + d[2 * k + 1];
        float dx = x - bx, dy = y - by;
        dx -= W * rintf(dx / W);
        dy -= H * rintf(dy / H);
        const float s = blobs[4 * k + 2];
        const float r2 = dx * dx + dy * dy;
        if (r2 > 16.f * s * s) continue;
        v += blobs[4 * k + 3] * expf(-r2 / (2.f * s * s));
    }
    v *= scale;
    if (poisson) {
        const unsigned long long key = mix64(seed ^ mix64(((unsigned long long)f << 40) ^
                                                          ((unsigned long long)y << 20) ^
                                                          (unsigned long long)x));
        const double u = (key >> 11) * (1.0 / 9007199254740992.0);
        double p = exp(-(double)v), c = p;
        int k = 0;
        while (u > c && k < 1000) {
            k++;
            p *= v / k;
            c += p;
        }
        v = (float)k;
    }
    frames[(long long)f * fstride + (long long)y * pitch + x] = v;
}

// ===========================================================================
// launch wrappers
// ===========================================================================
int mc_align_init_constants() {
    bool attr_ok = true;
    if (mc_upload_prime_table() != 0) return -1;
    const int bytes = (mc_lds_padded(MC_LDS_ELEMS) + 64) * (int)sizeof(cfloat);
    attr_ok &= hipSuccess == hipFuncSetAttribute((const void*)k_correlate, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    attr_ok &= hipSuccess == hipFuncSetAttribute((const void*)k_finalize, hipFuncAttributeMaxDynamicSharedMemorySize, 65536);
#define X(HH, CC, NT, WPE, BL)                                                                \
    attr_ok &= hipSuccess == hipFuncSetAttribute((const void*)k_correlate_rp<HH, CC, NT, 0, WPE, BL>,                 \
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);              \
    attr_ok &= hipSuccess == hipFuncSetAttribute((const void*)k_correlate_rp<HH, CC, NT, 1, WPE, BL>,                 \
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);              \
    attr_ok &= hipSuccess == hipFuncSetAttribute((const void*)k_correlate_rp<HH, CC, NT, 2, WPE, BL>,                 \
                        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    MC_RP_CORR_SPECS(X)
#undef X
    attr_ok &= hipSuccess == hipFuncSetAttribute((const void*)k_row_dft, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    attr_ok &= hipSuccess == hipFuncSetAttribute((const void*)k_feature_map, hipFuncAttributeMaxDynamicSharedMemorySize,
                        2 * 128 * 128 * (int)sizeof(float));
    attr_ok &= hipSuccess == hipFuncSetAttribute((const void*)k_local_warp, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)warp_lds_bytes(MC_WARP_MAXP));
    return attr_ok ? 0 : -1;
}

hipError_t mc_launch_phasors(const double* shifts, double mult, int skip, const unsigned char* done,
                             int per_job, int W, int cx, int H, int batch, cfloat* A, cfloat* B,
                             hipStream_t s, int h_norm) {
    dim3 grid(mc_div_up(cx + H, 256), batch);
    hipLaunchKernelGGL(k_phasors, grid, dim3(256), 0, s, shifts, mult, skip, done,
                       per_job > 0 ? per_job : 1, W, cx, H, h_norm > 0 ? h_norm : H, A, B);
    return hipGetLastError();
}

hipError_t mc_launch_phasors_jobs(const IterState& st, const double* shifts, int H, cfloat* A, cfloat* B,
                                  hipStream_t s) {
    if (!st.jobs || !st.fjob || st.NT < 1) return hipErrorInvalidValue;
    dim3 grid(mc_div_up(st.cx_max + H, 256), st.NT);
    hipLaunchKernelGGL(k_phasors_jobs, grid, dim3(256), 0, s, shifts, st.done, st.jobs, st.fjob, H, A, B);
    return hipGetLastError();
}

hipError_t mc_launch_frame_weights(cfloat* S, long long fs, int N, int cx, int H, const double* shifts,
                                   int motion, int corr_interp, hipStream_t s) {
    if (N < 2 || cx < 2 || H < 1 || (motion && !shifts)) return hipErrorInvalidValue;
    if (!motion && !corr_interp) return hipSuccess;
    hipLaunchKernelGGL(k_frame_weights, dim3(mc_div_up(H, 256), cx, N), dim3(256), 0, s, S, fs, N, cx, H,
                       shifts, motion, corr_interp);
    return hipGetLastError();
}

hipError_t mc_launch_aligned_sum(const cfloat* F, const cfloat* A, const cfloat* B,
                                 const unsigned char* done, int J, int N, int cx, int H,
                                 cfloat* S, hipStream_t s, int f_fmt, const float* f_inv,
                                 const JobGeo* jobs) {
    // with a job table, cx is the widest job's
    dim3 grid(mc_div_up(H, 256), cx, J);
    switch (f_fmt) {
        case MC_SPEC_F32:
            hipLaunchKernelGGL(k_aligned_sum<MC_SPEC_F32>, grid, dim3(256), 0, s, F, f_inv, A, B,
                               done, N, cx, H, S, jobs);
            break;
        case MC_SPEC_F16:
            hipLaunchKernelGGL(k_aligned_sum<MC_SPEC_F16>, grid, dim3(256), 0, s, F, f_inv, A, B,
                               done, N, cx, H, S, jobs);
            break;
        case MC_SPEC_BF16:
            hipLaunchKernelGGL(k_aligned_sum<MC_SPEC_BF16>, grid, dim3(256), 0, s, F, f_inv, A, B,
                               done, N, cx, H, S, jobs);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t mc_launch_correlate(const FftPlan1& colplan, const cfloat* ptw, int Cc, const cfloat* F,
                               const cfloat* S, const cfloat* A, const cfloat* B,
                               const IterState& st, int cx, int H, float bft, int phase_only,
                               int seay, cfloat* T1, hipStream_t s, int f_fmt,
                               const float* f_inv) {
    CorrArgs a;
    a.plan = colplan;
    a.F = F;
    a.S = S;
    a.A = A;
    a.B = B;
    a.gstart = st.gstart;
    a.gsize = st.gsize;
    a.done = st.done;
    a.N = st.N;
    a.G = st.G;
    a.cx = cx;
    a.H = H;
    a.Cc = Cc;
    a.filt0 = -0.25f * bft / ((cx - 1) * H);
    a.phase_only = phase_only;
    a.seay = seay;
    a.T1 = T1;
    a.finv = f_inv;
    a.jobs = st.jobs;
    a.J = st.J;
    a.nbq = -0.25f * bft;
    // a job table brings its tile count (st.tiles, for mc_correlate_rp_cols(H))
    // and its widest job (st.cx_max, st.G_max)
    if (st.jobs && (st.cx_max < 1 || st.G_max < 1)) return hipErrorInvalidValue;
#define X(HH, CC, NT, WPE, BL)                                                                \
    if (H == HH && ptw && seay <= HH) {                                                  \
        if (st.jobs && st.tiles < 1) return hipErrorInvalidValue;                        \
        const dim3 grid(st.jobs ? st.tiles : mc_div_up(cx, CC) * st.G * st.J);           \
        const size_t lds = (size_t)CC * rp_cs(HH, CC) * sizeof(cfloat);                  \
        if (f_fmt == MC_SPEC_F32)                                                        \
            hipLaunchKernelGGL((k_correlate_rp<HH, CC, NT, 0, WPE, BL>), grid, dim3(NT), lds, s, a, ptw); \
        else if (f_fmt == MC_SPEC_F16)                                                   \
            hipLaunchKernelGGL((k_correlate_rp<HH, CC, NT, 1, WPE, BL>), grid, dim3(NT), lds, s, a, ptw); \
        else if (f_fmt == MC_SPEC_BF16)                                                  \
            hipLaunchKernelGGL((k_correlate_rp<HH, CC, NT, 2, WPE, BL>), grid, dim3(NT), lds, s, a, ptw); \
        else                                                                             \
            return hipErrorInvalidValue;                                                 \
        return hipGetLastError();                                                        \
    }
    MC_RP_CORR_SPECS(X)
#undef X
    if (f_fmt != MC_SPEC_F32) return hipErrorInvalidValue;
    while (Cc > 1 && (long long)Cc * H > MC_LDS_ELEMS) Cc /= 2;
    if ((long long)Cc * H > MC_LDS_ELEMS) return hipErrorInvalidValue;
    a.Cc = Cc;
    dim3 grid(mc_div_up(st.jobs ? st.cx_max : cx, Cc), st.jobs ? st.G_max : st.G, st.J);
    const size_t lds = (size_t)mc_lds_padded(Cc * H) * sizeof(cfloat);
    hipLaunchKernelGGL(k_correlate, grid, dim3(MC_NTHR), lds, s, a);
    return hipGetLastError();
}

bool mc_correlate_has_rp(int H) {
#define X(HH, CC, NT, WPE, BL) \
    if (H == HH) return true;
    MC_RP_CORR_SPECS(X)
#undef X
    return false;
}

int mc_correlate_rp_cols(int H) {
#define X(HH, CC, NT, WPE, BL) \
    if (H == HH) return CC;
    MC_RP_CORR_SPECS(X)
#undef X
    return 0;
}

hipError_t mc_launch_row_dft(const cfloat* T1, const cfloat* twW, const unsigned char* done, int J,
                             int G, int cx, int W, int seax, int seay, float* win, hipStream_t s,
                             const JobGeo* jobs) {
    // with a job table: G, cx and W are the largest over the jobs
    dim3 grid(seay, G, J);
    const size_t lds = (size_t)(cx + W) * sizeof(cfloat) + 256 * sizeof(float);
    hipLaunchKernelGGL(k_row_dft, grid, dim3(256), lds, s, T1, twW, done, G, cx, W, seax, seay, win,
                       jobs);
    return hipGetLastError();
}

hipError_t mc_launch_peaks(const float* win, const unsigned char* done, int J, int G, int nx,
                           int ny, double* gshift, hipStream_t s, const JobGeo* jobs) {
    hipLaunchKernelGGL(k_peaks, dim3(G, J), dim3(256), 0, s, win, done, G, nx, ny, gshift, jobs);
    return hipGetLastError();
}

hipError_t mc_launch_peaks_f(const float* win, int nx, int ny, int count, double* out,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_peaks_f, dim3(count), dim3(256), 0, s, win, nx, ny, out);
    return hipGetLastError();
}

hipError_t mc_launch_update(const IterState& st, int it, double tol, hipStream_t s) {
    UpdArgs a;
    a.tmp = st.tmp;
    a.total = st.total;
    a.gshift = st.gshift;
    a.errors = st.errors;
    a.iters = st.iters;
    a.done = st.done;
    a.fgroup = st.fgroup;
    a.N = st.N;
    a.G = st.G;
    a.max_iter = st.max_iter;
    a.it = it;
    a.tol = tol;
    a.jobs = st.jobs;
    hipLaunchKernelGGL(k_update, dim3(st.J), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t mc_launch_finalize(const IterState& st, int binz, int quirks, hipStream_t s) {
    FinArgs a;
    a.total = st.total;
    a.final_ = st.final_;
    a.scratch = st.scratch;
    a.dummy = nullptr;
    a.gcent = st.gcent;
    a.fcent = st.fcent;
    a.N = st.N;
    a.G = st.G;
    a.binz = binz;
    a.quirks = quirks;
    a.jobs = st.jobs;
    // with a job table, sized for the job with most frames
    a.sstride = 16LL * (st.jobs ? st.N_max : st.N) + 128;
    const size_t bytes = (size_t)a.sstride * sizeof(double);
    a.use_lds = bytes <= 65536;
    hipLaunchKernelGGL(k_finalize, dim3(st.J), dim3(64), a.use_lds ? bytes : 0, s, a);
    return hipGetLastError();
}

hipError_t mc_launch_despike(double* traj, double* scratch, int n, int count, int quirks,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_despike, dim3(count), dim3(64), 0, s, traj, scratch, n, quirks);
    return hipGetLastError();
}

hipError_t mc_launch_feature_map(const float* img128, unsigned char* feat, hipStream_t s) {
    hipLaunchKernelGGL(k_feature_map, dim3(1), dim3(1024), 2 * 128 * 128 * sizeof(float), s,
                       img128, feat);
    return hipGetLastError();
}

hipError_t mc_launch_patch_stats(double* shifts, const double* centers, int N, int P, int W, int H,
                                 int quirks, unsigned char* bad, int* ref, hipStream_t s) {
    if (N < 1 || N > MC_PS_MAXN || P < 1 || P > MC_WARP_MAXP) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_patch_stats, dim3(1), dim3(256), 0, s, shifts, centers, N, P, W, H,
                       quirks, bad, ref);
    return hipGetLastError();
}

hipError_t mc_launch_local_warp(const float* R, float* Rp, long long pitch, long long fstride,
                                int W, int H, int N, int P, const float* cen, const float* sh,
                                const unsigned char* bad, unsigned long long bad_any, hipStream_t s) {
    if (P < 1 || P > MC_WARP_MAXP) return hipErrorInvalidValue;
    if (pitch * H >= (1LL << 31) || pitch < W) return hipErrorInvalidValue;
    WarpArgs a{R, Rp, pitch, fstride, W, H, N, P, cen, sh, bad, bad_any};
    dim3 grid(mc_div_up(W, MC_WARP_TX), mc_div_up(H, MC_WARP_TY));
    hipLaunchKernelGGL(k_local_warp, grid, dim3(MC_WARP_NT), warp_lds_bytes(P), s, a);
    return hipGetLastError();
}

hipError_t mc_launch_dw_sums(const cfloat* F, long long fstride, int N, int cx, int H,
                             const cfloat* A, const cfloat* B, int dw, double pix, double kvf,
                             double dose, int dws, double r0, double r1, int split, cfloat* sums,
                             long long sstride, hipStream_t s, const int* cum, int f_fmt,
                             const float* f_inv) {
    DwArgs a;
    a.F = F;
    a.finv = f_inv;
    a.fstride = fstride;
    a.N = N;
    a.cx = cx;
    a.H = H;
    a.A = A;
    a.B = B;
    a.dw = dw;
    a.pix = pix;
    a.kvf = kvf;
    a.dose = dose;
    a.dws = dws;
    a.r0 = r0;
    a.r1 = r1;
    a.split = split;
    a.sums = sums;
    a.sstride = sstride;
    a.cum = cum;
    dim3 grid(mc_div_up(H, 256), cx);
    switch (f_fmt) {
        case MC_SPEC_F32: hipLaunchKernelGGL(k_dw_sums<MC_SPEC_F32>, grid, dim3(256), 0, s, a); break;
        case MC_SPEC_F16: hipLaunchKernelGGL(k_dw_sums<MC_SPEC_F16>, grid, dim3(256), 0, s, a); break;
        case MC_SPEC_BF16: hipLaunchKernelGGL(k_dw_sums<MC_SPEC_BF16>, grid, dim3(256), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t mc_launch_dose_wsum(int cx, int H, int N, double pix, double kvf, double dose,
                               float* out, hipStream_t s) {
    dim3 grid(mc_div_up((long long)cx * H, 256));
    hipLaunchKernelGGL(k_dose_wsum, grid, dim3(256), 0, s, cx, H, N, pix, kvf, dose, out);
    return hipGetLastError();
}

hipError_t mc_launch_synth(float* frames, int W, int H, int N, long long pitch, long long fstride,
                           const float* blobs, const float* disp, int nblobs, const int* tile_off,
                           const int* tile_idx, int tile, int tiles_x, float scale,
                           unsigned long long seed, int poisson, hipStream_t s) {
    dim3 grid(mc_div_up(W, 256), H, N);
    hipLaunchKernelGGL(k_synth, grid, dim3(256), 0, s, frames, W, H, pitch, fstride, blobs, disp,
                       nblobs, tile_off, tile_idx, tile, tiles_x, scale, seed, poisson);
    return hipGetLastError();
}
