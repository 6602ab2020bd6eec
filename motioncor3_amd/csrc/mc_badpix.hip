// mc_badpix.hip -- bad-pixel detection and correction before the alignment.
//
// CDetectMain::DoIt (BadPixel/CDetectMain.cpp:54-75): the frame sum
// (MrcUtil/CSumFFTStack.cpp:65-117), hot pixels above
// mean + max(thr sigma, 10) of the sum (BadPixel/GDetectHot.cu:9-55), defect
// patches from a local correlation with a border template
// (BadPixel/CLocalCCMap.cpp, GLocalCC.cu:15-53, CTemplate.cpp:19-57,
// GDetectPatch.cu:9-105), then defect rectangles and an 8-bit defect map.
// CCorrectMain (BadPixel/CCorrectMain.cpp:18-81, GCorrectBad.cu:13-64)
// replaces every bad pixel by a good one found by a pseudo-random window
// search.  The reference indexes frames and map with the padded pitch
// (W/2+1)*2; the search's seeds and coordinates use that index space here
// too, so the choices are the same for any frame pitch.
#include <hip/hip_runtime.h>

#include "mc_badpix.h"

namespace {

// sum[y][x] = sum_n frame_n[y][x], frames added in order (FP32).
__global__ void k_frame_sum(const float* __restrict__ fr, int W, int H, int N, long long pitch,
                            long long fstride, float* __restrict__ sum) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    const float* p = fr + (long long)y * pitch + x;
    float s = p[0];
    for (int n = 1; n < N; n++) s = s + p[(long long)n * fstride];
    sum[(long long)y * W + x] = s;
}

// Per-workgroup sums of x and x^2 (FP64) over a W x H image.
__global__ __launch_bounds__(256) void k_moments(const float* __restrict__ img, long long n,
                                                 double* psum, double* psq) {
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const double v = img[i];
        a += v;
        b += v * v;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s1[threadIdx.x] += s1[threadIdx.x + s];
            s2[threadIdx.x] += s2[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        psum[blockIdx.x] = s1[0];
        psq[blockIdx.x] = s2[0];
    }
}

__global__ void k_hot(const float* __restrict__ sum, long long n, float thr, unsigned char* bad) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bad[i] = sum[i] <= thr ? 0 : 1;
}

// cc[y][x] = |mean(T * I)| / std(I) over the mw x mh window starting at
// (x, y), 0 where the window reaches the last column or row.
__global__ void k_local_cc(const float* __restrict__ img, int W, int H, const float* __restrict__ tpl,
                           int mw, int mh, float* __restrict__ cc) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    float out = 0.f;
    if (x + mw < W && y + mh < H) {
        const float* src = img + (long long)y * W + x;
        float mean = 0.f, sq = 0.f, c = 0.f;
        for (int j = 0; j < mh; j++)
            for (int i = 0; i < mw; i++) {
                const float v = src[(long long)j * W + i];
                mean += v;
                sq += v * v;
                c += tpl[j * mw + i] * v;
            }
        const float n = (float)(mw * mh);
        c /= n;
        mean /= n;
        float sd = sq / n - mean * mean;
        sd = sd < 0.f ? 0.f : sqrtf(sd);
        out = sd == 0.f ? 0.f : fabsf(c) / sd;
    }
    cc[(long long)y * W + x] = out;
}

// Mark the mw x mh block at every (x, y) whose correlation exceeds thr.
__global__ void k_cc_mark(const float* __restrict__ cc, int W, int H, int mw, int mh, float thr,
                          unsigned char* bad) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x + mw >= W || y + mh >= H) return;
    if (!(cc[(long long)y * W + x] > thr)) return;
    for (int j = 0; j < mh; j++)
        for (int i = 0; i < mw; i++) bad[(long long)(y + j) * W + x + i] = 1;
}

// GCorrectBad (BadPixel/GCorrectBad.cu:13-64) for frame blockIdx.z.
__global__ void k_correct_bad(float* fr, int W, int H, long long pitch, long long fstride,
                              const unsigned char* __restrict__ bad, int win0) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W || !bad[(long long)y * W + x]) return;
    float* f = fr + (long long)blockIdx.z * fstride;
    const unsigned int padx = (unsigned int)((W / 2 + 1) * 2);
    unsigned int next = ((unsigned int)y * padx + (unsigned int)x) * 109u + 619u;
    int win = win0;
    for (int j = 0; j < 10; j++) {
        unsigned int nk = next;
        const int w2 = win * win, half = win / 2;
        for (int k = 0; k < 500; k++) {
            nk *= 997u;
            int ix = (int)(nk % (unsigned int)w2);
            const int iy = ix / win - half + y;
            if (iy < 0 || iy >= H) continue;
            ix = ix % win - half + x;
            if (ix < 0 || ix >= W) continue;
            if (bad[(long long)iy * W + ix] == 1) continue;
            f[(long long)y * pitch + x] = f[(long long)iy * pitch + ix];
            return;
        }
        win += 50;
        next *= 997u;
    }
    const unsigned int idx = next % (unsigned int)(W * H);
    const int iy = (int)(idx / (unsigned int)W), ix = (int)(idx % (unsigned int)W);
    f[(long long)y * pitch + x] = f[(long long)iy * pitch + ix];
}

inline unsigned int nblk(long long n, int b) { return (unsigned int)((n + b - 1) / b); }

}  // namespace

hipError_t mc_launch_frame_sum(const float* fr, int W, int H, int N, long long pitch,
                               long long fstride, float* sum, hipStream_t s) {
    hipLaunchKernelGGL(k_frame_sum, dim3(nblk(W, 256), H), dim3(256), 0, s, fr, W, H, N, pitch, fstride, sum);
    return hipGetLastError();
}

hipError_t mc_launch_moments(const float* img, long long n, double* psum, double* psq, hipStream_t s) {
    hipLaunchKernelGGL(k_moments, dim3(MC_BAD_BLOCKS), dim3(256), 0, s, img, n, psum, psq);
    return hipGetLastError();
}

hipError_t mc_launch_hot(const float* sum, long long n, float thr, unsigned char* bad, hipStream_t s) {
    hipLaunchKernelGGL(k_hot, dim3(nblk(n, 256)), dim3(256), 0, s, sum, n, thr, bad);
    return hipGetLastError();
}

hipError_t mc_launch_local_cc(const float* img, int W, int H, const float* tpl, int mw, int mh,
                              float* cc, hipStream_t s) {
    hipLaunchKernelGGL(k_local_cc, dim3(nblk(W, 256), H), dim3(256), 0, s, img, W, H, tpl, mw, mh, cc);
    return hipGetLastError();
}

hipError_t mc_launch_cc_mark(const float* cc, int W, int H, int mw, int mh, float thr,
                             unsigned char* bad, hipStream_t s) {
    hipLaunchKernelGGL(k_cc_mark, dim3(nblk(W, 256), H), dim3(256), 0, s, cc, W, H, mw, mh, thr, bad);
    return hipGetLastError();
}

hipError_t mc_launch_correct_bad(float* fr, int W, int H, int N, long long pitch, long long fstride,
                                 const unsigned char* bad, int win, hipStream_t s) {
    if (N < 1 || N > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_correct_bad, dim3(nblk(W, 256), H, N), dim3(256), 0, s, fr, W, H, pitch,
                       fstride, bad, win);
    return hipGetLastError();
}
