// mc_prior.hip -- learned shift prior for the patch alignment.
//
// No reference counterpart: the reference starts every patch's
// CIterativeAlign from zero shifts (pStackShift->Reset(),
// Align/CIterativeAlign.cpp:68) and iterates until the largest group shift
// is below -Tol (:117-145, at least two iterations).  Here the first
// iteration runs as in the reference; a small MLP (in -> H -> H -> H -> out,
// ReLU) then reads that coarse pass (its peak windows and measured shifts)
// together with the global trajectory and the patch centre, and predicts the
// correction from the first iteration's shifts to the converged ones.  The
// iterations continue from the corrected guess.
//
// Weights are FP32 masters with BF16 copies; every layer product runs on
// MFMA (v_mfma_f32_32x32x16_bf16, FP32 accumulation).  Training (MSE loss,
// Adam) uses the kernels below, so the network is trained on the GPU by the
// same forward pass that the alignment uses.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "mc_prior.h"
#include "mc_spec.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short to_bf16(float f) { return (unsigned short)mc_bf16_bits(f); }

// C[i][j] = sum_k A(i,k) B(k,j), A(i,k) = A[i*sai + k*sak],
// B(k,j) = B[k*sbk + j*sbj], both BF16; C FP32 row-major with ldc.  One wave
// per 32x32 tile of C; operands outside [M) x [K) x [N) read as zero and C
// is written only inside [M) x [N).
// Operand layout of gfx950's v_mfma_f32_32x32x16_bf16 (lane l, 8 values
// each): A: row l % 32, k = 8 (l / 32) + 0..7;  B: column l % 32, same k.
//   D: 16 values per lane, value e: row 8 (e / 4) + 4 (l / 32) + e % 4,
//      column l % 32.
__global__ __launch_bounds__(64) void k_gemm_bf16(const unsigned short* __restrict__ A, long long sai,
                                                  long long sak, const unsigned short* __restrict__ B,
                                                  long long sbk, long long sbj, float* __restrict__ C,
                                                  int ldc, int M, int N, int K) {
    const int l = threadIdx.x;
    const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
    const int ra = i0 + (l & 31), cb = j0 + (l & 31), kq = 8 * (l >> 5);
    f32x16 acc;
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = 0.f;
    const bool rok = ra < M, cok = cb < N;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int k = k0 + kq + t;
            const unsigned short ua = (rok && k < K) ? A[ra * sai + k * sak] : (unsigned short)0;
            const unsigned short ub = (cok && k < K) ? B[k * sbk + cb * sbj] : (unsigned short)0;
            a[t] = __builtin_bit_cast(__bf16, ua);
            b[t] = __builtin_bit_cast(__bf16, ub);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int e = 0; e < 16; e++) {
        const int r = i0 + 8 * (e >> 2) + 4 * (l >> 5) + (e & 3);
        if (r < M && cok) C[(long long)r * ldc + cb] = acc[e];
    }
}

__global__ void k_to_bf16(const float* __restrict__ x, unsigned short* __restrict__ y, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// Z[r][j] += b[j]; H = relu(Z) (the output layer: H = Z), H16 = bf16(H).
__global__ void k_bias_act(float* Z, const float* __restrict__ b, float* H, unsigned short* H16,
                           int rows, int cols, int relu) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * cols) return;
    const float z = Z[i] + b[i % cols];
    Z[i] = z;
    const float h = relu ? fmaxf(z, 0.f) : z;
    H[i] = h;
    if (H16) H16[i] = to_bf16(h);
}

// dY = 2 (Y - T) / count with its BF16 copy, and the squared errors.
__global__ void k_mse_grad(const float* Y, const float* T, float* dY, unsigned short* dY16,
                           float* se, long long n, float inv_count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = Y[i] - T[i];
    const float g = 2.f * d * inv_count;
    dY[i] = g;
    dY16[i] = to_bf16(g);
    se[i] = d * d;
}

// dZ = dH * (Z > 0), with its BF16 copy.
__global__ void k_relu_back(const float* dH, const float* Z, float* dZ, unsigned short* dZ16,
                            long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = Z[i] > 0.f ? dH[i] : 0.f;
    dZ[i] = v;
    dZ16[i] = to_bf16(v);
}

// db[j] = sum_r dZ[r][j] in row order.
__global__ void k_colsum(const float* dZ, float* db, int rows, int cols) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; r++) s += dZ[(long long)r * cols + j];
    db[j] = s;
}

// acc[0] += sum x (one workgroup of 256, FP64, fixed order).
__global__ __launch_bounds__(256) void k_sum_add(const float* x, long long n, double* acc) {
    __shared__ double red[256];
    double s = 0.0;
    for (long long i = threadIdx.x; i < n; i += 256) s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += red[0];
}

// Adam with bias correction; refreshes the BF16 weight copy.
__global__ void k_adam(float* w, const float* g, float* m, float* v, unsigned short* w16, long long n,
                       float lr, float b1, float b2, float c1, float c2, float eps) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    w16[i] = to_bf16(wi);
}

// dst[r] = src[idx[r]] for rows of `cols` floats (shuffled minibatches).
__global__ void k_gather_rows(const float* __restrict__ src, const int* __restrict__ idx, float* dst,
                              int rows, int cols) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)rows * cols) return;
    const int r = (int)(i / cols), c = (int)(i - (long long)r * cols);
    dst[i] = src[(long long)idx[r] * cols + c];
}

// Features of job j (row j of X, D = ns + 2N + 64N):
//   [0, ns)           side[j] (global trajectory and patch centre, from the host)
//   [ns, ns + 2N)     the first iteration's frame shifts / unit
//   [ns + 2N, D)      per frame, the 8x8 2x2-average pool of the central
//                     16x16 of its group's correlation window, divided by the
//                     pool's largest |value|
// One workgroup of 64 per job.
__global__ __launch_bounds__(64) void k_prior_features(const float* __restrict__ side, int ns,
                                                       const double* __restrict__ tmp, float inv_unit,
                                                       const float* __restrict__ win, int N, int G,
                                                       const int* __restrict__ fgroup, int seax,
                                                       int seay, float* __restrict__ X, int D) {
    const int j = blockIdx.x, t = threadIdx.x;
    float* x = X + (long long)j * D;
    for (int i = t; i < ns; i += 64) x[i] = side[(long long)j * ns + i];
    for (int i = t; i < 2 * N; i += 64) x[ns + i] = (float)(tmp[(long long)j * 2 * N + i] * inv_unit);
    __shared__ float vmax[64];
    const int cy = seay / 2 - 8, cx = seax / 2 - 8;
    const int py = t >> 3, px = t & 7;
    for (int f = 0; f < N; f++) {
        const float* w = win + ((long long)j * G + fgroup[f]) * seay * seax;
        const float* q = w + (long long)(cy + 2 * py) * seax + cx + 2 * px;
        const float v = 0.25f * (q[0] + q[1] + q[seax] + q[seax + 1]);
        vmax[t] = fabsf(v);
        __syncthreads();
        for (int s = 32; s > 0; s >>= 1) {
            if (t < s) vmax[t] = fmaxf(vmax[t], vmax[t + s]);
            __syncthreads();
        }
        x[ns + 2 * N + f * 64 + t] = v / (vmax[0] + 1e-30f);
        __syncthreads();
    }
}

// Start the iterations of job j from the predicted trajectory: group total
// += unit * (mean prediction over the group's frames), and every frame's
// running shift becomes its group's total (as the iterations keep it).
__global__ void k_prior_apply(const float* __restrict__ Y, int N, int G, const int* __restrict__ gstart,
                              const int* __restrict__ gsize, const int* __restrict__ fgroup,
                              double unit, double* total, double* tmp) {
    const int j = blockIdx.x;
    if (threadIdx.x != 0) return;
    const float* y = Y + (long long)j * 2 * N;
    double* tj = total + (long long)j * G * 2;
    for (int g = 0; g < G; g++) {
        double sx = 0, sy = 0;
        for (int i = gstart[g]; i < gstart[g] + gsize[g]; i++) {
            sx += y[2 * i];
            sy += y[2 * i + 1];
        }
        tj[2 * g] += unit * sx / gsize[g];
        tj[2 * g + 1] += unit * sy / gsize[g];
    }
    for (int i = 0; i < N; i++) {
        tmp[((long long)j * N + i) * 2] = tj[2 * fgroup[i]];
        tmp[((long long)j * N + i) * 2 + 1] = tj[2 * fgroup[i] + 1];
    }
}

// After the guided iteration (index 1): a job whose largest group shift
// exceeds `fallback` restarts from zero shifts (a full plain solve follows);
// every other job stops.
__global__ void k_prior_gate(double* errors, int max_iter, unsigned char* done, int* iters,
                             double* tmp, double* total, int N, int G, double fallback,
                             unsigned char* fell) {
    const int j = blockIdx.x;
    if (threadIdx.x != 0) return;
    if (done[j] || errors[(long long)j * max_iter + 1] <= fallback) {
        done[j] = 1;
        fell[j] = 0;
        return;
    }
    fell[j] = 1;
    for (int i = 0; i < 2 * N; i++) tmp[(long long)j * 2 * N + i] = 0.0;
    for (int i = 0; i < 2 * G; i++) total[(long long)j * 2 * G + i] = 0.0;
    for (int i = 0; i < max_iter; i++) errors[(long long)j * max_iter + i] = 0.0;
    iters[j] = 0;
    done[j] = 0;
}

inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

hipError_t gemm(hipStream_t s, const unsigned short* A, long long sai, long long sak,
                const unsigned short* B, long long sbk, long long sbj, float* C, int ldc, int M,
                int N, int K) {
    if (M <= 0 || N <= 0 || K <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gemm_bf16, dim3(nblk(N, 32), nblk(M, 32)), dim3(64), 0, s, A, sai, sak, B,
                       sbk, sbj, C, ldc, M, N, K);
    return hipGetLastError();
}

}  // namespace

// ---------------------------------------------------------------------------
int Prior::init(int din, int hid, int dout, unsigned long long seed) {
    release();
    if (din <= 0 || hid <= 0 || dout <= 0) return -1;
    dims[0] = din;
    dims[1] = dims[2] = dims[3] = hid;
    dims[4] = dout;
    long long off = 0;
    for (int l = 0; l < kLayers; l++) {
        woff[l] = off;
        off += (long long)dims[l] * dims[l + 1];
        boff[l] = off;
        off += dims[l + 1];
    }
    nparam = off;
    std::vector<float> h(nparam, 0.f);
    // He-uniform weights from a seeded 64-bit LCG (top 53 bits), zero biases;
    // the output layer starts small so the first guess is near the first
    // iteration's own shifts
    unsigned long long st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    for (int l = 0; l < kLayers; l++) {
        const double lim = std::sqrt(6.0 / dims[l]) * (l == kLayers - 1 ? 0.1 : 1.0);
        for (long long i = 0; i < (long long)dims[l] * dims[l + 1]; i++) {
            st = st * 6364136223846793005ULL + 1442695040888963407ULL;
            const double u = (double)(st >> 11) / 9007199254740992.0;
            h[woff[l] + i] = (float)((2.0 * u - 1.0) * lim);
        }
    }
    const size_t fb = nparam * sizeof(float);
    if (hipMalloc(&w, fb) != hipSuccess || hipMalloc(&g, fb) != hipSuccess ||
        hipMalloc(&m, fb) != hipSuccess || hipMalloc(&v, fb) != hipSuccess ||
        hipMalloc(&w16, nparam * sizeof(unsigned short)) != hipSuccess || !scratch("acc", 8))
        return -1;
    if (hipMemcpy(w, h.data(), fb, hipMemcpyHostToDevice) != hipSuccess) return -1;
    ready = true;
    return reset_moments(nullptr) == hipSuccess && refresh16(nullptr) == hipSuccess &&
                   hipDeviceSynchronize() == hipSuccess
               ? 0
               : -1;
}

void Prior::release() {
    for (void* p : {(void*)w, (void*)g, (void*)m, (void*)v, (void*)w16})
        if (p) (void)hipFree(p);
    for (auto& kv : work)
        if (kv.second) (void)hipFree(kv.second);
    work.clear();
    worksz.clear();
    w = g = m = v = nullptr;
    w16 = nullptr;
    ready = false;
}

hipError_t Prior::refresh16(hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk(nparam, 256)), dim3(256), 0, s, w, w16, nparam);
    return hipGetLastError();
}

hipError_t Prior::reset_moments(hipStream_t s) {
    adam_t = 0;
    hipError_t e = hipMemsetAsync(m, 0, nparam * sizeof(float), s);
    if (e == hipSuccess) e = hipMemsetAsync(v, 0, nparam * sizeof(float), s);
    return e;
}

void* Prior::scratch(const char* name, size_t bytes) {
    auto it = work.find(name);
    if (it != work.end() && it->second && worksz[name] >= bytes) return it->second;
    if (it != work.end() && it->second) (void)hipFree(it->second);
    void* p = nullptr;
    if (hipMalloc(&p, bytes > 0 ? bytes : 256) != hipSuccess) p = nullptr;
    work[name] = p;
    worksz[name] = p ? bytes : 0;
    return p;
}

// Work buffers for passes over up to n rows.
bool Prior::reserve(int n) {
    const int hid = dims[1];
    const size_t big = (size_t)n * std::max(std::max(dims[0], hid), dims[4]);
    bool ok = scratch("x16", big * 2) && scratch("dz", big * 4) && scratch("dh", big * 4) &&
              scratch("dz16", big * 2) && scratch("se", (size_t)n * dims[4] * 4);
    for (int l = 0; l < kLayers; l++) {
        const char zn[3] = {'z', (char)('0' + l), 0}, hn[3] = {'h', (char)('0' + l), 0},
                   bn[3] = {'b', (char)('0' + l), 0};
        ok = ok && scratch(zn, (size_t)n * dims[l + 1] * 4) &&
             scratch(hn, (size_t)n * dims[l + 1] * 4) && scratch(bn, (size_t)n * dims[l + 1] * 2);
    }
    return ok;
}

// Forward over n rows of X (FP32 row-major, n x dims[0]) into Y
// (n x dims[4]); Z and H of every layer stay in the work buffers for a
// following backward pass.  reserve(n) must have succeeded.
hipError_t Prior::forward(const float* X, int n, float* Y, hipStream_t s) {
    unsigned short* x16 = (unsigned short*)work["x16"];
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk((long long)n * dims[0], 256)), dim3(256), 0, s, X, x16,
                       (long long)n * dims[0]);
    const unsigned short* in16 = x16;
    for (int l = 0; l < kLayers; l++) {
        const int din = dims[l], dout = dims[l + 1];
        const char zn[3] = {'z', (char)('0' + l), 0}, hn[3] = {'h', (char)('0' + l), 0},
                   bn[3] = {'b', (char)('0' + l), 0};
        float* Z = (float*)work[zn];
        const bool last = l == kLayers - 1;
        float* Hh = last ? Y : (float*)work[hn];
        unsigned short* H16 = last ? nullptr : (unsigned short*)work[bn];
        hipError_t e = gemm(s, in16, din, 1, w16 + woff[l], dout, 1, Z, dout, n, dout, din);
        if (e) return e;
        hipLaunchKernelGGL(k_bias_act, dim3(nblk((long long)n * dout, 256)), dim3(256), 0, s, Z,
                           w + boff[l], Hh, H16, n, dout, last ? 0 : 1);
        in16 = H16;
    }
    return hipGetLastError();
}

// Gradient of the MSE loss mean((Y - T)^2) for every parameter into g, after
// forward() on the same rows.  The squared errors go to work["se"].
hipError_t Prior::backward(const float* Y, const float* T, int n, hipStream_t s) {
    const int dout = dims[kLayers];
    float* dZ = (float*)work["dz"];
    float* dH = (float*)work["dh"];
    unsigned short* dZ16 = (unsigned short*)work["dz16"];
    float* se = (float*)work["se"];
    const long long no = (long long)n * dout;
    hipLaunchKernelGGL(k_mse_grad, dim3(nblk(no, 256)), dim3(256), 0, s, Y, T, dZ, dZ16, se, no,
                       1.f / (float)no);
    for (int l = kLayers - 1; l >= 0; l--) {
        const int din = dims[l], dO = dims[l + 1];
        const char bp[3] = {'b', (char)('0' + l - 1), 0}, zp[3] = {'z', (char)('0' + l - 1), 0};
        const unsigned short* in16 = l == 0 ? (const unsigned short*)work["x16"]
                                            : (const unsigned short*)work[bp];
        // dW = in^T dZ (din x dO), db = column sums of dZ
        hipError_t e = gemm(s, in16, 1, din, dZ16, dO, 1, g + woff[l], dO, din, dO, n);
        if (e) return e;
        hipLaunchKernelGGL(k_colsum, dim3(nblk(dO, 128)), dim3(128), 0, s, dZ, g + boff[l], n, dO);
        if (l == 0) break;
        // dH = dZ W^T (n x din), then through the ReLU of layer l-1
        e = gemm(s, dZ16, dO, 1, w16 + woff[l], 1, dO, dH, din, n, din, dO);
        if (e) return e;
        hipLaunchKernelGGL(k_relu_back, dim3(nblk((long long)n * din, 256)), dim3(256), 0, s, dH,
                           (const float*)work[zp], dZ, dZ16, (long long)n * din);
    }
    return hipGetLastError();
}

hipError_t Prior::loss_sum(int n, double* acc, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_add, dim3(1), dim3(256), 0, s, (const float*)work["se"],
                       (long long)n * dims[kLayers], acc);
    return hipGetLastError();
}

hipError_t Prior::adam_step(float lr, hipStream_t s) {
    adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)adam_t), c2 = 1.f - std::pow(b2, (float)adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nblk(nparam, 256)), dim3(256), 0, s, w, g, m, v, w16, nparam, lr,
                       b1, b2, c1, c2, 1e-8f);
    return hipGetLastError();
}

// Minibatch Adam over n rows of X / T (device), `epochs` passes, rows
// shuffled per epoch by a seeded Fisher-Yates permutation.  epoch_loss[e]
// (host, may be null) receives the mean squared error over the epoch's
// batches (each measured before its batch's step).
hipError_t Prior::train(const float* X, const float* T, int n, int epochs, int batch, float lr,
                        unsigned long long seed, float* epoch_loss, hipStream_t s) {
    if (n <= 0 || epochs <= 0 || batch <= 0) return hipErrorInvalidValue;
    const int nb = std::min(batch, n), din = dims[0], dout = dims[kLayers];
    if (!reserve(nb)) return hipErrorOutOfMemory;
    float* Xb = (float*)scratch("xb", (size_t)nb * din * 4);
    float* Tb = (float*)scratch("tb", (size_t)nb * dout * 4);
    float* Yb = (float*)scratch("yb", (size_t)nb * dout * 4);
    int* idx = (int*)scratch("idx", (size_t)n * 4);
    double* acc = (double*)work["acc"];
    if (!Xb || !Tb || !Yb || !idx || !acc) return hipErrorOutOfMemory;
    std::vector<int> perm(n);
    unsigned long long st = seed * 2862933555777941757ULL + 3037000493ULL;
    hipError_t e;
    for (int ep = 0; ep < epochs; ep++) {
        for (int i = 0; i < n; i++) perm[i] = i;
        for (int i = n - 1; i > 0; i--) {
            st = st * 6364136223846793005ULL + 1442695040888963407ULL;
            const int k = (int)((st >> 33) % (unsigned long long)(i + 1));
            std::swap(perm[i], perm[k]);
        }
        // perm stays unchanged until the stream synchronisation below
        if ((e = hipMemcpyAsync(idx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s))) return e;
        if ((e = hipMemsetAsync(acc, 0, sizeof(double), s))) return e;
        for (int b0 = 0; b0 < n; b0 += nb) {
            const int r = std::min(nb, n - b0);
            if ((e = prior_gather_rows(X, idx + b0, Xb, r, din, s))) return e;
            if ((e = prior_gather_rows(T, idx + b0, Tb, r, dout, s))) return e;
            if ((e = forward(Xb, r, Yb, s))) return e;
            if ((e = backward(Yb, Tb, r, s))) return e;
            if ((e = loss_sum(r, acc, s))) return e;
            if ((e = adam_step(lr, s))) return e;
        }
        double h = 0.0;
        if ((e = hipMemcpyAsync(&h, acc, sizeof(double), hipMemcpyDeviceToHost, s))) return e;
        if ((e = hipStreamSynchronize(s))) return e;
        if (epoch_loss) epoch_loss[ep] = (float)(h / ((double)n * dout));
    }
    return hipSuccess;
}

hipError_t prior_gather_rows(const float* src, const int* idx, float* dst, int rows, int cols,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_gather_rows, dim3(nblk((long long)rows * cols, 256)), dim3(256), 0, s, src,
                       idx, dst, rows, cols);
    return hipGetLastError();
}

hipError_t prior_features(const float* side, int ns, const double* tmp, float unit, const float* win,
                          int J, int N, int G, const int* fgroup, int seax, int seay, float* X,
                          int D, hipStream_t s) {
    if (seax < 16 || seay < 16 || D != ns + 66 * N) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_prior_features, dim3(J), dim3(64), 0, s, side, ns, tmp, 1.0f / unit, win, N,
                       G, fgroup, seax, seay, X, D);
    return hipGetLastError();
}

hipError_t prior_apply(const float* Y, int J, int N, int G, const int* gstart, const int* gsize,
                       const int* fgroup, double unit, double* total, double* tmp, hipStream_t s) {
    hipLaunchKernelGGL(k_prior_apply, dim3(J), dim3(64), 0, s, Y, N, G, gstart, gsize, fgroup, unit,
                       total, tmp);
    return hipGetLastError();
}

hipError_t prior_gate(double* errors, int max_iter, unsigned char* done, int* iters, double* tmp,
                      double* total, int J, int N, int G, double fallback, unsigned char* fell,
                      hipStream_t s) {
    if (max_iter < 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_prior_gate, dim3(J), dim3(64), 0, s, errors, max_iter, done, iters, tmp,
                       total, N, G, fallback, fell);
    return hipGetLastError();
}

hipError_t prior_gemm(const unsigned short* A, long long sai, long long sak, const unsigned short* B,
                      long long sbk, long long sbj, float* C, int ldc, int M, int N, int K,
                      hipStream_t s) {
    return gemm(s, A, sai, sak, B, sbk, sbj, C, ldc, M, N, K);
}
