// mc_prior.h -- learned shift prior (MLP) for the patch alignment.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>

// Four dense layers in -> H -> H -> H -> out, ReLU between them.
// Parameters (FP32, one buffer): W0 (in x H, row-major), b0, W1, b1, W2, b2,
// W3 (H x out), b3; the layer product of rows x is x W + b.
struct Prior {
    static constexpr int kLayers = 4;
    int dims[kLayers + 1] = {0, 0, 0, 0, 0};
    long long woff[kLayers] = {0, 0, 0, 0}, boff[kLayers] = {0, 0, 0, 0};
    long long nparam = 0;
    float* w = nullptr;             // FP32 master parameters
    float* g = nullptr;             // gradients, same layout
    float* m = nullptr;             // Adam moments
    float* v = nullptr;
    unsigned short* w16 = nullptr;  // BF16 copy (MFMA operands)
    int adam_t = 0;
    bool ready = false;
    std::map<std::string, void*> work;
    std::map<std::string, size_t> worksz;

    int init(int din, int hid, int dout, unsigned long long seed);
    void release();
    hipError_t refresh16(hipStream_t s);
    hipError_t reset_moments(hipStream_t s);
    void* scratch(const char* name, size_t bytes);
    bool reserve(int n);
    hipError_t forward(const float* X, int n, float* Y, hipStream_t s);
    hipError_t backward(const float* Y, const float* T, int n, hipStream_t s);
    hipError_t loss_sum(int n, double* acc, hipStream_t s);
    hipError_t adam_step(float lr, hipStream_t s);
    hipError_t train(const float* X, const float* T, int n, int epochs, int batch, float lr,
                     unsigned long long seed, float* epoch_loss, hipStream_t s);
    ~Prior() { release(); }
};

hipError_t prior_gather_rows(const float* src, const int* idx, float* dst, int rows, int cols,
                             hipStream_t s);
// Feature rows (D = ns + 66 N) of J patch jobs after the first iteration.
hipError_t prior_features(const float* side, int ns, const double* tmp, float unit, const float* win,
                          int J, int N, int G, const int* fgroup, int seax, int seay, float* X,
                          int D, hipStream_t s);
hipError_t prior_apply(const float* Y, int J, int N, int G, const int* gstart, const int* gsize,
                       const int* fgroup, double unit, double* total, double* tmp, hipStream_t s);
hipError_t prior_gate(double* errors, int max_iter, unsigned char* done, int* iters, double* tmp,
                      double* total, int J, int N, int G, double fallback, unsigned char* fell,
                      hipStream_t s);
hipError_t prior_gemm(const unsigned short* A, long long sai, long long sak, const unsigned short* B,
                      long long sbk, long long sbj, float* C, int ldc, int M, int N, int K,
                      hipStream_t s);
