// mc_badpix.h -- bad-pixel detection and correction before the alignment.
#pragma once
#include <hip/hip_runtime.h>

#define MC_BAD_BLOCKS 1024

hipError_t mc_launch_frame_sum(const float* fr, int W, int H, int N, long long pitch,
                               long long fstride, float* sum, hipStream_t s);
hipError_t mc_launch_moments(const float* img, long long n, double* psum, double* psq, hipStream_t s);
hipError_t mc_launch_hot(const float* sum, long long n, float thr, unsigned char* bad, hipStream_t s);
hipError_t mc_launch_local_cc(const float* img, int W, int H, const float* tpl, int mw, int mh,
                              float* cc, hipStream_t s);
hipError_t mc_launch_cc_mark(const float* cc, int W, int H, int mw, int mh, float thr,
                             unsigned char* bad, hipStream_t s);
hipError_t mc_launch_correct_bad(float* fr, int W, int H, int N, long long pitch, long long fstride,
                                 const unsigned char* bad, int win, hipStream_t s);
