// mc_output.h -- output files after the alignment.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#define MC_STATS_BLOCKS 1024

hipError_t mc_launch_image_stats(const float* img, int w, int h, long long pitch, float* pmin,
                                 float* pmax, double* psum, hipStream_t s);
std::string mc_output_name(const char* out_mrc, const char* suffix, const char* ext);
int mc_write_mrc_float(const char* path, const float* img, int w, int h, float pix, float amin,
                       float amax, float amean);
void mc_mrc_header(unsigned char* hdr, int w, int h, int nz, float pix, float amin, float amax,
                   float amean, int nreal);
hipError_t mc_launch_simple_sums(const float* frames, int W, int H, int N, long long pitch, long long fstride,
                                 int split, float* out, hipStream_t s);
hipError_t mc_launch_stack_add(const float* src, float* dst, long long n, int first, hipStream_t s);
int mc_write_aln(const char* path, int sx, int sy, int nstk, int px, int py, int throw0,
                 int throw1, int N, int P, const float* gs, const float* ps, const float* pc);
int mc_read_aln(const char* path, int* sx, int* sy, int* nfr, int* px, int* py, float* gs, float* ps,
                float* pc);
int mc_write_star(const char* path, const char* movie, const char* gain, int sx, int sy, int sz,
                  float pix, float ft_bin, float fm_dose, int kv, int throw0, const float* gs, int N);
