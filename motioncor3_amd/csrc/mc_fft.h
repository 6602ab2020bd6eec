// mc_fft.h -- batched 2-D real FFTs for gfx950 (host-side plan API).
//
// Layouts (all complex values are interleaved fp32):
//   real image   : rows of W floats, pitch `pitch` floats, H rows.
//   T (staging)  : column-blocked half spectrum [kx/C][y][kx%C], C = plan.C,
//                  so a column pass reads C*H contiguous values and a row pass
//                  writes R*C contiguous values per column block.
//   S (spectrum) : column-major half spectrum [kx][ky], kx < W/2+1, ky < H.
// Forward transforms use e^{-i}, inverse e^{+i}; both are unnormalised
// unless a scale is given (same conventions as cufftExecR2C/C2R).
#pragma once
#include <string>
#include <vector>
#include "mc_common.h"
#include "mc_spec.h"

// Tile shape of the two passes: C = the largest power of two <= MC_TILE_CMAX
// whose C columns fit in LDS, R = ceil(MC_TILE_RC / C) rows per row-pass
// workgroup (so a row pass writes R*C contiguous staging values).
#ifndef MC_TILE_CMAX
#define MC_TILE_CMAX 16
#endif
#ifndef MC_TILE_RC
#define MC_TILE_RC 16
#endif
// staging chunk width for heights with a register-pass column schedule
#ifndef MC_TILE_CT
#define MC_TILE_CT 4
#endif

struct Fft2D {
    int W = 0, H = 0, m = 0;   // m = W/2 (complex row length)
    int cx = 0;                // W/2 + 1
    int C = 0;                 // column block of the T layout
    int R = 0;                 // rows per row-pass workgroup
    FftPlan1 row{}, col{};
    cfloat* post_tw = nullptr; // m+1 entries e^{-2 pi i k / W}
    cfloat* row_tw = nullptr;  // owned device tables
    cfloat* col_tw = nullptr;
    cfloat* col_ptw = nullptr;  // register-pass twiddles of the column length (or null)
    cfloat* row_ptw = nullptr;  // register-pass twiddles of the row length m (or null)
    cfloat* col_ptw1 = nullptr; // the same for the alternative radix order (schedule 1)
    cfloat* row_ptw1 = nullptr;
    long long t_elems = 0;     // T elements per transform
    long long s_elems = 0;     // S elements per transform (cx * H)
    // tile variant of the register-pass row / column kernels (index among the
    // compiled variants for this length; 0 = the default, mc_autotune picks)
    int row_var = 0, col_var = 0;
    // the same for the inverse passes (C2R rows, inverse columns)
    int row_var_inv = 0, col_var_inv = 0;
};

// Per-transform phase factors e^{i(kx*ax + ky*ay)} applied on load (skip
// rule of Util/GPhaseShift2D.cu:68-79 already folded in: a transform whose
// shift is below 0.1 px gets unit phasors).
struct PhasorTabs {
    const cfloat* A;   // [batch][cx]
    const cfloat* B;   // [batch][H]
};

int mc_fft_init_constants();
int mc_fft2d_create(Fft2D* p, int W, int H);
void mc_fft2d_destroy(Fft2D* p);
int mc_fft_supported(int n);

// rows: real -> T (R2C along x), scale applied.  in + b*in_bstride is the
// first row of transform b.  Optional window offsets (x0,y0 per transform,
// int pairs) select a sub-window of a larger image.
hipError_t mc_rows_r2c(const Fft2D& p, const float* in, long long pitch, long long in_bstride,
                       const int* win_xy, cfloat* T, long long t_bstride, int batch, float scale,
                       hipStream_t s);
// cols: T -> S forward; optional xcf crop emission (Util/GFourierResize2D.cu
// mapping) into X with half-spectrum size (xcx, xcy), S may be null.
// S and X may be stored in a reduced format (mc_spec.h) with per-transform
// store factors s_g / x_g; strides are in elements of the format.
hipError_t mc_cols_fwd(const Fft2D& p, const cfloat* T, long long t_bstride, cfloat* S,
                       long long s_bstride, cfloat* X, int xcx, int xcy, long long x_bstride,
                       int batch, hipStream_t s, int s_fmt = MC_SPEC_F32,
                       const float* s_g = nullptr, int x_fmt = MC_SPEC_F32,
                       const float* x_g = nullptr);
// cols: S -> T inverse.  Optional phasors multiply the spectrum on load.  If
// src_cx/src_cy differ from the plan's (cx, H) the spectrum is read through
// the Fourier crop/pad mapping of Util/GFourierResize2D.cu:7-39.
hipError_t mc_cols_inv(const Fft2D& p, const cfloat* S, long long s_bstride, int src_cx,
                       int src_cy, const PhasorTabs* ph, cfloat* T, long long t_bstride, int batch,
                       hipStream_t s, int s_fmt = MC_SPEC_F32, const float* s_inv = nullptr);
// rows: T -> real (C2R along x), optional clamp at zero.
hipError_t mc_rows_c2r(const Fft2D& p, const cfloat* T, long long t_bstride, float* out,
                       long long pitch, long long out_bstride, int batch, int clamp,
                       hipStream_t s);

// As mc_rows_r2c; transform b reads image (b % in_mod) with window
// (b / in_mod) when in_mod > 0 (patch stacks: b = patch * N + frame).
hipError_t mc_rows_r2c_mod(const Fft2D& p, const float* in, long long pitch, long long in_bstride,
                           int in_mod, const int* win_xy, cfloat* T, long long t_bstride,
                           int batch, float scale, hipStream_t s, float* dc = nullptr);
// Per-transform factors of a reduced spectrum format from |row DC| values
// dc[batch][rows] (mc_spec.h).
hipError_t mc_launch_spec_scale(const float* dc, int rows, int batch, float* g, float* inv,
                                hipStream_t s);
// Whether the register-pass kernels (the ones that support reduced spectrum
// formats and row DC output) run the plan's rows / columns.
bool mc_fft_rp_rows(const Fft2D& p);
bool mc_fft_rp_cols(const Fft2D& p);
// Compiled tile variants of the plan's register-pass row (cols = 0) or
// column (cols = 1) kernels, with a printable name and the radix schedule
// of each (0 when the plan has no register-pass kernels on that side).
int mc_fft_variants(const Fft2D& p, int cols, std::vector<std::string>* names,
                    std::vector<int>* scheds = nullptr);
