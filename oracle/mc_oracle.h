/*
 * mc_oracle.h -- FP64 CPU restatement of MotionCor3's frame-alignment path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * HIP product path (libmc3.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in motioncor3_amd/ links
 * or calls it.
 *
 * Spectra use the reference's layout: half spectrum (W/2+1) x H, kx fastest
 * ("cmp" layout of Util/CCufft2D.cpp:23-53), as interleaved (re, im) doubles.
 * Real images are unpadded H x W, x fastest.
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Flag subset of CInput (CInput.cpp:82-128) that reaches the align path. */
typedef struct orc_params {
    int    patch[2];      /* -Patch x y          (default 0 0)   */
    int    iter;          /* -Iter               (15)            */
    double tol;           /* -Tol                (0.1)           */
    double bft[2];        /* -Bft global local   (500 100)       */
    int    phase_only;    /* -PhaseOnly          (0)             */
    double fm_dose;       /* -FmDose e/A^2/frame (0)             */
    double pix_size;      /* -PixSize A          (0)             */
    int    kv;            /* -kV                 (300)           */
    int    group[2];      /* -Group global local (1 4)           */
    int    fm_ref;        /* -FmRef              (-1 => N/2)     */
    double ft_bin;        /* -FtBin              (1)             */
    double sum_range[2];  /* -SumRange           (3 25)          */
    int    split_sum;     /* -SplitSum           (0)             */
    int    throw0;        /* -Throw first (only shifts frame centres) */
    int    ref_quirks;    /* 1: reproduce reference quirks (MakeRelative k+2 etc.) */
    const int* fm_int_start; /* -FmIntFile: raw start of each frame (NULL: i + throw0) */
    const int* fm_int_size;  /* raw frames in each frame (NULL: 1) */
    int    out_stack[2];  /* -OutStack: [0] 1 makes the aligned stack, [1] z-bin (1) */
    int    align;         /* -Align (1); 0: simple sums without alignment */
    int    in_fm_motion;  /* -InFmMotion (0) */
    int    corr_interp;   /* -CorrInterp (0) */
} orc_params;

/* Optional replacements for measured quantities.  Any NULL pointer means
 * "measure it".  xcf_shift replaces the global iterative alignment result
 * (xcf pixels, after smoothing); patch_starts replaces the feature-driven
 * patch placement; the replay_* fields follow -InAln semantics
 * (Align/CFullAlign.cpp:85-89, Align/CPatchAlign.cpp:74-78): measurement is
 * skipped and the given full-resolution shifts are applied. */
typedef struct orc_override {
    const double*        xcf_shift;      /* N x 2 */
    const int*           patch_starts;   /* P x 2, xcf pixels */
    const double*        replay_global;  /* N x 2, full-res px */
    const double*        replay_patch;   /* N x P x 2 relative local shifts */
    const double*        replay_centers; /* P x 2 full-res px */
    const unsigned char* replay_bad;     /* N x P (NULL => all good) */
} orc_override;

typedef struct orc_out {
    int     out_w, out_h, num_sums;  /* filled by the oracle */
    double* sums;            /* caller: >= 5 * W * H doubles, layout [sum][y][x] */
    double* global_shift;    /* N x 2 full-res px, as logged (-Full.log) */
    double* xcf_shift;       /* N x 2 xcf px (before Multiply), may be NULL */
    int*    global_iters;    /* 1 */
    double* global_errors;   /* >= iter, xcf px (unscaled) */
    double* patch_shift;     /* N x P x 2, after MakeRelative, [frame][patch][xy] */
    unsigned char* patch_bad;/* N x P */
    double* patch_center;    /* P x 2 full-res px */
    int*    patch_starts;    /* P x 2 xcf px */
    int*    patch_iters;     /* P */
    unsigned char* feature_map; /* 128 x 128, may be NULL */
    double* stack;           /* caller: >= N * W * H doubles when out_stack[0]; [frame][y][x] */
    int     stack_frames;    /* filled by the oracle */
} orc_out;

/* Geometry that CBufferPool derives from the stack size
 * (CBufferPool.cpp:185-262).  out[0..1] xcf image size, out[2..3] patch image
 * size, xcf_bin[2]. */
void orc_geometry(int W, int H, int px, int py, int* sizes, double* xcf_bin);

/* ---- FFT building blocks (double, reference conventions) ---- */
void orc_fft_c2c(double* data, int n, int howmany, int sign);
void orc_rfft2(const double* img, int W, int H, double* cmp);   /* unnormalised */
void orc_irfft2(const double* cmp, int W, int H, double* img);  /* unnormalised */

/* ---- path stages, exposed for unit checks ---- */
void orc_phase_shift(const double* cmp_in, int cmpx, int cmpy,
                     double sx, double sy, int accumulate, double* cmp_out);
void orc_fourier_resize(const double* cmp_in, int inx, int iny,
                        double* cmp_out, int outx, int outy, int accumulate);
void orc_correlate_window(const double* sum, const double* grp, int cmpx,
                          int cmpy, double bfactor, int phase_only,
                          int seax, int seay, double* window);
void orc_peak2d(const double* img, int nx, int ny, double* shift);
int  orc_lin_eqs(double* coeff, double* vals, int dim);
void orc_remove_spikes(double* sx, double* sy, int n, int single, int quirks);
void orc_smooth(double* sx, double* sy, int n, double weight);
/* Weight map [cy][cx] of a half spectrum: -CorrInterp's 1 / sinc^(1/8) when
 * corr_interp, times -InFmMotion's directional sinc for motion (mx, my). */
void orc_frame_weight(int cmpx, int cmpy, double mx, double my, int motion, int corr_interp,
                      double* w);
void orc_dose_weights(int cmpx, int cmpy, double pix, int kv,
                      const double* acc_dose, int nfrm, double* wsum);
void orc_make_relative(double* shifts, unsigned char* flags_region, int nfrm,
                       int npat, int quirks, int* ref_out);
void orc_detect_bads(const double* shifts, const double* centers, int nfrm,
                     int npat, int W, int H, unsigned char* bad);
void orc_local_warp(const double* img, int W, int H, const double* centers,
                    const double* shifts, const unsigned char* bad, int npat,
                    double* out);
void orc_feature_map(const double* xcf_sum_cmp, int xcfx, int xcfy,
                     unsigned char* features);
void orc_patch_centers(const unsigned char* features, int px, int py,
                       int xcfw, int xcfh, int* starts);

/* Whole path: frames are N x H x W float32 counts. Returns 0 on success. */
/* GNormByStd2D out of place (5 x 5 window, in-bounds neighbours only). */
void orc_norm_by_var(double* img, int w, int h);
/* CDetectFeatures thresholds on a normalised map of n values; thr gets
 * {mu - 4 sigma, mu - 0.08 sigma} of the clipped statistics (may be NULL). */
void orc_feature_threshold(const double* img, int n, unsigned char* feat, double* thr);
/* CInterpolateShift: group values (binz-frame groups of n frames with raw
 * sizes fmsize or NULL) to frames, nearest or linear with extrapolation. */
void orc_interpolate(const double* gx, const double* gy, int binz, int n, const int* fmsize,
                     int nearest, double* ox, double* oy, int* ngroups);

int orc_align_movie(const float* frames, int W, int H, int N,
                    const orc_params* prm, const orc_override* ov,
                    orc_out* out);

void orc_default_params(orc_params* prm);
int  orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
