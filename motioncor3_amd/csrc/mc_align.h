// mc_align.h -- launch wrappers for the alignment kernels (mc_align.hip).
#pragma once
#include "mc_fft.h"

// Geometry of one job of an iteration batch (device table, one entry per
// job).  The jobs of a batch share the column length H and the search
// window; they may differ in frame count, group count and row length, so
// movies of different sizes align in the same launches.  Frame-indexed
// arrays of the batch (running and final shifts, A and B phasor rows) are
// concatenated over jobs from frame `fo`, group-indexed ones (accumulated
// and per-iteration group shifts, windows) from group `go`; the group table
// of the job's movie (start, size, centre) starts at `gt`, its frame table
// (group of each frame, frame centre) at `ft`.
struct JobGeo {
    const cfloat* F;       // frame 0 of the job's spectra (stored format)
    const float* finv;     // the job's per-frame load factors (reduced formats)
    const cfloat* twW;     // e^{2 pi i k / W} for the job's row length W
    long long ao;          // first element of the job's A rows (N x cx)
    long long so;          // first element of the job's sum S (cx x H)
    long long t1o;         // first element of the job's correlation rows (G x seay x cx)
    long long sco;         // first element of the job's finalize scratch
    double tol;            // convergence tolerance (xcf px)
    int N, G, cx, W;
    int fo, go, gt, ft;
    int tile0;             // first tile of the job in the register-pass correlation
    int binz;
};

// Device-resident state of a batch of CIterativeAlign jobs (one job for the
// global alignment, one per patch).  Doubles, [job][...] (see JobGeo).
struct IterState {
    int J = 0, N = 0, G = 0, max_iter = 0;
    const JobGeo* jobs = nullptr;  // [J] device; null: J jobs of one geometry (N, G)
    const int* fjob = nullptr;     // [frames of the batch] job of each frame
    int NT = 0, cx_max = 0, G_max = 0, N_max = 0, W_max = 0;
    int tiles = 0;                 // register-pass correlation tiles of all jobs
    double* tmp = nullptr;     // [J][N][2] running frame shifts
    double* total = nullptr;   // [J][G][2] accumulated group shifts
    double* gshift = nullptr;  // [J][G][2] this iteration's group shifts
    double* final_ = nullptr;  // [J][N][2] after interpolation/despike/smooth
    double* errors = nullptr;  // [J][max_iter]
    double* scratch = nullptr; // [J][16*N + 128]
    int* iters = nullptr;      // [J]
    unsigned char* done = nullptr; // [J]
    int* gstart = nullptr;     // [G]
    int* gsize = nullptr;      // [G]
    int* fgroup = nullptr;     // [N] group of each frame
    double* gcent = nullptr;   // [G]
    double* fcent = nullptr;   // [N]
};

int mc_align_init_constants();

hipError_t mc_launch_phasors(const double* shifts, double mult, int skip, const unsigned char* done,
                             int per_job, int W, int cx, int H, int batch, cfloat* A, cfloat* B,
                             hipStream_t s, int h_norm = 0);  // h_norm: y frequency scale (0: H)
hipError_t mc_launch_fill_u8(unsigned char* p, long long n, unsigned char v, hipStream_t s);
hipError_t mc_launch_fourier_resize(const cfloat* in, int inx, int iny, long long in_stride, cfloat* out,
                                    int outx, int outy, long long out_stride, int batch, int sum,
                                    hipStream_t s);
hipError_t mc_launch_frame_weights(cfloat* S, long long fs, int N, int cx, int H, const double* shifts,
                                   int motion, int corr_interp, hipStream_t s);
// Running-shift phase factors of an iteration batch with a job table.
hipError_t mc_launch_phasors_jobs(const IterState& st, const double* shifts, int H, cfloat* A, cfloat* B,
                                  hipStream_t s);
hipError_t mc_launch_aligned_sum(const cfloat* F, const cfloat* A, const cfloat* B,
                                 const unsigned char* done, int J, int N, int cx, int H,
                                 cfloat* S, hipStream_t s, int f_fmt = MC_SPEC_F32,
                                 const float* f_inv = nullptr, const JobGeo* jobs = nullptr);
// Whether the correlation of column length H runs the register-pass kernel
// (the one that reads reduced spectrum formats), and its columns per
// workgroup (the tile width of JobGeo::tile0; 0 without one).
bool mc_correlate_has_rp(int H);
int mc_correlate_rp_cols(int H);
hipError_t mc_launch_correlate(const FftPlan1& colplan, const cfloat* ptw, int Cc, const cfloat* F,
                               const cfloat* S, const cfloat* A, const cfloat* B,
                               const IterState& st, int cx, int H, float bft, int phase_only,
                               int seay, cfloat* T1, hipStream_t s, int f_fmt = MC_SPEC_F32,
                               const float* f_inv = nullptr);
hipError_t mc_launch_row_dft(const cfloat* T1, const cfloat* twW, const unsigned char* done, int J,
                             int G, int cx, int W, int seax, int seay, float* win, hipStream_t s,
                             const JobGeo* jobs = nullptr);
hipError_t mc_launch_peaks(const float* win, const unsigned char* done, int J, int G, int nx,
                           int ny, double* gshift, hipStream_t s, const JobGeo* jobs = nullptr);
hipError_t mc_launch_update(const IterState& st, int it, double tol, hipStream_t s);
hipError_t mc_launch_finalize(const IterState& st, int binz, int quirks, hipStream_t s);
hipError_t mc_launch_feature_map(const float* img128, unsigned char* feat, hipStream_t s);
hipError_t mc_launch_patch_stats(double* shifts, const double* centers, int N, int P, int W, int H,
                                 int quirks, unsigned char* bad, int* ref, hipStream_t s);
hipError_t mc_launch_local_warp(const float* R, float* Rp, long long pitch, long long fstride,
                                int W, int H, int N, int P, const float* cen, const float* sh,
                                const unsigned char* bad, unsigned long long bad_any, hipStream_t s);
hipError_t mc_launch_dw_sums(const cfloat* F, long long fstride, int N, int cx, int H,
                             const cfloat* A, const cfloat* B, int dw, double pix, double kvf,
                             double dose, int dws, double r0, double r1, int split, cfloat* sums,
                             long long sstride, hipStream_t s, const int* cum = nullptr,
                             int f_fmt = MC_SPEC_F32, const float* f_inv = nullptr);
hipError_t mc_launch_dose_wsum(int cx, int H, int N, double pix, double kvf, double dose,
                               float* out, hipStream_t s);
hipError_t mc_launch_despike(double* traj, double* scratch, int n, int count, int quirks,
                             hipStream_t s);
hipError_t mc_launch_peaks_f(const float* win, int nx, int ny, int count, double* out,
                             hipStream_t s);
hipError_t mc_launch_synth(float* frames, int W, int H, int N, long long pitch, long long fstride,
                           const float* blobs, const float* disp, int nblobs, const int* tile_off,
                           const int* tile_idx, int tile, int tiles_x, float scale,
                           unsigned long long seed, int poisson, hipStream_t s);
