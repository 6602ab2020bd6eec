/*
 * mc_abi.h -- C ABI of libmc3, the MI355X frame-alignment path of MotionCor3.
 *
 * The reference has no plugin or FFI surface: the path is entered in-process
 * through Align::CAlignMain::DoIt(DU::CDataPackage*) (Align/CAlignInc.h:711-725,
 * Align/CAlignMain.cpp:39-75), called from CProcessThread::mAlignStack
 * (CProcessThread.cpp:178-182), with flags read from the CInput singleton
 * (CMainInc.h:14-144) and frames pre-loaded into the CBufferPool frm buffer
 * (CBufferPool.cpp:264-270).  mc_align_movie replaces that call; the stage
 * entry points below replace the reference's per-stage classes and are what a
 * caller (or a test) binds when it needs one stage on its own.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative MC_E* code on
 *     failure; mc_last_error() gives the message.  Nothing aborts
 *     (the reference prints and calls assert(0): Util/CSimpleFuncs.cpp:203-213).
 *   - "dev" pointers are device (HBM) addresses, "host" pointers are host
 *     memory.  All work is queued on the context's stream; functions that
 *     return host results synchronise that stream.
 *   - Half spectra are column-major: [kx][ky], kx < W/2+1, ky < H, complex
 *     interleaved float32.  The reference stores them row-major [ky][kx]
 *     (Util/CCufft2D.cpp:23-53); values are identical, only the order differs.
 *   - One context per host thread; contexts are independent (the reference's
 *     align classes keep file-static state and are not re-entrant,
 *     e.g. Align/CAlignedSum.cpp:13-15).
 */
#ifndef MC_ABI_H
#define MC_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

#define MC_OK 0
#define MC_E_ARG -1        /* invalid argument or unsupported size */
#define MC_E_HIP -2        /* HIP runtime error */
#define MC_E_NOMEM -3      /* device allocation failed */
#define MC_E_STATE -4      /* call sequence error */

typedef struct mc_ctx mc_ctx;

/* Flag subset of CInput that reaches the path (CInput.cpp:82-128 defaults,
 * :415-677 parsing).  mc_default_params fills the reference defaults. */
typedef struct mc_params {
    int    patch[2];      /* -Patch x y        (0 0: full-frame alignment only) */
    int    iter;          /* -Iter             (15) */
    float  tol;           /* -Tol              (0.1) */
    float  bft[2];        /* -Bft global local (500 100) */
    int    phase_only;    /* -PhaseOnly        (0) */
    float  fm_dose;       /* -FmDose           (0: no dose weighting) */
    float  pix_size;      /* -PixSize          (0) */
    int    kv;            /* -kV               (300) */
    int    group[2];      /* -Group            (1 4) */
    int    fm_ref;        /* -FmRef            (-1: middle frame) */
    float  ft_bin;        /* -FtBin            (1) */
    float  sum_range[2];  /* -SumRange         (3 25) */
    int    split_sum;     /* -SplitSum         (0) */
    int    throw0;        /* -Throw first      (0; only moves frame centres) */
    int    ref_quirks;    /* 1 (default): keep the reference's MakeRelative
                             indexing and spline-refit behaviour */
    const char* log_dir;  /* -LogDir, NULL/"" for no log files */
    const char* movie_name; /* input file name used to name the logs */
    int    verbose;       /* print the reference's progress lines */
    int    spectra_fmt;   /* storage of the xcf and patch spectrum stacks that
                             the alignment iterations read: 0 FP32 (default),
                             1 FP16, 2 BF16 (per-transform scale, FP32
                             arithmetic).  No reference counterpart: the
                             reference keeps cufftComplex stacks. */
    int    launch_graphs; /* 1 (default): run the alignment iterations as
                             captured HIP graphs, two iterations per graph
                             (same results as direct launches) */
    int    chunk_streams; /* 2 (default): frame-chunk FFT loops alternate over
                             two streams; 1: one stream (same results) */
    int    prior;         /* 1: start every patch's iterations from the learned
                             prior attached by mc_set_prior (0, default: from
                             zero shifts as the reference does).  No reference
                             counterpart (the reference resets the patch shifts,
                             Align/CIterativeAlign.cpp:68). */
    float  prior_fallback;/* full-resolution px (0.2): a patch whose largest
                             group shift in the guided second iteration exceeds
                             this restarts from zero shifts (plain solve) */
    const int* fm_int_start; /* -FmIntFile: raw start of each of the N frames
                                (NULL: i + throw0), from mc_fm_int_table */
    const int* fm_int_size;  /* raw frames integrated into each frame (NULL: 1);
                                frame centres, grouping by raw-frame count and
                                the accumulated dose FmDose x (start + size)
                                follow DataUtil/CFmIntParam.cpp:138-287 and
                                DataUtil/CFmGroupParam.cpp:26-97 */
    int    movie_pipeline;   /* mc_align_movies only.  1 (default): movies of one
                                size align one at a time while the next movie's
                                frames are transformed on a second stream;
                                0: they run their alignment iterations as one
                                batch of jobs.  Same results either way. */
    int    out_stack[2];     /* -OutStack (CInput.cpp:590-592): [0] 1 makes the
                                aligned stack (mc_result.stack), [1] frames summed
                                into each stack frame (z-bin, default 1) */
    int    align;            /* -Align (CInput.cpp:563-565): 1 (default) aligns;
                                0 makes simple sums without alignment
                                (Align/CSimpleSum.cpp:22-135): the sum and, with
                                split_sum, the _ODD / _EVN sums, no dose weighting */
    int    in_fm_motion;     /* -InFmMotion (0): weight each corrected frame by a
                                sinc along its motion to the next frame
                                (MotionDecon/GMotionWeight.cu:14-110); in patch
                                mode twice, as Correct/CGenRealStack.cpp:156-166
                                and the sums (Correct/CCorrectFullShift.cpp:311-320)
                                both apply it */
    int    corr_interp;      /* -CorrInterp (0), patch mode: divide the globally
                                corrected frames by sinc^(1/8)
                                (Util/GCorrLinearInterp.cu:7-24) */
    int    frame_fmt;        /* storage of the full-frame spectra (the frm
                                stack: forward transforms, local-correction
                                output, the spectra the sums read): 0 FP32
                                (default), 1 FP16, 2 BF16, each frame scaled
                                like the reduced stacks; arithmetic stays FP32.
                                No reference counterpart.  Kept at FP32 when the
                                frame stack is the xcf stack (xcf bin 1) or
                                -InFmMotion / -CorrInterp weigh it in place;
                                mc_result.frame_fmt_used reports it.  With a
                                reduced format and spectra_fmt reduced too, the
                                xcf stack takes the frame format. */
} mc_params;

/* N frames already gain/dark/defect corrected, resident in HBM
 * (the frm buffer of CBufferPool, filled by MrcUtil/CApplyRefs.cpp:186-212).
 * Rows of W floats with a pitch of `pitch` floats (the reference's padded
 * pitch (W/2+1)*2 or W), frames `frame_stride` floats apart.  Read only. */
typedef struct mc_movie {
    const float* frames;  /* dev */
    int W, H, N;
    long long pitch;
    long long frame_stride;
} mc_movie;

/* Outputs (DataUtil/CAlnSums.cpp:18-43 order: sum, _DW, _DWS, _ODD, _EVN).
 * sums[i] are device buffers of out_w*out_h floats (rows of out_w) supplied
 * by the caller for i < mc_num_sums(); all other fields may be NULL. */
typedef struct mc_result {
    float* sums[5];            /* dev */
    int    num_sums, out_w, out_h;
    float* global_shift;       /* host N x 2, as logged (-Full.log / -Patch-Full.log) */
    int*   global_iters;       /* host 1 */
    float* global_errors;      /* host iter, xcf pixels x xcf bin (as printed) */
    float* patch_shift;        /* host N x P x 2, [frame][patch][xy] */
    unsigned char* patch_bad;  /* host N x P */
    float* patch_center;       /* host P x 2, full-resolution pixels */
    int*   patch_starts;       /* host P x 2, xcf pixels */
    int*   patch_iters;        /* host P */
    double seconds_align;      /* wall time of measurement */
    double seconds_total;      /* wall time of the whole call */
    int    spectra_fmt_used[2];/* storage actually used for the xcf and patch
                                  stacks (a geometry without the register-pass
                                  kernels falls back to FP32) */
    float* prior_features;     /* host P x (68 N + 2): each patch's prior input
                                  after its first iteration (training data) */
    float* prior_targets;      /* host P x N x 2: converged frame shifts minus
                                  the first iteration's, in prior units */
    int    prior_fallbacks;    /* patches that fell back to the plain solve */
    float* stack;              /* dev, out_stack[0] only: mc_stack_frames() frames of
                                  out_w x out_h floats, each corrected frame before
                                  dose weighting at the -FtBin size
                                  (Correct/CCorrectFullShift.cpp:336-356), z-binned
                                  as Align/CZbinStack.cpp:27-70 */
    int    frame_fmt_used;     /* storage actually used for the full-frame spectra */
    double seconds_stage[5];   /* GPU time of each stage on the context stream,
                                  between events recorded at its ends: [0] forward
                                  transform, [1] global alignment, [2] patch alignment
                                  ("Patch alignment time", Align/CPatchAlign.cpp:35-39),
                                  [3] correction of the frames ("Global shifts are
                                  corrected" + "Correction of local motion",
                                  Align/CPatchAlign.cpp:59-63,
                                  Correct/GCorrectPatchShift.cu:133-134), [4] sums
                                  and output.  A stage that did not run reads 0; in a
                                  multi-movie call every movie gets the call's
                                  figures. */
} mc_result;

/* Replacements of measured quantities.  replay_* follow -InAln
 * (Align/CFullAlign.cpp:85-89, Align/CPatchAlign.cpp:74-78,
 * Align/CLoadAlign.cpp:108-249): measurement is skipped and the given
 * full-resolution shifts are applied.  patch_starts places patches at the
 * given xcf-pixel starts instead of the feature-driven centres. */
typedef struct mc_override {
    const float* replay_global;          /* host N x 2 */
    const float* replay_patch;           /* host N x P x 2 */
    const float* replay_centers;         /* host P x 2 */
    const unsigned char* replay_bad;     /* host N x P, NULL: none bad */
    const int* patch_starts;             /* host P x 2 */
} mc_override;

/* ---- context ---- */
int mc_ctx_create(int device, void* stream, mc_ctx** out);
int mc_ctx_destroy(mc_ctx* ctx);
const char* mc_last_error(mc_ctx* ctx);
int mc_sync(mc_ctx* ctx);
void mc_default_params(mc_params* p);
/* Number and size of output sums for a movie (CBufferPool.cpp:185-202,
 * Align/CAlignBase.cpp:41-47). */
int mc_output_geometry(const mc_params* p, int W, int H, int N, int* num_sums, int* out_w,
                       int* out_h);
/* Frames of the aligned stack (0 without -OutStack): N, or the number of
 * z-bin groups (Align/CZbinStack.cpp:36-52, DataUtil/CFmGroupParam.cpp:26-97). */
int mc_stack_frames(const mc_params* p, int N, int* frames);
/* xcf image size, patch image size and xcf bin (CBufferPool.cpp:220-262). */
int mc_geometry(int W, int H, int px, int py, int* sizes4, float* xcf_bin2);

/* ---- device memory for callers without their own HIP allocator ----
 * (the reference's CGpuBuffer / cudaMalloc role, CGpuBuffer.cpp:56-176).
 * Copies are ordered on the context stream and complete on return. */
int mc_malloc(mc_ctx* ctx, unsigned long long bytes, void** out);
int mc_free(mc_ctx* ctx, void* ptr);
/* dir: 0 host->device, 1 device->host, 2 device->device */
int mc_memcpy(mc_ctx* ctx, void* dst, const void* src, unsigned long long bytes, int dir);

/* ---- whole path: Align::CAlignMain::DoIt (Align/CAlignMain.cpp:39-75) ---- */
int mc_align_movie(mc_ctx* ctx, const mc_params* p, const mc_movie* movie,
                   const mc_override* ov, mc_result* res);

/* ---- several movies: no reference counterpart (the reference aligns one
 * movie per CProcessThread pass, CProcessThread.cpp:82-107, while CMain
 * loads the next one, CMain.cpp:193-210) ----
 * Movies of equal size and frame count are grouped.  With movie_pipeline
 * (default) a group's movies align one after another and each movie's
 * forward transform runs on a second stream while the previous movie aligns
 * (two stack sets).  Without it, a group runs its alignment iterations as one
 * batch of jobs (the global alignment of each movie and every movie's
 * patches share the iteration launches), holding the per-movie stacks of at
 * most half the device memory per batch.  Either way results are
 * bit-identical to separate mc_align_movie calls.
 * params[m] names movie m's logs; every other field must equal params[0]'s.
 * results[m] as for mc_align_movie.  No overrides. */
int mc_align_movies(mc_ctx* ctx, const mc_params* params, const mc_movie* movies, int count,
                    mc_result* results);

/* ---- stage entry points ---- */
/* Batched 2-D R2C (cufftExecR2C, Util/CCufft2D.cpp:65-87): `batch` real
 * images (rows of W, pitch, image stride) -> half spectra [kx][ky], each
 * scaled by `scale` (1/(W*H) is the reference's normalised forward). */
int mc_rfft2(mc_ctx* ctx, const float* in, int W, int H, long long pitch, long long in_stride,
             int batch, float scale, float* out_spec);
/* Batched 2-D C2R (cufftExecC2R, Util/CCufft2D.cpp:104-119), unnormalised;
 * clamp != 0 applies max(0, .) (Util/GPositivity2D.cu:7-15). */
int mc_irfft2(mc_ctx* ctx, const float* in_spec, int W, int H, int batch, float* out,
              long long pitch, long long out_stride, int clamp);
/* GPhaseShift2D (Util/GPhaseShift2D.cu:12-79) fused with the aligned sum of
 * CAlignedSum (Align/CAlignedSum.cpp:25-126): out = sum_i in_i *
 * e^{i(kx sx_i' + ky sy_i')}, shifts in pixels (host N x 2, applied as given,
 * no skip rule). */
int mc_phase_shift_accumulate(mc_ctx* ctx, const float* in_spec, int W, int H, int N,
                              const float* shifts, float* out_spec);
/* GCorrelateSum2D (Align/GCorrelateSum2D.cu:13-140) for one group: filtered
 * cross spectrum conj(G)(S-G), inverse transform restricted to the central
 * sea_x x sea_y window (host result, row-major). */
int mc_group_correlate_window(mc_ctx* ctx, const float* sum_spec, const float* grp_spec, int W,
                              int H, float bfactor, int phase_only, int sea_x, int sea_y,
                              float* window_host);
/* CPeak2D (Align/CPeak2D.cpp:39-109) on `count` device windows. */
int mc_peak_fit(mc_ctx* ctx, const float* windows, int nx, int ny, int count,
                float* shifts_host);
/* CRemoveSpikes1D (Util/CRemoveSpikes1D.cpp:31-148) in FP64 on the GPU for
 * `count` trajectories of n points (host in/out, [traj][2][n]). */
int mc_spline_despike_fp64(mc_ctx* ctx, double* traj_host, int n, int count, int quirks);
/* CPatchShifts::MakeRelative + DetectBads (Align/CPatchShifts.cpp:263-355)
 * in FP64 on the GPU.  shifts host N x P x 2 in/out, centers P x 2. */
int mc_patch_stats_fp64(mc_ctx* ctx, double* shifts_host, const double* centers_host, int N,
                        int P, int W, int H, int quirks, unsigned char* bad_host,
                        int* ref_host);
/* mGCorrect3D (Correct/GCorrectPatchShift.cu:17-66) on N real frames
 * (rows of W floats, dense) with host tables; in and out must not overlap
 * (the reference gathers from the upsampled copy into a separate tmp frame,
 * Correct/GCorrectPatchShift.cu:268-281). */
int mc_local_warp(mc_ctx* ctx, const float* in, float* out, int W, int H, int N, int P,
                  const float* centers_host, const float* shifts_host,
                  const unsigned char* bad_host);
/* As mc_local_warp, with no device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * patch tables (centres, shifts, bad flags) go to a file-scope device array
 * of 4 MiB per GPU (held until the process ends) instead of an allocation,
 * so N * P * 9 bytes plus about 1 KiB must fit in it (e.g. 25 patches x
 * 18000 frames); larger tables return MC_E_ARG.  Every call on one GPU
 * shares that array and leaves its tables there: calls that overlap on one
 * GPU (other streams or contexts) race.  Same kernel and results as
 * mc_local_warp. */
int mc_local_warp_noalloc(mc_ctx* ctx, const float* in, float* out, int W, int H, int N, int P,
                          const float* centers_host, const float* shifts_host,
                          const unsigned char* bad_host);
/* GWeightFrame weight normaliser sqrt(sum_i w_i^2) (Correct/GWeightFrame.cu:
 * 19-30, 79-133) as a [kx][ky] table on the device. */
int mc_dose_weight_sum(mc_ctx* ctx, int W, int H, int N, float pix, int kv, float fm_dose,
                       float* out_table);

/* GFourierResize2D (Util/GFourierResize2D.cu:7-39, 110-125): `batch` half
 * spectra [kx][ky] of in_w x in_h images cropped or padded to out_w x out_h,
 * rows of negative frequencies kept at the end.  Output elements with no
 * source are left as they are (zero the output first when padding);
 * sum != 0 adds into the output (bSum). */
int mc_fourier_crop(mc_ctx* ctx, const float* in_spec, int in_w, int in_h, float* out_spec,
                    int out_w, int out_h, int batch, int sum);
/* CDetectFeatures::DoIt + CPatchCenters::Calculate (Align/CDetectFeatures.cpp:
 * 58-396, Align/CPatchCenters.cpp:42-105): from the aligned xcf sum (device
 * half spectrum [kx][ky] of an xw x xh image) the 128 x 128 feature map
 * (host, row-major, may be NULL) and the px x py patch starts in xcf pixels
 * (host P x 2, patch-size rule of mc_geometry). */
int mc_feature_centres(mc_ctx* ctx, const float* xcf_sum_spec, int xw, int xh, int px, int py,
                       unsigned char* feature_map_host, int* starts_host);
/* CExtractPatch + the patch R2C of CMeasurePatches (Align/CExtractPatch.cpp:
 * 27-106, Align/CMeasurePatches.cpp:54-55): N real xcf frames (device, rows of
 * xw floats, frames xw * xh apart) -> P * N unnormalised half spectra [kx][ky]
 * of the pw x ph windows at starts_host (P x 2 xcf pixels); transform
 * q * N + i is patch q of frame i. */
int mc_extract_patches(mc_ctx* ctx, const float* xcf_frames, int xw, int xh, int N,
                       const int* starts_host, int P, int pw, int ph, float* out_spec);
/* CCorrectFullShift::mUnpadSums (Correct/CCorrectFullShift.cpp:146-185):
 * `count` sum spectra [kx][ky] of W x H frames (device, W/2+1 by H each)
 * Fourier-cropped to the -FtBin size (out_w x out_h of mc_output_geometry),
 * C2R and clamped at zero into out (device, count images of out_w x out_h). */
int mc_finalize_sums(mc_ctx* ctx, const float* sum_specs, int W, int H, int count, float ft_bin,
                     float* out);

/* ---- instrumentation (used by the benchmark) ---- */
/* Record HIP events around every launch of the named kernel class
 * ("fft_cols_fwd", "fft_rows_r2c", ...); NULL disables. */
int mc_set_kernel_timing(mc_ctx* ctx, const char* kernel);
int mc_get_kernel_timing(mc_ctx* ctx, double* total_ms, int* launches);

/* ---- synthetic input (benchmark / tests, not part of the path) ---- */
/* Render N noiseless frames of Gaussian blobs (x, y, sigma, amp per blob,
 * per-frame blob displacement dx, dy [frame][blob]) on a unit background,
 * scale, then Poisson-sample with a counter-based generator (seed). */
int mc_synth_render(mc_ctx* ctx, float* frames, int W, int H, int N, long long pitch,
                    long long frame_stride, int nblobs, const float* blob_host,
                    const float* disp_host, float scale, unsigned long long seed, int poisson);

/* ---- ingest: raw camera frames to the corrected frame buffer ----
 * The step before the boundary: CApplyRefs::DoIt (MrcUtil/CApplyRefs.cpp:34-
 * 212) with GApplyRefsToFrame (MrcUtil/GApplyRefsToFrame.cu:15-225) turns each
 * raw frame into padded FP32, cropped to the frame size from the centre of
 * the stack frame, v = (raw - dark) * gain.  Raw frames arrive in MRC modes
 * 0 and 5 (unsigned byte), 1 (int16), 6 (uint16), 2 (float) and 101 (4-bit
 * packed, rows of (sx + 1) / 2 bytes, even x in the low nibble). */
typedef struct mc_raw_stack {
    const void* data;     /* n raw frames, host or dev; host memory from
                             mc_host_alloc makes the copy asynchronous */
    int mode;             /* MRC mode */
    int sx, sy, n;        /* stack frame size and frame count */
    int n_int;            /* frame integration (-FmIntFile, modes 0, 5, 101):
                             0 = none, else output frame k is the sum of raw
                             frames int_start[k] .. + int_size[k] - 1 (host
                             arrays), added as unsigned bytes as
                             MrcUtil/CLoadCryoEMStack.cpp:184-259 does
                             (Util/GAddFrames.cu:32-41, wrapping at 256) */
    const int* int_start;
    const int* int_size;
} mc_raw_stack;
/* Pinned host memory (the reference's cudaMallocHost staging,
 * CBufferPool.cpp:159-165,204-218), e.g. to read frames from disk into. */
int mc_host_alloc(mc_ctx* ctx, unsigned long long bytes, void** out);
int mc_host_free(mc_ctx* ctx, void* ptr);
/* Gain and dark references (host, ref_w x ref_h floats, either may be NULL)
 * with -RotGain (0..3 x 90 degrees), -FlipGain (1 upside down, 2 left-right)
 * and -InvGain as CLoadRefs::PostProcess (CLoadRefs.cpp:204-228); quirks = 1
 * keeps its rule for the dark (rotated and flipped only when rot != 0 or
 * flip == 0). */
int mc_set_refs(mc_ctx* ctx, const float* gain, const float* dark, int ref_w, int ref_h, int rot,
                int flip, int inv, int quirks);
/* The references as post-processed (host outputs may be NULL). */
int mc_get_refs(mc_ctx* ctx, float* gain, float* dark, int* ref_w, int* ref_h);
/* Correct raw frames into `frames` (dev, W x H, rows of `pitch` floats,
 * `frame_stride` floats apart); W <= sx, H <= sy.  async = 0 queues the work
 * on the context stream; async = 1 queues it on a separate ingest stream and
 * returns at once, so that the copy of the next movie overlaps the current
 * alignment.  An alignment whose frames an asynchronous load writes waits
 * for that load; mc_load_wait makes the context stream wait for all loads.
 * References of the stack frame size apply as they are; references f >= 2
 * times smaller in both directions are enlarged by repeating every pixel
 * f x f (CLoadRefs::AugmentRefs, CLoadRefs.cpp:231-266); other sizes are an
 * error. */
int mc_load_frames(mc_ctx* ctx, const mc_raw_stack* raw, float* frames, int W, int H,
                   long long pitch, long long frame_stride, int async);
int mc_load_wait(mc_ctx* ctx);
/* MRC stack files (LibSrc/Mrcfile/Header/CLoadMainHeader.cpp:18-40,
 * LibSrc/Mrcfile/Image/CLoadImage.cpp:45-90): header size, frame count,
 * mode and data offset (1024 + extended header); frames [first, first+count)
 * read into dst (host).  Byte-swapped files are swapped on reading. */
int mc_mrc_info(const char* path, int* sx, int* sy, int* n, int* mode, long long* data_offset);
/* Frame-integration table of a movie of n_raw raw frames:
 * CReadFmIntFile::DoIt (DataUtil/CReadFmIntFile.cpp:110-173) reads path
 * (NULL or "": no file), CFmIntParam::mSetup / mCalcIntFms / mThrowIntFrames
 * (DataUtil/CFmIntParam.cpp:118-262) give the n_out frames' raw start, size
 * and accumulated dose (arrays of cap entries, or NULL to query n_out) and
 * the raw-frame dose; integrated = 1 when the file applies. */
int mc_fm_int_table(const char* path, int n_raw, float fm_dose, int throw0, int throw1, int cap,
                    int* n_out, int* start, int* size, float* acc_dose, float* raw_dose,
                    int* integrated);
int mc_mrc_read(const char* path, int first, int count, void* dst);
/* Movies in one MRC file (MrcUtil/CLoadStack.cpp:40-55,95-112): each has
 * `frames` frames, movie s being frames [s * frames, (s + 1) * frames).
 * frames = stack_z (-StackZ) when above 1, else the header's mz when above 1,
 * else nz.  stacks = nz / stack_z with -StackZ, else 1. */
int mc_mrc_stacks(const char* path, int stack_z, int* frames, int* stacks);

/* TIFF movies (TiffUtil/CLoadTiffHeader.cpp:18-147): size of the first
 * directory, directory count, MRC mode of the samples (8 bits -> 0, 16 bits ->
 * 1 signed or 6 unsigned, 32-bit float -> 2; -1 otherwise, e.g. EER), the
 * compression code and the pixel size XResolution / 2.54e8 (0 if absent). */
int mc_tiff_info(const char* path, int* sx, int* sy, int* n, int* mode, int* compression,
                 float* pix_size);
/* Frames first..first+count-1 decoded (uncompressed, LZW, Deflate, PackBits;
 * strips or tiles; predictor 1 or 2) into dst in MRC-mode layout, rows stored
 * bottom-up as TiffUtil/CLoadTiffImage.cpp:72-160 stores them.  The frames
 * then go through mc_load_frames like an MRC stack. */
int mc_tiff_read(const char* path, int first, int count, void* dst);

/* EER files (EerUtil/CLoadEerHeader.cpp:22-47): camera size, fraction count,
 * code width (compression 65001 -> 7, 65000 -> 8) and the total bytes of the
 * fractions' raw strips. */
int mc_eer_info(const char* path, int* cam_w, int* cam_h, int* n, int* bits, long long* bytes);
/* The fractions' raw strips back to back (EerUtil/CLoadEerFrames.cpp:72-95):
 * fraction k at data[offsets[k], offsets[k+1]); cap bytes of data, n + 1
 * offsets. */
int mc_eer_read(const char* path, unsigned char* data, long long cap, long long* offsets);

/* An EER movie for mc_load_eer. */
typedef struct mc_eer_stack {
    const unsigned char* data;  /* compressed fractions (host; pinned from
                                   mc_host_alloc makes an asynchronous load) */
    const long long* offsets;   /* host, n + 1 entries, offsets[0] = 0 */
    int n;                      /* EER fractions */
    int cam_w, cam_h;           /* camera size */
    int bits;                   /* 7 or 8 */
    int sampling;               /* -EerSampling: 1, 2 (2x) or 3 (4x) */
    int n_int;                  /* rendered frames (0: one per fraction) */
    const int* int_start;       /* fraction start of each rendered frame */
    const int* int_size;        /* fractions summed into each rendered frame */
} mc_eer_stack;
/* CRenderMrcStack (EerUtil/CRenderMrcStack.cpp:22-101) + CDecodeEerFrame
 * (EerUtil/CDecodeEerFrame.cpp:46-215) on the GPU: every rendered frame is
 * the unsigned-byte event count of its fractions at (cam x 1, 2 or 4) size,
 * then the gain/dark correction and centre crop of mc_load_frames (references
 * of the rendered size, or enlarged to it as in mc_load_frames: a camera-size
 * gain for 2x or 4x rendering).  async as for mc_load_frames. */
int mc_load_eer(mc_ctx* ctx, const mc_eer_stack* eer, float* frames, int W, int H, long long pitch,
                long long frame_stride, int async);

/* ---- bad pixels (before the boundary) ----
 * mc_detect_bad follows CDetectMain::DoIt (BadPixel/CDetectMain.cpp:54-245):
 * on the sum of the N corrected frames, hot pixels above
 * mean + max(threshold x sigma, 10) (GDetectHot.cu:25-55); defect patches
 * where the local correlation with a defect_w x defect_h border template
 * exceeds mean + threshold x sigma of the correlation map, marking the
 * template's block (GLocalCC.cu, GDetectPatch.cu, CTemplate.cpp); then the
 * "x y w h" rectangles of defect_file and the low bit of defect_map (host,
 * W x H bytes), both optional.  bad (dev, W x H bytes) receives the map;
 * n_bad and thresholds (hot, correlation) are optional host outputs.  The
 * reference uses threshold 6 and a 6 x 6 template.  mc_correct_bad
 * replaces every bad pixel of every frame by a good pixel of the same frame
 * found by GCorrectBad's pseudo-random window search (GCorrectBad.cu:13-64,
 * window max(win, 31)). */
int mc_detect_bad(mc_ctx* ctx, const float* frames, int W, int H, int N, long long pitch,
                  long long frame_stride, float threshold, int defect_w, int defect_h,
                  const char* defect_file, const unsigned char* defect_map, unsigned char* bad,
                  int* n_bad, float* thresholds);
int mc_correct_bad(mc_ctx* ctx, float* frames, int W, int H, int N, long long pitch,
                   long long frame_stride, const unsigned char* bad, int win);

/* ---- output files after the alignment ----
 * mc_save_sums writes every output sum of `result` as a float MRC image
 * named like sGenOutputName (CSaveSerialCryoEM.cpp:13-32): out_mrc without
 * ".mrc" / ".st", + "" / "_DW" / "_DWS" / "_ODD" / "_EVN"
 * (DataUtil/CAlnSums.cpp:18-44) + ".mrc", with the main header of
 * CSaveSerialCryoEM::mSaveImage (:208-240) and pixel size PixSize x FtBin.
 * The header min / max / mean are those of the sum (a GPU reduction); the
 * reference takes them from a device buffer that the sum is never copied
 * into.  mc_save_aln / mc_load_aln write and read the -OutAln / -InAln text
 * format (Align/CSaveAlign.cpp:84-151, Align/CLoadAlign.cpp:108-250): the
 * setting block, N global shifts and, in patch mode, every patch's centre
 * and local shifts ([frame][patch][xy], as mc_result), the input of
 * mc_override's replay; mc_load_aln with NULL arrays returns only the sizes
 * (n = stack frames minus throws).  mc_save_star writes the Relion star file
 * of CGenStarFile (CGenStarFile.cpp:39-120), global shifts relative to the
 * first frame. */
int mc_save_sums(mc_ctx* ctx, const mc_params* p, const mc_result* result, const char* out_mrc);
/* The aligned stack of `result` (`frames` frames) as out_mrc + "_Stk.mrc"
 * (CSaveSerialCryoEM.cpp:146-206): nz = frames, 32 floats of extended header
 * per frame holding the pixel size in float 11, header min / max = extremes
 * over the frames, mean = mean of the frame means. */
int mc_save_stack(mc_ctx* ctx, const mc_params* p, const mc_result* result, int frames,
                  const char* out_mrc);
int mc_save_aln(const char* path, const mc_params* p, int sx, int sy, int n_stack, int N,
                const float* global_shift, const float* patch_shift, const float* patch_center);
int mc_load_aln(const char* path, int* sx, int* sy, int* n, int* px, int* py, float* global_shift,
                float* patch_shift, float* patch_center);
int mc_save_star(const char* path, const mc_params* p, int sx, int sy, int n_stack,
                 const char* gain_name, const float* global_shift, int N);

/* ---- FFT tile autotuner (no reference counterpart: cuFFT plans its own
 * kernels, Util/CCufft2D.cpp:23-53) ----
 * The register-pass FFT kernels have several compiled tile variants per
 * length (rows or columns per workgroup, threads), all bit-identical in
 * results, and for some lengths a second radix order (the same transform up
 * to FP32 rounding).  mc_autotune times the tile variants of each length's
 * default radix order on the batch shapes an alignment of a W x H x N movie
 * with px x py patches uses and keeps the fastest in the context's plans;
 * mc_autotune_schedules with radix = 1 also times the other radix orders.
 * report (may be NULL) receives the timings as text.  mc_fft_variant queries
 * (set < 0) or forces (set >= 0) the variant of the W x H plan's rows
 * (cols = 0) or columns (cols = 1); mc_fft_variant_name gives variant v's
 * tile and radix order as text. */
int mc_autotune(mc_ctx* ctx, int W, int H, int N, int px, int py, float ft_bin, char* report,
                int report_len);
int mc_autotune_schedules(mc_ctx* ctx, int W, int H, int N, int px, int py, float ft_bin, int radix,
                          char* report, int report_len);
int mc_fft_variant(mc_ctx* ctx, int W, int H, int cols, int set, int* count, int* current);
int mc_fft_variant_name(mc_ctx* ctx, int W, int H, int cols, int v, char* name, int len);

/* ---- learned shift prior (no reference counterpart) ----
 * An MLP in -> H -> H -> H -> out (ReLU) over rows of
 *   in  = 68 N + 2 features of one patch after its first alignment iteration:
 *         global trajectory (2N, xcf px / unit), patch centre (2, in [-1, 1]),
 *         first-iteration frame shifts (2N, / unit), and per frame the 8x8
 *         2x2-pooled centre of its group's correlation window (64N),
 *   out = 2N: converged minus first-iteration frame shifts, / unit
 *         (unit = 2 xcf px).
 * Layer products are BF16 MFMA with FP32 accumulation, FP32 master weights.
 * Parameters in one FP32 vector: W0 (in x H row-major), b0, W1, b1, W2, b2,
 * W3 (H x out), b3.  x, t, y below are device row-major FP32 arrays. */
typedef struct mc_prior mc_prior;
int mc_prior_create(mc_ctx* ctx, int frames, int hidden, unsigned long long seed,
                    mc_prior** out);
int mc_prior_destroy(mc_ctx* ctx, mc_prior* prior);
int mc_prior_dims(const mc_prior* prior, int* in_dim, int* out_dim, long long* n_params);
/* copy the FP32 parameters out (get_host) and/or in (set_host), either NULL */
int mc_prior_params(mc_ctx* ctx, mc_prior* prior, float* get_host, const float* set_host);
int mc_prior_forward(mc_ctx* ctx, mc_prior* prior, const float* x, int n, float* y);
/* loss = mean((y - t)^2) over n rows; grad_host (n_params floats) or NULL */
int mc_prior_loss_grad(mc_ctx* ctx, mc_prior* prior, const float* x, const float* t, int n,
                       double* loss, float* grad_host);
/* minibatch Adam; epoch_loss_host (epochs floats) or NULL */
int mc_prior_train(mc_ctx* ctx, mc_prior* prior, const float* x, const float* t, int n,
                   int epochs, int batch, float lr, unsigned long long seed,
                   float* epoch_loss_host);
/* the prior used by mc_align_movie(s) when params.prior = 1 (NULL detaches);
 * the context does not own it */
int mc_set_prior(mc_ctx* ctx, mc_prior* prior);
/* c[i][j] = sum_k a[i*sai + k*sak] b[k*sbk + j*sbj] (BF16 bit patterns in,
 * FP32 out with row pitch ldc): the MFMA product the prior's layers use */
int mc_gemm_bf16(mc_ctx* ctx, const unsigned short* a, long long sai, long long sak,
                 const unsigned short* b, long long sbk, long long sbj, float* c, int ldc, int m,
                 int n, int k);

#ifdef __cplusplus
}
#endif
#endif
