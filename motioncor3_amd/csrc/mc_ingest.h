// mc_ingest.h -- raw camera frames -> gain/dark-corrected FP32 frames.
#pragma once
#include <hip/hip_runtime.h>

struct IngestArgs {
    const unsigned char* raw;  // n frames of frame_bytes, stack size sx x sy
    const int* tab;            // integration: (start, size) per output frame, or null
    int n_out;                 // output frames (n without integration)
    long long frame_bytes;     // bytes per raw frame
    long long row_bytes;       // bytes per raw row (4-bit packed rows)
    int sx, sy, n;
    int x0, y0;                // crop start in stack coordinates
    int W, H;                  // output frame size
    const float* gain;         // sx x sy or null
    const float* dark;         // sx x sy or null
    float* out;                // frames: rows of `pitch` floats, `frame_stride` apart
    long long pitch, frame_stride;
};

// Raw mode of rendered EER frames: 32-bit event counters, read modulo 256.
#define MC_MODE_EER_COUNTS 1032

#define MC_EER_PHASES 11  // possible offsets (bits) of a segment's first 7-bit code
// Segments of a fraction's codes: 2048 bits (7-bit codes) or 255 bytes
// (8-bit codes, 85 code pairs); at least one per fraction.
#define MC_EER_SEG_BITS 2048
#define MC_EER_SEG_BYTES8 255

struct EerArgs {
    const unsigned char* data;  // compressed EER fractions
    const long long* off;       // fraction k occupies data[off[k], off[k+1])
    const int* job;             // (fraction, rendered frame, first segment) per job
    int njob;
    const int* seg_job;         // job of every segment
    int nseg;
    int* tab;                   // per segment and start offset: (end offset, pixels advanced)
    int* start;                 // per segment: (start offset, start pixel)
    int cam_w, cam_h;           // camera size
    int bits;                   // 7 (compression 65001) or 8 (65000)
    int sampling;               // 1 camera size, 2 and 3: 2x and 4x super-resolution
    int fw, fh;                 // rendered frame size
    unsigned int* counts;       // rendered frames, fw x fh counters each
};

// Bytes per raw row of an MRC mode (0, 1, 2, 5, 6, 101), or -1.
long long mc_raw_row_bytes(int mode, int sx);
hipError_t mc_launch_apply_refs(const IngestArgs& a, int mode, hipStream_t s);
hipError_t mc_launch_eer_render(const EerArgs& a, hipStream_t s);
// w x h reference -> (w f) x (h f), every pixel repeated f x f
hipError_t mc_launch_augment_ref(const float* in, int w, int h, int f, float* out, hipStream_t s);
// n 32-bit zeros by a kernel (no memset node if the caller captures a graph)
hipError_t mc_launch_zero_u32(unsigned int* p, long long n, hipStream_t s);
