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

// Bytes per raw row of an MRC mode (0, 1, 2, 5, 6, 101), or -1.
long long mc_raw_row_bytes(int mode, int sx);
hipError_t mc_launch_apply_refs(const IngestArgs& a, int mode, hipStream_t s);
