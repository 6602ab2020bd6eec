// mc_ingest.hip -- raw camera frames -> gain/dark-corrected FP32 frames.
//
// Follows GApplyRefsToFrame (MrcUtil/GApplyRefsToFrame.cu:15-77,130-225)
// as driven by CApplyRefs (MrcUtil/CApplyRefs.cpp:101-212): a frame of the
// output size W x H is cut from the centre of the sx x sy stack frame
// (start = ((sx - W) / 2, (sy - H) / 2)), each pixel becomes
//   v = (float)raw;  v -= dark (if any);  v *= gain (if any)
// with the gain and dark references in stack coordinates, and is written to
// rows of `pitch` floats (the padding columns are left untouched, as in the
// reference).  MRC modes: 0 and 5 unsigned bytes, 1 int16, 6 uint16,
// 2 float, 101 4-bit packed (rows of (sx + 1) / 2 bytes, even x in the low
// nibble).
//
// One thread per output pixel loops over all frames, so the references are
// read once per call instead of once per frame; loads and stores of
// consecutive lanes are consecutive.
#include <hip/hip_runtime.h>

#include "mc_ingest.h"

namespace {

template <int MODE>
__device__ __forceinline__ float raw_at(const unsigned char* __restrict__ f, long long idx, int x,
                                        long long row_bytes, int y) {
    if constexpr (MODE == 0 || MODE == 5) return (float)f[idx];
    if constexpr (MODE == 1) return (float)reinterpret_cast<const short*>(f)[idx];
    if constexpr (MODE == 6) return (float)reinterpret_cast<const unsigned short*>(f)[idx];
    if constexpr (MODE == 2) return reinterpret_cast<const float*>(f)[idx];
    if constexpr (MODE == 101) {
        const unsigned int b = f[(long long)y * row_bytes + (x >> 1)];
        return (float)((x & 1) ? ((b >> 4) & 0xf) : (b & 0xf));
    }
    return 0.f;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_apply_refs(IngestArgs a) {
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.W) return;
    const int xs = a.x0 + x, ys = a.y0 + y;
    const long long ri = (long long)ys * a.sx + xs;  // stack coordinates
    const float dk = a.dark ? a.dark[ri] : 0.f;
    const float gn = a.gain ? a.gain[ri] : 1.f;
    const bool has_d = a.dark != nullptr, has_g = a.gain != nullptr;
    float* out = a.out + (long long)y * a.pitch + x;
    for (int n = 0; n < a.n_out; n++) {
        float v;
        if (a.tab) {
            // frame integration: unsigned-byte sum of the raw frames
            // (Util/GAddFrames.cu:32-41 wraps at 256)
            const int s0 = a.tab[2 * n], ns = a.tab[2 * n + 1];
            unsigned int acc = 0;
            for (int k = 0; k < ns; k++)
                acc += (unsigned int)raw_at<MODE>(a.raw + (long long)(s0 + k) * a.frame_bytes, ri, xs,
                                                  a.row_bytes, ys);
            v = (float)(acc & 0xffu);
        } else {
            v = raw_at<MODE>(a.raw + (long long)n * a.frame_bytes, ri, xs, a.row_bytes, ys);
        }
        if (has_d) v -= dk;
        if (has_g) v *= gn;
        out[(long long)n * a.frame_stride] = v;
    }
}

}  // namespace

long long mc_raw_row_bytes(int mode, int sx) {
    switch (mode) {
        case 0:
        case 5: return sx;
        case 1:
        case 6: return 2LL * sx;
        case 2: return 4LL * sx;
        case 101: return (sx + 1) / 2;
        default: return -1;
    }
}

hipError_t mc_launch_apply_refs(const IngestArgs& a, int mode, hipStream_t s) {
    if (a.W < 1 || a.H < 1 || a.n < 1 || a.n_out < 1 || a.x0 < 0 || a.y0 < 0 || a.x0 + a.W > a.sx ||
        a.y0 + a.H > a.sy || a.pitch < a.W || a.frame_stride < a.pitch * a.H)
        return hipErrorInvalidValue;
    const dim3 grid((a.W + 255) / 256, a.H), blk(256);
    switch (mode) {
        case 0: hipLaunchKernelGGL(k_apply_refs<0>, grid, blk, 0, s, a); break;
        case 5: hipLaunchKernelGGL(k_apply_refs<5>, grid, blk, 0, s, a); break;
        case 1: hipLaunchKernelGGL(k_apply_refs<1>, grid, blk, 0, s, a); break;
        case 6: hipLaunchKernelGGL(k_apply_refs<6>, grid, blk, 0, s, a); break;
        case 2: hipLaunchKernelGGL(k_apply_refs<2>, grid, blk, 0, s, a); break;
        case 101: hipLaunchKernelGGL(k_apply_refs<101>, grid, blk, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
