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

#include <algorithm>

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
    // rendered EER counts: 32-bit counters read as the reference's unsigned
    // bytes (each event adds 1 to an unsigned char, EerUtil/CDecodeEerFrame.cpp)
    if constexpr (MODE == MC_MODE_EER_COUNTS)
        return (float)(reinterpret_cast<const unsigned int*>(f)[idx] & 0xffu);
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

// ---------------------------------------------------------------------------
// EER rendering (EerUtil/CDecodeEerFrame.cpp:46-215, CRenderMrcStack.cpp:36-101):
// one thread per (EER fraction, rendered frame) job walks the fraction's
// run-length codes in order and adds 1 to the rendered frame's counter of
// every electron.  7-bit streams (compression 65001): a 7-bit zero run p,
// then, unless p == 127, 4 sub-pixel bits; two codes are taken from one
// 32-bit little-endian window.  8-bit streams (65000): three bytes hold two
// 8-bit runs with their sub-pixel nibbles.  Sampling 1 renders at camera
// size; 2 and 3 render 2x and 4x super-resolution from the sub-pixel bits
// (s ^ 0x0A), with the reference's 4096-wide camera rows
// (x = (pixel & 4095) << k, y = (pixel >> 12) << k).  The 8-bit
// super-resolution walk reads its second code's high bits and sub-pixel
// nibble from byte 0 of the fraction and takes the first code's sub-pixel
// bits as byte1 & 0x0A, as the reference does.  Reads stop at the end of the
// fraction and events outside the rendered frame are dropped (the reference
// reads and writes past them).
__device__ __forceinline__ unsigned int eer_byte(const unsigned char* p, long long k, long long nb) {
    return k < nb ? p[k] : 0u;
}

__global__ __launch_bounds__(64) void k_eer_render(EerArgs a) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= a.njob) return;
    const int fr = a.job[2 * i], o = a.job[2 * i + 1];
    const unsigned char* p = a.data + a.off[fr];
    const long long nb = a.off[fr + 1] - a.off[fr];
    unsigned int* cnt = a.counts + (long long)o * a.fw * a.fh;
    const unsigned int cam = (unsigned int)a.cam_w * (unsigned int)a.cam_h;
    const unsigned int fw = (unsigned int)a.fw, fh = (unsigned int)a.fh;
    unsigned int and0 = 0, and1 = 0, sh0 = 0, sh1 = 0, sh2 = 0;
    if (a.sampling == 2) { and0 = 2; and1 = 8; sh0 = 1; sh1 = 1; sh2 = 3; }
    else if (a.sampling == 3) { and0 = 3; and1 = 12; sh0 = 2; sh1 = 0; sh2 = 2; }
    const bool sr = a.sampling != 1;
    auto hit = [&](unsigned int num, unsigned int sub) {
        unsigned int x, y;
        if (sr) {
            x = ((num & 4095u) << sh0) | ((sub & and0) >> sh1);
            y = ((num >> 12) << sh0) | ((sub & and1) >> sh2);
        } else {
            x = num % fw;
            y = num / fw;
        }
        if (x < fw && y < fh) atomicAdd(cnt + (long long)y * fw + x, 1u);
    };
    unsigned int num = 0;
    if (a.bits == 7) {
        unsigned long long bit = 0;
        while (true) {
            const long long b0 = (long long)(bit >> 3);
            if (b0 >= nb) break;
            const unsigned int chunk = (eer_byte(p, b0, nb) | (eer_byte(p, b0 + 1, nb) << 8) |
                                        (eer_byte(p, b0 + 2, nb) << 16) | (eer_byte(p, b0 + 3, nb) << 24)) >>
                                       (unsigned int)(bit & 7);
            unsigned int r = chunk & 127u;
            bit += 7;
            num += r;
            if (num >= cam) break;
            if (r == 127u) continue;
            bit += 4;
            hit(num, ((chunk >> 7) & 15u) ^ 0x0Au);
            num += 1;
            r = (chunk >> 11) & 127u;
            bit += 7;
            num += r;
            if (num >= cam) break;
            if (r == 127u) continue;
            bit += 4;
            hit(num, ((chunk >> 18) & 15u) ^ 0x0Au);
            num += 1;
        }
    } else {
        const unsigned int b0 = eer_byte(p, 0, nb);
        for (long long pos = 0; pos < nb; pos += 3) {
            const unsigned int c0 = eer_byte(p, pos, nb), c1 = eer_byte(p, pos + 1, nb);
            const unsigned int r1 = c0;
            num += r1;
            if (num >= cam) break;
            if (r1 < 255u) {
                hit(num, c1 & 0x0Au);
                num += 1;
            }
            unsigned int r2, s2;
            if (sr) {
                r2 = ((c1 >> 4) | b0) & 0xffu;
                s2 = (b0 >> 4) ^ 0x0Au;
            } else {
                r2 = ((c1 >> 4) | (eer_byte(p, pos + 2, nb) << 4)) & 0xffu;
                s2 = 0;
            }
            num += r2;
            if (num >= cam) break;
            if (r2 < 255u) {
                hit(num, s2);
                num += 1;
            }
        }
    }
}

// GAugmentRef (MrcUtil/GAugmentRef.cu): out[y][x] = in[y / f][x / f].
__global__ void k_augment_ref(const float* __restrict__ in, int w, int h, int f, float* __restrict__ out) {
    const long long ow = (long long)w * f;
    const long long n = ow * h * f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long y = i / ow, x = i - y * ow;
        out[i] = in[(y / f) * w + x / f];
    }
}

__global__ void k_fill_u32(unsigned int* p, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

}  // namespace

hipError_t mc_launch_augment_ref(const float* in, int w, int h, int f, float* out, hipStream_t s) {
    if (w < 1 || h < 1 || f < 1) return hipErrorInvalidValue;
    const long long n = (long long)w * h * f * f;
    const long long blocks = std::min<long long>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_augment_ref, dim3((unsigned int)blocks), dim3(256), 0, s, in, w, h, f, out);
    return hipGetLastError();
}

hipError_t mc_launch_zero_u32(unsigned int* p, long long n, hipStream_t s) {
    if (n < 1) return hipSuccess;
    const long long blocks = std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned int)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t mc_launch_eer_render(const EerArgs& a, hipStream_t s) {
    if (a.njob < 1 || (a.bits != 7 && a.bits != 8) || a.sampling < 1 || a.sampling > 3 || a.fw < 1 ||
        a.fh < 1)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_eer_render, dim3((a.njob + 63) / 64), dim3(64), 0, s, a);
    return hipGetLastError();
}

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
        case MC_MODE_EER_COUNTS:
            hipLaunchKernelGGL(k_apply_refs<MC_MODE_EER_COUNTS>, grid, blk, 0, s, a);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
