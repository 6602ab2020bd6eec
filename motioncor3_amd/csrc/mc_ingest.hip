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
// EER rendering (EerUtil/CDecodeEerFrame.cpp:46-215, CRenderMrcStack.cpp:36-101).
// A fraction is a chain of run codes: 7-bit streams (compression 65001) hold
// a 7-bit zero run r and, unless r == 127, 4 sub-pixel bits (the reference
// reads them two codes per 32-bit little-endian window, which parses the
// same); 8-bit streams (65000) hold two 8-bit runs and their sub-pixel
// nibbles in every three bytes.  Each electron adds 1 to its rendered
// frame's counter at the pixel reached by the runs.  Sampling 1 renders at
// camera size; 2 and 3 render 2x and 4x super-resolution from the sub-pixel
// bits (s ^ 0x0A), with the reference's 4096-wide camera rows
// (x = (pixel & 4095) << k, y = (pixel >> 12) << k).  The 8-bit
// super-resolution walk reads its second run's high bits and sub-pixel
// nibble from byte 0 of the fraction and takes the first code's sub-pixel
// bits as byte1 & 0x0A, as the reference does.
//
// The walk is split into segments of MC_EER_SEG_BITS bits (8-bit streams:
// 255 bytes, 85 code pairs), decoded in three launches:
//  1. k_eer_seg_table: for every segment and every possible offset of its
//     first code (0..10 bits for 7-bit codes, none for 8-bit), the offset at
//     which the walk leaves it and the pixels it advances;
//  2. k_eer_scan: one thread per fraction chains the offsets and prefix-sums
//     the advances, giving each segment its true start offset and pixel;
//  3. k_eer_emit: every segment walks again from there and adds its
//     electrons, stopping where the walk reaches the end of the camera.
// Each thread copies its segment to its own slice of LDS first.  Reads stop
// at the end of a fraction (bytes past it read as 0) and events outside the
// rendered frame are dropped (the reference reads and writes past both,
// which matters only for malformed streams).
#define MC_EER_SLICE 67  // words per thread slice (odd: no LDS bank conflicts)

struct EerSeg {
    long long nb;       // fraction bytes
    long long S;        // first logical bit of the segment
    long long E;        // end bit (exclusive) of the segment's code starts
    long long nbits;    // 8 * nb
    int off0;           // bit of S within slice word 0
    const unsigned int* sl;
    unsigned int b0;    // byte 0 of the fraction (8-bit super-resolution)
};

// Segment `k` of fraction `fr`: its words to this thread's LDS slice.
__device__ EerSeg eer_seg(const EerArgs& a, int fr, long long k, unsigned int* slice) {
    EerSeg g;
    const unsigned char* p = a.data + a.off[fr];
    g.nb = a.off[fr + 1] - a.off[fr];
    g.nbits = 8 * g.nb;
    g.S = a.bits == 7 ? k * MC_EER_SEG_BITS : k * (8LL * MC_EER_SEG_BYTES8);
    g.E = g.S + (a.bits == 7 ? MC_EER_SEG_BITS : 8LL * MC_EER_SEG_BYTES8);
    const uintptr_t ad = (uintptr_t)p;
    const int mis = (int)(ad & 3u);
    const unsigned int* w = reinterpret_cast<const unsigned int*>(ad - mis);
    const long long lim = g.nb + mis;  // physical end byte
    const long long nw = (lim + 3) / 4;
    const long long pb = g.S + 8LL * mis;  // physical bit of S
    const long long w0 = pb >> 5;
    g.off0 = (int)(pb & 31);
#pragma unroll 8
    for (int i = 0; i < MC_EER_SLICE; i++) {
        const long long wi = w0 + i;
        unsigned int v = 0u;
        if (wi < nw) {
            v = w[wi];
            const long long b = lim - 4 * wi;
            if (b < 4) v &= (1u << (8 * b)) - 1u;
        }
        slice[i] = v;
    }
    g.sl = slice;
    g.b0 = g.nb > 0 ? p[0] : 0u;
    return g;
}

// n (<= 24) bits at logical bit q (S <= q < E + 24)
__device__ __forceinline__ unsigned int eer_bits(const EerSeg& g, long long q, int n) {
    const int l = (int)(q - g.S) + g.off0;
    const unsigned long long v =
        ((unsigned long long)g.sl[(l >> 5) + 1] << 32 | g.sl[l >> 5]) >> (l & 31);
    return (unsigned int)v & ((1u << n) - 1u);
}

__device__ __forceinline__ void eer_hit(const EerArgs& a, unsigned int* cnt, unsigned int num,
                                        unsigned int sub) {
    const unsigned int fw = (unsigned int)a.fw, fh = (unsigned int)a.fh;
    unsigned int x, y;
    if (a.sampling == 1) {
        x = num % fw;
        y = num / fw;
    } else if (a.sampling == 2) {
        x = ((num & 4095u) << 1) | ((sub & 2u) >> 1);
        y = ((num >> 12) << 1) | ((sub & 8u) >> 3);
    } else {
        x = ((num & 4095u) << 2) | (sub & 3u);
        y = ((num >> 12) << 2) | ((sub & 12u) >> 2);
    }
    if (x < fw && y < fh) atomicAdd(cnt + (long long)y * fw + x, 1u);
}

__global__ __launch_bounds__(64) void k_eer_seg_table(EerArgs a) {
    __shared__ unsigned int slices[64 * MC_EER_SLICE];
    const int s = blockIdx.x * 64 + threadIdx.x;
    if (s >= a.nseg) return;
    const int j = a.seg_job[s];
    const int fr = a.job[3 * j];
    const EerSeg g = eer_seg(a, fr, s - a.job[3 * j + 2], slices + threadIdx.x * MC_EER_SLICE);
    if (a.bits == 7) {
        const long long end = min(g.E, g.nbits);
        for (int ph = 0; ph < MC_EER_PHASES; ph++) {
            long long q = g.S + ph;
            unsigned int dn = 0;
            while (q < end) {
                const unsigned int r = eer_bits(g, q, 7);
                q += 7;
                dn += r;
                if (r != 127u) {
                    q += 4;
                    dn += 1;
                }
            }
            a.tab[2 * ((long long)s * MC_EER_PHASES + ph)] = (int)(q - g.E);
            a.tab[2 * ((long long)s * MC_EER_PHASES + ph) + 1] = (int)dn;
        }
    } else {
        const long long kb = g.S >> 3, end = min(kb + MC_EER_SEG_BYTES8, g.nb);
        const bool sr = a.sampling != 1;
        unsigned int dn = 0;
        for (long long pos = kb; pos < end; pos += 3) {
            const unsigned int v = eer_bits(g, 8 * pos, 24);
            const unsigned int c0 = v & 0xffu, c1 = (v >> 8) & 0xffu, c2 = (v >> 16) & 0xffu;
            dn += c0 + (c0 < 255u ? 1u : 0u);
            const unsigned int r2 = (sr ? ((c1 >> 4) | g.b0) : ((c1 >> 4) | (c2 << 4))) & 0xffu;
            dn += r2 + (r2 < 255u ? 1u : 0u);
        }
        a.tab[2 * ((long long)s * MC_EER_PHASES)] = 0;
        a.tab[2 * ((long long)s * MC_EER_PHASES) + 1] = (int)dn;
    }
}

__global__ __launch_bounds__(64) void k_eer_scan(EerArgs a) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    if (j >= a.njob) return;
    const int s0 = a.job[3 * j + 2];
    const int s1 = j + 1 < a.njob ? a.job[3 * (j + 1) + 2] : a.nseg;
    const unsigned int cam = (unsigned int)a.cam_w * (unsigned int)a.cam_h;
    unsigned int num = 0;
    int ph = 0;
    for (int s = s0; s < s1; s++) {
        a.start[2 * s] = ph;
        a.start[2 * s + 1] = (int)num;
        if (num >= cam) continue;
        const long long t = 2 * ((long long)s * MC_EER_PHASES + ph);
        num += (unsigned int)a.tab[t + 1];
        if (num > cam) num = cam;
        ph = a.bits == 7 ? a.tab[t] : 0;
        if (ph < 0 || ph >= MC_EER_PHASES) ph = 0;  // past the stream's end
    }
}

__global__ __launch_bounds__(64) void k_eer_emit(EerArgs a) {
    __shared__ unsigned int slices[64 * MC_EER_SLICE];
    const int s = blockIdx.x * 64 + threadIdx.x;
    if (s >= a.nseg) return;
    const unsigned int cam = (unsigned int)a.cam_w * (unsigned int)a.cam_h;
    unsigned int num = (unsigned int)a.start[2 * s + 1];
    if (num >= cam) return;
    const int j = a.seg_job[s];
    const int fr = a.job[3 * j], o = a.job[3 * j + 1];
    const EerSeg g = eer_seg(a, fr, s - a.job[3 * j + 2], slices + threadIdx.x * MC_EER_SLICE);
    unsigned int* cnt = a.counts + (long long)o * a.fw * a.fh;
    if (a.bits == 7) {
        const long long end = min(g.E, g.nbits);
        long long q = g.S + a.start[2 * s];
        while (q < end) {
            const unsigned int r = eer_bits(g, q, 7);
            q += 7;
            num += r;
            if (num >= cam) return;
            if (r == 127u) continue;
            eer_hit(a, cnt, num, eer_bits(g, q, 4) ^ 0x0Au);
            q += 4;
            num += 1;
        }
    } else {
        const long long kb = g.S >> 3, end = min(kb + MC_EER_SEG_BYTES8, g.nb);
        const bool sr = a.sampling != 1;
        for (long long pos = kb; pos < end; pos += 3) {
            const unsigned int v = eer_bits(g, 8 * pos, 24);
            const unsigned int c0 = v & 0xffu, c1 = (v >> 8) & 0xffu, c2 = (v >> 16) & 0xffu;
            num += c0;
            if (num >= cam) return;
            if (c0 < 255u) {
                eer_hit(a, cnt, num, c1 & 0x0Au);
                num += 1;
            }
            unsigned int r2, s2;
            if (sr) {
                r2 = ((c1 >> 4) | g.b0) & 0xffu;
                s2 = (g.b0 >> 4) ^ 0x0Au;
            } else {
                r2 = ((c1 >> 4) | (c2 << 4)) & 0xffu;
                s2 = 0;
            }
            num += r2;
            if (num >= cam) return;
            if (r2 < 255u) {
                eer_hit(a, cnt, num, s2);
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
    if (a.njob < 1 || a.nseg < a.njob || (a.bits != 7 && a.bits != 8) || a.sampling < 1 ||
        a.sampling > 3 || a.fw < 1 || a.fh < 1)
        return hipErrorInvalidValue;
    const dim3 gs((a.nseg + 63) / 64), gj((a.njob + 63) / 64);
    hipLaunchKernelGGL(k_eer_seg_table, gs, dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_eer_scan, gj, dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_eer_emit, gs, dim3(64), 0, s, a);
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
