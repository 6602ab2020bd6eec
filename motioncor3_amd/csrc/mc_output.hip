// mc_output.hip -- outputs after the alignment: MRC sums and the aligned
// stack with header statistics, the z-binning of the stack, the -Align 0
// frame sums, the .aln alignment file (save and load) and the Relion .star
// file of the global shifts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mc_output.h"

namespace {

// Per-workgroup min, max and sum of an image with rows of `pitch` floats.
__global__ __launch_bounds__(256) void k_image_stats(const float* __restrict__ img, int w, int h,
                                                     long long pitch, float* pmin, float* pmax,
                                                     double* psum) {
    __shared__ float smin[256], smax[256];
    __shared__ double ssum[256];
    float lo = INFINITY, hi = -INFINITY;
    double acc = 0.0;
    const long long n = (long long)w * h;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const int y = (int)(i / w), x = (int)(i - (long long)y * w);
        const float v = img[(long long)y * pitch + x];
        lo = fminf(lo, v);
        hi = fmaxf(hi, v);
        acc += v;
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    ssum[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pmin[blockIdx.x] = smin[0];
        pmax[blockIdx.x] = smax[0];
        psum[blockIdx.x] = ssum[0];
    }
}

// One frame of the aligned stack into its z-bin: dst = src for the first
// frame of a group, dst += src for the others (Align/CZbinStack.cpp:59-70).
__global__ __launch_bounds__(256) void k_stack_add(const float* __restrict__ src, float* __restrict__ dst,
                                                   long long n, int first) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        dst[i] = first ? src[i] : dst[i] + src[i];
}

// -Align 0 sums (MrcUtil/CSumFFTStack.cpp:60-210): the frames added in frame
// order; with split, also the even-indexed frames, and all minus those
// (sDoSplitSum's GAddFrames(sum0, 1, sum1, -1)).
__global__ __launch_bounds__(256) void k_simple_sums(const float* __restrict__ frames, int W, int H, int N,
                                                     long long pitch, long long fstride, int split,
                                                     float* __restrict__ out) {
    const long long np = (long long)W * H;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < np; q += (long long)gridDim.x * 256) {
        const int y = (int)(q / W), x = (int)(q - (long long)y * W);
        const float* f = frames + (long long)y * pitch + x;
        float s0 = 0.f, s1 = 0.f;
        for (int i = 0; i < N; i++) {
            const float v = f[(long long)i * fstride];
            s0 += v;
            if (i % 2 == 0) s1 += v;
        }
        out[q] = s0;
        if (split) {
            out[np + q] = s1;
            out[2 * np + q] = s0 - s1;
        }
    }
}

}  // namespace

hipError_t mc_launch_simple_sums(const float* frames, int W, int H, int N, long long pitch, long long fstride,
                                 int split, float* out, hipStream_t s) {
    if (W < 1 || H < 1 || N < 1 || pitch < W) return hipErrorInvalidValue;
    const long long blocks = std::min<long long>(((long long)W * H + 255) / 256, 8192);
    hipLaunchKernelGGL(k_simple_sums, dim3((unsigned)blocks), dim3(256), 0, s, frames, W, H, N, pitch, fstride,
                       split, out);
    return hipGetLastError();
}

hipError_t mc_launch_stack_add(const float* src, float* dst, long long n, int first, hipStream_t s) {
    if (n < 1) return hipErrorInvalidValue;
    const long long blocks = std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_stack_add, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n, first);
    return hipGetLastError();
}

hipError_t mc_launch_image_stats(const float* img, int w, int h, long long pitch, float* pmin,
                                 float* pmax, double* psum, hipStream_t s) {
    if (w < 1 || h < 1 || pitch < w) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_image_stats, dim3(MC_STATS_BLOCKS), dim3(256), 0, s, img, w, h, pitch, pmin,
                       pmax, psum);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// sGenOutputName (CSaveSerialCryoEM.cpp:13-32): -OutMrc with ".mrc" and
// ".st" (case-insensitive) cut off, then the suffix and the extension.
std::string mc_output_name(const char* out_mrc, const char* suffix, const char* ext) {
    std::string s(out_mrc ? out_mrc : "");
    auto cut = [&](const char* pat) {
        std::string low = s;
        for (auto& c : low) c = (char)tolower(c);
        const size_t p = low.find(pat);
        if (p != std::string::npos) s = s.substr(0, p);
    };
    cut(".mrc");
    cut(".st");
    return s + (suffix ? suffix : "") + (ext ? ext : "");
}

// CSaveMainHeader defaults and SetImgSize / SetMinMaxMean
// (LibSrc/Mrcfile/Header/CSaveMainHeader.cpp:39-81,183-222): nz float
// images, cell lengths = size x pixel size; `nreal` floats of extended
// header per image (CSaveMrc::SetExtHeader, LibSrc/Mrcfile/App/CSaveMrc.cpp:132-151).
void mc_mrc_header(unsigned char* hdr, int w, int h, int nz, float pix, float amin, float amax,
                   float amean, int nreal) {
    std::memset(hdr, 0, 1024);
    auto put_i = [&](int off, int v) { std::memcpy(hdr + off, &v, 4); };
    auto put_f = [&](int off, float v) { std::memcpy(hdr + off, &v, 4); };
    auto put_s = [&](int off, short v) { std::memcpy(hdr + off, &v, 2); };
    put_i(0, w);
    put_i(4, h);
    put_i(8, nz);
    put_i(12, 2);  // mode: float
    put_i(28, w);
    put_i(32, h);
    put_i(36, nz);
    put_f(40, w * pix);
    put_f(44, h * pix);
    put_f(48, nz * pix);
    put_i(92, nreal * 4 * nz);  // nsymbt
    put_s(130, (short)nreal);   // numfloats
    put_f(52, 90.f);
    put_f(56, 90.f);
    put_f(60, 90.f);
    put_i(64, 1);
    put_i(68, 2);
    put_i(72, 3);
    put_f(76, amin);
    put_f(80, amax);
    put_f(84, amean);
    put_s(96, -16224);  // dvid
    std::memcpy(hdr + 104, "AGAR", 4);  // EXTTYPE and NVERSION 20140
    put_i(108, 2014 * 10);
    put_s(132, 1);  // sub
    put_s(160, 5);  // type
    put_f(184, 90.f);
    put_f(188, 90.f);
    put_f(192, 90.f);
    std::memcpy(hdr + 208, "MAP", 4);
    hdr[212] = 0x44;  // machine stamp: little-endian
    hdr[213] = 0x44;
}

int mc_write_mrc_float(const char* path, const float* img, int w, int h, float pix, float amin,
                       float amax, float amean) {
    unsigned char hdr[1024];
    mc_mrc_header(hdr, w, h, 1, pix, amin, amax, amean, 0);
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    bool ok = fwrite(hdr, 1, 1024, f) == 1024;
    ok = ok && fwrite(img, sizeof(float), (size_t)w * h, f) == (size_t)w * h;
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : -1;
}

// ---------------------------------------------------------------------------
// CSaveAlign (Align/CSaveAlign.cpp:84-151): the setting block, the global
// shifts and, in patch mode, every patch's centre and local shifts.
int mc_write_aln(const char* path, int sx, int sy, int nstk, int px, int py, int throw0,
                 int throw1, int N, int P, const float* gs, const float* ps, const float* pc) {
    FILE* f = fopen(path, "wt");
    if (!f) return -1;
    fprintf(f, "%s\n", "setting");
    fprintf(f, "   %s: %d %d %d %d\n", "stackSize", sx, sy, nstk, 1);
    fprintf(f, "   %s: %d %d %d\n", "patches", px, py, 0);
    fprintf(f, "   %s: %d %d\n", "throw", throw0, throw1);
    fprintf(f, "#    Global Shift\n#---------------------\n%s\n   %s %d\n", "globalShift", "stackID", 0);
    for (int i = 0; i < N; i++) fprintf(f, "%6d  %+8.3f  %+8.3f\n", i, gs[2 * i], gs[2 * i + 1]);
    if (P > 0) {
        fprintf(f, "#    Local Shift\n#---------------------\n%s\n", "localShift");
        for (int q = 0; q < P; q++) {
            fprintf(f, "   %s: %d\n", "patchID", q);
            fprintf(f, "   %s: %d\n", "Converge", 1);
            for (int i = 0; i < N; i++) {
                const int k = i * P + q;
                fprintf(f, "%6d  %8.2f  %8.2f  %+8.3f  %+8.3f\n", i, pc[2 * q], pc[2 * q + 1], ps[2 * k],
                        ps[2 * k + 1]);
            }
            if (q < P - 1) fprintf(f, "\n");
        }
    }
    return fclose(f) == 0 ? 0 : -1;
}

// CLoadAlign (Align/CLoadAlign.cpp:108-250): the setting block by keyword,
// N = stack frames - throws, then "%d %f %f" global lines and, per patch,
// "%d %f %f %f %f" lines.  With null arrays only the sizes are returned.
int mc_read_aln(const char* path, int* sx, int* sy, int* nfr, int* px, int* py, float* gs, float* ps,
                float* pc) {
    FILE* f = fopen(path, "rt");
    if (!f) return -1;
    char line[256];
    auto find = [&](const char* key) -> bool {
        while (fgets(line, sizeof(line), f))
            if (strstr(line, key)) return true;
        return false;
    };
    char tag[64];
    int stk[4] = {0, 0, 0, 0}, pat[3] = {0, 0, 0}, thr[2] = {0, 0};
    bool ok = find("setting") && find("stackSize") &&
              sscanf(line, "%63s %d %d %d %d", tag, stk, stk + 1, stk + 2, stk + 3) == 5 &&
              find("patches") && sscanf(line, "%63s %d %d %d", tag, pat, pat + 1, pat + 2) == 4 &&
              find("throw") && sscanf(line, "%63s %d %d", tag, thr, thr + 1) == 3;
    const int N = stk[2] - thr[0] - thr[1];
    const int P = pat[0] * pat[1];
    if (!ok || N < 1) {
        fclose(f);
        return -1;
    }
    if (sx) *sx = stk[0];
    if (sy) *sy = stk[1];
    if (nfr) *nfr = N;
    if (px) *px = pat[0];
    if (py) *py = pat[1];
    if (!gs) {
        fclose(f);
        return 0;
    }
    int cnt = 0, id;
    float a, b, c2, d;
    while (cnt < N && fgets(line, sizeof(line), f)) {
        if (sscanf(line, "%d %f %f", &id, &a, &b) != 3) continue;
        gs[2 * cnt] = a;
        gs[2 * cnt + 1] = b;
        cnt++;
    }
    ok = cnt == N;
    if (ok && P > 0 && ps && pc) {
        int q = 0, fm = 0;
        while (q < P && fgets(line, sizeof(line), f)) {
            if (line[0] == '#') continue;
            if (sscanf(line, "%d %f %f %f %f", &id, &a, &b, &c2, &d) != 5) continue;
            ps[2 * (fm * P + q)] = c2;
            ps[2 * (fm * P + q) + 1] = d;
            pc[2 * q] = a;
            pc[2 * q + 1] = b;
            if (++fm == N) {
                fm = 0;
                q++;
            }
        }
        ok = q == P;
    }
    fclose(f);
    return ok ? 0 : -1;
}

// CGenStarFile (CGenStarFile.cpp:39-120): general block, then the global
// shifts relative to the first frame, frames numbered from 1.
int mc_write_star(const char* path, const char* movie, const char* gain, int sx, int sy, int sz,
                  float pix, float ft_bin, float fm_dose, int kv, int throw0, const float* gs, int N) {
    FILE* f = fopen(path, "wt");
    if (!f) return -1;
    auto base = [](const char* p) -> const char* {
        if (!p) return nullptr;
        const char* s = strrchr(p, '/');
        return s ? s + 1 : p;
    };
    fprintf(f, "\n# version 30001\n\n");
    fprintf(f, "data_general\n\n");
    fprintf(f, "%40s %-16d\n", "_rlnImageSizeX", sx);
    fprintf(f, "%40s %-16d\n", "_rlnImageSizeY", sy);
    fprintf(f, "%40s %-16d\n", "_rlnImageSizeZ", sz);
    fprintf(f, "%40s %s\n", "_rlnMicrographMovieName", base(movie) ? base(movie) : "");
    if (!gain || !gain[0]) fprintf(f, "%40s\n", "_rlnMicrographGainName");
    else fprintf(f, "%40s %s\n", "_rlnMicrographGainName", base(gain));
    fprintf(f, "%40s %-16.6f\n", "_rlnMicrographOriginalPixelSize", pix);
    fprintf(f, "%40s %-16.6f\n", "rlnMicrographBinning", ft_bin);
    fprintf(f, "%40s %-16.6f\n", "_rlnMicrographDoseRate", fm_dose);
    fprintf(f, "%40s %-16.6f\n", "_rlnMicrographPreExposure", 0.0f);
    fprintf(f, "%40s %-16.6f\n", "_rlnVoltage", (float)kv);
    fprintf(f, "%40s %-16d\n", "_rlnMicrographStartFrame", throw0 + 1);
    fprintf(f, "%40s %-16d\n", "_rlnMotionModelVersion", 0);
    fprintf(f, "\n");
    fprintf(f, "# version 30001\n\n");
    fprintf(f, "data_global_shift\n\n");
    fprintf(f, "loop\n");
    fprintf(f, "_rlnMicrographFrameNumber #1\n");
    fprintf(f, "_rlnMicrographShiftX #2\n");
    fprintf(f, "_rlnMicrographShiftY #3\n");
    for (int i = 0; i < N; i++)
        fprintf(f, "%13d %13.6f %13.6f\n", i + 1, gs[2 * i] - gs[0], gs[2 * i + 1] - gs[1]);
    fprintf(f, "\n");
    return fclose(f) == 0 ? 0 : -1;
}
