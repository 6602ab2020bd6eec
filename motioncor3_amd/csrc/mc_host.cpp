// mc_host.cpp -- host side of libmc3: Align::CAlignMain for one MI355X.
//
// Mirrors the reference's control flow (Align/CAlignMain.cpp:39-75,
// Align/CFullAlign.cpp:70-171, Align/CPatchAlign.cpp:25-99,
// Correct/CCorrectFullShift.cpp, Correct/GCorrectPatchShift.cu) and its flag
// semantics (CInput, CAlignParam, CFmIntParam, CFmGroupParam), but every
// per-frame and per-patch loop of the reference is a batched launch here and
// all intermediate data stays in HBM.  The host only sees the shift tables.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mc_abi.h"
#include "mc_align.h"
#include "mc_fft.h"
#include "mc_badpix.h"
#include "mc_ingest.h"
#include "mc_output.h"
#include "mc_prior.h"
#include "mc_tiff.h"

// A learned shift prior on the context's device (mc_prior_create).
constexpr float kPriorUnit = 2.0f;  // xcf px per unit of the prior's shift inputs and outputs
struct mc_prior {
    Prior net;
    int frames = 0, hidden = 0;
    int device = 0;
};

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct TimedEv {
    hipEvent_t a, b;
};

}  // namespace

struct mc_ctx {
    int device = 0;
    // this device's address of the file-scope `scratch` array
    unsigned char* scratch = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    std::map<std::pair<int, int>, Fft2D*> plans;
    std::map<int, cfloat*> rowtw;
    std::map<std::string, DevBuf> bufs;
    std::string timed;
    std::vector<TimedEv> evs;
    size_t nev = 0;
    // captured alignment-iteration blocks, keyed by everything they bake in
    std::map<std::string, hipGraphExec_t> graphs;
    // second stream for the frame-chunk loops, and its fork/join events
    hipStream_t s2 = nullptr;
    hipEvent_t evf = nullptr, evj = nullptr;
    // end of a local-correction inverse queued early on the second stream,
    // its phase factors (and frame weights) being ready, and the pinned host
    // copy of its shifts
    hipEvent_t evi = nullptr, evw = nullptr;
    double* pin_shifts = nullptr;
    size_t pin_shifts_n = 0;
    // pinned staging of an iteration state's host tables (one upload per
    // alignment stage) and the event of its last upload
    unsigned char* pin_tab = nullptr;
    size_t pin_tab_n = 0;
    hipEvent_t evt = nullptr;
    // learned prior for guided patch alignment (not owned)
    mc_prior* prior = nullptr;
    // timing events of the stage clock (mc_result.seconds_stage), reused
    std::vector<hipEvent_t> stage_ev;
    // ingest: stream and completion event of asynchronous frame loads, and
    // the gain / dark references (device, ref_w x ref_h)
    hipStream_t s_in = nullptr;
    struct PendingLoad {
        const char* lo;
        const char* hi;  // output byte range of the load
        hipEvent_t ev;   // recorded on s_in after the load
    };
    std::vector<PendingLoad> loads;
    std::vector<hipEvent_t> free_ev;
    // host integration tables of queued loads, kept until their copy ran
    std::vector<std::pair<hipEvent_t, std::vector<int>>> tab_keep;
    float* gain = nullptr;
    float* dark = nullptr;
    int ref_w = 0, ref_h = 0;
    // references enlarged by an integer factor for larger frames
    // (CLoadRefs::AugmentRefs), built on first use of that factor
    float* gain_aug = nullptr;
    float* dark_aug = nullptr;
    int aug_fact = 0;
};

// Device memory of the allocation-free entry point (mc_local_warp_noalloc):
// its patch tables live in this array, one copy per GPU for the life of the
// process, instead of in an allocation.  Host code never passes `scratch`
// itself (on the host it names a stand-in, not device memory): each device's
// address comes from hipGetSymbolAddress, taken once per device when a
// context is created there.
#define MC_SCRATCH_BYTES (4u << 20)
#define MC_SCRATCH_DEVICES 64
static __device__ __align__(256) unsigned char scratch[MC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[MC_SCRATCH_DEVICES] = {};

// With `device` current: its address of `scratch`, looked up on first use.
static hipError_t scratch_for_device(int device, unsigned char** out) {
    if (device < 0 || device >= MC_SCRATCH_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// Chunks of the early local-correction inverse that the second stream takes
// (it also runs the next movie's forward transform); the correction stream
// transforms the rest once its patch alignment is done.  Frames per chunk of
// the full-frame FFT loops: larger launches leave less of each pass to its
// tail.  At 50 frames, alternating on one box: 10 frames per chunk with 3
// early chunks 38.72-38.75 movies/s, 8 with 4 (or 3) 38.1-38.6, 12 with 2
// 38.4-38.5, 16 with 2 38.4, against 37.3-37.8 for 4 with 7 (which had
// measured 37.3-37.5 against 36.0-37.0 with all 13 on the second stream).
#ifndef MC_EARLY_INV_CHUNKS
#define MC_EARLY_INV_CHUNKS 3
#endif
#ifndef MC_CHUNK_FRAMES
#define MC_CHUNK_FRAMES 10
#endif

namespace {

#define MC_TRY(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            c->err = std::string(#x) + ": " + hipGetErrorString(e_);               \
            return MC_E_HIP;                                                       \
        }                                                                          \
    } while (0)

#define MC_CHECK(x)              \
    do {                         \
        int r_ = (x);            \
        if (r_ != MC_OK) return r_; \
    } while (0)

int fail(mc_ctx* c, int code, const std::string& msg) {
    c->err = msg;
    return code;
}

// Buffer names of the calling host thread carry this prefix: the threads
// of one mc_align_movies call that align movies of different sizes side by
// side (IterMeet) keep their stacks apart.  Empty on every other thread.
thread_local std::string tl_buf_prefix;

void* buf(mc_ctx* c, const char* name, size_t bytes) {
    DevBuf& b = tl_buf_prefix.empty() ? c->bufs[name] : c->bufs[tl_buf_prefix + name];
    if (b.bytes >= bytes && b.p) return b.p;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (hipMalloc(&b.p, bytes < 256 ? 256 : bytes) != hipSuccess) {
        b.p = nullptr;
        return nullptr;
    }
    b.bytes = bytes;
    return b.p;
}

#define MC_BUF(var, type, name, count)                                                    \
    type* var = (type*)buf(c, name, sizeof(type) * (size_t)(count));                       \
    if (!var) return fail(c, MC_E_NOMEM, std::string("hipMalloc failed for ") + name)

Fft2D* plan(mc_ctx* c, int W, int H) {
    auto key = std::make_pair(W, H);
    auto it = c->plans.find(key);
    if (it != c->plans.end()) return it->second;
    Fft2D* p = new Fft2D;
    if (mc_fft2d_create(p, W, H) != 0) {
        delete p;
        return nullptr;
    }
    c->plans[key] = p;
    return p;
}

cfloat* row_twiddles(mc_ctx* c, int W) {
    auto it = c->rowtw.find(W);
    if (it != c->rowtw.end()) return it->second;
    std::vector<cfloat> h(W);
    for (int q = 0; q < W; q++) {
        const double a = 2.0 * M_PI * q / W;
        h[q].x = (float)std::cos(a);
        h[q].y = (float)std::sin(a);
    }
    cfloat* d = nullptr;
    if (hipMalloc(&d, W * sizeof(cfloat)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), W * sizeof(cfloat), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    c->rowtw[W] = d;
    return d;
}

// kernel timing hooks for the benchmark
struct KTimer {
    mc_ctx* c;
    bool on;
    size_t idx;
    hipStream_t st;
    KTimer(mc_ctx* ctx, const char* name, hipStream_t stream = nullptr)
        : c(ctx), on(false), idx(0), st(stream ? stream : ctx->stream) {
        if (c->timed.empty() || c->timed != name) return;
        if (c->nev >= c->evs.size()) {
            TimedEv e;
            (void)hipEventCreate(&e.a);
            (void)hipEventCreate(&e.b);
            c->evs.push_back(e);
        }
        idx = c->nev++;
        (void)hipEventRecord(c->evs[idx].a, st);
        on = true;
    }
    ~KTimer() {
        if (on) (void)hipEventRecord(c->evs[idx].b, st);
    }
};

// Host work the caller wants done once the first iterations of the next
// iterative alignment are queued (and the GPU is busy with them), instead
// of before: it then does not hold back the alignment's first launch.
thread_local std::function<int()>* tl_after_launch = nullptr;

// Stage clock of one call: events recorded on the context stream at the
// ends of each stage interval; after the call's final synchronisation the
// intervals are summed per stage (mc_result.seconds_stage).  The events are
// the context's, created on first use.
struct StageClock {
    mc_ctx* c;
    hipStream_t s;
    size_t used = 0;
    struct Span {
        int stage;
        size_t a, b;
    };
    std::vector<Span> spans;
    size_t open_at = 0;
    int open_stage = -1;
    StageClock(mc_ctx* ctx, hipStream_t st) : c(ctx), s(st) {}
    hipError_t mark(size_t* idx) {
        if (used >= c->stage_ev.size()) {
            hipEvent_t e;
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            c->stage_ev.push_back(e);
        }
        *idx = used++;
        return hipEventRecord(c->stage_ev[*idx], s);
    }
    // close the open interval (if any) and open one for `stage` (-1: none)
    hipError_t begin(int stage) {
        size_t i = 0;
        const hipError_t r = mark(&i);
        if (r != hipSuccess) return r;
        if (open_stage >= 0) spans.push_back({open_stage, open_at, i});
        open_stage = stage;
        open_at = i;
        return hipSuccess;
    }
    // after the stream has drained: seconds per stage
    void read(double* out5) const {
        for (int k = 0; k < 5; k++) out5[k] = 0.0;
        for (const Span& sp : spans) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, c->stage_ev[sp.a], c->stage_ev[sp.b]) == hipSuccess)
                out5[sp.stage] += ms * 1e-3;
        }
    }
};

// roctx range for the profiler's timeline (the reference's NVTX ranges,
// e.g. Align/CAlignMain.cpp:41,48, Align/CPatchAlign.cpp:49,69)
struct Range {
    explicit Range(const char* name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
    Range(const Range&) = delete;
    Range& operator=(const Range&) = delete;
};

// Patches flagged bad in any frame of an N x P host table (bit p), for the
// local warp's patch selection.
unsigned long long bad_any_of(const unsigned char* bad, int N, int P) {
    unsigned long long m = 0;
    for (long long e = 0; bad && e < (long long)N * P; e++)
        if (bad[e]) m |= 1ull << (e % P);
    return m;
}

// ---------------------------------------------------------------------------
// geometry: CBufferPool.cpp:185-262, Align/CAlignBase.cpp:41-47
struct Geo {
    int W, H, N, cx;
    int xcx, xh, xw;
    int pw, ph, pcx;
    float xbin[2];
    int ocx, oh, ow;
};

Geo make_geo(int W, int H, int N, int px, int py, float ftbin) {
    Geo g;
    g.W = W;
    g.H = H;
    g.N = N;
    g.cx = W / 2 + 1;
    float fb = std::fmin(W / 2048.0f, H / 2048.0f);
    if (fb < 1) fb = 1.0f;
    const int ixx = (int)(W / fb + 0.5f), ixy = (int)(H / fb + 0.5f);
    g.xcx = ixx / 2 + 1;
    g.xh = ixy / 2 * 2;
    g.xw = (g.xcx - 1) * 2;
    g.xbin[0] = (g.cx - 1.0f) / (g.xcx - 1.0f);
    g.xbin[1] = g.H * 1.0f / g.xh;
    g.pw = g.ph = g.pcx = 0;
    if (px > 1 && py > 1) {
        const int ipx = g.xw / px, ipy = g.xh / py;
        g.pcx = ipx / 2 + 1;
        g.pw = (g.pcx - 1) * 2;
        g.ph = ipy / 2 * 2;
    }
    g.ocx = g.cx;
    g.oh = g.H;
    if (ftbin != 1.0f) {
        int nx = (int)((g.cx - 1) * 2 / ftbin), ny = (int)(g.H / ftbin);
        nx = nx / 2 * 2;
        ny = ny / 2 * 2;
        g.ocx = nx / 2 + 1;
        g.oh = ny;
    }
    g.ow = (g.ocx - 1) * 2;
    return g;
}

// DataUtil/CFmGroupParam.cpp:26-78 with one raw frame per integrated frame;
// frame centres DataUtil/CFmIntParam.cpp:280-287.
struct Groups {
    int G = 0, binz = 1;
    std::vector<int> start, size, fgroup;
    std::vector<double> gcent, fcent;
};

// Frames are integrated frames: frame i starts at raw frame fst[i] and holds
// fsz[i] raw frames (null: fst[i] = i + throw0, fsz[i] = 1).  Frame centres
// (DataUtil/CFmIntParam.cpp:281-287), groups of at most binz x (smallest
// frame) raw frames (DataUtil/CFmGroupParam.cpp:26-97).
Groups make_groups(int binz, int N, int throw0, const int* fst = nullptr, const int* fsz = nullptr) {
    Groups g;
    g.binz = binz;
    g.fgroup.assign(N, 0);
    int minraw = 1 << 30;
    for (int i = 0; i < N; i++) {
        const int sz = fsz ? fsz[i] : 1;
        const int st = fst ? fst[i] : i + throw0;
        g.fcent.push_back((double)((float)st + 0.5f * ((float)sz - 1.0f)));
        minraw = std::min(minraw, sz);
    }
    const int maxraw = binz * minraw;
    int f = 0;
    for (int i = 0; i < N; i++) {
        g.start.push_back(f);
        int cnt = 0, raw = 0;
        while (true) {
            raw += fsz ? fsz[f] : 1;
            cnt += 1;
            g.fgroup[f] = g.G;
            f += 1;
            if (f >= N) break;
            if (raw >= maxraw) break;
        }
        g.size.push_back(cnt);
        g.G += 1;
        if (f >= N) break;
    }
    for (int k = 0; k < g.G; k++) {
        double s = 0;
        for (int j = 0; j < g.size[k]; j++) s += g.fcent[g.start[k] + j];
        g.gcent.push_back(s / g.size[k]);
    }
    return g;
}

// CDetectFeatures::mFindCenters/mFindCenter/mSetUsed/mCheckFeature
// (Align/CDetectFeatures.cpp:289-396) and CPatchCenters::Calculate
// (Align/CPatchCenters.cpp:42-105): patch starts from the feature map.
void patch_starts(const unsigned char* feat, int px, int py, int xw, int xh, int* starts) {
    const int B = 128;
    std::vector<unsigned char> used(B * B, 0);
    const float pat[2] = {B * 1.0f / px, B * 1.0f / py};
    int sea[4];
    sea[0] = std::max(4, (int)(pat[0] * 0.5f - 0.5f));
    sea[1] = B - sea[0];
    sea[2] = std::max(4, (int)(pat[1] * 0.5f - 0.5f));
    sea[3] = B - sea[2];
    auto range = [](int start, int size, const int* r) {
        if (start < r[0]) return r[0];
        if (start + size > r[1]) return r[1] - size;
        return start;
    };
    auto set_used = [&](float cxv, float cyv) {
        const int wx = (int)(pat[0] * 0.8f), wy = (int)(pat[1] * 0.8f);
        int sx = (int)(cxv - wx / 2), sy = (int)(cyv - wy / 2);
        sx = range(sx, wx, &sea[0]);
        sy = range(sy, wy, &sea[2]);
        for (int y = 0; y < wy; y++)
            for (int x = 0; x < wx; x++) used[(sy + y) * B + sx + x] = 1;
    };
    const int np = px * py;
    std::vector<float> cen(3 * np);
    for (int i = 0; i < np; i++) {
        const float fx = (i % px + 0.5f) * pat[0], fy = (i / px + 0.5f) * pat[1];
        const int ix = (int)fx, iy = (int)fy;
        bool has = !(ix < sea[0] || ix > sea[1] || iy < sea[2] || iy > sea[3]) && feat[iy * B + ix];
        cen[3 * i] = fx;
        cen[3 * i + 1] = fy;
        cen[3 * i + 2] = has ? 1.0f : -1.0f;
        if (has) set_used(fx, fy);
    }
    for (int i = 0; i < np; i++) {
        float* c = &cen[3 * i];
        if (c[2] > 0) continue;
        const int wx = B / 3, wy = B / 3;
        int sx = (int)(c[0] - wx * 0.5f + 0.5f), sy = (int)(c[1] - wy * 0.5f + 0.5f);
        sx = range(sx, wx, &sea[0]);
        sy = range(sy, wy, &sea[2]);
        float minr = 1e20f, ac[2] = {-1.0f, 0.0f};
        for (int y = 0; y < wy; y++)
            for (int x = 0; x < wx; x++) {
                const int j = (sy + y) * B + sx + x;
                if (used[j] || !feat[j]) continue;
                const float fx = x + sx + 0.5f, fy = y + sy + 0.5f;
                const float dx = fx - c[0], dy = fy - c[1];
                const float r = sqrtf(dx * dx + dy * dy);
                if (r >= minr) continue;
                minr = r;
                ac[0] = fx;
                ac[1] = fy;
            }
        if (ac[0] > 0) {
            c[0] = ac[0];
            c[1] = ac[1];
        }
        set_used(c[0], c[1]);
    }
    const int psx = xw / px / 2 * 2, psy = xh / py / 2 * 2;
    auto calc = [](float cv, int ps, int xs) {
        int s = (int)(cv - ps / 2);
        if (s < 0) s = 0;
        else if (s + ps > xs) s = xs - ps;
        return s;
    };
    for (int i = 0; i < np; i++) {
        starts[2 * i] = calc(cen[3 * i] * (xw * 1.0f / B), psx, xw);
        starts[2 * i + 1] = calc(cen[3 * i + 1] * (xh * 1.0f / B), psy, xh);
    }
}

// ---------------------------------------------------------------------------
// CIterativeAlign for J jobs at once (Align/CIterativeAlign.cpp:32-175).
struct IterOut {
    std::vector<double> shifts;  // [J][N][2]
    std::vector<double> errors;  // [J][max_iter]
    std::vector<int> iters;      // [J]
    std::vector<double> state;   // [J][N][2] running shifts at the end (with a PriorRun)
    std::vector<unsigned char> fell;  // [J] guided jobs that restarted plainly
};

// Learned-prior hooks of run_iterative.  With X set, the first iteration
// also writes every job's prior features to X.  With net set as well, the
// iterations are guided: after the first iteration the network's correction
// is added to the shifts, the second iteration measures the residual, and a
// job whose residual exceeds `fallback` restarts from zero shifts (the plain
// solve); all others stop after these two iterations.
struct PriorRun {
    mc_prior* net = nullptr;
    const float* side = nullptr;  // [J][ns] device
    int ns = 0;
    float* X = nullptr;           // [J][ns + 66 N] device
    float* Y = nullptr;           // [J][2N] device (guided)
    unsigned char* fell = nullptr;  // [J] device (guided)
    double fallback = 0.2;        // xcf px
    float unit = 2.0f;
};

// Element `elems` of a spectrum stack stored in format fmt (mc_spec.h).
static cfloat* spec_at(void* base, long long elems, int fmt) {
    return (cfloat*)((char*)base + elems * (fmt == MC_SPEC_F32 ? 8 : 4));
}

// One class of jobs of an iteration batch: J jobs of one movie geometry
// (N frames of row length W), their spectra at F [job][frame] in the
// batch's stored format, the movie's frame groups and its tolerance.
struct IterClass {
    const cfloat* F = nullptr;
    const float* f_inv = nullptr;  // [job][frame] load factors of a reduced format, or null
    int J = 0, N = 0, W = 0;
    const Groups* grp = nullptr;
    double tol = 0.0;               // xcf px (-Tol over the movie's xcf bin)
};

// CIterativeAlign for all jobs of all classes at once: every iteration stage
// is one launch over every job (the classes share the column length H; each
// job reads its own geometry from the job table).  A job's arithmetic does
// not depend on the other jobs, so results equal separate runs.  Outputs are
// concatenated over the classes in order.
int run_iterative_classes(mc_ctx* c, const char* tag, const IterClass* cls, int ncls, int H, Fft2D* pl,
                          float bft, int phase_only, int max_iter, int quirks, IterOut* out,
                          int f_fmt = MC_SPEC_F32, bool graphs = false, PriorRun* pr = nullptr) {
    if (ncls < 1) return fail(c, MC_E_ARG, "no iteration jobs");
    if (pr && ncls != 1) return fail(c, MC_E_ARG, "the learned prior needs movies of one geometry");
    const int seay = std::min(64, H);
    const int seax = std::min(64, cls[0].W);
    const int C = mc_correlate_rp_cols(H);
    const size_t fb = f_fmt == MC_SPEC_F32 ? 8 : 4;
    std::vector<JobGeo> jg;
    std::vector<int> fjob, gstart, gsize, fgroup;
    std::vector<double> gcent, fcent;
    long long ao = 0, so = 0, t1o = 0, sco = 0;
    int fo = 0, go = 0, tiles = 0;
    int cx_max = 0, G_max = 0, N_max = 0, W_max = 0;
    std::string sig;
    for (int k = 0; k < ncls; k++) {
        const IterClass& q = cls[k];
        const Groups& grp = *q.grp;
        const int cx = q.W / 2 + 1, G = grp.G, N = q.N;
        // linear group->frame interpolation needs two groups
        // (Align/CInterpolateShift.cpp:129-142 reads group 1)
        if (G < 2 && G != N) return fail(c, MC_E_ARG, "-Group leaves fewer than two frame groups");
        if (std::min(64, q.W) != seax) return fail(c, MC_E_ARG, "batched jobs differ in search window");
        cfloat* twW = row_twiddles(c, q.W);
        if (!twW) return fail(c, MC_E_NOMEM, "twiddle table");
        const int gt = (int)gstart.size(), ft = (int)fgroup.size();
        gstart.insert(gstart.end(), grp.start.begin(), grp.start.end());
        gsize.insert(gsize.end(), grp.size.begin(), grp.size.end());
        gcent.insert(gcent.end(), grp.gcent.begin(), grp.gcent.end());
        fgroup.insert(fgroup.end(), grp.fgroup.begin(), grp.fgroup.end());
        fcent.insert(fcent.end(), grp.fcent.begin(), grp.fcent.end());
        for (int j = 0; j < q.J; j++) {
            JobGeo e;
            e.F = (const cfloat*)((const char*)q.F + (size_t)j * N * cx * H * fb);
            e.finv = q.f_inv ? q.f_inv + (size_t)j * N : nullptr;
            e.twW = twW;
            e.ao = ao;
            e.so = so;
            e.t1o = t1o;
            e.sco = sco;
            e.tol = q.tol;
            e.N = N;
            e.G = G;
            e.cx = cx;
            e.W = q.W;
            e.fo = fo;
            e.go = go;
            e.gt = gt;
            e.ft = ft;
            e.tile0 = tiles;
            e.binz = grp.binz;
            jg.push_back(e);
            for (int i = 0; i < N; i++) fjob.push_back((int)jg.size() - 1);
            ao += (long long)N * cx;
            so += (long long)cx * H;
            t1o += (long long)G * seay * cx;
            sco += 16LL * N + 128;
            fo += N;
            go += G;
            if (C > 0) tiles += ((cx + C - 1) / C) * G;
        }
        cx_max = std::max(cx_max, cx);
        G_max = std::max(G_max, G);
        N_max = std::max(N_max, N);
        W_max = std::max(W_max, q.W);
        char b[160];
        snprintf(b, sizeof(b), "%d:%d:%d:%d:%.17g:%p:%p;", q.J, N, q.W, G, q.tol, (const void*)q.F,
                 (const void*)q.f_inv);
        sig += b;
    }
    const int J = (int)jg.size(), NT = fo, GT = go;
    std::string t(tag);
    auto nm = [&](const char* s) { return (t + s); };
    IterState st;
    st.J = J;
    st.N = cls[0].N;
    st.G = cls[0].grp->G;
    st.max_iter = max_iter;
    st.NT = NT;
    st.cx_max = cx_max;
    st.G_max = G_max;
    st.N_max = N_max;
    st.W_max = W_max;
    st.tiles = tiles;
    // The iteration state is one device block: the part that starts at zero
    // (tmp, total, errors, iters, done), the rest of the state, then the
    // host tables, each region 256-byte aligned.  The tables go up in one
    // copy from pinned memory and the zeroed part is cleared by one fill
    // kernel: five small memsets and seven pageable copies each held the
    // stream (and the host) for about 60 us.
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_tmp = take(sizeof(double) * NT * 2), o_total = take(sizeof(double) * GT * 2),
                 o_errors = take(sizeof(double) * J * max_iter), o_iters = take(sizeof(int) * J),
                 o_done = take(J);
    const size_t clear_end = off;
    const size_t o_gshift = take(sizeof(double) * GT * 2), o_final = take(sizeof(double) * NT * 2),
                 o_scratch = take(sizeof(double) * sco);
    const size_t tab0 = off;
    const size_t o_gstart = take(sizeof(int) * gstart.size()), o_gsize = take(sizeof(int) * gsize.size()),
                 o_fgroup = take(sizeof(int) * fgroup.size()), o_gcent = take(sizeof(double) * gcent.size()),
                 o_fcent = take(sizeof(double) * fcent.size()), o_jobs = take(sizeof(JobGeo) * J),
                 o_fjob = take(sizeof(int) * NT);
    const size_t tab_bytes = off - tab0;
    unsigned char* blob = (unsigned char*)buf(c, nm("state").c_str(), off);
    if (!blob) return fail(c, MC_E_NOMEM, "iteration state allocation failed");
    st.tmp = (double*)(blob + o_tmp);
    st.total = (double*)(blob + o_total);
    st.errors = (double*)(blob + o_errors);
    st.iters = (int*)(blob + o_iters);
    st.done = blob + o_done;
    st.gshift = (double*)(blob + o_gshift);
    st.final_ = (double*)(blob + o_final);
    st.scratch = (double*)(blob + o_scratch);
    st.gstart = (int*)(blob + o_gstart);
    st.gsize = (int*)(blob + o_gsize);
    st.fgroup = (int*)(blob + o_fgroup);
    st.gcent = (double*)(blob + o_gcent);
    st.fcent = (double*)(blob + o_fcent);
    JobGeo* djobs = (JobGeo*)(blob + o_jobs);
    int* dfjob = (int*)(blob + o_fjob);
    st.jobs = djobs;
    st.fjob = dfjob;
    MC_BUF(A, cfloat, nm("phA").c_str(), (size_t)ao);
    MC_BUF(B, cfloat, nm("phB").c_str(), (size_t)NT * H);
    MC_BUF(S, cfloat, nm("sum").c_str(), (size_t)so);
    MC_BUF(T1, cfloat, nm("t1").c_str(), (size_t)t1o);
    MC_BUF(win, float, nm("win").c_str(), (size_t)GT * seay * seax);
    hipStream_t s = c->stream;
    // the staging is rewritten once its previous upload has run
    MC_TRY(hipEventSynchronize(c->evt));
    if (c->pin_tab_n < tab_bytes) {
        if (c->pin_tab) MC_TRY(hipHostFree(c->pin_tab));
        c->pin_tab = nullptr;
        c->pin_tab_n = 0;
        MC_TRY(hipHostMalloc((void**)&c->pin_tab, tab_bytes));
        c->pin_tab_n = tab_bytes;
    }
    auto stage = [&](size_t o, const void* src, size_t n) {
        if (n) std::memcpy(c->pin_tab + (o - tab0), src, n);
    };
    stage(o_gstart, gstart.data(), sizeof(int) * gstart.size());
    stage(o_gsize, gsize.data(), sizeof(int) * gsize.size());
    stage(o_fgroup, fgroup.data(), sizeof(int) * fgroup.size());
    stage(o_gcent, gcent.data(), sizeof(double) * gcent.size());
    stage(o_fcent, fcent.data(), sizeof(double) * fcent.size());
    stage(o_jobs, jg.data(), sizeof(JobGeo) * J);
    stage(o_fjob, fjob.data(), sizeof(int) * NT);
    MC_TRY(hipMemcpyAsync(blob + tab0, c->pin_tab, tab_bytes, hipMemcpyHostToDevice, s));
    MC_TRY(hipEventRecord(c->evt, s));
    MC_TRY(mc_launch_fill_u8(blob, (long long)clear_end, 0, s));
    const int N = st.N, G = st.G;  // the prior hooks (one class)
    const double tol0 = cls[0].tol;
    bool feat = pr && pr->X;  // write the prior features after iteration 0
    // One iteration (Align/CIterativeAlign.cpp:117-140): B-factor -2 per
    // iteration with a floor of 20.
    auto bfac = [&](int it) {
        float b = bft;
        for (int i = 0; i < it; i++) {
            b -= 2;
            if (b < 20) b = 20.0f;
        }
        return b;
    };
    auto enqueue = [&](int it, bool timed) -> hipError_t {
        const float b = bfac(it);
        hipError_t e;
        {
            KTimer kt(c, timed ? "phasors" : "");
            if ((e = mc_launch_phasors_jobs(st, st.tmp, H, A, B, s))) return e;
        }
        {
            KTimer kt(c, timed ? "aligned_sum" : "");
            if ((e = mc_launch_aligned_sum(nullptr, A, B, st.done, J, 0, cx_max, H, S, s, f_fmt, nullptr,
                                           st.jobs)))
                return e;
        }
        {
            KTimer kt(c, timed ? "correlate" : "");
            if ((e = mc_launch_correlate(pl->col, pl->col_ptw, pl->C, nullptr, S, A, B, st, cx_max, H, b,
                                         phase_only, seay, T1, s, f_fmt, nullptr)))
                return e;
        }
        {
            KTimer kt(c, timed ? "row_dft" : "");
            if ((e = mc_launch_row_dft(T1, nullptr, st.done, J, G_max, cx_max, W_max, seax, seay, win, s,
                                       st.jobs)))
                return e;
        }
        if ((e = mc_launch_peaks(win, st.done, J, G_max, seax, seay, st.gshift, s, st.jobs))) return e;
        if ((e = mc_launch_update(st, it, tol0, s))) return e;
        if (it == 0 && feat)
            return prior_features(pr->side, pr->ns, st.tmp, pr->unit, win, J, N, G, st.fgroup, seax,
                                  seay, pr->X, pr->ns + 66 * N, s);
        return hipSuccess;
    };
    auto after_launch = [&]() -> int {
        std::function<int()>* h = tl_after_launch;
        tl_after_launch = nullptr;
        return (h && *h) ? (*h)() : MC_OK;
    };
    // Guided start: iterations 0 and 1 with the prior's correction between
    // them, then the residual gate; the block loop below then runs only for
    // the jobs that fell back (all others are done).
    std::vector<unsigned char> hdone(J);
    bool skip_loop = false;
    if (pr && pr->net) {
        if (max_iter < 2) return fail(c, MC_E_ARG, "the learned prior needs -Iter >= 2");
        Prior& net = pr->net->net;
        if (!net.reserve(J)) return fail(c, MC_E_NOMEM, "prior work buffers");
        MC_TRY(enqueue(0, true));
        MC_TRY(net.forward(pr->X, J, pr->Y, s));
        MC_TRY(prior_apply(pr->Y, J, N, G, st.gstart, st.gsize, st.fgroup, pr->unit, st.total,
                           st.tmp, s));
        MC_TRY(enqueue(1, true));
        MC_TRY(prior_gate(st.errors, max_iter, st.done, st.iters, st.tmp, st.total, J, N, G,
                          pr->fallback, pr->fell, s));
        MC_CHECK(after_launch());
        MC_TRY(hipMemcpyAsync(hdone.data(), st.done, J, hipMemcpyDeviceToHost, s));
        MC_TRY(hipStreamSynchronize(s));
        skip_loop = std::all_of(hdone.begin(), hdone.end(), [](unsigned char d) { return d != 0; });
        feat = false;  // the restarted jobs keep the features of the guided pass
    }
    // Blocks of two iterations; with `graphs` each block is captured once
    // into a HIP graph (no memset nodes) and replayed.  Timing one of the
    // block's kernels needs the direct launches.
    const bool use_graph = graphs && c->timed != "phasors" && c->timed != "aligned_sum" &&
                           c->timed != "correlate" && c->timed != "row_dft";
    for (int it0 = 0; it0 < max_iter && !skip_loop; it0 += 2) {
        const int it1 = std::min(it0 + 2, max_iter);
        if (use_graph) {
            char key0[1024];
            snprintf(key0, sizeof(key0),
                     "%s|%d|%d|%d %d %d %d %d %d %d|%.9g %d %d %d %d|%p %p %p %p %p %p %p %p %p %p "
                     "%p %p %p %p %p %p %p %p %p %p %p %d %p|%d %p %p %d|",
                     tag, it0, it1, J, NT, H, max_iter, cx_max, G_max, tiles, (double)bft, phase_only,
                     seax, seay, pl->C, (void*)A, (void*)B, (void*)S, (void*)T1, (void*)win,
                     (void*)st.tmp, (void*)st.total, (void*)st.gshift, (void*)st.final_,
                     (void*)st.errors, (void*)st.iters, (void*)st.done, (void*)st.gstart,
                     (void*)st.gsize, (void*)st.fgroup, (void*)st.gcent, (void*)st.fcent,
                     (void*)djobs, (void*)dfjob, (const void*)pl->col_ptw, (const void*)pl->col.tw,
                     f_fmt, (const void*)pl, feat ? 1 : 0, feat ? (const void*)pr->side : nullptr,
                     feat ? (void*)pr->X : nullptr, feat ? pr->ns : 0);
            const std::string key = std::string(key0) + sig;
            auto itg = c->graphs.find(key);
            hipGraphExec_t ex = nullptr;
            if (itg == c->graphs.end()) {
                MC_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                hipError_t e = hipSuccess;
                for (int it = it0; it < it1 && e == hipSuccess; it++) e = enqueue(it, false);
                hipGraph_t gr = nullptr;
                const hipError_t e2 = hipStreamEndCapture(s, &gr);
                if (e != hipSuccess || e2 != hipSuccess || !gr) {
                    if (gr) (void)hipGraphDestroy(gr);
                    return fail(c, MC_E_HIP, std::string("iteration graph capture: ") +
                                                 hipGetErrorString(e != hipSuccess ? e : e2));
                }
                const hipError_t e3 = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (e3 != hipSuccess) return fail(c, MC_E_HIP, "iteration graph instantiation");
                c->graphs[key] = ex;
            } else {
                ex = itg->second;
            }
            MC_TRY(hipGraphLaunch(ex, s));
        } else {
            for (int it = it0; it < it1; it++) MC_TRY(enqueue(it, true));
        }
        MC_CHECK(after_launch());
        // A converged job skips every kernel of the later iterations, so once
        // all jobs are done the remaining launches would change nothing: look
        // at the flags after every block of two iterations (2 is the
        // reference's minimum, Align/CIterativeAlign.cpp:144) and stop.
        if (it1 < max_iter) {
            MC_TRY(hipMemcpyAsync(hdone.data(), st.done, J, hipMemcpyDeviceToHost, s));
            MC_TRY(hipStreamSynchronize(s));
            bool all = true;
            for (int j = 0; j < J; j++) all = all && hdone[j];
            if (all) break;
        }
    }
    MC_TRY(mc_launch_finalize(st, cls[0].grp->binz, quirks, s));
    out->shifts.resize((size_t)NT * 2);
    out->errors.resize((size_t)J * max_iter);
    out->iters.resize(J);
    MC_TRY(hipMemcpyAsync(out->shifts.data(), st.final_, sizeof(double) * NT * 2,
                          hipMemcpyDeviceToHost, s));
    MC_TRY(hipMemcpyAsync(out->errors.data(), st.errors, sizeof(double) * J * max_iter,
                          hipMemcpyDeviceToHost, s));
    MC_TRY(hipMemcpyAsync(out->iters.data(), st.iters, sizeof(int) * J, hipMemcpyDeviceToHost, s));
    if (pr) {
        out->state.resize((size_t)J * N * 2);
        MC_TRY(hipMemcpyAsync(out->state.data(), st.tmp, sizeof(double) * J * N * 2,
                              hipMemcpyDeviceToHost, s));
        out->fell.assign(J, 0);
        if (pr->net)
            MC_TRY(hipMemcpyAsync(out->fell.data(), pr->fell, J, hipMemcpyDeviceToHost, s));
    }
    MC_TRY(hipStreamSynchronize(s));
    return MC_OK;
}

// J jobs of one geometry.
int run_iterative(mc_ctx* c, const char* tag, const cfloat* F, int J, int N, int W, int H,
                  Fft2D* pl, const Groups& grp, float bft, int phase_only, int max_iter,
                  double tol, int quirks, IterOut* out, int f_fmt = MC_SPEC_F32,
                  const float* f_inv = nullptr, bool graphs = false, PriorRun* pr = nullptr) {
    IterClass k;
    k.F = F;
    k.f_inv = f_inv;
    k.J = J;
    k.N = N;
    k.W = W;
    k.grp = &grp;
    k.tol = tol;
    return run_iterative_classes(c, tag, &k, 1, H, pl, bft, phase_only, max_iter, quirks, out, f_fmt,
                                 graphs, pr);
}

// ---------------------------------------------------------------------------
// Alignment iterations shared by movies of different sizes.  mc_align_movies
// runs one align_batch per movie geometry, each on its own host thread; the
// threads take turns (a thread runs only while it holds `mu`), so they never
// touch the context at the same time.  At each iteration stage a thread
// leaves its request and waits; when every live thread has asked, the last
// one runs all requests as one batch per column length (one launch per
// iteration stage over the jobs of every movie) and wakes the others.
struct IterMeet {
    struct Req {
        std::string tag;
        IterClass cls;
        int H = 0;
        Fft2D* pl = nullptr;
        float bft = 0.f;
        int phase_only = 0, max_iter = 0, quirks = 0, f_fmt = 0;
        bool graphs = false;
        IterOut* out = nullptr;
        int rc = MC_OK;
        std::string err;
        bool done = false;
    };
    std::mutex mu;
    std::condition_variable cv;
    int live = 0;
    std::vector<Req*> pending;

    // With mu held: every pending request, in batches of requests that can
    // share launches, then wake the waiting threads.
    void run_pending(mc_ctx* c) {
        std::vector<Req*> todo;
        todo.swap(pending);
        std::vector<char> used(todo.size(), 0);
        for (size_t a = 0; a < todo.size(); a++) {
            if (used[a]) continue;
            std::vector<Req*> part;
            for (size_t b = a; b < todo.size(); b++) {
                const Req& x = *todo[a];
                const Req& y = *todo[b];
                if (!used[b] && x.tag == y.tag && x.H == y.H && x.bft == y.bft &&
                    x.phase_only == y.phase_only && x.max_iter == y.max_iter && x.quirks == y.quirks &&
                    x.f_fmt == y.f_fmt && x.graphs == y.graphs && std::min(64, x.cls.W) == std::min(64, y.cls.W)) {
                    used[b] = 1;
                    part.push_back(todo[b]);
                }
            }
            std::vector<IterClass> cls;
            for (Req* r : part) cls.push_back(r->cls);
            const Req& r0 = *part[0];
            IterOut all;
            const int rc = run_iterative_classes(c, r0.tag.c_str(), cls.data(), (int)cls.size(), r0.H, r0.pl,
                                                 r0.bft, r0.phase_only, r0.max_iter, r0.quirks, &all,
                                                 r0.f_fmt, r0.graphs);
            size_t fo = 0, jo = 0;
            for (Req* r : part) {
                r->rc = rc;
                r->err = rc == MC_OK ? std::string() : c->err;
                if (rc == MC_OK) {
                    const size_t nf = (size_t)r->cls.J * r->cls.N, nj = (size_t)r->cls.J;
                    r->out->shifts.assign(all.shifts.begin() + 2 * fo, all.shifts.begin() + 2 * (fo + nf));
                    r->out->errors.assign(all.errors.begin() + jo * r0.max_iter,
                                          all.errors.begin() + (jo + nj) * r0.max_iter);
                    r->out->iters.assign(all.iters.begin() + jo, all.iters.begin() + jo + nj);
                    fo += nf;
                    jo += nj;
                }
                r->done = true;
            }
        }
        cv.notify_all();
    }
};

// The meeting of the calling thread and its lock (null: not in a meeting).
thread_local IterMeet* tl_meet = nullptr;
thread_local std::unique_lock<std::mutex>* tl_lock = nullptr;

// run_iterative for one class of jobs, shared with the other threads of a
// meeting when the calling thread is in one.
int run_iter_meet(mc_ctx* c, const char* tag, const cfloat* F, int J, int N, int W, int H, Fft2D* pl,
                  const Groups& grp, float bft, int phase_only, int max_iter, double tol, int quirks,
                  IterOut* out, int f_fmt, const float* f_inv, bool graphs) {
    if (!tl_meet)
        return run_iterative(c, tag, F, J, N, W, H, pl, grp, bft, phase_only, max_iter, tol, quirks, out,
                             f_fmt, f_inv, graphs);
    IterMeet& m = *tl_meet;
    IterMeet::Req r;
    r.tag = tag;
    r.cls.F = F;
    r.cls.f_inv = f_inv;
    r.cls.J = J;
    r.cls.N = N;
    r.cls.W = W;
    r.cls.grp = &grp;
    r.cls.tol = tol;
    r.H = H;
    r.pl = pl;
    r.bft = bft;
    r.phase_only = phase_only;
    r.max_iter = max_iter;
    r.quirks = quirks;
    r.f_fmt = f_fmt;
    r.graphs = graphs;
    r.out = out;
    m.pending.push_back(&r);
    if ((int)m.pending.size() == m.live) m.run_pending(c);
    else m.cv.wait(*tl_lock, [&] { return r.done; });
    if (r.rc != MC_OK) c->err = r.err;
    return r.rc;
}

// ---------------------------------------------------------------------------
// logs: Align/CAlignMain.cpp:77-107 (name), Align/CFullAlign.cpp:173-194,
// Align/CPatchShifts.cpp:155-249 (formats)
std::string log_stem(const mc_params* p) {
    if (!p->log_dir || !p->log_dir[0] || !p->movie_name) return "";
    const char* base = strrchr(p->movie_name, '/');
    base = base ? base + 1 : p->movie_name;
    std::string f = std::string(p->log_dir) + base;
    std::string low = f;
    for (auto& ch : low) ch = (char)tolower(ch);
    for (const char* ext : {".mrc", ".tif", ".eer"}) {
        const size_t pos = low.find(ext);
        if (pos != std::string::npos) return f.substr(0, pos);
    }
    return "";
}

void write_logs(const mc_params* p, int W, int H, int N, int P, const std::vector<float>& gs,
                const std::vector<float>& ps, const std::vector<unsigned char>& bad,
                const std::vector<float>& pc) {
    const std::string stem = log_stem(p);
    if (stem.empty()) return;
    if (P == 0) {
        FILE* f = fopen((stem + "-Full.log").c_str(), "wt");
        if (!f) return;
        fprintf(f, "# full frame alignment\n");
        for (int i = 0; i < N; i++) fprintf(f, "%4d  %8.2f  %8.2f\n", i + 1, gs[2 * i], gs[2 * i + 1]);
        fclose(f);
        return;
    }
    FILE* f = fopen((stem + "-Patch-Full.log").c_str(), "wt");
    if (f) {
        fprintf(f, "# Patch based alignment\n");
        fprintf(f, "# Initial alignment based upon full frames\n\n");
        for (int i = 0; i < N; i++) fprintf(f, "%4d  %8.2f  %8.2f\n", i + 1, gs[2 * i], gs[2 * i + 1]);
        fclose(f);
    }
    for (int mode = 0; mode < 2; mode++) {
        f = fopen((stem + (mode == 0 ? "-Patch-Patch.log" : "-Patch-Frame.log")).c_str(), "wt");
        if (!f) continue;
        fprintf(f, "# Patch based alignment\n");
        fprintf(f, "# Number of patches: %d\n", P);
        fprintf(f, mode == 0 ? "# Shifts are listed per patch.\n" : "# Shifts are listed per frame.\n");
        fprintf(f, "# Movie size: %d  %d  %d\n\n", W, H, N);
        if (mode == 0) {
            for (int i = 0; i < P; i++) {
                fprintf(f, "#Patch %03d raw and fit shifts\n", i + 1);
                for (int j = 0; j < N; j++) {
                    const int k = j * P + i;
                    fprintf(f, "%4d %8.2f %8.2f ", j + 1, pc[2 * i], pc[2 * i + 1]);
                    fprintf(f, "%8.2f %8.2f %4d\n", ps[2 * k], ps[2 * k + 1], (int)bad[k]);
                }
                fprintf(f, "\n");
            }
        } else {
            for (int j = 0; j < N; j++) {
                for (int i = 0; i < P; i++) {
                    const int k = j * P + i;
                    fprintf(f, "%4d %8.2f %8.2f ", j + 1, pc[2 * i], pc[2 * i + 1]);
                    fprintf(f, "%8.2f %8.2f %4d\n", ps[2 * k], ps[2 * k + 1], (int)bad[k]);
                }
                fprintf(f, "\n");
            }
        }
        fclose(f);
    }
}

double kv_factor(int kv) {
    if (kv >= 300) return 1.0;
    if (kv >= 200) return 0.002 * (kv - 200) + 0.8;
    if (kv >= 120) return 0.004375 * (kv - 120) + 0.45;
    return 0.45;
}

// ---------------------------------------------------------------------------
// Make the context stream wait for the asynchronous loads (mc_load_frames)
// whose output overlaps [lo, hi), or for all of them when lo is null.
int wait_loads(mc_ctx* c, const char* lo, const char* hi, hipStream_t st = nullptr) {
    std::vector<mc_ctx::PendingLoad> keep;
    for (auto& l : c->loads) {
        if (!lo || (l.lo < hi && lo < l.hi)) {
            MC_TRY(hipStreamWaitEvent(st ? st : c->stream, l.ev, 0));
            c->free_ev.push_back(l.ev);
        } else {
            keep.push_back(l);
        }
    }
    c->loads.swap(keep);
    // events of finished work can go
    std::vector<hipEvent_t> still;
    for (hipEvent_t e : c->free_ev) {
        if (hipEventQuery(e) == hipSuccess) (void)hipEventDestroy(e);
        else still.push_back(e);
    }
    (void)hipGetLastError();
    c->free_ev.swap(still);
    return MC_OK;
}

// Make `st` wait for the loads that write any of movie `mv`'s frames.
int wait_movie_loads(mc_ctx* c, const mc_movie* mv, hipStream_t st) {
    const char* lo = (const char*)mv->frames;
    const size_t bytes = sizeof(float) * (size_t)(mv->frame_stride * (mv->N - 1) +
                                                  mv->pitch * (mv->H - 1) + mv->W);
    return wait_loads(c, lo, lo + bytes, st);
}

// ---------------------------------------------------------------------------
// Align/CTransformStack.cpp + Align/CGenXcfStack.cpp: normalised R2C of every
// frame of `count` movies and the Fourier crop to the xcf size (Sx == null:
// none) in the same pass, CH frames per chunk.  Chunk k runs on lanes[k %
// nl] with that lane's staging buffers; the caller orders the lanes.
struct FwdLane {
    hipStream_t s;
    cfloat* T;   // CH * t_elems
    float* dcf;  // CH * H, reduced stack formats only
};
int enqueue_forward(mc_ctx* c, Fft2D* pf, const mc_movie* const* mvs, int count, const Geo& g,
                    int fx, cfloat* Sf_all, void* Sx_all, float* gx, float* ix, const FwdLane* lanes,
                    int nl, int CH, int ff = MC_SPEC_F32, float* gf = nullptr, float* iff = nullptr) {
    const int W = mvs[0]->W, H = mvs[0]->H, N = mvs[0]->N;
    const long long fs = (long long)g.cx * H, xfs = (long long)g.xcx * g.xh;
    const float scale = (float)(1.0 / ((double)W * H));
    int chunk = 0;
    for (int m = 0; m < count; m++) {
        const mc_movie* mv = mvs[m];
        for (int f0 = 0; f0 < N; f0 += CH, chunk++) {
            const int nb = std::min(CH, N - f0);
            const long long fi = (long long)m * N + f0;
            const FwdLane& L = lanes[chunk % nl];
            {
                KTimer kt(c, "fft_rows_r2c", L.s);
                MC_TRY(mc_rows_r2c_mod(*pf, mv->frames + (long long)f0 * mv->frame_stride, mv->pitch,
                                       mv->frame_stride, 0, nullptr, L.T, pf->t_elems, nb, scale, L.s,
                                       L.dcf));
            }
            if (fx != MC_SPEC_F32) MC_TRY(mc_launch_spec_scale(L.dcf, H, nb, gx + fi, ix + fi, L.s));
            if (ff != MC_SPEC_F32) MC_TRY(mc_launch_spec_scale(L.dcf, H, nb, gf + fi, iff + fi, L.s));
            {
                KTimer kt(c, "fft_cols_fwd", L.s);
                MC_TRY(mc_cols_fwd(*pf, L.T, pf->t_elems, spec_at(Sf_all, fi * fs, ff), fs,
                                   Sx_all ? spec_at(Sx_all, fi * xfs, fx) : nullptr, g.xcx, g.xh, xfs,
                                   nb, L.s, ff, ff != MC_SPEC_F32 ? gf + fi : nullptr, fx,
                                   fx != MC_SPEC_F32 ? gx + fi : nullptr));
            }
        }
    }
    return MC_OK;
}

// Align/CFullAlign.cpp:160-163,96: global shifts in full-resolution pixels
// (xcf bin multiply), truncated toward zero at 0.01 px.
void global_shifts(const std::vector<double>& xs, const Geo& g, int N, double* out) {
    for (int i = 0; i < N; i++) {
        const float sx = (float)xs[2 * i] * g.xbin[0];
        const float sy = (float)xs[2 * i + 1] * g.xbin[1];
        out[2 * i] = 0.01 * (double)(long long)((double)sx * 100.0);
        out[2 * i + 1] = 0.01 * (double)(long long)((double)sy * 100.0);
    }
}

// Align/CPatchAlign.cpp:47-65: shifts relative to the reference frame
// (-FmRef, 1-based; the middle frame when negative).
void relative_to_ref(double* sh, int N, int fm_ref) {
    const int ref = (fm_ref < 0) ? N / 2 : std::min(std::max(fm_ref - 1, 0), N - 1);
    const double rx = sh[2 * ref], ry = sh[2 * ref + 1];
    for (int i = 0; i < N; i++) {
        sh[2 * i] -= rx;
        sh[2 * i + 1] -= ry;
    }
}

// Movie pipelining across align_batch calls of one geometry: the stacks of
// the movie being aligned live in buffer slot `slot`; when `ready`, its
// spectra were transformed by the previous call on the second stream and
// `ev[slot]` marks their completion.  `next` (or null) is transformed into
// the other slot on the second stream while this movie aligns.
struct Pipe {
    int slot = 0;
    bool ready = false;
    const mc_movie* next = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
};

// ---------------------------------------------------------------------------
// `count` movies of one geometry.  The alignment iterations of all movies
// run as one batch of jobs (the global alignment of movie m is job m, patch
// q of movie m is job m * P + q); every job's arithmetic is the same as in a
// batch of one, so results do not depend on the batching.  Overrides need
// count == 1.  pp[m] supplies movie m's log name; everything else comes
// from pp[0].
// -Align 0 (Align/CSimpleSum.cpp:22-135, MrcUtil/CSumFFTStack.cpp:60-210):
// no alignment.  The frames are summed as they are (with -SplitSum also the
// even-indexed frames, and all minus those), and each sum, and with
// -OutStack each frame, is cropped to the -FtBin size through a normalised
// R2C and a C2R, with no clamp at zero (CSimpleSum::mCropFrame, mUnpad).
int simple_sums(mc_ctx* c, const mc_params* p, const mc_movie* mv, mc_result* res) {
    const int W = mv->W, H = mv->H, N = mv->N;
    const Geo g = make_geo(W, H, N, 0, 0, p->ft_bin);
    const int nsum = p->split_sum ? 3 : 1;
    res->num_sums = nsum;
    res->out_w = g.ow;
    res->out_h = g.oh;
    for (int i = 0; i < nsum; i++)
        if (!res->sums[i]) return fail(c, MC_E_ARG, "missing output sum buffer");
    if (p->out_stack[0] && !res->stack) return fail(c, MC_E_ARG, "out_stack needs a result stack buffer");
    Fft2D* pf = plan(c, W, H);
    Fft2D* po = (g.ow == W && g.oh == H) ? pf : plan(c, g.ow, g.oh);
    if (!pf || !po) return fail(c, MC_E_ARG, "unsupported FFT size");
    hipStream_t s = c->stream;
    MC_CHECK(wait_movie_loads(c, mv, s));
    const long long np = (long long)W * H, fs = (long long)g.cx * H, ofs = (long long)g.ow * g.oh;
    const float scale = (float)(1.0 / ((double)W * H));
    const int SC = std::min(N, 8);
    MC_BUF(R, float, "ss_sums", (size_t)nsum * np);
    MC_BUF(T, cfloat, "ss_T", (size_t)SC * pf->t_elems);
    MC_BUF(S, cfloat, "ss_S", (size_t)SC * fs);
    MC_BUF(To, cfloat, "ss_To", (size_t)SC * po->t_elems);
    auto crop = [&](const float* in, long long pitch, long long bstride, int nb, float* out) -> int {
        MC_TRY(mc_rows_r2c(*pf, in, pitch, bstride, nullptr, T, pf->t_elems, nb, scale, s));
        MC_TRY(mc_cols_fwd(*pf, T, pf->t_elems, S, fs, nullptr, 0, 0, 0, nb, s));
        MC_TRY(mc_cols_inv(*po, S, fs, g.cx, H, nullptr, To, po->t_elems, nb, s));
        MC_TRY(mc_rows_c2r(*po, To, po->t_elems, out, g.ow, ofs, nb, 0, s));
        return MC_OK;
    };
    {
        KTimer kt(c, "simple_sums");
        MC_TRY(mc_launch_simple_sums(mv->frames, W, H, N, mv->pitch, mv->frame_stride, p->split_sum, R, s));
    }
    for (int i = 0; i < nsum; i++) MC_CHECK(crop(R + (long long)i * np, W, np, 1, res->sums[i]));
    if (p->out_stack[0]) {
        // then Align/CZbinStack.cpp:27-70 as after an alignment
        const int zb = p->out_stack[1];
        Groups zg;
        if (zb > 1) zg = make_groups(zb, N, p->throw0, p->fm_int_start, p->fm_int_size);
        float* Rst = nullptr;
        if (zb > 1) {
            MC_BUF(r_, float, "stk_chunk", (size_t)SC * ofs);
            Rst = r_;
        }
        for (int f0 = 0; f0 < N; f0 += SC) {
            const int nb = std::min(SC, N - f0);
            float* dst = zb > 1 ? Rst : res->stack + (long long)f0 * ofs;
            MC_CHECK(crop(mv->frames + (long long)f0 * mv->frame_stride, mv->pitch, mv->frame_stride, nb, dst));
            for (int k = 0; zb > 1 && k < nb; k++) {
                const int f = f0 + k, gi = zg.fgroup[f];
                MC_TRY(mc_launch_stack_add(Rst + (long long)k * ofs, res->stack + (long long)gi * ofs, ofs,
                                           f == zg.start[gi], s));
            }
        }
    }
    MC_TRY(hipStreamSynchronize(s));
    if (res->global_shift) std::fill(res->global_shift, res->global_shift + 2 * N, 0.0f);
    if (res->global_iters) res->global_iters[0] = 0;
    return MC_OK;
}

int align_batch(mc_ctx* c, const mc_params* const* pp_, const mc_movie* const* mvs,
                const mc_override* ov, mc_result* const* rs, int count, Pipe* pipe = nullptr) {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    if (count < 1 || (count > 1 && ov)) return fail(c, MC_E_ARG, "overrides need a single movie");
    const mc_params* p = pp_[0];
    const int W = mvs[0]->W, H = mvs[0]->H, N = mvs[0]->N;
    for (int m = 0; m < count; m++) {
        const mc_movie* mv = mvs[m];
        if (!mv->frames || mv->W != W || mv->H != H || mv->N != N || W < 8 || (W % 2) || H < 8 ||
            N < 2 || mv->pitch < W || mv->frame_stride < mv->pitch * H)
            return fail(c, MC_E_ARG, "invalid movie geometry");
    }
    for (int m = 0; m < count; m++) std::fill(rs[m]->seconds_stage, rs[m]->seconds_stage + 5, 0.0);
    if (p->align == 0) {
        for (int m = 0; m < count; m++) MC_CHECK(simple_sums(c, pp_[m], mvs[m], rs[m]));
        return MC_OK;
    }
    if (p->iter < 1) return fail(c, MC_E_ARG, "-Iter must be >= 1");
    const bool patch = p->patch[0] > 1 && p->patch[1] > 1;
    const Geo g = make_geo(W, H, N, p->patch[0], p->patch[1], p->ft_bin);
    const int P = patch ? p->patch[0] * p->patch[1] : 0;
    if (patch && (g.pw > g.xw / 2 || g.ph > g.xh / 2 || g.pw < 8 || g.ph < 8))
        return fail(c, MC_E_ARG, "patch grid too fine for the xcf size");
    if (P > 64) return fail(c, MC_E_ARG, "at most 64 patches");
    const bool replay = ov && ov->replay_global;
    if (patch && replay && (!ov->replay_patch || !ov->replay_centers))
        return fail(c, MC_E_ARG, "patch replay needs local shifts and centres");
    const int quirks = p->ref_quirks;
    hipStream_t s = c->stream;
    if (pipe && (count != 1 || ov)) return fail(c, MC_E_ARG, "pipelined movies go one at a time");
    // frames written by asynchronous loads: wait for exactly those loads
    if (!(pipe && pipe->ready))
        for (int m = 0; m < count; m++) MC_CHECK(wait_movie_loads(c, mvs[m], s));

    // DataUtil/CFmIntParam.cpp:80-123 and CBufferPool.cpp:185-202
    const bool dw = !(p->kv < 80 || p->kv > 300) && p->pix_size > 0 && p->fm_dose > 0;
    const bool dws = dw && !(p->sum_range[0] < 0) && !(p->sum_range[1] <= 0) &&
                     (p->sum_range[1] - p->sum_range[0] > 1.0f);
    const int nsum = 1 + (p->split_sum ? 2 : 0) + (dw ? 1 + (dws ? 1 : 0) : 0);
    for (int m = 0; m < count; m++) {
        rs[m]->num_sums = nsum;
        rs[m]->out_w = g.ow;
        rs[m]->out_h = g.oh;
        for (int i = 0; i < nsum; i++)
            if (!rs[m]->sums[i]) return fail(c, MC_E_ARG, "missing output sum buffer");
    }

    Fft2D* pf = plan(c, W, H);
    Fft2D* px = plan(c, g.xw, g.xh);
    Fft2D* pp = patch ? plan(c, g.pw, g.ph) : nullptr;
    Fft2D* po = (g.ow == W && g.oh == H) ? pf : plan(c, g.ow, g.oh);
    if (!pf || !px || (patch && !pp) || !po) return fail(c, MC_E_ARG, "unsupported FFT size");
    if (p->spectra_fmt < MC_SPEC_F32 || p->spectra_fmt > MC_SPEC_BF16)
        return fail(c, MC_E_ARG, "spectra_fmt must be 0 (FP32), 1 (FP16) or 2 (BF16)");

    const long long fs = (long long)g.cx * H;
    const long long xfs = (long long)g.xcx * g.xh;
    const bool xcf_alias = (g.xcx == g.cx && g.xh == H);
    // A reduced stack format needs the register-pass kernels on every stage
    // that writes or reads the stack; otherwise the stack stays FP32.
    int fx = p->spectra_fmt, fpch = p->spectra_fmt;
    if (xcf_alias || replay || !mc_fft_rp_rows(*pf) || !mc_fft_rp_cols(*pf) ||
        !mc_correlate_has_rp(g.xh) || (patch && !mc_fft_rp_cols(*px)))
        fx = MC_SPEC_F32;
    if (!patch || replay || !mc_fft_rp_rows(*pp) || !mc_fft_rp_cols(*pp) ||
        !mc_correlate_has_rp(g.ph))
        fpch = MC_SPEC_F32;
    // Full-frame spectra in a reduced format: every stage that reads or
    // writes them must run the register-pass kernels; frame weights
    // (in place, FP32) and a frame stack that is also the xcf stack keep FP32.
    if (p->frame_fmt < MC_SPEC_F32 || p->frame_fmt > MC_SPEC_BF16)
        return fail(c, MC_E_ARG, "frame_fmt must be 0 (FP32), 1 (FP16) or 2 (BF16)");
    int ff = p->frame_fmt;
    if (xcf_alias || !mc_fft_rp_rows(*pf) || !mc_fft_rp_cols(*pf) || p->in_fm_motion || p->corr_interp ||
        (po != pf && !mc_fft_rp_cols(*po)))
        ff = MC_SPEC_F32;
    // the forward column pass stores both the frames and the xcf crop; mixed
    // reduced formats are not compiled
    if (ff != MC_SPEC_F32 && fx != MC_SPEC_F32 && ff != fx) fx = ff;
    for (int m = 0; m < count; m++) {
        rs[m]->spectra_fmt_used[0] = fx;
        rs[m]->spectra_fmt_used[1] = fpch;
        rs[m]->frame_fmt_used = ff;
        rs[m]->prior_fallbacks = 0;
    }
    const int CH = MC_CHUNK_FRAMES;  // frames per forward/inverse chunk
    // movie m's full spectra at Sf_all + m * N * fs, its xcf stack at element
    // m * N * xfs of Sx_all, its patch stack at element m * P * N * pfs of Sp_all
    // the stacks of a movie slot: full spectra, xcf stack, per-frame factors
    struct Stacks {
        cfloat* Sf = nullptr;   // stored in format ff
        cfloat* Sx = nullptr;
        float* gx = nullptr;
        float* ix = nullptr;
        float* gf = nullptr;    // per-frame store / load factors of a reduced ff
        float* iff = nullptr;
    };
    auto stacks = [&](int slot, Stacks& o) -> bool {
        const std::string k = slot ? "1" : "";
        o.Sf = (cfloat*)buf(c, ("Sf" + k).c_str(), (ff == MC_SPEC_F32 ? 8 : 4) * (size_t)count * N * fs);
        if (o.Sf && ff != MC_SPEC_F32) {
            o.gf = (float*)buf(c, ("gf" + k).c_str(), sizeof(float) * (size_t)count * N);
            o.iff = (float*)buf(c, ("iff" + k).c_str(), sizeof(float) * (size_t)count * N);
            if (!o.gf || !o.iff) return false;
        }
        o.Sx = o.Sf;
        if (o.Sf && !xcf_alias && !replay)
            o.Sx = (cfloat*)buf(c, ("Sx" + k).c_str(), (fx == MC_SPEC_F32 ? 8 : 4) * (size_t)count * N * xfs);
        if (o.Sx && fx != MC_SPEC_F32) {
            o.gx = (float*)buf(c, ("gx" + k).c_str(), sizeof(float) * (size_t)count * N);
            o.ix = (float*)buf(c, ("ix" + k).c_str(), sizeof(float) * (size_t)count * N);
            return o.gx && o.ix;
        }
        return o.Sx != nullptr;
    };
    Stacks cur;
    if (!stacks(pipe ? pipe->slot : 0, cur)) return fail(c, MC_E_NOMEM, "movie stacks");
    cfloat* Sf_all = cur.Sf;
    cfloat* Sx_all = cur.Sx;
    float* gx = cur.gx;
    float* ix = cur.ix;
    float* gf = cur.gf;
    float* iff = cur.iff;
    MC_BUF(T, cfloat, "T", (size_t)CH * pf->t_elems);
    float* dcf = nullptr;   // |row DC| of the chunk's frames
    if (fx != MC_SPEC_F32 || ff != MC_SPEC_F32) {
        dcf = (float*)buf(c, "dcf", sizeof(float) * (size_t)CH * H);
        if (!dcf) return fail(c, MC_E_NOMEM, "stack scale factors");
    }

    // Frame-chunk loops alternate between the context stream and a second
    // stream, each with its own staging buffer, so that one chunk's column
    // pass can run beside the next chunk's row pass (their workgroups are
    // LDS-limited to two per CU and latency-bound).  Fork and join by events.
    const bool two = p->chunk_streams >= 2;
    MC_BUF(T2, cfloat, "T2", (size_t)CH * pf->t_elems);
    float* dcf2 = nullptr;
    if (fx != MC_SPEC_F32 || ff != MC_SPEC_F32) {
        dcf2 = (float*)buf(c, "dcf2", sizeof(float) * (size_t)CH * H);
        if (!dcf2) return fail(c, MC_E_NOMEM, "stack scale factors");
    }
    auto fork = [&]() -> hipError_t {
        hipError_t e = hipEventRecord(c->evf, s);
        return e ? e : hipStreamWaitEvent(c->s2, c->evf, 0);
    };
    auto join = [&]() -> hipError_t {
        hipError_t e = hipEventRecord(c->evj, c->s2);
        return e ? e : hipStreamWaitEvent(s, c->evj, 0);
    };

    const float scale = (float)(1.0 / ((double)W * H));
    void* Sx_fwd = (xcf_alias || replay) ? nullptr : (void*)Sx_all;
    Range r_call("mc::align_batch");
    StageClock sc(c, s);
    MC_TRY(sc.begin(0));
    if (pipe && pipe->ready) {
        // transformed on the second stream while the previous movie aligned
        MC_TRY(hipStreamWaitEvent(s, pipe->ev[pipe->slot], 0));
    } else {
        const FwdLane lanes[2] = {{s, T, dcf}, {c->s2, T2, dcf2}};
        MC_TRY(fork());
        MC_CHECK(enqueue_forward(c, pf, mvs, count, g, fx, Sf_all, Sx_fwd, gx, ix, lanes, two ? 2 : 1, CH,
                                 ff, gf, iff));
        MC_TRY(join());
    }
    // The next movie's forward transform on the second stream, behind
    // everything already queued there; the frame-chunk loops below queue
    // their second-stream chunks after it.  Its launches are queued once the
    // global alignment's first iterations are (queue_next through
    // tl_after_launch, or right after that alignment): launching them first
    // kept the main stream idle for about a millisecond per movie.
    Stacks nx;
    bool next_queued = true;
    if (pipe && pipe->next) {
        if (!stacks(pipe->slot ^ 1, nx)) return fail(c, MC_E_NOMEM, "next movie's stacks");
        MC_CHECK(wait_movie_loads(c, pipe->next, c->s2));
        next_queued = false;
    }
    std::function<int()> queue_next = [&]() -> int {
        if (next_queued) return MC_OK;
        next_queued = true;
        const FwdLane lane{c->s2, T2, dcf2};
        MC_CHECK(enqueue_forward(c, pf, &pipe->next, 1, g, fx, nx.Sf, Sx_fwd ? (void*)nx.Sx : nullptr,
                                 nx.gx, nx.ix, &lane, 1, CH, ff, nx.gf, nx.iff));
        MC_TRY(hipEventRecord(pipe->ev[pipe->slot ^ 1], c->s2));
        return MC_OK;
    };
    if (replay) MC_CHECK(queue_next());

    struct MovieState {
        std::vector<double> gshift, ps, pc, gerr, xs;
        std::vector<float> gsf;
        std::vector<unsigned char> bad;
        std::vector<int> starts, piters;
        int giters = 0;
    };
    std::vector<MovieState> ms(count);
    for (auto& q : ms) {
        q.gshift.assign(2 * N, 0.0);
        q.gsf.assign(2 * N, 0.f);
        q.starts.assign(2 * P, 0);
        q.piters.assign(P, 0);
    }
    const double tol = p->tol / std::max(g.xbin[0], g.xbin[1]);
    const auto ta = clk::now();
    bool early_inverse = false;
    int early_split = 0;  // frames the early inverse transforms (the rest: the correction)

    if (!replay) {
        Groups g0 = make_groups(p->group[0], N, p->throw0, p->fm_int_start, p->fm_int_size);
        IterOut go;
        MC_TRY(sc.begin(1));
        Range r_g("mc::global_align");
        tl_after_launch = &queue_next;
        const int rc_g = run_iter_meet(c, "g_", Sx_all, count, N, g.xw, g.xh, px, g0, p->bft[0],
                                       p->phase_only, p->iter, tol, quirks, &go, fx, ix,
                                       p->launch_graphs != 0);
        tl_after_launch = nullptr;
        MC_CHECK(rc_g);
        MC_CHECK(queue_next());  // (no-op once the hook has run)
        for (int m = 0; m < count; m++) {
            ms[m].giters = go.iters[m];
            ms[m].gerr.assign(go.errors.begin() + (size_t)m * p->iter,
                              go.errors.begin() + (size_t)m * p->iter + ms[m].giters);
            ms[m].xs.assign(go.shifts.begin() + (size_t)m * N * 2,
                            go.shifts.begin() + (size_t)(m + 1) * N * 2);
        }
        // The local-correction inverse of a single movie needs only the
        // global shifts (relative to the reference frame): it is queued on
        // the second stream once the patch preparation is queued (its 50
        // launches then cost host time while the GPU is busy), so that its
        // HBM-bound passes run beside the latency-bound patch alignment
        // instead of after it.  The correction below waits for it.  Same
        // kernels and inputs, so the same results.
        auto early_inv = [&]() -> int {
            std::vector<double> rel(2 * N);
            global_shifts(ms[0].xs, g, N, rel.data());
            relative_to_ref(rel.data(), N, p->fm_ref);
            MC_BUF(gsh_e, double, "gsh", 2 * N);
            MC_BUF(fA_e, cfloat, "fphA", (size_t)N * g.cx);
            MC_BUF(fB_e, cfloat, "fphB", (size_t)N * H);
            MC_BUF(Rr_e, float, "Rr", (size_t)N * W * H);
            // everything on the second stream, behind this movie's forward
            // chunks (or its prefetch); the shifts go through pinned host
            // memory, so the copy does not hold the host while the stream
            // is busy; the previous copy out of it has run once its
            // inverse's event has
            MC_TRY(hipEventSynchronize(c->evi));
            if (c->pin_shifts_n < (size_t)(2 * N)) {
                if (c->pin_shifts) MC_TRY(hipHostFree(c->pin_shifts));
                c->pin_shifts = nullptr;
                c->pin_shifts_n = 0;
                MC_TRY(hipHostMalloc((void**)&c->pin_shifts, sizeof(double) * 2 * N));
                c->pin_shifts_n = (size_t)(2 * N);
            }
            std::copy(rel.begin(), rel.end(), c->pin_shifts);
            MC_TRY(hipMemcpyAsync(gsh_e, c->pin_shifts, sizeof(double) * 2 * N, hipMemcpyHostToDevice,
                                  c->s2));
            MC_TRY(mc_launch_phasors(gsh_e, -1.0, 1, nullptr, N, W, g.cx, H, N, fA_e, fB_e, c->s2));
            // Correct/CGenRealStack.cpp:132-166: -CorrInterp, -InFmMotion.
            // The weights rewrite the frame spectra in place.  When the xcf
            // stack is the frame stack itself (xcf bin 1), the patch
            // preparation queued on the main stream still reads it
            // unweighted (Align/CFullAlign.cpp:157-158), so the second
            // stream waits for everything queued there so far.
            if (p->in_fm_motion || p->corr_interp) MC_TRY(fork());
            MC_TRY(mc_launch_frame_weights(Sf_all, fs, N, g.cx, H, gsh_e, p->in_fm_motion, p->corr_interp,
                                           c->s2));
            MC_TRY(hipEventRecord(c->evw, c->s2));
            const long long rfs = (long long)W * H;
            early_split = std::min(N, CH * MC_EARLY_INV_CHUNKS);
            for (int f0 = 0; f0 < early_split; f0 += CH) {
                const int nb = std::min(CH, N - f0);
                PhasorTabs ph{fA_e + (long long)f0 * g.cx, fB_e + (long long)f0 * H};
                {
                    KTimer kt(c, "fft_cols_inv", c->s2);
                    MC_TRY(mc_cols_inv(*pf, spec_at(Sf_all, (long long)f0 * fs, ff), fs, g.cx, H, &ph, T2,
                                       pf->t_elems, nb, c->s2, ff, iff ? iff + f0 : nullptr));
                }
                {
                    KTimer kt(c, "fft_rows_c2r", c->s2);
                    MC_TRY(mc_rows_c2r(*pf, T2, pf->t_elems, Rr_e + f0 * rfs, W, rfs, nb, 0, c->s2));
                }
            }
            MC_TRY(hipEventRecord(c->evi, c->s2));
            early_inverse = true;
            return MC_OK;
        };
        if (patch) {
            MC_TRY(sc.begin(2));
            Range r_p("mc::patch_align");
            const long long pfs = (long long)g.pcx * g.ph;
            MC_BUF(Tp, cfloat, "Tp", (size_t)P * N * pp->t_elems);
            cfloat* Sp_all = (cfloat*)buf(c, "Sp", (fpch == MC_SPEC_F32 ? 8 : 4) * (size_t)count * P * N * pfs);
            if (!Sp_all) return fail(c, MC_E_NOMEM, "patch stack");
            float* dcp = nullptr;
            float* gp = nullptr;
            float* ip = nullptr;
            if (fpch != MC_SPEC_F32) {
                dcp = (float*)buf(c, "dcp", sizeof(float) * (size_t)P * N * g.ph);
                gp = (float*)buf(c, "gp", sizeof(float) * (size_t)count * P * N);
                ip = (float*)buf(c, "ip", sizeof(float) * (size_t)count * P * N);
                if (!dcp || !gp || !ip) return fail(c, MC_E_NOMEM, "patch scale factors");
            }
          for (int m = 0; m < count; m++) {
            std::vector<double>& xs = ms[m].xs;
            std::vector<int>& starts = ms[m].starts;
            cfloat* Sx = spec_at(Sx_all, (long long)m * N * xfs, fx);
            const float* ixm = ix ? ix + (long long)m * N : nullptr;
            // Align/CFullAlign.cpp:147-159
            MC_BUF(fin, double, "xfinal", 2 * N);
            MC_BUF(A, cfloat, "xphA", (size_t)N * g.xcx);
            MC_BUF(B, cfloat, "xphB", (size_t)N * g.xh);
            MC_BUF(xsum, cfloat, "xsum", (size_t)xfs);
            MC_TRY(hipMemcpyAsync(fin, xs.data(), sizeof(double) * 2 * N, hipMemcpyHostToDevice, s));
            MC_TRY(mc_launch_phasors(fin, -1.0, 0, nullptr, N, g.xw, g.xcx, g.xh, N, A, B, s));
            MC_TRY(mc_launch_aligned_sum(Sx, A, B, nullptr, 1, N, g.xcx, g.xh, xsum, s, fx, ixm));
            Fft2D* p128 = plan(c, 128, 128);
            if (!p128) return fail(c, MC_E_ARG, "128x128 plan");
            MC_BUF(T128, cfloat, "T128", (size_t)p128->t_elems);
            MC_BUF(img128, float, "img128", 128 * 128);
            MC_BUF(feat, unsigned char, "feat", 128 * 128);
            MC_TRY(mc_cols_inv(*p128, xsum, xfs, g.xcx, g.xh, nullptr, T128, p128->t_elems, 1, s));
            MC_TRY(mc_rows_c2r(*p128, T128, p128->t_elems, img128, 128, 128 * 128, 1, 0, s));
            MC_TRY(mc_launch_feature_map(img128, feat, s));
            std::vector<unsigned char> hf(128 * 128);
            MC_TRY(hipMemcpyAsync(hf.data(), feat, hf.size(), hipMemcpyDeviceToHost, s));
            MC_TRY(hipStreamSynchronize(s));
            if (ov && ov->patch_starts) std::copy(ov->patch_starts, ov->patch_starts + 2 * P, starts.begin());
            else patch_starts(hf.data(), p->patch[0], p->patch[1], g.xw, g.xh, starts.data());
            // xcf frames: shift (|s| >= 0.1) then C2R (Correct/CGenRealStack.cpp:147-166)
            MC_BUF(Tx, cfloat, "Tx", (size_t)N * px->t_elems);
            MC_BUF(Rx, float, "Rx", (size_t)N * g.xw * g.xh);
            MC_TRY(mc_launch_phasors(fin, -1.0, 1, nullptr, N, g.xw, g.xcx, g.xh, N, A, B, s));
            PhasorTabs ph{A, B};
            MC_TRY(mc_cols_inv(*px, Sx, xfs, g.xcx, g.xh, &ph, Tx, px->t_elems, N, s, fx, ixm));
            MC_TRY(mc_rows_c2r(*px, Tx, px->t_elems, Rx, g.xw, (long long)g.xw * g.xh, N, 0, s));
            // patches: CExtractPatch + unnormalised R2C, all patches and frames at once
            MC_BUF(dstart, int, "starts", 2 * P);
            MC_TRY(hipMemcpyAsync(dstart, starts.data(), sizeof(int) * 2 * P, hipMemcpyHostToDevice, s));
            const long long pm = (long long)m * P * N;
            MC_TRY(mc_rows_r2c_mod(*pp, Rx, g.xw, (long long)g.xw * g.xh, N, dstart, Tp, pp->t_elems,
                                   P * N, 1.0f, s, dcp));
            if (fpch != MC_SPEC_F32) MC_TRY(mc_launch_spec_scale(dcp, g.ph, P * N, gp + pm, ip + pm, s));
            MC_TRY(mc_cols_fwd(*pp, Tp, pp->t_elems, spec_at(Sp_all, pm * pfs, fpch), pfs, nullptr, 0,
                               0, 0, P * N, s, fpch, gp ? gp + pm : nullptr));
            if (count == 1 && two) MC_CHECK(early_inv());
          }
            Groups g1 = make_groups(p->group[1], N, p->throw0, p->fm_int_start, p->fm_int_size);
            // Learned prior: features of every patch after its first iteration
            // (training data, when a result asks for them) and/or iterations
            // guided by the attached network.
            const bool want_feat = std::any_of(rs, rs + count, [](mc_result* r) {
                return r->prior_features != nullptr || r->prior_targets != nullptr;
            });
            const bool guided = p->prior != 0;
            const int J = count * P, ns = 2 * N + 2, D = ns + 66 * N;
            PriorRun prr;
            if (guided || want_feat) {
                if (guided && (!c->prior || c->prior->frames != N))
                    return fail(c, MC_E_ARG, "no learned prior for this frame count attached");
                if (g.pw < 16 || g.ph < 16)
                    return fail(c, MC_E_ARG, "the learned prior needs patches of >= 16 xcf px");
                std::vector<float> side((size_t)J * ns);
                for (int m = 0; m < count; m++)
                    for (int q = 0; q < P; q++) {
                        float* sd = side.data() + ((size_t)m * P + q) * ns;
                        for (int i = 0; i < 2 * N; i++) sd[i] = (float)(ms[m].xs[i] / kPriorUnit);
                        sd[2 * N] = (ms[m].starts[2 * q] + g.pw / 2) / (float)g.xw * 2.0f - 1.0f;
                        sd[2 * N + 1] = (ms[m].starts[2 * q + 1] + g.ph / 2) / (float)g.xh * 2.0f - 1.0f;
                    }
                MC_BUF(dside, float, "prior_side", side.size());
                MC_BUF(dX, float, "prior_X", (size_t)J * D);
                MC_BUF(dY, float, "prior_Y", (size_t)J * 2 * N);
                MC_BUF(dfell, unsigned char, "prior_fell", (size_t)J);
                MC_TRY(hipMemcpyAsync(dside, side.data(), sizeof(float) * side.size(),
                                      hipMemcpyHostToDevice, s));
                MC_TRY(hipStreamSynchronize(s));
                prr.net = guided ? c->prior : nullptr;
                prr.side = dside;
                prr.ns = ns;
                prr.X = dX;
                prr.Y = dY;
                prr.fell = dfell;
                prr.fallback = p->prior_fallback / std::max(g.xbin[0], g.xbin[1]);
                prr.unit = kPriorUnit;
            }
            IterOut po_;
            if (guided || want_feat)
                MC_CHECK(run_iterative(c, "p_", Sp_all, J, N, g.pw, g.ph, pp, g1, p->bft[1],
                                       p->phase_only, p->iter, tol, quirks, &po_, fpch, ip,
                                       p->launch_graphs != 0, &prr));
            else
                MC_CHECK(run_iter_meet(c, "p_", Sp_all, J, N, g.pw, g.ph, pp, g1, p->bft[1],
                                       p->phase_only, p->iter, tol, quirks, &po_, fpch, ip,
                                       p->launch_graphs != 0));
            std::vector<float> hx;
            if (want_feat) {
                hx.resize((size_t)J * D);
                MC_TRY(hipMemcpyAsync(hx.data(), prr.X, sizeof(float) * hx.size(),
                                      hipMemcpyDeviceToHost, s));
                MC_TRY(hipStreamSynchronize(s));
            }
            for (int m = 0; m < count; m++) {
                mc_result* r = rs[m];
                if (r->prior_features)
                    std::copy(hx.begin() + (size_t)m * P * D, hx.begin() + (size_t)(m + 1) * P * D,
                              r->prior_features);
                if (r->prior_targets)
                    for (int q = 0; q < P; q++) {
                        const size_t jq = (size_t)m * P + q;
                        for (int i = 0; i < 2 * N; i++)
                            r->prior_targets[(size_t)q * 2 * N + i] =
                                (float)(po_.state[jq * 2 * N + i] / kPriorUnit) - hx[jq * D + ns + i];
                    }
                if (guided)
                    for (int q = 0; q < P; q++) {
                        const size_t jq = (size_t)m * P + q;
                        if (po_.fell[jq]) {
                            r->prior_fallbacks++;
                            po_.iters[jq] += 2;  // the guided pass before the restart
                        }
                    }
            }
          for (int m = 0; m < count; m++) {
            MovieState& q0 = ms[m];
            // Align/CMeasurePatches.cpp:66-82: centre and shifts in full-res px
            q0.ps.assign((size_t)N * P * 2, 0.0);
            q0.pc.assign(2 * P, 0.0);
            for (int q = 0; q < P; q++) {
                const size_t jq = (size_t)m * P + q;
                q0.piters[q] = po_.iters[jq];
                const int cxp = q0.starts[2 * q] + g.pw / 2, cyp = q0.starts[2 * q + 1] + g.ph / 2;
                q0.pc[2 * q] = (double)(cxp * g.xbin[0]);
                q0.pc[2 * q + 1] = (double)(cyp * g.xbin[1]);
                for (int i = 0; i < N; i++) {
                    q0.ps[((size_t)i * P + q) * 2] = po_.shifts[(jq * N + i) * 2] * g.xbin[0];
                    q0.ps[((size_t)i * P + q) * 2 + 1] = po_.shifts[(jq * N + i) * 2 + 1] * g.xbin[1];
                }
            }
            MC_BUF(dps, double, "ps", (size_t)N * P * 2);
            MC_BUF(dpc, double, "pc", 2 * P);
            MC_BUF(dbad, unsigned char, "bad", (size_t)N * P);
            MC_TRY(hipMemcpyAsync(dps, q0.ps.data(), sizeof(double) * q0.ps.size(), hipMemcpyHostToDevice, s));
            MC_TRY(hipMemcpyAsync(dpc, q0.pc.data(), sizeof(double) * q0.pc.size(), hipMemcpyHostToDevice, s));
            MC_TRY(mc_launch_patch_stats(dps, dpc, N, P, W, H, quirks, dbad, nullptr, s));
            q0.bad.assign((size_t)N * P, 0);
            MC_TRY(hipMemcpyAsync(q0.ps.data(), dps, sizeof(double) * q0.ps.size(), hipMemcpyDeviceToHost, s));
            MC_TRY(hipMemcpyAsync(q0.bad.data(), dbad, q0.bad.size(), hipMemcpyDeviceToHost, s));
            MC_TRY(hipStreamSynchronize(s));
          }
        }
        for (int m = 0; m < count; m++) global_shifts(ms[m].xs, g, N, ms[m].gshift.data());
    } else {
        MovieState& q0 = ms[0];
        for (int i = 0; i < 2 * N; i++) q0.gshift[i] = ov->replay_global[i];
        if (patch) {
            q0.ps.assign(ov->replay_patch, ov->replay_patch + (size_t)N * P * 2);
            q0.pc.assign(ov->replay_centers, ov->replay_centers + 2 * P);
            q0.bad.assign((size_t)N * P, 0);
            if (ov->replay_bad) q0.bad.assign(ov->replay_bad, ov->replay_bad + (size_t)N * P);
        }
    }
    const auto tb = clk::now();

    // ---- correction and sums, one movie at a time ----
    for (int m = 0; m < count; m++) {
        MC_TRY(sc.begin(3));
        Range r_c("mc::correct");
        mc_result* res = rs[m];
        std::vector<double>& gshift = ms[m].gshift;
        const std::vector<double>& ps = ms[m].ps;
        const std::vector<double>& pc = ms[m].pc;
        const std::vector<unsigned char>& bad = ms[m].bad;
        cfloat* Sf = spec_at(Sf_all, (long long)m * N * fs, ff);
        float* gfm = gf ? gf + (long long)m * N : nullptr;
        float* ifm = iff ? iff + (long long)m * N : nullptr;
        MC_BUF(sums, cfloat, "sums", (size_t)nsum * fs);
        MC_BUF(gsh, double, "gsh", 2 * N);
        MC_BUF(fA, cfloat, "fphA", (size_t)N * g.cx);
        MC_BUF(fB, cfloat, "fphB", (size_t)N * H);
        const double kvf = kv_factor(p->kv);
        // integrated frames: raw frames counted into each frame's dose
        int* dcum = nullptr;
        if (p->fm_int_size) {
            std::vector<int> cum(N);
            for (int i = 0; i < N; i++)
                cum[i] = (p->fm_int_start ? p->fm_int_start[i] : i + p->throw0) + p->fm_int_size[i];
            dcum = (int*)buf(c, "dcum", sizeof(int) * N);
            if (!dcum) return fail(c, MC_E_NOMEM, "dose table");
            MC_TRY(hipMemcpyAsync(dcum, cum.data(), sizeof(int) * N, hipMemcpyHostToDevice, s));
            MC_TRY(hipStreamSynchronize(s));
        }
        if (patch) {
            relative_to_ref(gshift.data(), N, p->fm_ref);
            // Correct/GCorrectPatchShift.cu:214-285: C2R, warp, normalised R2C
            MC_BUF(Rr, float, "Rr", (size_t)N * W * H);
            MC_BUF(Rw, float, "Rw", (size_t)N * W * H);
            const long long rfs = (long long)W * H;
            if (early_inverse) {
                // the chunks the early inverse left, on this stream, which
                // is free once the patch alignment is done
                if (early_split < N) {
                    MC_TRY(hipStreamWaitEvent(s, c->evw, 0));
                    for (int f0 = early_split; f0 < N; f0 += CH) {
                        const int nb = std::min(CH, N - f0);
                        PhasorTabs ph{fA + (long long)f0 * g.cx, fB + (long long)f0 * H};
                        {
                            KTimer kt(c, "fft_cols_inv", s);
                            MC_TRY(mc_cols_inv(*pf, spec_at(Sf, (long long)f0 * fs, ff), fs, g.cx, H, &ph, T,
                                               pf->t_elems, nb, s, ff, ifm ? ifm + f0 : nullptr));
                        }
                        {
                            KTimer kt(c, "fft_rows_c2r", s);
                            MC_TRY(mc_rows_c2r(*pf, T, pf->t_elems, Rr + f0 * rfs, W, rfs, nb, 0, s));
                        }
                    }
                }
                MC_TRY(hipStreamWaitEvent(s, c->evi, 0));
            } else {
            MC_TRY(hipMemcpyAsync(gsh, gshift.data(), sizeof(double) * 2 * N, hipMemcpyHostToDevice, s));
            MC_TRY(mc_launch_phasors(gsh, -1.0, 1, nullptr, N, W, g.cx, H, N, fA, fB, s));
            // Correct/CGenRealStack.cpp:132-166: -CorrInterp, -InFmMotion
            MC_TRY(mc_launch_frame_weights(Sf, fs, N, g.cx, H, gsh, p->in_fm_motion, p->corr_interp, s));
            MC_TRY(fork());
            for (int f0 = 0; f0 < N; f0 += CH) {
                const int nb = std::min(CH, N - f0);
                const bool odd = two && ((f0 / CH) & 1);
                hipStream_t sk = odd ? c->s2 : s;
                cfloat* Tk = odd ? T2 : T;
                PhasorTabs ph{fA + (long long)f0 * g.cx, fB + (long long)f0 * H};
                {
                    KTimer kt(c, "fft_cols_inv", sk);
                    MC_TRY(mc_cols_inv(*pf, spec_at(Sf, (long long)f0 * fs, ff), fs, g.cx, H, &ph, Tk, pf->t_elems,
                                       nb, sk, ff, ifm ? ifm + f0 : nullptr));
                }
                {
                    KTimer kt(c, "fft_rows_c2r", sk);
                    MC_TRY(mc_rows_c2r(*pf, Tk, pf->t_elems, Rr + f0 * rfs, W, rfs, nb, 0, sk));
                }
            }
            MC_TRY(join());
            }
            std::vector<float> psf(ps.size()), pcf(pc.size());
            for (size_t i = 0; i < ps.size(); i++) psf[i] = (float)ps[i];
            for (size_t i = 0; i < pc.size(); i++) pcf[i] = (float)pc[i];
            MC_BUF(dpsf, float, "psf", psf.size());
            MC_BUF(dpcf, float, "pcf", pcf.size());
            MC_BUF(dbad2, unsigned char, "bad2", bad.size());
            MC_TRY(hipMemcpyAsync(dpsf, psf.data(), sizeof(float) * psf.size(), hipMemcpyHostToDevice, s));
            MC_TRY(hipMemcpyAsync(dpcf, pcf.data(), sizeof(float) * pcf.size(), hipMemcpyHostToDevice, s));
            MC_TRY(hipMemcpyAsync(dbad2, bad.data(), bad.size(), hipMemcpyHostToDevice, s));
            {
                KTimer kt(c, "local_warp");
                MC_TRY(mc_launch_local_warp(Rr, Rw, W, rfs, W, H, N, P, dpcf, dpsf, dbad2,
                                            bad_any_of(bad.data(), N, P), s));
            }
            MC_TRY(fork());
            for (int f0 = 0; f0 < N; f0 += CH) {
                const int nb = std::min(CH, N - f0);
                const bool odd = two && ((f0 / CH) & 1);
                hipStream_t sk = odd ? c->s2 : s;
                cfloat* Tk = odd ? T2 : T;
                float* dck = odd ? dcf2 : dcf;
                {
                    KTimer kt(c, "fft_rows_r2c", sk);
                    MC_TRY(mc_rows_r2c_mod(*pf, Rw + f0 * rfs, W, rfs, 0, nullptr, Tk, pf->t_elems, nb, scale, sk,
                                           ff != MC_SPEC_F32 ? dck : nullptr));
                }
                // a reduced format: new factors from the warped frames' row
                // DCs (the inverse above has read this chunk's old ones)
                if (ff != MC_SPEC_F32)
                    MC_TRY(mc_launch_spec_scale(dck, H, nb, gfm + f0, ifm + f0, sk));
                {
                    KTimer kt(c, "fft_cols_fwd", sk);
                    MC_TRY(mc_cols_fwd(*pf, Tk, pf->t_elems, spec_at(Sf, (long long)f0 * fs, ff), fs, nullptr, 0,
                                       0, 0, nb, sk, ff, gfm ? gfm + f0 : nullptr));
                }
            }
            MC_TRY(join());
            // mGenSums' mMotionDecon (Correct/CCorrectFullShift.cpp:271-320) with
            // the same full shifts (Correct/GCorrectPatchShift.cu:106): the
            // motion weight a second time
            MC_TRY(sc.begin(4));
            MC_TRY(mc_launch_frame_weights(Sf, fs, N, g.cx, H, gsh, p->in_fm_motion, 0, s));
            KTimer kt(c, "dw_sums");
            MC_TRY(mc_launch_dw_sums(Sf, fs, N, g.cx, H, nullptr, nullptr, dw, p->pix_size, kvf,
                                     p->fm_dose, dws, p->sum_range[0], p->sum_range[1], p->split_sum,
                                     sums, fs, s, dcum, ff, ifm));
        } else {
            // Correct/CCorrectFullShift.cpp:229-309 with the absolute shifts
            MC_TRY(hipMemcpyAsync(gsh, gshift.data(), sizeof(double) * 2 * N, hipMemcpyHostToDevice, s));
            MC_TRY(mc_launch_phasors(gsh, -1.0, 1, nullptr, N, W, g.cx, H, N, fA, fB, s));
            MC_TRY(sc.begin(4));
            // mGenSums' mMotionDecon (Correct/CCorrectFullShift.cpp:271-320);
            // the weight commutes with the phase shift, so it goes on the
            // uncorrected spectra
            MC_TRY(mc_launch_frame_weights(Sf, fs, N, g.cx, H, gsh, p->in_fm_motion, 0, s));
            KTimer kt(c, "dw_sums");
            MC_TRY(mc_launch_dw_sums(Sf, fs, N, g.cx, H, fA, fB, dw, p->pix_size, kvf, p->fm_dose, dws,
                                     p->sum_range[0], p->sum_range[1], p->split_sum, sums, fs, s, dcum,
                                     ff, ifm));
        }
        // Correct/CCorrectFullShift.cpp:146-185: crop to -FtBin size, C2R, >= 0
        MC_BUF(To, cfloat, "To", (size_t)po->t_elems);
        for (int i = 0; i < nsum; i++) {
            MC_TRY(mc_cols_inv(*po, sums + (long long)i * fs, fs, g.cx, H, nullptr, To, po->t_elems, 1, s));
            MC_TRY(mc_rows_c2r(*po, To, po->t_elems, res->sums[i], g.ow, (long long)g.ow * g.oh, 1, 1, s));
        }
        // -OutStack: Correct/CCorrectFullShift.cpp:336-356 (mCropFrame), every
        // corrected frame before dose weighting cropped to the -FtBin size,
        // C2R and unpadded, with no positivity clamp; then
        // Align/CZbinStack.cpp:27-70 sums each z-bin group in frame order.
        if (p->out_stack[0]) {
            if (!res->stack) return fail(c, MC_E_ARG, "out_stack needs a result stack buffer");
            const int zb = p->out_stack[1];
            Groups zg;
            if (zb > 1) zg = make_groups(zb, N, p->throw0, p->fm_int_start, p->fm_int_size);
            const long long ofs = (long long)g.ow * g.oh;
            const int SC = std::min(N, 8);
            MC_BUF(Tst, cfloat, "To_stk", (size_t)SC * po->t_elems);
            float* Rst = nullptr;
            if (zb > 1) {
                MC_BUF(r_, float, "stk_chunk", (size_t)SC * ofs);
                Rst = r_;
            }
            // the global correction's phasors, at the output size when cropped
            // (rows of the crop keep their source frequency)
            const cfloat *oA = fA, *oB = fB;
            if (!patch && (g.ow != W || g.oh != H)) {
                MC_BUF(a_, cfloat, "stk_phA", (size_t)N * g.ocx);
                MC_BUF(b_, cfloat, "stk_phB", (size_t)N * g.oh);
                MC_TRY(mc_launch_phasors(gsh, -1.0, 1, nullptr, N, W, g.ocx, g.oh, N, a_, b_, s, H));
                oA = a_;
                oB = b_;
            }
            KTimer kt(c, "out_stack");
            for (int f0 = 0; f0 < N; f0 += SC) {
                const int nb = std::min(SC, N - f0);
                PhasorTabs ph{oA + (long long)f0 * g.ocx, oB + (long long)f0 * g.oh};
                MC_TRY(mc_cols_inv(*po, spec_at(Sf, (long long)f0 * fs, ff), fs, g.cx, H, patch ? nullptr : &ph,
                                   Tst, po->t_elems, nb, s, ff, ifm ? ifm + f0 : nullptr));
                float* dst = zb > 1 ? Rst : res->stack + (long long)f0 * ofs;
                MC_TRY(mc_rows_c2r(*po, Tst, po->t_elems, dst, g.ow, ofs, nb, 0, s));
                for (int k = 0; zb > 1 && k < nb; k++) {
                    const int f = f0 + k, gi = zg.fgroup[f];
                    MC_TRY(mc_launch_stack_add(Rst + (long long)k * ofs, res->stack + (long long)gi * ofs, ofs,
                                               f == zg.start[gi], s));
                }
            }
        }
    }
    MC_TRY(sc.begin(-1));
    MC_TRY(hipStreamSynchronize(s));
    const auto tc = clk::now();
    double stage_s[5];
    sc.read(stage_s);

    for (int m = 0; m < count; m++) {
        mc_result* res = rs[m];
        MovieState& q0 = ms[m];
        std::vector<float>& gsf = q0.gsf;
        const std::vector<double>& gshift = q0.gshift;
        const std::vector<double>& gerr = q0.gerr;
        const std::vector<double>& ps = q0.ps;
        const std::vector<double>& pc = q0.pc;
        const std::vector<unsigned char>& bad = q0.bad;
        const std::vector<int>& starts = q0.starts;
        const std::vector<int>& piters = q0.piters;
        const int giters = q0.giters;
        const mc_params* pm = pp_[m];
        for (int i = 0; i < 2 * N; i++) gsf[i] = (float)gshift[i];
        if (res->global_shift) std::copy(gsf.begin(), gsf.end(), res->global_shift);
        if (res->global_iters) res->global_iters[0] = giters;
        if (res->global_errors) {
            const float xb = std::max(g.xbin[0], g.xbin[1]);
            for (int i = 0; i < giters; i++) res->global_errors[i] = (float)(gerr[i] * xb);
        }
        std::vector<float> psf(ps.size()), pcf(pc.size());
        for (size_t i = 0; i < ps.size(); i++) psf[i] = (float)ps[i];
        for (size_t i = 0; i < pc.size(); i++) pcf[i] = (float)pc[i];
        if (patch) {
            if (res->patch_shift) std::copy(psf.begin(), psf.end(), res->patch_shift);
            if (res->patch_bad) std::copy(bad.begin(), bad.end(), res->patch_bad);
            if (res->patch_center) std::copy(pcf.begin(), pcf.end(), res->patch_center);
            if (res->patch_starts) std::copy(starts.begin(), starts.end(), res->patch_starts);
            if (res->patch_iters) std::copy(piters.begin(), piters.end(), res->patch_iters);
        }
        write_logs(pm, W, H, N, P, gsf, psf, bad, pcf);
        if (p->verbose) {
            for (int i = 0; i < giters; i++)
                printf("Iteration %2d  Error: %9.3f\n", i + 1, gerr[i] * std::max(g.xbin[0], g.xbin[1]));
            printf("Full-frame alignment shift\n");
            for (int i = 0; i < N; i++)
                printf("...... Frame (%3d) shift: %9.3f  %9.3f\n", i + 1, gsf[2 * i] - gsf[0],
                       gsf[2 * i + 1] - gsf[1]);
            printf("\n");
        }
        res->seconds_align = std::chrono::duration<double>(tb - ta).count();
        res->seconds_total = std::chrono::duration<double>(tc - t0).count();
        std::copy(stage_s, stage_s + 5, res->seconds_stage);
        if (p->verbose && m == 0) {
            // the reference's stage lines (Align/CPatchAlign.cpp:38,61,
            // Correct/GCorrectPatchShift.cu:134), GPU time of the call
            if (patch) printf("Patch alignment time: %.2f(sec)\n\n", stage_s[2]);
            printf("%s: %f sec\n\n", patch ? "Correction of local motion" : "Global shifts are corrected",
                   stage_s[3]);
        }
    }
    return MC_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

void mc_default_params(mc_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->iter = 15;
    p->tol = 0.1f;
    p->bft[0] = 500.0f;
    p->bft[1] = 100.0f;
    p->kv = 300;
    p->group[0] = 1;
    p->group[1] = 4;
    p->fm_ref = -1;
    p->ft_bin = 1.0f;
    p->sum_range[0] = 3.0f;
    p->sum_range[1] = 25.0f;
    p->ref_quirks = 1;
    p->launch_graphs = 1;
    p->chunk_streams = 2;
    p->prior = 0;
    p->prior_fallback = 0.2f;
    p->movie_pipeline = 1;
    p->frame_fmt = 0;
    p->out_stack[1] = 1;
    p->align = 1;
}

int mc_ctx_create(int device, void* stream, mc_ctx** out) {
    if (!out) return MC_E_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return MC_E_HIP;
    mc_ctx* c = new mc_ctx;
    c->device = device;
    if (scratch_for_device(device, &c->scratch) != hipSuccess) {
        delete c;
        return MC_E_HIP;
    }
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreate(&c->stream) != hipSuccess) {
            delete c;
            return MC_E_HIP;
        }
        c->own_stream = true;
    }
    if (mc_fft_init_constants() != 0 || mc_align_init_constants() != 0 ||
        hipStreamCreate(&c->s2) != hipSuccess ||
        hipEventCreateWithFlags(&c->evf, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->evj, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->evi, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->evw, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->evt, hipEventDisableTiming) != hipSuccess) {
        if (c->s2) (void)hipStreamDestroy(c->s2);
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return MC_E_HIP;
    }
    *out = c;
    return MC_OK;
}

int mc_ctx_destroy(mc_ctx* c) {
    if (!c) return MC_E_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->plans) {
        mc_fft2d_destroy(kv.second);
        delete kv.second;
    }
    for (auto& kv : c->rowtw) (void)hipFree(kv.second);
    for (auto& kv : c->bufs)
        if (kv.second.p) (void)hipFree(kv.second.p);
    for (auto& e : c->evs) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
    for (hipEvent_t e : c->stage_ev) (void)hipEventDestroy(e);
    if (c->s2) {
        (void)hipStreamSynchronize(c->s2);
        (void)hipStreamDestroy(c->s2);
    }
    if (c->evf) (void)hipEventDestroy(c->evf);
    if (c->evj) (void)hipEventDestroy(c->evj);
    if (c->evi) (void)hipEventDestroy(c->evi);
    if (c->evw) (void)hipEventDestroy(c->evw);
    if (c->evt) (void)hipEventDestroy(c->evt);
    if (c->pin_shifts) (void)hipHostFree(c->pin_shifts);
    if (c->pin_tab) (void)hipHostFree(c->pin_tab);
    if (c->s_in) {
        (void)hipStreamSynchronize(c->s_in);
        (void)hipStreamDestroy(c->s_in);
    }
    for (auto& l : c->loads) (void)hipEventDestroy(l.ev);
    for (hipEvent_t e : c->free_ev) (void)hipEventDestroy(e);
    for (auto& t : c->tab_keep) (void)hipEventDestroy(t.first);
    for (float* p : {c->gain, c->dark, c->gain_aug, c->dark_aug})
        if (p) (void)hipFree(p);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MC_OK;
}

const char* mc_last_error(mc_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mc_sync(mc_ctx* c) {
    if (!c) return MC_E_ARG;
    MC_TRY(hipStreamSynchronize(c->stream));
    return MC_OK;
}

int mc_malloc(mc_ctx* c, unsigned long long bytes, void** out) {
    if (!c || !out) return MC_E_ARG;
    *out = nullptr;
    MC_TRY(hipSetDevice(c->device));
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) return fail(c, MC_E_NOMEM, "hipMalloc");
    return MC_OK;
}

int mc_free(mc_ctx* c, void* ptr) {
    if (!c) return MC_E_ARG;
    if (!ptr) return MC_OK;
    MC_TRY(hipStreamSynchronize(c->stream));
    MC_TRY(hipFree(ptr));
    return MC_OK;
}

int mc_memcpy(mc_ctx* c, void* dst, const void* src, unsigned long long bytes, int dir) {
    if (!c || (!dst && bytes) || (!src && bytes) || dir < 0 || dir > 2) return MC_E_ARG;
    const hipMemcpyKind k = dir == 0 ? hipMemcpyHostToDevice
                                     : (dir == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    MC_TRY(hipMemcpyAsync(dst, src, bytes, k, c->stream));
    MC_TRY(hipStreamSynchronize(c->stream));
    return MC_OK;
}

int mc_geometry(int W, int H, int px, int py, int* sizes, float* xbin) {
    if (W < 8 || H < 8 || (W % 2)) return MC_E_ARG;
    const Geo g = make_geo(W, H, 2, px, py, 1.0f);
    sizes[0] = g.xw;
    sizes[1] = g.xh;
    sizes[2] = g.pw;
    sizes[3] = g.ph;
    xbin[0] = g.xbin[0];
    xbin[1] = g.xbin[1];
    return MC_OK;
}

int mc_output_geometry(const mc_params* p, int W, int H, int N, int* ns, int* ow, int* oh) {
    if (!p || W < 8 || H < 8 || (W % 2)) return MC_E_ARG;
    const Geo g = make_geo(W, H, N, p->patch[0], p->patch[1], p->ft_bin);
    const bool dw = !(p->kv < 80 || p->kv > 300) && p->pix_size > 0 && p->fm_dose > 0;
    const bool dws = dw && !(p->sum_range[0] < 0) && !(p->sum_range[1] <= 0) &&
                     (p->sum_range[1] - p->sum_range[0] > 1.0f);
    // -Align 0 makes the sum and, with -SplitSum, the odd and even sums only
    // (DataUtil/CAlnSums.cpp:23-28)
    *ns = p->align == 0 ? 1 + (p->split_sum ? 2 : 0)
                        : 1 + (p->split_sum ? 2 : 0) + (dw ? 1 + (dws ? 1 : 0) : 0);
    *ow = g.ow;
    *oh = g.oh;
    return MC_OK;
}

int mc_stack_frames(const mc_params* p, int N, int* frames) {
    if (!p || !frames || N < 1) return MC_E_ARG;
    if (!p->out_stack[0]) *frames = 0;
    else if (p->out_stack[1] <= 1) *frames = N;
    else *frames = make_groups(p->out_stack[1], N, p->throw0, p->fm_int_start, p->fm_int_size).G;
    return MC_OK;
}

static int align_movie(mc_ctx* c, const mc_params* p, const mc_movie* mv, const mc_override* ov,
                       mc_result* res) {
    const mc_params* pa[1] = {p};
    const mc_movie* ma[1] = {mv};
    mc_result* ra[1] = {res};
    return align_batch(c, pa, ma, ov, ra, 1);
}

static bool same_alignment_flags(const mc_params& a, const mc_params& b) {
    return a.patch[0] == b.patch[0] && a.patch[1] == b.patch[1] && a.iter == b.iter &&
           a.tol == b.tol && a.bft[0] == b.bft[0] && a.bft[1] == b.bft[1] &&
           a.phase_only == b.phase_only && a.fm_dose == b.fm_dose && a.pix_size == b.pix_size &&
           a.kv == b.kv && a.group[0] == b.group[0] && a.group[1] == b.group[1] &&
           a.fm_ref == b.fm_ref && a.ft_bin == b.ft_bin && a.sum_range[0] == b.sum_range[0] &&
           a.sum_range[1] == b.sum_range[1] && a.split_sum == b.split_sum &&
           a.throw0 == b.throw0 && a.ref_quirks == b.ref_quirks &&
           a.spectra_fmt == b.spectra_fmt && a.launch_graphs == b.launch_graphs &&
           a.chunk_streams == b.chunk_streams && a.prior == b.prior &&
           a.fm_int_start == b.fm_int_start && a.fm_int_size == b.fm_int_size &&
           a.prior_fallback == b.prior_fallback && a.verbose == b.verbose &&
           a.movie_pipeline == b.movie_pipeline && a.out_stack[0] == b.out_stack[0] &&
           a.out_stack[1] == b.out_stack[1] && a.align == b.align && a.in_fm_motion == b.in_fm_motion &&
           a.corr_interp == b.corr_interp && a.frame_fmt == b.frame_fmt;
}

// The movies `grp` (one geometry) of an mc_align_movies call: pipelined one
// at a time, or in batches of at most what `budget` bytes of device memory
// hold.
static int align_group(mc_ctx* c, const mc_params* params, const mc_movie* movies, mc_result* results,
                       const std::vector<int>& grp, double budget) {
    const mc_params& p = params[grp[0]];
    const mc_movie& mv = movies[grp[0]];
    if (mv.W < 8 || mv.H < 8 || mv.N < 2) return fail(c, MC_E_ARG, "invalid movie geometry");
    if (p.movie_pipeline && grp.size() > 1) {
        // one movie at a time; movie k+1's forward transform runs on the
        // second stream while movie k aligns (stack slots alternate)
        struct Events {
            hipEvent_t e[2] = {nullptr, nullptr};
            ~Events() {
                for (hipEvent_t x : e)
                    if (x) (void)hipEventDestroy(x);
            }
        } evs;
        for (hipEvent_t& e : evs.e)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                return fail(c, MC_E_HIP, "hipEventCreate");
        Pipe pipe;
        pipe.ev[0] = evs.e[0];
        pipe.ev[1] = evs.e[1];
        for (size_t k = 0; k < grp.size(); k++) {
            pipe.slot = (int)(k & 1);
            pipe.ready = k > 0;
            pipe.next = k + 1 < grp.size() ? &movies[grp[k + 1]] : nullptr;
            const mc_params* pa = &params[grp[k]];
            const mc_movie* ma = &movies[grp[k]];
            mc_result* ra = &results[grp[k]];
            const int rc = align_batch(c, &pa, &ma, nullptr, &ra, 1, &pipe);
            if (rc != MC_OK) {
                (void)hipStreamSynchronize(c->s2);
                return rc;
            }
        }
        return MC_OK;
    }
    const bool patch = p.patch[0] > 1 && p.patch[1] > 1;
    const Geo g = make_geo(mv.W, mv.H, mv.N, p.patch[0], p.patch[1], p.ft_bin);
    const double per_movie = (double)mv.N * ((double)g.cx * mv.H * 8 + (double)g.xcx * g.xh * 8 +
                                             (patch ? (double)p.patch[0] * p.patch[1] * g.pcx * g.ph * 8
                                                    : 0.0));
    const int cap = std::max(1, (int)std::min<double>(grp.size(), budget / per_movie));
    for (size_t b0 = 0; b0 < grp.size(); b0 += cap) {
        const int nb = (int)std::min<size_t>(cap, grp.size() - b0);
        std::vector<const mc_params*> pa(nb);
        std::vector<const mc_movie*> ma(nb);
        std::vector<mc_result*> ra(nb);
        for (int k = 0; k < nb; k++) {
            pa[k] = &params[grp[b0 + k]];
            ma[k] = &movies[grp[b0 + k]];
            ra[k] = &results[grp[b0 + k]];
        }
        MC_CHECK(align_batch(c, pa.data(), ma.data(), nullptr, ra.data(), nb));
    }
    return MC_OK;
}

int mc_align_movies(mc_ctx* c, const mc_params* params, const mc_movie* movies, int count,
                    mc_result* results) {
    if (!c || !params || !movies || !results || count < 1) return MC_E_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MC_E_HIP, "hipSetDevice");
    for (int m = 1; m < count; m++)
        if (!same_alignment_flags(params[0], params[m]))
            return fail(c, MC_E_ARG, "mc_align_movies: movies differ in alignment flags");
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return fail(c, MC_E_HIP, "hipMemGetInfo");
    // movies of equal size, groups in order of their first movie
    std::vector<std::vector<int>> groups;
    std::vector<char> taken(count, 0);
    for (int m0 = 0; m0 < count; m0++) {
        if (taken[m0]) continue;
        std::vector<int> grp;
        for (int m = m0; m < count; m++)
            if (!taken[m] && movies[m].W == movies[m0].W && movies[m].H == movies[m0].H &&
                movies[m].N == movies[m0].N) {
                grp.push_back(m);
                taken[m] = 1;
            }
        groups.push_back(grp);
    }
    const mc_params& p0 = params[0];
    if (groups.size() == 1 || p0.movie_pipeline || p0.prior || p0.align == 0) {
        for (const auto& grp : groups) MC_CHECK(align_group(c, params, movies, results, grp, 0.5 * (double)tot));
        return MC_OK;
    }
    // Movies of different sizes with batched iterations: one host thread per
    // size, taking turns; their alignment iterations meet (IterMeet), so each
    // iteration stage is one launch over the jobs of all sizes that share a
    // column length.  The device memory budget is split between the sizes.
    IterMeet meet;
    meet.live = (int)groups.size();
    const double share = 0.5 * (double)tot / (double)groups.size();
    std::vector<int> rcs(groups.size(), MC_OK);
    std::vector<std::string> errs(groups.size());
    std::vector<std::thread> th;
    for (size_t k = 0; k < groups.size(); k++) {
        th.emplace_back([&, k]() {
            std::unique_lock<std::mutex> lk(meet.mu);
            if (hipSetDevice(c->device) != hipSuccess) {
                rcs[k] = fail(c, MC_E_HIP, "hipSetDevice");
            } else {
                tl_buf_prefix = k ? "m" + std::to_string(k) + ":" : std::string();
                tl_meet = &meet;
                tl_lock = &lk;
                rcs[k] = align_group(c, params, movies, results, groups[k], share);
                tl_meet = nullptr;
                tl_lock = nullptr;
                tl_buf_prefix.clear();
            }
            if (rcs[k] != MC_OK) errs[k] = c->err;
            meet.live--;
            if (!meet.pending.empty() && (int)meet.pending.size() == meet.live) meet.run_pending(c);
        });
    }
    for (auto& t : th) t.join();
    for (size_t k = 0; k < groups.size(); k++)
        if (rcs[k] != MC_OK) {
            c->err = errs[k];
            return rcs[k];
        }
    return MC_OK;
}

int mc_align_movie(mc_ctx* c, const mc_params* p, const mc_movie* mv, const mc_override* ov,
                   mc_result* res) {
    if (!c || !p || !mv || !res) return MC_E_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MC_E_HIP, "hipSetDevice");
    return align_movie(c, p, mv, ov, res);
}

int mc_rfft2(mc_ctx* c, const float* in, int W, int H, long long pitch, long long in_stride,
             int batch, float scale, float* out_spec) {
    if (!c || !in || !out_spec || batch < 1 || pitch < W) return MC_E_ARG;
    Fft2D* pl = plan(c, W, H);
    if (!pl) return fail(c, MC_E_ARG, "unsupported FFT size");
    MC_BUF(T, cfloat, "api_T", (size_t)batch * pl->t_elems);
    MC_TRY(mc_rows_r2c(*pl, in, pitch, in_stride, nullptr, T, pl->t_elems, batch, scale, c->stream));
    MC_TRY(mc_cols_fwd(*pl, T, pl->t_elems, (cfloat*)out_spec, pl->s_elems, nullptr, 0, 0, 0, batch,
                       c->stream));
    return MC_OK;
}

int mc_irfft2(mc_ctx* c, const float* in_spec, int W, int H, int batch, float* out,
              long long pitch, long long out_stride, int clamp) {
    if (!c || !in_spec || !out || batch < 1 || pitch < W) return MC_E_ARG;
    Fft2D* pl = plan(c, W, H);
    if (!pl) return fail(c, MC_E_ARG, "unsupported FFT size");
    MC_BUF(T, cfloat, "api_T", (size_t)batch * pl->t_elems);
    MC_TRY(mc_cols_inv(*pl, (const cfloat*)in_spec, pl->s_elems, pl->cx, H, nullptr, T, pl->t_elems,
                       batch, c->stream));
    MC_TRY(mc_rows_c2r(*pl, T, pl->t_elems, out, pitch, out_stride, batch, clamp, c->stream));
    return MC_OK;
}

int mc_phase_shift_accumulate(mc_ctx* c, const float* in_spec, int W, int H, int N,
                              const float* shifts, float* out_spec) {
    if (!c || !in_spec || !shifts || !out_spec || N < 1 || W < 2 || (W % 2)) return MC_E_ARG;
    const int cx = W / 2 + 1;
    std::vector<double> sh(shifts, shifts + 2 * N);
    MC_BUF(ds, double, "api_sh", 2 * N);
    MC_BUF(A, cfloat, "api_A", (size_t)N * cx);
    MC_BUF(B, cfloat, "api_B", (size_t)N * H);
    MC_TRY(hipMemcpyAsync(ds, sh.data(), sizeof(double) * 2 * N, hipMemcpyHostToDevice, c->stream));
    MC_TRY(mc_launch_phasors(ds, 1.0, 0, nullptr, N, W, cx, H, N, A, B, c->stream));
    MC_TRY(mc_launch_aligned_sum((const cfloat*)in_spec, A, B, nullptr, 1, N, cx, H,
                                 (cfloat*)out_spec, c->stream));
    return MC_OK;
}

int mc_group_correlate_window(mc_ctx* c, const float* sum_spec, const float* grp_spec, int W,
                              int H, float bfactor, int phase_only, int sea_x, int sea_y,
                              float* window_host) {
    if (!c || !sum_spec || !grp_spec || !window_host || W < 4 || (W % 2) || sea_x < 3 ||
        sea_y < 3 || sea_x > 64 || sea_y > std::min(64, H) || sea_x > W)
        return MC_E_ARG;
    Fft2D* pl = plan(c, W, H);
    if (!pl) return fail(c, MC_E_ARG, "unsupported FFT size");
    const int cx = W / 2 + 1;
    IterState st;
    st.J = 1;
    st.N = 1;
    st.G = 1;
    MC_BUF(gs, int, "api_g", 2);
    const int hg[2] = {0, 1};
    MC_TRY(hipMemcpyAsync(gs, hg, sizeof(hg), hipMemcpyHostToDevice, c->stream));
    st.gstart = gs;
    st.gsize = gs + 1;
    st.done = nullptr;
    MC_BUF(zero, double, "api_z", 2);
    MC_TRY(hipMemsetAsync(zero, 0, 2 * sizeof(double), c->stream));
    MC_BUF(A, cfloat, "api_A", cx);
    MC_BUF(B, cfloat, "api_B", H);
    MC_TRY(mc_launch_phasors(zero, 1.0, 0, nullptr, 1, W, cx, H, 1, A, B, c->stream));
    MC_BUF(T1, cfloat, "api_T1", (size_t)sea_y * cx);
    MC_BUF(win, float, "api_win", (size_t)sea_x * sea_y);
    cfloat* twW = row_twiddles(c, W);
    MC_TRY(mc_launch_correlate(pl->col, pl->col_ptw, pl->C, (const cfloat*)grp_spec, (const cfloat*)sum_spec, A,
                               B, st, cx, H, bfactor, phase_only, sea_y, T1, c->stream));
    MC_TRY(mc_launch_row_dft(T1, twW, nullptr, 1, 1, cx, W, sea_x, sea_y, win, c->stream));
    MC_TRY(hipMemcpyAsync(window_host, win, sizeof(float) * sea_x * sea_y, hipMemcpyDeviceToHost,
                          c->stream));
    MC_TRY(hipStreamSynchronize(c->stream));
    return MC_OK;
}

int mc_peak_fit(mc_ctx* c, const float* windows, int nx, int ny, int count, float* shifts_host) {
    if (!c || !windows || !shifts_host || nx < 3 || ny < 3 || count < 1) return MC_E_ARG;
    MC_BUF(d, double, "api_pk", 2 * count);
    MC_TRY(mc_launch_peaks_f(windows, nx, ny, count, d, c->stream));
    std::vector<double> h(2 * count);
    MC_TRY(hipMemcpyAsync(h.data(), d, sizeof(double) * 2 * count, hipMemcpyDeviceToHost, c->stream));
    MC_TRY(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 2 * count; i++) shifts_host[i] = (float)h[i];
    return MC_OK;
}

int mc_spline_despike_fp64(mc_ctx* c, double* traj, int n, int count, int quirks) {
    if (!c || !traj || n < 3 || count < 1) return MC_E_ARG;
    MC_BUF(d, double, "api_traj", (size_t)2 * n * count);
    MC_BUF(sc, double, "api_trs", (size_t)5 * n * count);
    MC_TRY(hipMemcpyAsync(d, traj, sizeof(double) * 2 * n * count, hipMemcpyHostToDevice, c->stream));
    MC_TRY(mc_launch_despike(d, sc, n, count, quirks, c->stream));
    MC_TRY(hipMemcpyAsync(traj, d, sizeof(double) * 2 * n * count, hipMemcpyDeviceToHost, c->stream));
    MC_TRY(hipStreamSynchronize(c->stream));
    return MC_OK;
}

int mc_patch_stats_fp64(mc_ctx* c, double* shifts, const double* centers, int N, int P, int W,
                        int H, int quirks, unsigned char* bad_host, int* ref_host) {
    if (!c || !shifts || !centers || !bad_host || N < 2 || P < 2) return MC_E_ARG;
    MC_BUF(ds, double, "api_ps", (size_t)N * P * 2);
    MC_BUF(dc, double, "api_pc", 2 * P);
    MC_BUF(db, unsigned char, "api_bad", (size_t)N * P);
    MC_BUF(dr, int, "api_ref", 1);
    MC_TRY(hipMemcpyAsync(ds, shifts, sizeof(double) * N * P * 2, hipMemcpyHostToDevice, c->stream));
    MC_TRY(hipMemcpyAsync(dc, centers, sizeof(double) * 2 * P, hipMemcpyHostToDevice, c->stream));
    MC_TRY(mc_launch_patch_stats(ds, dc, N, P, W, H, quirks, db, dr, c->stream));
    MC_TRY(hipMemcpyAsync(shifts, ds, sizeof(double) * N * P * 2, hipMemcpyDeviceToHost, c->stream));
    MC_TRY(hipMemcpyAsync(bad_host, db, (size_t)N * P, hipMemcpyDeviceToHost, c->stream));
    int r = 0;
    MC_TRY(hipMemcpyAsync(&r, dr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    MC_TRY(hipStreamSynchronize(c->stream));
    if (ref_host) *ref_host = r;
    return MC_OK;
}

int mc_local_warp(mc_ctx* c, const float* in, float* out, int W, int H, int N, int P,
                  const float* centers, const float* shifts, const unsigned char* bad) {
    if (!c || !in || !out || !centers || !shifts || W < 2 || H < 2 || N < 1 || P < 1 || P > 64)
        return MC_E_ARG;
    {
        const char* i0 = (const char*)in;
        const char* o0 = (const char*)out;
        const size_t bytes = sizeof(float) * (size_t)W * H * N;
        if (i0 < o0 + bytes && o0 < i0 + bytes) return fail(c, MC_E_ARG, "mc_local_warp: in and out overlap");
    }
    MC_BUF(dc, float, "api_wc", 2 * P);
    MC_BUF(ds, float, "api_ws", (size_t)2 * N * P);
    MC_BUF(db, unsigned char, "api_wb", (size_t)N * P);
    MC_TRY(hipMemcpyAsync(dc, centers, sizeof(float) * 2 * P, hipMemcpyHostToDevice, c->stream));
    MC_TRY(hipMemcpyAsync(ds, shifts, sizeof(float) * 2 * N * P, hipMemcpyHostToDevice, c->stream));
    if (bad) MC_TRY(hipMemcpyAsync(db, bad, (size_t)N * P, hipMemcpyHostToDevice, c->stream));
    else MC_TRY(hipMemsetAsync(db, 0, (size_t)N * P, c->stream));
    MC_TRY(mc_launch_local_warp(in, out, W, (long long)W * H, W, H, N, P, dc, ds, db, bad_any_of(bad, N, P),
                                c->stream));
    return MC_OK;
}

int mc_local_warp_noalloc(mc_ctx* c, const float* in, float* out, int W, int H, int N, int P,
                          const float* centers, const float* shifts, const unsigned char* bad) {
    if (!c || !in || !out || !centers || !shifts || W < 2 || H < 2 || N < 1 || P < 1 || P > 64)
        return MC_E_ARG;
    {
        const char* i0 = (const char*)in;
        const char* o0 = (const char*)out;
        const size_t bytes = sizeof(float) * (size_t)W * H * N;
        if (i0 < o0 + bytes && o0 < i0 + bytes)
            return fail(c, MC_E_ARG, "mc_local_warp_noalloc: in and out overlap");
    }
    // tables in the scratch array: centres, shifts, bad flags, each at a
    // 256-byte boundary
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t oc = 0, os = up(sizeof(float) * 2 * (size_t)P);
    const size_t ob = os + up(sizeof(float) * 2 * (size_t)N * P);
    const size_t need = ob + up((size_t)N * P);
    if (need > MC_SCRATCH_BYTES)
        return fail(c, MC_E_ARG, "mc_local_warp_noalloc: tables of N x P = " + std::to_string(N) + " x " +
                                     std::to_string(P) + " need " + std::to_string(need) +
                                     " bytes of scratch, more than its " +
                                     std::to_string(MC_SCRATCH_BYTES));
    if (!c->scratch) return fail(c, MC_E_STATE, "mc_local_warp_noalloc: no scratch address");
    float* dc = (float*)(c->scratch + oc);
    float* ds = (float*)(c->scratch + os);
    unsigned char* db = c->scratch + ob;
    MC_TRY(hipMemcpyAsync(dc, centers, sizeof(float) * 2 * P, hipMemcpyHostToDevice, c->stream));
    MC_TRY(hipMemcpyAsync(ds, shifts, sizeof(float) * 2 * (size_t)N * P, hipMemcpyHostToDevice, c->stream));
    // the previous call's flags are still there: write or clear all of them
    if (bad) MC_TRY(hipMemcpyAsync(db, bad, (size_t)N * P, hipMemcpyHostToDevice, c->stream));
    else MC_TRY(mc_launch_fill_u8(db, (long long)N * P, 0, c->stream));
    MC_TRY(mc_launch_local_warp(in, out, W, (long long)W * H, W, H, N, P, dc, ds, db, bad_any_of(bad, N, P),
                                c->stream));
    return MC_OK;
}

int mc_dose_weight_sum(mc_ctx* c, int W, int H, int N, float pix, int kv, float fm_dose,
                       float* out) {
    if (!c || !out || W < 2 || (W % 2) || H < 2 || N < 1 || pix <= 0) return MC_E_ARG;
    MC_TRY(mc_launch_dose_wsum(W / 2 + 1, H, N, pix, kv_factor(kv), fm_dose, out, c->stream));
    return MC_OK;
}

int mc_fourier_crop(mc_ctx* c, const float* in_spec, int in_w, int in_h, float* out_spec, int out_w,
                    int out_h, int batch, int sum) {
    if (!c || !in_spec || !out_spec || batch < 1 || in_w < 2 || in_h < 1 || out_w < 2 || out_h < 1 ||
        (in_w % 2) || (out_w % 2))
        return MC_E_ARG;
    const int inx = in_w / 2 + 1, outx = out_w / 2 + 1;
    MC_TRY(mc_launch_fourier_resize((const cfloat*)in_spec, inx, in_h, (long long)inx * in_h,
                                    (cfloat*)out_spec, outx, out_h, (long long)outx * out_h, batch,
                                    sum != 0, c->stream));
    return MC_OK;
}

int mc_feature_centres(mc_ctx* c, const float* xcf_sum_spec, int xw, int xh, int px, int py,
                       unsigned char* feature_map_host, int* starts_host) {
    if (!c || !xcf_sum_spec || !starts_host || xw < 128 || xh < 128 || (xw % 2) || px < 2 ||
        py < 2)
        return MC_E_ARG;
    Fft2D* p128 = plan(c, 128, 128);
    if (!p128) return fail(c, MC_E_ARG, "128x128 plan");
    MC_BUF(T128, cfloat, "api_T128", (size_t)p128->t_elems);
    MC_BUF(img, float, "api_img128", 128 * 128);
    MC_BUF(feat, unsigned char, "api_feat", 128 * 128);
    const int xcx = xw / 2 + 1;
    // Align/CDetectFeatures.cpp:58-100: the crop to 128 x 128 and its C2R
    MC_TRY(mc_cols_inv(*p128, (const cfloat*)xcf_sum_spec, (long long)xcx * xh, xcx, xh, nullptr, T128,
                       p128->t_elems, 1, c->stream));
    MC_TRY(mc_rows_c2r(*p128, T128, p128->t_elems, img, 128, 128 * 128, 1, 0, c->stream));
    MC_TRY(mc_launch_feature_map(img, feat, c->stream));
    std::vector<unsigned char> hf(128 * 128);
    MC_TRY(hipMemcpyAsync(hf.data(), feat, hf.size(), hipMemcpyDeviceToHost, c->stream));
    MC_TRY(hipStreamSynchronize(c->stream));
    if (feature_map_host) std::copy(hf.begin(), hf.end(), feature_map_host);
    patch_starts(hf.data(), px, py, xw, xh, starts_host);
    return MC_OK;
}

int mc_extract_patches(mc_ctx* c, const float* xcf_frames, int xw, int xh, int N, const int* starts_host,
                       int P, int pw, int ph, float* out_spec) {
    if (!c || !xcf_frames || !starts_host || !out_spec || N < 1 || P < 1 || pw < 4 || ph < 4 ||
        (pw % 2) || pw > xw || ph > xh)
        return MC_E_ARG;
    for (int q = 0; q < P; q++) {
        const int x0 = starts_host[2 * q], y0 = starts_host[2 * q + 1];
        if (x0 < 0 || y0 < 0 || x0 + pw > xw || y0 + ph > xh)
            return fail(c, MC_E_ARG, "patch window outside the xcf frame");
    }
    Fft2D* pp = plan(c, pw, ph);
    if (!pp) return fail(c, MC_E_ARG, "unsupported patch FFT size");
    MC_BUF(Tp, cfloat, "api_Tp", (size_t)P * N * pp->t_elems);
    MC_BUF(dst, int, "api_starts", 2 * P);
    MC_TRY(hipMemcpyAsync(dst, starts_host, sizeof(int) * 2 * P, hipMemcpyHostToDevice, c->stream));
    // Align/CExtractPatch.cpp:27-106 fused into the row pass (transform
    // q * N + i reads frame i through window q), then the unnormalised R2C
    // of Align/CMeasurePatches.cpp:54-55
    MC_TRY(mc_rows_r2c_mod(*pp, xcf_frames, xw, (long long)xw * xh, N, dst, Tp, pp->t_elems, P * N, 1.0f,
                           c->stream));
    MC_TRY(mc_cols_fwd(*pp, Tp, pp->t_elems, (cfloat*)out_spec, pp->s_elems, nullptr, 0, 0, 0, P * N,
                       c->stream));
    return MC_OK;
}

int mc_finalize_sums(mc_ctx* c, const float* sum_specs, int W, int H, int count, float ft_bin, float* out) {
    if (!c || !sum_specs || !out || count < 1 || W < 8 || H < 8 || (W % 2) || !(ft_bin >= 1.0f))
        return MC_E_ARG;
    const Geo g = make_geo(W, H, 1, 0, 0, ft_bin);
    Fft2D* po = plan(c, g.ow, g.oh);
    if (!po) return fail(c, MC_E_ARG, "unsupported output FFT size");
    MC_BUF(To, cfloat, "api_To", (size_t)po->t_elems);
    const long long fs = (long long)g.cx * H;
    for (int i = 0; i < count; i++) {
        MC_TRY(mc_cols_inv(*po, (const cfloat*)sum_specs + i * fs, fs, g.cx, H, nullptr, To, po->t_elems, 1,
                           c->stream));
        MC_TRY(mc_rows_c2r(*po, To, po->t_elems, out + (long long)i * g.ow * g.oh, g.ow,
                           (long long)g.ow * g.oh, 1, 1, c->stream));
    }
    return MC_OK;
}

int mc_set_kernel_timing(mc_ctx* c, const char* kernel) {
    if (!c) return MC_E_ARG;
    c->timed = kernel ? kernel : "";
    c->nev = 0;
    return MC_OK;
}

int mc_get_kernel_timing(mc_ctx* c, double* total_ms, int* launches) {
    if (!c || !total_ms || !launches) return MC_E_ARG;
    MC_TRY(hipStreamSynchronize(c->stream));
    double t = 0;
    for (size_t i = 0; i < c->nev; i++) {
        float ms = 0;
        MC_TRY(hipEventElapsedTime(&ms, c->evs[i].a, c->evs[i].b));
        t += ms;
    }
    *total_ms = t;
    *launches = (int)c->nev;
    c->nev = 0;
    return MC_OK;
}

int mc_synth_render(mc_ctx* c, float* frames, int W, int H, int N, long long pitch,
                    long long fstride, int nblobs, const float* blobs, const float* disp,
                    float scale, unsigned long long seed, int poisson) {
    if (!c || !frames || !blobs || !disp || W < 2 || H < 2 || N < 1 || nblobs < 1 || pitch < W)
        return MC_E_ARG;
    const int tile = 32, tx = (W + tile - 1) / tile, ty = (H + tile - 1) / tile;
    std::vector<std::vector<int>> lists((size_t)tx * ty);
    for (int k = 0; k < nblobs; k++) {
        float dmax = 0;
        for (int f = 0; f < N; f++) {
            const float dx = disp[((size_t)f * nblobs + k) * 2], dy = disp[((size_t)f * nblobs + k) * 2 + 1];
            dmax = std::max(dmax, std::sqrt(dx * dx + dy * dy));
        }
        const float r = 4.0f * blobs[4 * k + 2] + dmax + 1.0f;
        const int x0 = (int)std::floor((blobs[4 * k] - r) / tile), x1 = (int)std::floor((blobs[4 * k] + r) / tile);
        const int y0 = (int)std::floor((blobs[4 * k + 1] - r) / tile), y1 = (int)std::floor((blobs[4 * k + 1] + r) / tile);
        for (int yy = y0; yy <= y1; yy++)
            for (int xx = x0; xx <= x1; xx++) {
                const int wx = ((xx % tx) + tx) % tx, wy = ((yy % ty) + ty) % ty;
                std::vector<int>& L = lists[(size_t)wy * tx + wx];
                if (L.empty() || L.back() != k) L.push_back(k);
            }
    }
    std::vector<int> off(1, 0), idx;
    for (auto& L : lists) {
        idx.insert(idx.end(), L.begin(), L.end());
        off.push_back((int)idx.size());
    }
    if (idx.empty()) idx.push_back(0);
    MC_BUF(db, float, "syn_b", 4 * (size_t)nblobs);
    MC_BUF(dd, float, "syn_d", 2 * (size_t)nblobs * N);
    MC_BUF(doff, int, "syn_o", off.size());
    MC_BUF(didx, int, "syn_i", idx.size());
    MC_TRY(hipMemcpyAsync(db, blobs, sizeof(float) * 4 * nblobs, hipMemcpyHostToDevice, c->stream));
    MC_TRY(hipMemcpyAsync(dd, disp, sizeof(float) * 2 * (size_t)nblobs * N, hipMemcpyHostToDevice, c->stream));
    MC_TRY(hipMemcpyAsync(doff, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, c->stream));
    MC_TRY(hipMemcpyAsync(didx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, c->stream));
    MC_TRY(mc_launch_synth(frames, W, H, N, pitch, fstride, db, dd, nblobs, doff, didx, tile, tx,
                           scale, seed, poisson, c->stream));
    MC_TRY(hipStreamSynchronize(c->stream));
    return MC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Learned shift prior (no reference counterpart; mc_prior.hip).
int mc_prior_create(mc_ctx* c, int frames, int hidden, unsigned long long seed, mc_prior** out) {
    if (!c || !out) return MC_E_ARG;
    *out = nullptr;
    if (frames < 2 || hidden < 1 || hidden > 4096) return fail(c, MC_E_ARG, "prior: frames >= 2, 1 <= hidden <= 4096");
    MC_TRY(hipSetDevice(c->device));
    mc_prior* pr = new mc_prior;
    pr->frames = frames;
    pr->hidden = hidden;
    pr->device = c->device;
    if (pr->net.init(68 * frames + 2, hidden, 2 * frames, seed) != 0) {
        delete pr;
        return fail(c, MC_E_NOMEM, "prior parameter allocation failed");
    }
    *out = pr;
    return MC_OK;
}

int mc_prior_destroy(mc_ctx* c, mc_prior* pr) {
    if (!c || !pr) return MC_E_ARG;
    (void)hipStreamSynchronize(c->stream);
    if (c->prior == pr) c->prior = nullptr;
    delete pr;
    return MC_OK;
}

int mc_prior_dims(const mc_prior* pr, int* in_dim, int* out_dim, long long* n_params) {
    if (!pr) return MC_E_ARG;
    if (in_dim) *in_dim = pr->net.dims[0];
    if (out_dim) *out_dim = pr->net.dims[Prior::kLayers];
    if (n_params) *n_params = pr->net.nparam;
    return MC_OK;
}

int mc_prior_params(mc_ctx* c, mc_prior* pr, float* get_host, const float* set_host) {
    if (!c || !pr) return MC_E_ARG;
    const size_t b = sizeof(float) * pr->net.nparam;
    if (get_host) {
        MC_TRY(hipMemcpyAsync(get_host, pr->net.w, b, hipMemcpyDeviceToHost, c->stream));
        MC_TRY(hipStreamSynchronize(c->stream));
    }
    if (set_host) {
        MC_TRY(hipMemcpyAsync(pr->net.w, set_host, b, hipMemcpyHostToDevice, c->stream));
        MC_TRY(pr->net.refresh16(c->stream));
        MC_TRY(pr->net.reset_moments(c->stream));
        MC_TRY(hipStreamSynchronize(c->stream));
    }
    return MC_OK;
}

int mc_prior_forward(mc_ctx* c, mc_prior* pr, const float* x, int n, float* y) {
    if (!c || !pr || !x || !y || n < 1) return MC_E_ARG;
    if (!pr->net.reserve(n)) return fail(c, MC_E_NOMEM, "prior work buffers");
    MC_TRY(pr->net.forward(x, n, y, c->stream));
    return MC_OK;
}

int mc_prior_loss_grad(mc_ctx* c, mc_prior* pr, const float* x, const float* t, int n, double* loss,
                       float* grad_host) {
    if (!c || !pr || !x || !t || n < 1) return MC_E_ARG;
    Prior& net = pr->net;
    if (!net.reserve(n)) return fail(c, MC_E_NOMEM, "prior work buffers");
    float* y = (float*)net.scratch("yl", sizeof(float) * (size_t)n * net.dims[Prior::kLayers]);
    double* acc = (double*)net.work["acc"];
    if (!y || !acc) return fail(c, MC_E_NOMEM, "prior work buffers");
    hipStream_t s = c->stream;
    MC_TRY(net.forward(x, n, y, s));
    MC_TRY(net.backward(y, t, n, s));
    MC_TRY(hipMemsetAsync(acc, 0, sizeof(double), s));
    MC_TRY(net.loss_sum(n, acc, s));
    double h = 0.0;
    MC_TRY(hipMemcpyAsync(&h, acc, sizeof(double), hipMemcpyDeviceToHost, s));
    if (grad_host)
        MC_TRY(hipMemcpyAsync(grad_host, net.g, sizeof(float) * net.nparam, hipMemcpyDeviceToHost, s));
    MC_TRY(hipStreamSynchronize(s));
    if (loss) *loss = h / ((double)n * net.dims[Prior::kLayers]);
    return MC_OK;
}

int mc_prior_train(mc_ctx* c, mc_prior* pr, const float* x, const float* t, int n, int epochs,
                   int batch, float lr, unsigned long long seed, float* epoch_loss) {
    if (!c || !pr || !x || !t || n < 1 || epochs < 1 || batch < 1 || !(lr > 0)) return MC_E_ARG;
    MC_TRY(pr->net.train(x, t, n, epochs, batch, lr, seed, epoch_loss, c->stream));
    return MC_OK;
}

int mc_set_prior(mc_ctx* c, mc_prior* pr) {
    if (!c) return MC_E_ARG;
    if (pr && pr->device != c->device) return fail(c, MC_E_ARG, "prior lives on another device");
    c->prior = pr;
    return MC_OK;
}

int mc_gemm_bf16(mc_ctx* c, const unsigned short* a, long long sai, long long sak,
                 const unsigned short* b, long long sbk, long long sbj, float* out, int ldc, int m,
                 int n, int k) {
    if (!c || !a || !b || !out || m < 1 || n < 1 || k < 1 || ldc < n) return MC_E_ARG;
    MC_TRY(prior_gemm(a, sai, sak, b, sbk, sbj, out, ldc, m, n, k, c->stream));
    return MC_OK;
}

// ---------------------------------------------------------------------------
// FFT tile autotuner (no reference counterpart: cuFFT plans its own kernels,
// Util/CCufft2D.cpp:23-53).  Every register-pass row and column kernel has a
// few compiled tile variants (rows or columns per workgroup, threads); all
// of them compute each row or column with the same arithmetic, so results
// are bit-identical whichever runs.  mc_autotune times them on the batch
// shapes the alignment uses and keeps the fastest in the plan.
namespace {

// Median of `reps` timed runs of fn() on the context stream, in ms.
template <class F>
int time_median(mc_ctx* c, F fn, int reps, double* ms) {
    hipEvent_t a, b;
    MC_TRY(hipEventCreate(&a));
    MC_TRY(hipEventCreate(&b));
    std::vector<double> t;
    hipError_t e = fn();  // warm-up
    for (int r = 0; r < reps && e == hipSuccess; r++) {
        e = hipEventRecord(a, c->stream);
        if (e == hipSuccess) e = fn();
        if (e == hipSuccess) e = hipEventRecord(b, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float f = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&f, a, b);
        t.push_back(f);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return fail(c, MC_E_HIP, std::string("autotune timing: ") + hipGetErrorString(e));
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return MC_OK;
}

// Tune one plan's rows and columns for `batch` transforms.
// Time the plan's row and column variants on `batch` transforms and keep
// the fastest; variants of another radix schedule only when `radix`.
int tune_plan(mc_ctx* c, Fft2D* p, int batch, std::string* rep, bool radix) {
    const long long te = p->t_elems, se = p->s_elems;
    float* real = nullptr;
    cfloat* T = nullptr;
    cfloat* S = nullptr;
    const size_t rb = sizeof(float) * (size_t)batch * p->W * p->H;
    if (hipMalloc(&real, rb) != hipSuccess || hipMalloc(&T, sizeof(cfloat) * (size_t)batch * te) != hipSuccess ||
        hipMalloc(&S, sizeof(cfloat) * (size_t)batch * se) != hipSuccess) {
        (void)hipFree(real);
        (void)hipFree(T);
        (void)hipFree(S);
        return fail(c, MC_E_NOMEM, "autotune buffers");
    }
    int rc = MC_OK;
    hipStream_t s = c->stream;
    if (hipMemsetAsync(real, 0, rb, s) != hipSuccess ||
        hipMemsetAsync(T, 0, sizeof(cfloat) * (size_t)batch * te, s) != hipSuccess)
        rc = fail(c, MC_E_HIP, "autotune buffer clear");
    char line[256];
    for (int cols = 0; cols < 2 && rc == MC_OK; cols++) {
        std::vector<std::string> names;
        std::vector<int> scheds;
        const int nv = mc_fft_variants(*p, cols, &names, &scheds);
        if (nv < 2) continue;
        // forward (R2C rows / forward columns) and inverse passes separately
        for (int inv = 0; inv < 2 && rc == MC_OK; inv++) {
            int& var = cols ? (inv ? p->col_var_inv : p->col_var) : (inv ? p->row_var_inv : p->row_var);
            int best = 0;
            double best_ms = 1e30;
            for (int v = 0; v < nv && rc == MC_OK; v++) {
                if (scheds[v] != 0 && !radix) continue;
                var = v;
                double ms = 0;
                rc = time_median(c, [&]() -> hipError_t {
                    if (cols)
                        return inv ? mc_cols_inv(*p, S, se, p->cx, p->H, nullptr, T, te, batch, s)
                                   : mc_cols_fwd(*p, T, te, S, se, nullptr, 0, 0, 0, batch, s);
                    return inv ? mc_rows_c2r(*p, T, te, real, p->W, (long long)p->W * p->H, batch, 0, s)
                               : mc_rows_r2c(*p, real, p->W, (long long)p->W * p->H, nullptr, T, te, batch,
                                             1.0f, s);
                }, 5, &ms);
                snprintf(line, sizeof(line), "%dx%d x%d %s: %.4f ms (%s)\n", p->W, p->H, batch,
                         names[v].c_str(), ms, inv ? "inverse" : "forward");
                *rep += line;
                if (ms < best_ms) {
                    best_ms = ms;
                    best = v;
                }
            }
            var = best;
            snprintf(line, sizeof(line), "%dx%d %s%s -> %s\n", p->W, p->H, cols ? "columns" : "rows",
                     inv ? " (inverse)" : "", names[best].c_str());
            *rep += line;
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(real);
    (void)hipFree(T);
    (void)hipFree(S);
    return rc;
}

}  // namespace

int mc_autotune_schedules(mc_ctx* c, int W, int H, int N, int px, int py, float ft_bin, int radix,
                          char* report, int report_len) {
    if (!c || W < 8 || H < 8 || (W % 2) || N < 1) return c ? fail(c, MC_E_ARG, "autotune: bad geometry") : MC_E_ARG;
    MC_TRY(hipSetDevice(c->device));
    const bool patch = px > 1 && py > 1;
    const Geo g = make_geo(W, H, N, px, py, ft_bin > 0 ? ft_bin : 1.0f);
    // the batches the alignment launches: frame chunks of full frames (and
    // output sums), N xcf frames, P * N patch frames
    struct Item { int w, h, batch; };
    std::vector<Item> items = {{W, H, std::min(N, MC_CHUNK_FRAMES)}, {g.xw, g.xh, N}};
    if (patch) items.push_back({g.pw, g.ph, px * py * N});
    if (g.ow != W || g.oh != H) items.push_back({g.ow, g.oh, 2});
    std::string rep;
    std::vector<Fft2D*> done;
    for (const Item& it : items) {
        Fft2D* p = plan(c, it.w, it.h);
        if (!p) return fail(c, MC_E_ARG, "autotune: unsupported FFT size");
        if (std::find(done.begin(), done.end(), p) != done.end()) continue;
        done.push_back(p);
        MC_CHECK(tune_plan(c, p, it.batch, &rep, radix != 0));
    }
    // captured iteration graphs bake in launch shapes: drop them
    for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
    c->graphs.clear();
    if (report && report_len > 0) {
        const size_t n = std::min((size_t)report_len - 1, rep.size());
        std::memcpy(report, rep.data(), n);
        report[n] = 0;
    }
    return MC_OK;
}

int mc_autotune(mc_ctx* c, int W, int H, int N, int px, int py, float ft_bin, char* report,
                int report_len) {
    return mc_autotune_schedules(c, W, H, N, px, py, ft_bin, 0, report, report_len);
}
int mc_fft_variant_name(mc_ctx* c, int W, int H, int cols, int v, char* name, int len) {
    if (!c || !name || len < 1) return MC_E_ARG;
    Fft2D* p = plan(c, W, H);
    if (!p) return fail(c, MC_E_ARG, "unsupported FFT size");
    std::vector<std::string> names;
    const int n = mc_fft_variants(*p, cols ? 1 : 0, &names);
    if (v < 0 || v >= n) return fail(c, MC_E_ARG, "no such FFT tile variant");
    const size_t k = std::min((size_t)len - 1, names[v].size());
    std::memcpy(name, names[v].data(), k);
    name[k] = 0;
    return MC_OK;
}
int mc_fft_variant(mc_ctx* c, int W, int H, int cols, int set, int* count, int* current) {
    if (!c) return MC_E_ARG;
    Fft2D* p = plan(c, W, H);
    if (!p) return fail(c, MC_E_ARG, "unsupported FFT size");
    const int n = mc_fft_variants(*p, cols ? 1 : 0, nullptr);
    if (set >= 0) {
        if (set >= std::max(n, 1)) return fail(c, MC_E_ARG, "no such FFT tile variant");
        MC_TRY(hipStreamSynchronize(c->stream));
        (cols ? p->col_var : p->row_var) = set;
        (cols ? p->col_var_inv : p->row_var_inv) = set;
    }
    if (count) *count = n;
    if (current) *current = cols ? p->col_var : p->row_var;
    return MC_OK;
}

// ---------------------------------------------------------------------------
// Ingest: raw frames to the corrected FP32 frame buffer (MrcUtil/CApplyRefs.cpp,
// MrcUtil/GApplyRefsToFrame.cu), gain/dark post-processing (CLoadRefs.cpp:204-
// 318) and a minimal MRC reader (LibSrc/Mrcfile header layout).
namespace {

// In-place 90-degree rotation by k (G90Rotate2D, MrcUtil/G90Rotate2D.cu:17-
// 33,80-91): nearest source pixel of the rotated grid about the image centre.
void ref_rotate(std::vector<float>& img, int& w, int& h, int k) {
    if (k == 0) return;
    const double ang = std::atan(1.0) * 4 / 180.0 * 90.0 * k;
    const float fc = (float)std::cos(ang), fs = (float)std::sin(ang);
    const int ic = fc > 0 ? (int)(fc + 0.1f) : (int)(fc - 0.1f);
    const int is = fs > 0 ? (int)(fs + 0.1f) : (int)(fs - 0.1f);
    const int rw = (k % 2 == 0) ? w : h, rh = (k % 2 == 0) ? h : w;
    std::vector<float> out((size_t)rw * rh);
    for (int y = 0; y < rh; y++)
        for (int x = 0; x < rw; x++) {
            const float nx = x - (rw - 1) * 0.5f, ny = y - (rh - 1) * 0.5f;
            const float fx = nx * ic + ny * is, fy = -nx * is + ny * ic;
            const int sx = (int)(fx + (w - 1) * 0.5f + 0.5f), sy = (int)(fy + (h - 1) * 0.5f + 0.5f);
            out[(size_t)y * rw + x] = img[(size_t)sy * w + sx];
        }
    img.swap(out);
    w = rw;
    h = rh;
}

// 1: upside down (rows reversed), 2: left-right (GFlip2D, MrcUtil/GFlip2D.cu).
void ref_flip(std::vector<float>& img, int w, int h, int f) {
    if (f == 1)
        for (int y = 0; y < h / 2; y++)
            std::swap_ranges(img.begin() + (size_t)y * w, img.begin() + (size_t)(y + 1) * w,
                             img.begin() + (size_t)(h - 1 - y) * w);
    else if (f == 2)
        for (int y = 0; y < h; y++) std::reverse(img.begin() + (size_t)y * w, img.begin() + (size_t)(y + 1) * w);
}

}  // namespace

int mc_host_alloc(mc_ctx* c, unsigned long long bytes, void** out) {
    if (!c || !out || bytes == 0) return MC_E_ARG;
    *out = nullptr;
    MC_TRY(hipSetDevice(c->device));
    if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return fail(c, MC_E_NOMEM, "pinned host allocation failed");
    }
    return MC_OK;
}

int mc_host_free(mc_ctx* c, void* p) {
    if (!c) return MC_E_ARG;
    if (p) MC_TRY(hipHostFree(p));
    return MC_OK;
}

int mc_set_refs(mc_ctx* c, const float* gain, const float* dark, int rw, int rh, int rot, int flip,
                int inv, int quirks) {
    if (!c) return MC_E_ARG;
    if ((gain || dark) && (rw < 1 || rh < 1)) return fail(c, MC_E_ARG, "reference size");
    if (rot < 0 || rot > 3 || flip < 0 || flip > 2) return fail(c, MC_E_ARG, "-RotGain 0..3, -FlipGain 0..2");
    MC_TRY(hipSetDevice(c->device));
    if (c->s_in) MC_TRY(hipStreamSynchronize(c->s_in));
    MC_TRY(hipStreamSynchronize(c->stream));
    for (float** p : {&c->gain, &c->dark, &c->gain_aug, &c->dark_aug})
        if (*p) {
            MC_TRY(hipFree(*p));
            *p = nullptr;
        }
    c->ref_w = c->ref_h = 0;
    c->aug_fact = 0;
    if (!gain && !dark) return MC_OK;
    const size_t n = (size_t)rw * rh;
    std::vector<float> g, d;
    int gw = rw, gh = rh, dw = rw, dh = rh;
    if (gain) g.assign(gain, gain + n);
    if (dark) d.assign(dark, dark + n);
    // CLoadRefs::PostProcess (CLoadRefs.cpp:204-228): nothing without a gain
    // or without any operation; the gain is rotated, flipped, inverted; the
    // dark is rotated and flipped when (rot != 0 || flip == 0) -- the
    // reference's condition, kept with quirks -- or when any of the two is
    // set otherwise
    if (gain && (rot || flip || inv)) {
        ref_rotate(g, gw, gh, rot);
        ref_flip(g, gw, gh, flip);
        if (inv)
            for (float& v : g) v = (v == 0.f) ? 0.f : 1.0f / v;
        if (dark && (quirks ? (rot != 0 || flip == 0) : (rot != 0 || flip != 0))) {
            ref_rotate(d, dw, dh, rot);
            ref_flip(d, dw, dh, flip);
        }
    }
    if (gain && dark && (gw != dw || gh != dh)) return fail(c, MC_E_ARG, "gain and dark sizes differ");
    const int fw = gain ? gw : dw, fh = gain ? gh : dh;
    if (gain) {
        MC_TRY(hipMalloc(&c->gain, sizeof(float) * n));
        MC_TRY(hipMemcpy(c->gain, g.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    }
    if (dark) {
        MC_TRY(hipMalloc(&c->dark, sizeof(float) * n));
        MC_TRY(hipMemcpy(c->dark, d.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    }
    c->ref_w = fw;
    c->ref_h = fh;
    return MC_OK;
}

int mc_get_refs(mc_ctx* c, float* gain, float* dark, int* rw, int* rh) {
    if (!c) return MC_E_ARG;
    if (rw) *rw = c->ref_w;
    if (rh) *rh = c->ref_h;
    const size_t n = (size_t)c->ref_w * c->ref_h;
    if (gain && c->gain) MC_TRY(hipMemcpy(gain, c->gain, sizeof(float) * n, hipMemcpyDeviceToHost));
    if (dark && c->dark) MC_TRY(hipMemcpy(dark, c->dark, sizeof(float) * n, hipMemcpyDeviceToHost));
    return MC_OK;
}

namespace {

// The gain and dark references for sx x sy stack frames: the references
// themselves at their own size; for frames an integer factor f >= 2 larger
// in both directions, copies with every reference pixel repeated f x f
// (CLoadRefs::AugmentRefs, CLoadRefs.cpp:231-266, MrcUtil/GAugmentRef.cu).
int refs_for(mc_ctx* c, long long sx, long long sy, const float** g, const float** d) {
    *g = c->gain;
    *d = c->dark;
    if ((!c->gain && !c->dark) || (sx == c->ref_w && sy == c->ref_h)) return MC_OK;
    const long long f = sx / c->ref_w;
    if (f < 2 || sx != f * c->ref_w || sy != f * c->ref_h)
        return fail(c, MC_E_ARG, "gain/dark reference size is not the frame size or an integer "
                                 "fraction of it (CLoadRefs::AugmentRefs)");
    if (c->aug_fact != f) {
        // earlier loads may still read the previous copies
        if (c->s_in) MC_TRY(hipStreamSynchronize(c->s_in));
        MC_TRY(hipStreamSynchronize(c->stream));
        for (float** p : {&c->gain_aug, &c->dark_aug})
            if (*p) {
                MC_TRY(hipFree(*p));
                *p = nullptr;
            }
        c->aug_fact = 0;
        const size_t n = (size_t)(sx * sy);
        if (c->gain) {
            MC_TRY(hipMalloc(&c->gain_aug, sizeof(float) * n));
            MC_TRY(mc_launch_augment_ref(c->gain, c->ref_w, c->ref_h, (int)f, c->gain_aug, c->stream));
        }
        if (c->dark) {
            MC_TRY(hipMalloc(&c->dark_aug, sizeof(float) * n));
            MC_TRY(mc_launch_augment_ref(c->dark, c->ref_w, c->ref_h, (int)f, c->dark_aug, c->stream));
        }
        MC_TRY(hipStreamSynchronize(c->stream));
        c->aug_fact = (int)f;
    }
    *g = c->gain_aug;
    *d = c->dark_aug;
    return MC_OK;
}

}  // namespace

int mc_load_frames(mc_ctx* c, const mc_raw_stack* r, float* frames, int W, int H, long long pitch,
                   long long frame_stride, int async) {
    if (!c || !r || !r->data || !frames) return MC_E_ARG;
    const long long rb = mc_raw_row_bytes(r->mode, r->sx);
    if (rb < 0) return fail(c, MC_E_ARG, "MRC mode must be 0, 1, 2, 5, 6 or 101");
    if (r->sx < 1 || r->sy < 1 || r->n < 1 || W < 1 || H < 1 || W > r->sx || H > r->sy ||
        pitch < W || frame_stride < pitch * H)
        return fail(c, MC_E_ARG, "invalid raw stack or frame geometry");
    MC_TRY(hipSetDevice(c->device));
    const float* rg = nullptr;
    const float* rdk = nullptr;
    MC_CHECK(refs_for(c, r->sx, r->sy, &rg, &rdk));
    if (async && !c->s_in) MC_TRY(hipStreamCreate(&c->s_in));
    hipStream_t s = async ? c->s_in : c->stream;
    if (async) {
        // the load runs after everything already queued on the context stream
        // (which may still read the destination frames)
        hipEvent_t e = nullptr;
        MC_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        MC_TRY(hipEventRecord(e, c->stream));
        MC_TRY(hipStreamWaitEvent(c->s_in, e, 0));
        c->free_ev.push_back(e);
    }
    const long long fb = rb * r->sy, bytes = fb * r->n;
    const bool integ = r->n_int > 0;
    if (integ && (!r->int_start || !r->int_size || !(r->mode == 0 || r->mode == 5 || r->mode == 101)))
        return fail(c, MC_E_ARG, "frame integration needs modes 0, 5 or 101 and start/size tables");
    int* dtab = nullptr;
    if (integ) {
        for (int k = 0; k < r->n_int; k++)
            if (r->int_size[k] < 1 || r->int_start[k] < 0 || r->int_start[k] + r->int_size[k] > r->n)
                return fail(c, MC_E_ARG, "frame integration table exceeds the raw frames");
        dtab = (int*)buf(c, async ? "int_tab_a" : "int_tab", sizeof(int) * 2 * (size_t)r->n_int);
        if (!dtab) return fail(c, MC_E_NOMEM, "integration table");
        // retire host tables whose copies have run
        std::vector<std::pair<hipEvent_t, std::vector<int>>> keep;
        for (auto& t : c->tab_keep) {
            if (hipEventQuery(t.first) == hipSuccess) (void)hipEventDestroy(t.first);
            else keep.push_back(std::move(t));
        }
        (void)hipGetLastError();
        c->tab_keep.swap(keep);
        std::vector<int> tab(2 * (size_t)r->n_int);
        for (int k = 0; k < r->n_int; k++) {
            tab[2 * k] = r->int_start[k];
            tab[2 * k + 1] = r->int_size[k];
        }
        hipEvent_t te = nullptr;
        MC_TRY(hipEventCreateWithFlags(&te, hipEventDisableTiming));
        c->tab_keep.emplace_back(te, std::move(tab));
        const std::vector<int>& kt = c->tab_keep.back().second;
        MC_TRY(hipMemcpyAsync(dtab, kt.data(), sizeof(int) * kt.size(), hipMemcpyHostToDevice, s));
        MC_TRY(hipEventRecord(te, s));
    }
    // device input is read in place; host input is copied by DMA (pinned
    // host memory from mc_host_alloc makes the copy asynchronous)
    const unsigned char* src = nullptr;
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, r->data) == hipSuccess && at.type == hipMemoryTypeDevice) {
        src = (const unsigned char*)r->data;
    } else {
        (void)hipGetLastError();  // a plain host pointer is not an error here
        // asynchronous loads (in order on the ingest stream) and synchronous
        // ones (context stream) stage through different buffers
        unsigned char* staging = (unsigned char*)buf(c, async ? "raw_in_a" : "raw_in", (size_t)bytes);
        if (!staging) return fail(c, MC_E_NOMEM, "raw frame staging");
        MC_TRY(hipMemcpyAsync(staging, r->data, (size_t)bytes, hipMemcpyHostToDevice, s));
        src = staging;
    }
    IngestArgs a;
    a.raw = src;
    a.tab = dtab;
    a.n_out = integ ? r->n_int : r->n;
    a.frame_bytes = fb;
    a.row_bytes = rb;
    a.sx = r->sx;
    a.sy = r->sy;
    a.n = r->n;
    a.x0 = (r->sx - W) / 2;
    a.y0 = (r->sy - H) / 2;
    a.W = W;
    a.H = H;
    a.gain = rg;
    a.dark = rdk;
    a.out = frames;
    a.pitch = pitch;
    a.frame_stride = frame_stride;
    {
        KTimer kt(c, "apply_refs", s);
        MC_TRY(mc_launch_apply_refs(a, r->mode, s));
    }
    if (async) {
        hipEvent_t e = nullptr;
        MC_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        MC_TRY(hipEventRecord(e, c->s_in));
        const char* lo = (const char*)frames;
        const int nout = integ ? r->n_int : r->n;
        c->loads.push_back({lo, lo + sizeof(float) * (size_t)(frame_stride * (nout - 1) + pitch * (H - 1) + W), e});
    }
    return MC_OK;
}

int mc_load_wait(mc_ctx* c) {
    if (!c) return MC_E_ARG;
    return wait_loads(c, nullptr, nullptr);
}

namespace {

unsigned int bswap32(unsigned int v) {
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

// Header words 1-4 (nx ny nz mode), 24 (nsymbt) and 10 (mz); a file written with the
// other byte order shows a negative or oversized value there
// (LibSrc/Mrcfile/Header/CLoadMainHeader.cpp:27-35).
int mrc_header(const char* path, int* w, bool* swapped) {
    FILE* f = fopen(path, "rb");
    if (!f) return MC_E_ARG;
    unsigned int hdr[256];
    const size_t got = fread(hdr, 4, 256, f);
    fclose(f);
    if (got != 256) return MC_E_ARG;
    int v[6] = {(int)hdr[0], (int)hdr[1], (int)hdr[2], (int)hdr[3], (int)hdr[23], (int)hdr[9]};
    const int lim = 32 * 1024;
    bool sw = v[3] < 0 || v[0] < 0 || v[1] < 0 || v[2] < 0 || v[0] > lim || v[1] > lim || v[2] > lim;
    if (sw)
        for (int i = 0; i < 6; i++) v[i] = (int)bswap32((unsigned int)v[i]);
    for (int i = 0; i < 6; i++) w[i] = v[i];
    *swapped = sw;
    if (w[0] < 1 || w[1] < 1 || w[2] < 1 || mc_raw_row_bytes(w[3], w[0]) < 0 || w[4] < 0) return MC_E_ARG;
    return MC_OK;
}

}  // namespace

int mc_mrc_info(const char* path, int* sx, int* sy, int* n, int* mode, long long* data_offset) {
    if (!path) return MC_E_ARG;
    int w[6];
    bool sw;
    if (mrc_header(path, w, &sw) != MC_OK) return MC_E_ARG;
    if (sx) *sx = w[0];
    if (sy) *sy = w[1];
    if (n) *n = w[2];
    if (mode) *mode = w[3];
    if (data_offset) *data_offset = 1024LL + w[4];
    return MC_OK;
}

int mc_mrc_stacks(const char* path, int stack_z, int* frames, int* stacks) {
    if (!path || !frames || !stacks) return MC_E_ARG;
    int w[6];
    bool sw;
    if (mrc_header(path, w, &sw) != MC_OK) return MC_E_ARG;
    // MrcUtil/CLoadStack.cpp:49-53,100-107: -StackZ, else the header's mz
    // when above 1, else every section; only -StackZ makes several movies
    int z = w[2];
    if (stack_z > 1) z = stack_z;
    else if (w[5] > 1) z = w[5];
    if (z > w[2]) return MC_E_ARG;
    *frames = z;
    *stacks = stack_z > 1 ? w[2] / stack_z : 1;
    return MC_OK;
}

int mc_mrc_read(const char* path, int first, int count, void* dst) {
    int w[6];
    bool sw;
    if (!path || !dst || mrc_header(path, w, &sw) != MC_OK) return MC_E_ARG;
    if (first < 0 || count < 1 || first + count > w[2]) return MC_E_ARG;
    const long long fb = mc_raw_row_bytes(w[3], w[0]) * w[1];
    const long long off = 1024LL + w[4];
    FILE* f = fopen(path, "rb");
    if (!f) return MC_E_ARG;
    int rc = MC_OK;
    const size_t want = (size_t)(fb * count);
    if (fseeko(f, (off_t)(off + fb * first), SEEK_SET) != 0 || fread(dst, 1, want, f) != want) rc = MC_E_ARG;
    fclose(f);
    if (rc == MC_OK && sw) {
        const int es = (w[3] == 1 || w[3] == 6) ? 2 : (w[3] == 2 ? 4 : 1);
        unsigned char* b = (unsigned char*)dst;
        if (es == 2)
            for (size_t i = 0; i + 1 < want; i += 2) std::swap(b[i], b[i + 1]);
        else if (es == 4)
            for (size_t i = 0; i + 3 < want; i += 4) {
                std::swap(b[i], b[i + 3]);
                std::swap(b[i + 1], b[i + 2]);
            }
    }
    return rc;
}

static int mc_tiff_info_impl(const char* path, int* sx, int* sy, int* n, int* mode, int* compression,
                 float* pix_size) {
    if (!path) return MC_E_ARG;
    TiffFile t;
    if (!tiff_open(path, t)) return MC_E_ARG;
    const TiffDir& d = t.dirs[0];
    if (sx) *sx = d.w;
    if (sy) *sy = d.h;
    if (n) *n = (int)t.dirs.size();
    if (mode) *mode = tiff_mode(d);
    if (compression) *compression = d.comp;
    // TiffUtil/CLoadTiffHeader.cpp:137-147
    if (pix_size)
        *pix_size = d.has_xres ? (float)(d.xres / 2.54e8) : (d.has_yres ? (float)(d.yres / 2.54e8) : 0.f);
    return MC_OK;
}

int mc_tiff_info(const char* path, int* sx, int* sy, int* n, int* mode, int* compression,
                 float* pix_size) {
    // malformed files end in an error code, never an exception through the C ABI
    try {
        return mc_tiff_info_impl(path, sx, sy, n, mode, compression, pix_size);
    } catch (...) {
        return MC_E_ARG;
    }
}

static int mc_tiff_read_impl(const char* path, int first, int count, void* dst) {
    if (!path || !dst || first < 0 || count < 1) return MC_E_ARG;
    TiffFile t;
    if (!tiff_open(path, t) || first + count > (int)t.dirs.size()) return MC_E_ARG;
    const TiffDir& d0 = t.dirs[0];
    const int mode = tiff_mode(d0);
    if (mode < 0) return MC_E_ARG;
    const size_t fb = (size_t)d0.w * d0.h * (d0.bits / 8);
    for (int k = 0; k < count; k++) {
        const TiffDir& d = t.dirs[first + k];
        if (d.w != d0.w || d.h != d0.h || tiff_mode(d) != mode) return MC_E_ARG;
        if (!tiff_decode(t, d, (unsigned char*)dst + fb * k)) return MC_E_ARG;
    }
    return MC_OK;
}

int mc_tiff_read(const char* path, int first, int count, void* dst) {
    // malformed files end in an error code, never an exception through the C ABI
    try {
        return mc_tiff_read_impl(path, first, count, dst);
    } catch (...) {
        return MC_E_ARG;
    }
}

static int mc_eer_info_impl(const char* path, int* cam_w, int* cam_h, int* n, int* bits, long long* bytes) {
    if (!path) return MC_E_ARG;
    TiffFile t;
    if (!tiff_open(path, t)) return MC_E_ARG;
    const TiffDir& d = t.dirs[0];
    // EerUtil/CLoadEerHeader.cpp:31-36,50-66
    const int b = d.comp == 65000 ? 8 : (d.comp == 65001 ? 7 : -1);
    if (b < 0 || d.w < 1 || d.h < 1) return MC_E_ARG;
    uint64_t tot = 0;
    for (const TiffDir& e : t.dirs) {
        uint64_t k = 0;
        if (!tiff_raw(t, e, nullptr, &k)) return MC_E_ARG;
        tot += k;
    }
    if (cam_w) *cam_w = d.w;
    if (cam_h) *cam_h = d.h;
    if (n) *n = (int)t.dirs.size();
    if (bits) *bits = b;
    if (bytes) *bytes = (long long)tot;
    return MC_OK;
}

int mc_eer_info(const char* path, int* cam_w, int* cam_h, int* n, int* bits, long long* bytes) {
    // malformed files end in an error code, never an exception through the C ABI
    try {
        return mc_eer_info_impl(path, cam_w, cam_h, n, bits, bytes);
    } catch (...) {
        return MC_E_ARG;
    }
}

static int mc_eer_read_impl(const char* path, unsigned char* data, long long cap, long long* offsets) {
    if (!path || !data || !offsets || cap < 0) return MC_E_ARG;
    TiffFile t;
    if (!tiff_open(path, t)) return MC_E_ARG;
    uint64_t tot = 0;
    for (size_t k = 0; k < t.dirs.size(); k++) {
        uint64_t nb = 0;
        if (!tiff_raw(t, t.dirs[k], nullptr, &nb) || tot + nb > (uint64_t)cap) return MC_E_ARG;
        offsets[k] = (long long)tot;
        if (!tiff_raw(t, t.dirs[k], data + tot, &nb)) return MC_E_ARG;
        tot += nb;
    }
    offsets[t.dirs.size()] = (long long)tot;
    return MC_OK;
}

int mc_eer_read(const char* path, unsigned char* data, long long cap, long long* offsets) {
    // malformed files end in an error code, never an exception through the C ABI
    try {
        return mc_eer_read_impl(path, data, cap, offsets);
    } catch (...) {
        return MC_E_ARG;
    }
}

int mc_load_eer(mc_ctx* c, const mc_eer_stack* e, float* frames, int W, int H, long long pitch,
                long long frame_stride, int async) {
    if (!c || !e || !e->data || !e->offsets || !frames) return MC_E_ARG;
    if (e->n < 1 || e->cam_w < 1 || e->cam_h < 1 || (e->bits != 7 && e->bits != 8) ||
        e->sampling < 1 || e->sampling > 3 || e->offsets[0] != 0)
        return fail(c, MC_E_ARG, "invalid EER stack");
    for (int k = 0; k < e->n; k++)
        if (e->offsets[k + 1] < e->offsets[k]) return fail(c, MC_E_ARG, "EER offsets must not decrease");
    const int fact = e->sampling == 1 ? 1 : (e->sampling == 2 ? 2 : 4);
    const long long fw = (long long)e->cam_w * fact, fh = (long long)e->cam_h * fact;
    if (fw * fh >= (1LL << 31)) return fail(c, MC_E_ARG, "rendered EER frame too large");
    if (W < 1 || H < 1 || W > fw || H > fh || pitch < W || frame_stride < pitch * H)
        return fail(c, MC_E_ARG, "invalid frame geometry");
    const int nout = e->n_int > 0 ? e->n_int : e->n;
    // offsets (two ints each), then (fraction, frame, first segment) per job,
    // then the job of every segment
    std::vector<int> blob(2 * (size_t)(e->n + 1));
    memcpy(blob.data(), e->offsets, sizeof(long long) * (size_t)(e->n + 1));
    std::vector<int> jobs, segjob;
    for (int o = 0; o < nout; o++) {
        const int st = e->n_int > 0 ? (e->int_start ? e->int_start[o] : -1) : o;
        const int sz = e->n_int > 0 ? (e->int_size ? e->int_size[o] : -1) : 1;
        if (st < 0 || sz < 1 || st + sz > e->n)
            return fail(c, MC_E_ARG, "EER integration table exceeds the fractions");
        for (int k = 0; k < sz; k++) {
            const long long nb = e->offsets[st + k + 1] - e->offsets[st + k];
            const long long ns = std::max<long long>(
                1, e->bits == 7 ? (8 * nb + MC_EER_SEG_BITS - 1) / MC_EER_SEG_BITS
                                : (nb + MC_EER_SEG_BYTES8 - 1) / MC_EER_SEG_BYTES8);
            if ((long long)segjob.size() + ns >= (1LL << 30))
                return fail(c, MC_E_ARG, "EER movie too large");
            const int jn = (int)(jobs.size() / 3);
            jobs.push_back(st + k);
            jobs.push_back(o);
            jobs.push_back((int)segjob.size());
            segjob.insert(segjob.end(), (size_t)ns, jn);
        }
    }
    const int njob = (int)(jobs.size() / 3), nseg = (int)segjob.size();
    const size_t jofs = blob.size();
    blob.insert(blob.end(), jobs.begin(), jobs.end());
    const size_t sofs = blob.size();
    blob.insert(blob.end(), segjob.begin(), segjob.end());
    MC_TRY(hipSetDevice(c->device));
    const float* rg = nullptr;
    const float* rdk = nullptr;
    MC_CHECK(refs_for(c, fw, fh, &rg, &rdk));
    if (async && !c->s_in) MC_TRY(hipStreamCreate(&c->s_in));
    hipStream_t s = async ? c->s_in : c->stream;
    if (async) {
        hipEvent_t ev = nullptr;
        MC_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        MC_TRY(hipEventRecord(ev, c->stream));
        MC_TRY(hipStreamWaitEvent(c->s_in, ev, 0));
        c->free_ev.push_back(ev);
    }
    const std::string sfx = async ? "_a" : "";
    const size_t nbytes = (size_t)e->offsets[e->n];
    unsigned char* din = (unsigned char*)buf(c, ("eer_in" + sfx).c_str(), nbytes + 4);
    int* dtab = (int*)buf(c, ("eer_tab" + sfx).c_str(), sizeof(int) * blob.size());
    unsigned int* cnt =
        (unsigned int*)buf(c, ("eer_counts" + sfx).c_str(), sizeof(unsigned int) * (size_t)(nout * fw * fh));
    int* dseg = (int*)buf(c, ("eer_seg" + sfx).c_str(), sizeof(int) * 2 * (size_t)nseg * (MC_EER_PHASES + 1));
    if (!din || !dtab || !cnt || !dseg) return fail(c, MC_E_NOMEM, "EER staging");
    // the host tables stay alive until their copy has run
    std::vector<std::pair<hipEvent_t, std::vector<int>>> keep;
    for (auto& t : c->tab_keep) {
        if (hipEventQuery(t.first) == hipSuccess) (void)hipEventDestroy(t.first);
        else keep.push_back(std::move(t));
    }
    (void)hipGetLastError();
    c->tab_keep.swap(keep);
    hipEvent_t te = nullptr;
    MC_TRY(hipEventCreateWithFlags(&te, hipEventDisableTiming));
    c->tab_keep.emplace_back(te, std::move(blob));
    const std::vector<int>& kt = c->tab_keep.back().second;
    if (nbytes) MC_TRY(hipMemcpyAsync(din, e->data, nbytes, hipMemcpyHostToDevice, s));
    MC_TRY(hipMemcpyAsync(dtab, kt.data(), sizeof(int) * kt.size(), hipMemcpyHostToDevice, s));
    MC_TRY(hipEventRecord(te, s));
    MC_TRY(mc_launch_zero_u32(cnt, nout * fw * fh, s));
    EerArgs ea;
    ea.data = din;
    ea.off = reinterpret_cast<const long long*>(dtab);
    ea.job = dtab + jofs;
    ea.njob = njob;
    ea.seg_job = dtab + sofs;
    ea.nseg = nseg;
    ea.tab = dseg;
    ea.start = dseg + 2 * (size_t)nseg * MC_EER_PHASES;
    ea.cam_w = e->cam_w;
    ea.cam_h = e->cam_h;
    ea.bits = e->bits;
    ea.sampling = e->sampling;
    ea.fw = (int)fw;
    ea.fh = (int)fh;
    ea.counts = cnt;
    {
        KTimer kt2(c, "eer_render", s);
        MC_TRY(mc_launch_eer_render(ea, s));
    }
    IngestArgs a;
    a.raw = reinterpret_cast<const unsigned char*>(cnt);
    a.tab = nullptr;
    a.n_out = nout;
    a.frame_bytes = 4 * fw * fh;
    a.row_bytes = 4 * fw;
    a.sx = (int)fw;
    a.sy = (int)fh;
    a.n = nout;
    a.x0 = (int)((fw - W) / 2);
    a.y0 = (int)((fh - H) / 2);
    a.W = W;
    a.H = H;
    a.gain = rg;
    a.dark = rdk;
    a.out = frames;
    a.pitch = pitch;
    a.frame_stride = frame_stride;
    {
        KTimer kt2(c, "apply_refs", s);
        MC_TRY(mc_launch_apply_refs(a, MC_MODE_EER_COUNTS, s));
    }
    if (async) {
        hipEvent_t ev = nullptr;
        MC_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        MC_TRY(hipEventRecord(ev, c->s_in));
        const char* lo = (const char*)frames;
        c->loads.push_back(
            {lo, lo + sizeof(float) * (size_t)(frame_stride * (nout - 1) + pitch * (H - 1) + W), ev});
    }
    return MC_OK;
}

// ---------------------------------------------------------------------------
// Output files (CSaveSerialCryoEM.cpp:126-240, Align/CSaveAlign.cpp:84-151,
// Align/CLoadAlign.cpp:108-250, CGenStarFile.cpp:39-120).
int mc_save_sums(mc_ctx* c, const mc_params* p, const mc_result* r, const char* out_mrc) {
    if (!c || !p || !r || !out_mrc || !out_mrc[0]) return c ? fail(c, MC_E_ARG, "save sums: arguments") : MC_E_ARG;
    // CAlnSums::Setup (DataUtil/CAlnSums.cpp:18-44) file suffixes, in the
    // order of the sums
    const bool dw = !(p->kv < 80 || p->kv > 300) && p->pix_size > 0 && p->fm_dose > 0;
    const bool dws = dw && !(p->sum_range[0] < 0) && !(p->sum_range[1] <= 0) &&
                     (p->sum_range[1] - p->sum_range[0] > 1.0f);
    std::vector<const char*> sfx = {""};
    if (dw && p->align != 0) sfx.push_back("_DW");
    if (dws && p->align != 0) sfx.push_back("_DWS");
    if (p->split_sum) {
        sfx.push_back("_ODD");
        sfx.push_back("_EVN");
    }
    if ((int)sfx.size() != r->num_sums) return fail(c, MC_E_ARG, "save sums: result does not match the flags");
    const int w = r->out_w, h = r->out_h;
    MC_BUF(pmin, float, "st_min", MC_STATS_BLOCKS);
    MC_BUF(pmax, float, "st_max", MC_STATS_BLOCKS);
    MC_BUF(psum, double, "st_sum", MC_STATS_BLOCKS);
    std::vector<float> hmin(MC_STATS_BLOCKS), hmax(MC_STATS_BLOCKS), img((size_t)w * h);
    std::vector<double> hsum(MC_STATS_BLOCKS);
    // GetFinalPixelSize (CInput.cpp:709-715) with unit magnification
    const float pix = p->pix_size * p->ft_bin;
    for (int i = 0; i < r->num_sums; i++) {
        if (!r->sums[i]) return fail(c, MC_E_ARG, "save sums: missing sum buffer");
        MC_TRY(mc_launch_image_stats(r->sums[i], w, h, w, pmin, pmax, psum, c->stream));
        MC_TRY(hipMemcpyAsync(hmin.data(), pmin, sizeof(float) * MC_STATS_BLOCKS, hipMemcpyDeviceToHost, c->stream));
        MC_TRY(hipMemcpyAsync(hmax.data(), pmax, sizeof(float) * MC_STATS_BLOCKS, hipMemcpyDeviceToHost, c->stream));
        MC_TRY(hipMemcpyAsync(hsum.data(), psum, sizeof(double) * MC_STATS_BLOCKS, hipMemcpyDeviceToHost, c->stream));
        MC_TRY(hipMemcpyAsync(img.data(), r->sums[i], sizeof(float) * img.size(), hipMemcpyDeviceToHost, c->stream));
        MC_TRY(hipStreamSynchronize(c->stream));
        float lo = hmin[0], hi = hmax[0];
        double tot = 0.0;
        for (int b = 0; b < MC_STATS_BLOCKS; b++) {
            lo = std::min(lo, hmin[b]);
            hi = std::max(hi, hmax[b]);
            tot += hsum[b];
        }
        const std::string path = mc_output_name(out_mrc, sfx[i], ".mrc");
        if (mc_write_mrc_float(path.c_str(), img.data(), w, h, pix, lo, hi, (float)(tot / ((double)w * h))) != 0)
            return fail(c, MC_E_ARG, "cannot write " + path);
    }
    return MC_OK;
}

int mc_save_stack(mc_ctx* c, const mc_params* p, const mc_result* r, int frames, const char* out_mrc) {
    if (!c || !p || !r || !out_mrc || !out_mrc[0]) return c ? fail(c, MC_E_ARG, "save stack: arguments") : MC_E_ARG;
    if (!r->stack || frames < 1 || r->out_w < 1 || r->out_h < 1)
        return fail(c, MC_E_ARG, "save stack: no stack in the result");
    const int w = r->out_w, h = r->out_h;
    const long long npx = (long long)w * h;
    MC_BUF(pmin, float, "st_min", MC_STATS_BLOCKS);
    MC_BUF(pmax, float, "st_max", MC_STATS_BLOCKS);
    MC_BUF(psum, double, "st_sum", MC_STATS_BLOCKS);
    std::vector<float> hmin(MC_STATS_BLOCKS), hmax(MC_STATS_BLOCKS);
    std::vector<double> hsum(MC_STATS_BLOCKS);
    // CSaveSerialCryoEM::mSaveStack (:177-204): extremes over the frames and
    // the mean of the frame means
    float lo = 1e20f, hi = -1e20f;
    double mean = 0.0;
    for (int i = 0; i < frames; i++) {
        MC_TRY(mc_launch_image_stats(r->stack + (long long)i * npx, w, h, w, pmin, pmax, psum, c->stream));
        MC_TRY(hipMemcpyAsync(hmin.data(), pmin, sizeof(float) * MC_STATS_BLOCKS, hipMemcpyDeviceToHost, c->stream));
        MC_TRY(hipMemcpyAsync(hmax.data(), pmax, sizeof(float) * MC_STATS_BLOCKS, hipMemcpyDeviceToHost, c->stream));
        MC_TRY(hipMemcpyAsync(hsum.data(), psum, sizeof(double) * MC_STATS_BLOCKS, hipMemcpyDeviceToHost, c->stream));
        MC_TRY(hipStreamSynchronize(c->stream));
        double tot = 0.0;
        for (int b = 0; b < MC_STATS_BLOCKS; b++) {
            lo = std::min(lo, hmin[b]);
            hi = std::max(hi, hmax[b]);
            tot += hsum[b];
        }
        mean += (float)(tot / (double)npx);
    }
    mean /= frames;
    // GetFinalPixelSize (CInput.cpp:709-715) with unit magnification
    const float pix = p->pix_size * p->ft_bin;
    unsigned char hdr[1024];
    mc_mrc_header(hdr, w, h, frames, pix, lo, hi, (float)mean, 32);
    std::vector<float> ext((size_t)32 * frames, 0.0f);
    for (int i = 0; i < frames; i++) ext[(size_t)32 * i + 11] = pix;  // CSaveExtHeader::SetPixelSize
    const std::string path = mc_output_name(out_mrc, "_Stk", ".mrc");
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return fail(c, MC_E_ARG, "cannot write " + path);
    bool ok = fwrite(hdr, 1, 1024, f) == 1024;
    ok = ok && fwrite(ext.data(), sizeof(float), ext.size(), f) == ext.size();
    std::vector<float> img((size_t)npx);
    for (int i = 0; ok && i < frames; i++) {
        if (hipMemcpyAsync(img.data(), r->stack + (long long)i * npx, sizeof(float) * npx, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            fclose(f);
            return fail(c, MC_E_HIP, "save stack: copy");
        }
        ok = fwrite(img.data(), sizeof(float), img.size(), f) == img.size();
    }
    ok = (fclose(f) == 0) && ok;
    return ok ? MC_OK : fail(c, MC_E_ARG, "cannot write " + path);
}

int mc_save_aln(const char* path, const mc_params* p, int sx, int sy, int n_stack, int N,
                const float* global_shift, const float* patch_shift, const float* patch_center) {
    if (!path || !p || !global_shift || N < 1) return MC_E_ARG;
    const bool patch = p->patch[0] > 1 && p->patch[1] > 1;
    const int P = patch ? p->patch[0] * p->patch[1] : 0;
    if (P && (!patch_shift || !patch_center)) return MC_E_ARG;
    return mc_write_aln(path, sx, sy, n_stack, p->patch[0], p->patch[1], p->throw0,
                        n_stack - N - p->throw0 > 0 ? n_stack - N - p->throw0 : 0, N, P, global_shift,
                        patch_shift, patch_center) == 0
               ? MC_OK
               : MC_E_ARG;
}

int mc_load_aln(const char* path, int* sx, int* sy, int* n, int* px, int* py, float* global_shift,
                float* patch_shift, float* patch_center) {
    if (!path) return MC_E_ARG;
    return mc_read_aln(path, sx, sy, n, px, py, global_shift, patch_shift, patch_center) == 0 ? MC_OK
                                                                                             : MC_E_ARG;
}

int mc_save_star(const char* path, const mc_params* p, int sx, int sy, int n_stack,
                 const char* gain_name, const float* global_shift, int N) {
    if (!path || !p || !global_shift || N < 1) return MC_E_ARG;
    return mc_write_star(path, p->movie_name, gain_name, sx, sy, n_stack, p->pix_size, p->ft_bin,
                         p->fm_dose, p->kv, p->throw0, global_shift, N) == 0
               ? MC_OK
               : MC_E_ARG;
}

// ---------------------------------------------------------------------------
// Frame integration (-FmIntFile): CReadFmIntFile::DoIt
// (DataUtil/CReadFmIntFile.cpp:110-173) and CFmIntParam::Setup / mSetup /
// mCalcIntFms / mThrowIntFrames (DataUtil/CFmIntParam.cpp:29-46,118-262).
int mc_fm_int_table(const char* path, int n_raw, float fm_dose, int throw0, int throw1, int cap,
                    int* n_out, int* start, int* size, float* acc_dose, float* raw_dose,
                    int* integrated) {
    if (n_raw < 1 || !n_out) return MC_E_ARG;
    struct Entry { int group, isz; float dose; };
    std::vector<Entry> ent;
    FILE* f = path && path[0] ? fopen(path, "rt") : nullptr;
    if (f) {
        char buf[256] = {0};
        int g = 0, isz = 0;
        float d = 0.f;
        if (fgets(buf, sizeof(buf), f)) {
            int items = sscanf(buf, "%d  %d  %f", &g, &isz, &d);
            bool ok = items >= 2;
            const bool two = items == 2;
            if (two) d = fm_dose;
            while (ok) {
                ent.push_back({g, isz, d});
                if (feof(f)) break;
                if (!fgets(buf, sizeof(buf), f)) break;
                if (two) {
                    if (sscanf(buf, "%d  %d", &g, &isz) != 2) continue;
                } else {
                    // the reference compares instead of assigning here
                    // (`iItems == sscanf(...)`), so a line that does not
                    // parse repeats the previous entry
                    (void)sscanf(buf, "%d  %d  %f", &g, &isz, &d);
                    if (d < 0) ok = false;
                }
            }
        }
        fclose(f);
        for (const Entry& e : ent)
            if (e.dose != ent[0].dose) {
                ent.clear();  // variable frame dose: no integration
                break;
            }
    }
    const bool integrate =
        !ent.empty() && std::fabs(ent.front().dose - ent.back().dose) < 0.0001f;
    std::vector<int> st, sz;
    std::vector<float> acc;
    float rd = fm_dose;
    if (!integrate) {
        const int n = n_raw - throw0 - throw1;
        if (n < 1) return MC_E_ARG;
        for (int i = 0; i < n; i++) {
            st.push_back(i + throw0);
            sz.push_back(1);
            acc.push_back(fm_dose * (i + 1));
        }
    } else {
        const int ne = (int)ent.size();
        std::vector<int> gs(ne), ni(ne);
        int count = 0;
        for (int i = 0; i < ne; i++) {
            gs[i] = ent[i].group;
            count += gs[i];
        }
        int extra = count - n_raw;
        if (extra > 0) {
            for (int i = ne - 1; i >= 0; i--) {
                const int g = gs[i];
                gs[i] -= extra;
                if (gs[i] <= 0) gs[i] = 0;
                else extra = extra - g;
                if (extra <= 0) break;
            }
        } else if (extra < 0) {
            gs[ne - 1] -= extra;
        }
        for (int i = 0; i < ne - 1; i++) {
            const int isz = std::max(ent[i].isz, 1);
            ni[i] = gs[i] / isz;
            gs[i + 1] += gs[i] % isz;
        }
        const int lsz = std::max(ent[ne - 1].isz, 1);
        extra = gs[ne - 1] % lsz;
        ni[ne - 1] = gs[ne - 1] / lsz;
        for (int i = 0; i < ne; i++)
            for (int k = 0; k < ni[i]; k++) sz.push_back(std::max(ent[i].isz, 1));
        const int nint = (int)sz.size();
        if (nint < 1) return MC_E_ARG;
        for (int i = 0; i < extra && i < nint; i++) sz[nint - 1 - i] += 1;
        st.assign(nint, 0);
        for (int i = 1; i < nint; i++) st[i] = st[i - 1] + sz[i - 1];
        rd = ent[0].dose;
        std::vector<float> idose(nint);
        for (int i = 0; i < nint; i++) idose[i] = sz[i] * rd;
        acc.assign(nint, 0.f);
        acc[0] = idose[0];
        for (int i = 1; i < nint; i++) acc[i] = acc[i - 1] + idose[i];
        // mThrowIntFrames: ignored when five or fewer frames would remain
        const int nt = throw0 + throw1;
        if (nt > 0 && nint - nt > 5) {
            const int keep = nint - nt;
            std::vector<int> st2(st.begin() + throw0, st.begin() + throw0 + keep);
            std::vector<int> sz2(sz.begin() + throw0, sz.begin() + throw0 + keep);
            std::vector<float> acc2(acc.begin() + throw0, acc.begin() + throw0 + keep);
            st.swap(st2);
            sz.swap(sz2);
            acc.swap(acc2);
        }
    }
    const int n = (int)st.size();
    *n_out = n;
    if (raw_dose) *raw_dose = rd;
    if (integrated) *integrated = integrate ? 1 : 0;
    if (start || size || acc_dose) {
        if (cap < n) return MC_E_ARG;
        for (int i = 0; i < n; i++) {
            if (start) start[i] = st[i];
            if (size) size[i] = sz[i];
            if (acc_dose) acc_dose[i] = acc[i];
        }
    }
    return MC_OK;
}

// ---------------------------------------------------------------------------
// Bad pixels (BadPixel/CDetectMain.cpp:54-245, BadPixel/CCorrectMain.cpp).
namespace {

// mean and E[x^2] of a W x H device image (FP64 partial sums, host total)
int image_moments(mc_ctx* c, const float* img, long long n, double* mean, double* e2) {
    MC_BUF(ps, double, "bp_ps", MC_BAD_BLOCKS);
    MC_BUF(pq, double, "bp_pq", MC_BAD_BLOCKS);
    MC_TRY(mc_launch_moments(img, n, ps, pq, c->stream));
    std::vector<double> a(MC_BAD_BLOCKS), b(MC_BAD_BLOCKS);
    MC_TRY(hipMemcpyAsync(a.data(), ps, sizeof(double) * MC_BAD_BLOCKS, hipMemcpyDeviceToHost, c->stream));
    MC_TRY(hipMemcpyAsync(b.data(), pq, sizeof(double) * MC_BAD_BLOCKS, hipMemcpyDeviceToHost, c->stream));
    MC_TRY(hipStreamSynchronize(c->stream));
    double s1 = 0, s2 = 0;
    for (int i = 0; i < MC_BAD_BLOCKS; i++) {
        s1 += a[i];
        s2 += b[i];
    }
    *mean = s1 / (double)n;
    *e2 = s2 / (double)n;
    return MC_OK;
}

}  // namespace

int mc_detect_bad(mc_ctx* c, const float* frames, int W, int H, int N, long long pitch,
                  long long frame_stride, float threshold, int defect_w, int defect_h,
                  const char* defect_file, const unsigned char* defect_map, unsigned char* bad,
                  int* n_bad, float* thresholds) {
    if (!c) return MC_E_ARG;
    if (!frames || !bad || W < 2 || H < 2 || N < 1 || pitch < W || frame_stride < pitch * H ||
        defect_w < 1 || defect_h < 1 || defect_w * defect_h > 4096)
        return fail(c, MC_E_ARG, "detect bad pixels: arguments");
    MC_TRY(hipSetDevice(c->device));
    MC_CHECK(wait_loads(c, nullptr, nullptr));
    hipStream_t s = c->stream;
    const long long n = (long long)W * H;
    MC_BUF(sum, float, "bp_sum", (size_t)n);
    MC_BUF(cc, float, "bp_cc", (size_t)n);
    MC_BUF(tpl, float, "bp_tpl", (size_t)defect_w * defect_h);
    // CSumFFTStack: the frame sum
    MC_TRY(mc_launch_frame_sum(frames, W, H, N, pitch, frame_stride, sum, s));
    // GDetectHot::DoIt (BadPixel/GDetectHot.cu:25-55), FP32 threshold
    double m = 0, e2 = 0;
    MC_CHECK(image_moments(c, sum, n, &m, &e2));
    const float fm = (float)m, fe2 = (float)e2;
    const float sd = sqrtf(std::max(fe2 - fm * fm, 0.0f));
    float delta = threshold * sd;
    if (delta < 10) delta = 10;
    const float hot = fm + delta;
    MC_TRY(mc_launch_hot(sum, n, hot, bad, s));
    // CTemplate::Create (BadPixel/CTemplate.cpp:19-57): border -1, inside +1,
    // normalised to zero mean and unit deviation
    const int ts = defect_w * defect_h;
    std::vector<float> t(ts);
    for (int i = 0; i < ts; i++) {
        const int x = i % defect_w, y = i / defect_w;
        t[i] = (x == 0 || x == defect_w - 1 || y == 0 || y == defect_h - 1) ? -1.0f : 1.0f;
    }
    double dm = 0, ds = 0;
    for (int i = 0; i < ts; i++) {
        dm += t[i];
        ds += t[i] * t[i];
    }
    const float tm = (float)(dm / ts);
    ds = ds / ts - tm * tm;
    float tsd = ds <= 0 ? 0.0f : (float)std::sqrt(ds);
    tsd += (float)1e-30;
    for (int i = 0; i < ts; i++) t[i] = (t[i] - tm) / tsd;
    MC_TRY(hipMemcpyAsync(tpl, t.data(), sizeof(float) * ts, hipMemcpyHostToDevice, s));
    // GLocalCC + GDetectPatch (BadPixel/GDetectPatch.cu:55-105)
    MC_TRY(mc_launch_local_cc(sum, W, H, tpl, defect_w, defect_h, cc, s));
    MC_CHECK(image_moments(c, cc, n, &m, &e2));
    const float cm = (float)m, ce2 = (float)e2;
    const float csd = sqrtf(std::max(0.0f, ce2 - cm * cm));
    const float cthr = cm + threshold * csd;
    MC_TRY(mc_launch_cc_mark(cc, W, H, defect_w, defect_h, cthr, bad, s));
    // defect rectangles "x y w h" and an 8-bit defect map (CDetectMain.cpp:146-232)
    std::vector<unsigned char> hm;
    const bool extra = (defect_file && defect_file[0]) || defect_map;
    if (extra || n_bad) {
        hm.resize((size_t)n);
        MC_TRY(hipMemcpyAsync(hm.data(), bad, (size_t)n, hipMemcpyDeviceToHost, s));
        MC_TRY(hipStreamSynchronize(s));
    }
    if (defect_file && defect_file[0]) {
        FILE* f = fopen(defect_file, "rt");
        if (!f) return fail(c, MC_E_ARG, std::string("cannot read ") + defect_file);
        char line[256];
        int e[4];
        while (fgets(line, sizeof(line), f)) {
            if (sscanf(line, "%d  %d  %d  %d", e, e + 1, e + 2, e + 3) != 4) continue;
            const int x0 = std::max(e[0], 0), y0 = std::max(e[1], 0);
            const int x1 = std::min(e[0] + e[2], W), y1 = std::min(e[1] + e[3], H);
            for (int y = y0; y < y1; y++)
                for (int x = x0; x < x1; x++) hm[(size_t)y * W + x] = 1;
        }
        fclose(f);
    }
    if (defect_map)
        for (long long i = 0; i < n; i++) hm[i] = (unsigned char)(hm[i] | (defect_map[i] & 1));
    if (extra) MC_TRY(hipMemcpyAsync(bad, hm.data(), (size_t)n, hipMemcpyHostToDevice, s));
    if (n_bad) {
        long long cnt = 0;
        for (long long i = 0; i < n; i++) cnt += hm[i];
        *n_bad = (int)cnt;
    }
    if (thresholds) {
        thresholds[0] = hot;
        thresholds[1] = cthr;
    }
    MC_TRY(hipStreamSynchronize(s));
    return MC_OK;
}

int mc_correct_bad(mc_ctx* c, float* frames, int W, int H, int N, long long pitch,
                   long long frame_stride, const unsigned char* bad, int win) {
    if (!c) return MC_E_ARG;
    if (!frames || !bad || W < 2 || H < 2 || N < 1 || pitch < W || frame_stride < pitch * H)
        return fail(c, MC_E_ARG, "correct bad pixels: arguments");
    MC_TRY(hipSetDevice(c->device));
    MC_CHECK(wait_loads(c, nullptr, nullptr));
    // GCorrectBad::SetWinSize: the window is at least 31 pixels
    MC_TRY(mc_launch_correct_bad(frames, W, H, N, pitch, frame_stride, bad, std::max(win, 31), c->stream));
    MC_TRY(hipStreamSynchronize(c->stream));
    return MC_OK;
}
