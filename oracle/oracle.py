"""ctypes loader for the FP64 oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It wraps oracle/liboracle_mc.so (built by oracle/Makefile) and
the optional oracle/_ref/libref_lineqs.so compiled from the reference's own
LibSrc/Util/Util_LinEqs.cpp.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_ubyte)
_fp = C.POINTER(C.c_float)


class Params(C.Structure):
    """orc_params: the CInput flag subset (CInput.cpp:82-128)."""
    _fields_ = [
        ("patch", C.c_int * 2), ("iter", C.c_int), ("tol", C.c_double),
        ("bft", C.c_double * 2), ("phase_only", C.c_int),
        ("fm_dose", C.c_double), ("pix_size", C.c_double), ("kv", C.c_int),
        ("group", C.c_int * 2), ("fm_ref", C.c_int), ("ft_bin", C.c_double),
        ("sum_range", C.c_double * 2), ("split_sum", C.c_int),
        ("throw0", C.c_int), ("ref_quirks", C.c_int),
        ("fm_int_start", _ip), ("fm_int_size", _ip), ("out_stack", C.c_int * 2),
        ("align", C.c_int), ("in_fm_motion", C.c_int), ("corr_interp", C.c_int),
    ]


class Override(C.Structure):
    _fields_ = [
        ("xcf_shift", _dp), ("patch_starts", _ip), ("replay_global", _dp),
        ("replay_patch", _dp), ("replay_centers", _dp), ("replay_bad", _up),
    ]


class Out(C.Structure):
    _fields_ = [
        ("out_w", C.c_int), ("out_h", C.c_int), ("num_sums", C.c_int),
        ("sums", _dp), ("global_shift", _dp), ("xcf_shift", _dp),
        ("global_iters", _ip), ("global_errors", _dp), ("patch_shift", _dp),
        ("patch_bad", _up), ("patch_center", _dp), ("patch_starts", _ip),
        ("patch_iters", _ip), ("feature_map", _up), ("stack", _dp), ("stack_frames", C.c_int),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_mc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _LIB = C.CDLL(path)
        _LIB.orc_align_movie.restype = C.c_int
        _LIB.orc_lin_eqs.restype = C.c_int
        _LIB.orc_num_threads.restype = C.c_int
    return _LIB


def ref_lib():
    """The reference's Util_LinEqs built from source, or None if absent."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_lineqs.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path)
        _REF.ref_lin_eqs.restype = C.c_int
    return _REF


def _d(a):
    return a.ctypes.data_as(_dp)


def default_params(**kw):
    p = Params()
    lib().orc_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("patch", "bft", "group", "sum_range", "out_stack"):
            arr = getattr(p, k)
            arr[0], arr[1] = v[0], v[1]
        elif k in ("fm_int_start", "fm_int_size"):
            a = np.ascontiguousarray(v, np.int32)
            p.__dict__.setdefault("_keep", []).append(a)
            setattr(p, k, a.ctypes.data_as(_ip))
        else:
            setattr(p, k, v)
    return p


def geometry(W, H, px=0, py=0):
    sizes = (C.c_int * 4)()
    xb = (C.c_double * 2)()
    lib().orc_geometry(W, H, px, py, sizes, xb)
    return dict(xcf=(sizes[0], sizes[1]), patch=(sizes[2], sizes[3]),
                xcf_bin=(xb[0], xb[1]))


def fft_c2c(x, sign=-1):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n = a.shape[-1]
    lib().orc_fft_c2c(a.ctypes.data_as(_dp), n, a.size // n, sign)
    return a


def rfft2(img):
    img = np.ascontiguousarray(img, dtype=np.float64)
    H, W = img.shape
    out = np.zeros((H, W // 2 + 1), np.complex128)
    lib().orc_rfft2(_d(img), W, H, out.ctypes.data_as(_dp))
    return out


def irfft2(cmp, W):
    cmp = np.ascontiguousarray(cmp, dtype=np.complex128)
    H = cmp.shape[0]
    out = np.zeros((H, W), np.float64)
    lib().orc_irfft2(cmp.ctypes.data_as(_dp), W, H, _d(out))
    return out


def phase_shift(cmp, sx, sy):
    cmp = np.ascontiguousarray(cmp, dtype=np.complex128)
    out = np.zeros_like(cmp)
    lib().orc_phase_shift(cmp.ctypes.data_as(_dp), cmp.shape[1], cmp.shape[0],
                          C.c_double(sx), C.c_double(sy), 0, out.ctypes.data_as(_dp))
    return out


def fourier_resize(cmp, outx, outy):
    cmp = np.ascontiguousarray(cmp, dtype=np.complex128)
    out = np.zeros((outy, outx), np.complex128)
    lib().orc_fourier_resize(cmp.ctypes.data_as(_dp), cmp.shape[1], cmp.shape[0],
                             out.ctypes.data_as(_dp), outx, outy, 0)
    return out


def correlate_window(sum_cmp, grp_cmp, bft, phase_only=False, sea=(64, 64)):
    s = np.ascontiguousarray(sum_cmp, dtype=np.complex128)
    g = np.ascontiguousarray(grp_cmp, dtype=np.complex128)
    win = np.zeros((sea[1], sea[0]), np.float64)
    lib().orc_correlate_window(s.ctypes.data_as(_dp), g.ctypes.data_as(_dp),
                               s.shape[1], s.shape[0], C.c_double(bft),
                               int(phase_only), sea[0], sea[1], _d(win))
    return win


def peak2d(img):
    img = np.ascontiguousarray(img, dtype=np.float64)
    sh = np.zeros(2)
    lib().orc_peak2d(_d(img), img.shape[1], img.shape[0], _d(sh))
    return sh


def lin_eqs(a, v):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    v = np.ascontiguousarray(v, dtype=np.float64).copy()
    ok = lib().orc_lin_eqs(_d(a), _d(v), len(v))
    return bool(ok), v


def ref_lin_eqs(a, v):
    r = ref_lib()
    if r is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32).copy()
    v = np.ascontiguousarray(v, dtype=np.float32).copy()
    ok = r.ref_lin_eqs(a.ctypes.data_as(_fp), v.ctypes.data_as(_fp), len(v))
    return bool(ok), v


def remove_spikes(sx, sy, single=False, quirks=True):
    x = np.ascontiguousarray(sx, dtype=np.float64).copy()
    y = np.ascontiguousarray(sy, dtype=np.float64).copy()
    lib().orc_remove_spikes(_d(x), _d(y), len(x), int(single), int(quirks))
    return x, y


def frame_weight(cx, cy, mx, my, motion=True, corr_interp=False):
    """Weight map [cy, cx] of -InFmMotion / -CorrInterp (see mc_oracle.cpp)."""
    w = np.zeros((cy, cx))
    lib().orc_frame_weight(cx, cy, C.c_double(mx), C.c_double(my), int(motion), int(corr_interp), _d(w))
    return w


def smooth(sx, sy, weight):
    """CStackShift::Smooth (Align/CStackShift.cpp:180-198)."""
    x = np.ascontiguousarray(sx, dtype=np.float64).copy()
    y = np.ascontiguousarray(sy, dtype=np.float64).copy()
    lib().orc_smooth(_d(x), _d(y), len(x), C.c_double(weight))
    return x, y


def dose_weight_sum(cmpx, cmpy, pix, kv, acc_dose):
    acc = np.ascontiguousarray(acc_dose, dtype=np.float64)
    w = np.zeros((cmpy, cmpx))
    lib().orc_dose_weights(cmpx, cmpy, C.c_double(pix), kv, _d(acc), len(acc), _d(w))
    return w


def make_relative(shifts, quirks=True):
    s = np.ascontiguousarray(shifts, dtype=np.float64).copy()
    nf, npat = s.shape[0], s.shape[1]
    ref = C.c_int(0)
    lib().orc_make_relative(_d(s), None, nf, npat, int(quirks), C.byref(ref))
    return s, ref.value


def detect_bads(shifts, centers, W, H):
    s = np.ascontiguousarray(shifts, dtype=np.float64)
    c = np.ascontiguousarray(centers, dtype=np.float64)
    nf, npat = s.shape[0], s.shape[1]
    bad = np.zeros((nf, npat), np.uint8)
    lib().orc_detect_bads(_d(s), _d(c), nf, npat, W, H, bad.ctypes.data_as(_up))
    return bad


def local_warp(img, centers, shifts, bad=None):
    img = np.ascontiguousarray(img, dtype=np.float64)
    c = np.ascontiguousarray(centers, dtype=np.float64)
    s = np.ascontiguousarray(shifts, dtype=np.float64)
    H, W = img.shape
    out = np.zeros_like(img)
    b = None
    if bad is not None:
        bad = np.ascontiguousarray(bad, dtype=np.uint8)
        b = bad.ctypes.data_as(_up)
    lib().orc_local_warp(_d(img), W, H, _d(c), _d(s), b, len(c), _d(out))
    return out


def feature_map(xcf_sum_cmp, xw, xh):
    s = np.ascontiguousarray(xcf_sum_cmp, dtype=np.complex128)
    f = np.zeros((128, 128), np.uint8)
    lib().orc_feature_map(s.ctypes.data_as(_dp), xw, xh, f.ctypes.data_as(_up))
    return f


def patch_starts(features, px, py, xw, xh):
    f = np.ascontiguousarray(features, dtype=np.uint8)
    st = np.zeros((px * py, 2), np.int32)
    lib().orc_patch_centers(f.ctypes.data_as(_up), px, py, xw, xh,
                            st.ctypes.data_as(_ip))
    return st


def norm_by_var(img):
    """GNormByStd2D out of place: each pixel over the variance of its 5 x 5
    in-bounds neighbourhood (Align/GNormByStd2D.cu:7-42)."""
    a = np.ascontiguousarray(img, dtype=np.float64).copy()
    lib().orc_norm_by_var(_d(a), a.shape[1], a.shape[0])
    return a


def feature_threshold(img):
    """CDetectFeatures thresholds (Align/CDetectFeatures.cpp:134-159) on a
    normalised map: (feature mask, (lo, hi))."""
    a = np.ascontiguousarray(img, dtype=np.float64)
    f = np.zeros(a.shape, np.uint8)
    thr = np.zeros(2)
    lib().orc_feature_threshold(_d(a), a.size, f.ctypes.data_as(_up), _d(thr))
    return f, (thr[0], thr[1])


def interpolate(gx, gy, binz, n, fmsize=None, nearest=False):
    """CInterpolateShift (Align/CInterpolateShift.cpp:41-142): group values to
    n frames grouped by binz (raw frame sizes fmsize)."""
    gx = np.ascontiguousarray(gx, dtype=np.float64)
    gy = np.ascontiguousarray(gy, dtype=np.float64)
    ox, oy = np.zeros(n), np.zeros(n)
    ng = C.c_int(0)
    fs = None
    if fmsize is not None:
        fmsize = np.ascontiguousarray(fmsize, np.int32)
        fs = fmsize.ctypes.data_as(_ip)
    lib().orc_interpolate(_d(gx), _d(gy), binz, n, fs, int(nearest), _d(ox), _d(oy), C.byref(ng))
    return ox, oy, ng.value


def align_movie(frames, params=None, xcf_shift=None, patch_starts=None,
                replay_global=None, replay_patch=None, replay_centers=None,
                replay_bad=None):
    """Run the whole path on N x H x W float32 frames. Returns a dict."""
    frames = np.ascontiguousarray(frames, dtype=np.float32)
    N, H, W = frames.shape
    p = params if params is not None else default_params()
    npat = p.patch[0] * p.patch[1] if (p.patch[0] > 1 and p.patch[1] > 1) else 0
    keep = []

    def dp(a, dt=np.float64):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data_as(_dp if dt == np.float64 else (_ip if dt == np.int32 else _up))

    ov = Override(dp(xcf_shift), dp(patch_starts, np.int32), dp(replay_global),
                  dp(replay_patch), dp(replay_centers), dp(replay_bad, np.uint8))
    res = dict(
        sums=np.zeros((5, H, W)), global_shift=np.zeros((N, 2)),
        xcf_shift=np.zeros((N, 2)), global_iters=np.zeros(1, np.int32),
        global_errors=np.zeros(max(p.iter, 1)),
        patch_shift=np.zeros((N, max(npat, 1), 2)),
        patch_bad=np.zeros((N, max(npat, 1)), np.uint8),
        patch_center=np.zeros((max(npat, 1), 2)),
        patch_starts=np.zeros((max(npat, 1), 2), np.int32),
        patch_iters=np.zeros(max(npat, 1), np.int32),
        feature_map=np.zeros((128, 128), np.uint8),
    )
    o = Out()
    o.sums = _d(res["sums"])
    o.global_shift = _d(res["global_shift"])
    o.xcf_shift = _d(res["xcf_shift"])
    o.global_iters = res["global_iters"].ctypes.data_as(_ip)
    o.global_errors = _d(res["global_errors"])
    o.patch_shift = _d(res["patch_shift"])
    o.patch_bad = res["patch_bad"].ctypes.data_as(_up)
    o.patch_center = _d(res["patch_center"])
    o.patch_starts = res["patch_starts"].ctypes.data_as(_ip)
    o.patch_iters = res["patch_iters"].ctypes.data_as(_ip)
    o.feature_map = res["feature_map"].ctypes.data_as(_up)
    if p.out_stack[0]:
        res["stack"] = np.zeros((N, H, W))
        o.stack = _d(res["stack"])
    rc = lib().orc_align_movie(frames.ctypes.data_as(_fp), W, H, N, C.byref(p),
                               C.byref(ov), C.byref(o))
    if rc != 0:
        raise RuntimeError("orc_align_movie failed: %d" % rc)
    ns, ow, oh = o.num_sums, o.out_w, o.out_h
    res["sums"] = res["sums"].reshape(-1)[: ns * ow * oh].reshape(ns, oh, ow)
    if p.out_stack[0]:
        nk = o.stack_frames
        res["stack"] = res["stack"].reshape(-1)[: nk * ow * oh].reshape(nk, oh, ow)
    res["global_iters"] = int(res["global_iters"][0])
    res["global_errors"] = res["global_errors"][: res["global_iters"]]
    if npat == 0:
        for k in ("patch_shift", "patch_bad", "patch_center", "patch_starts",
                  "patch_iters", "feature_map"):
            res.pop(k)
    return res
