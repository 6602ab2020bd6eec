"""ctypes binding of libmc3.so (the C ABI in include/mc_abi.h).

Device buffers are DevArray objects allocated through the library's own
mc_malloc/mc_memcpy, with numpy arrays on the host side.  (torch's wheel
bundles a different HIP runtime; loading both in one process leaves the
second without a device, so torch is not used for device memory here.)
Every call goes to the HIP library; a missing library raises and there is
no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmc3.so")
_LIB = None

MC_OK = 0
MC_E_ARG = -1
# names of mc_result.seconds_stage[0..4]
STAGES = ("forward", "global_align", "patch_align", "correct", "sums")


class McParams(C.Structure):
    """mc_params: the CInput flag subset (CInput.cpp:82-128)."""
    _fields_ = [
        ("patch", C.c_int * 2), ("iter", C.c_int), ("tol", C.c_float),
        ("bft", C.c_float * 2), ("phase_only", C.c_int), ("fm_dose", C.c_float),
        ("pix_size", C.c_float), ("kv", C.c_int), ("group", C.c_int * 2),
        ("fm_ref", C.c_int), ("ft_bin", C.c_float), ("sum_range", C.c_float * 2),
        ("split_sum", C.c_int), ("throw0", C.c_int), ("ref_quirks", C.c_int),
        ("log_dir", C.c_char_p), ("movie_name", C.c_char_p), ("verbose", C.c_int),
        ("spectra_fmt", C.c_int), ("launch_graphs", C.c_int), ("chunk_streams", C.c_int),
        ("prior", C.c_int), ("prior_fallback", C.c_float),
        ("fm_int_start", C.POINTER(C.c_int)), ("fm_int_size", C.POINTER(C.c_int)),
        ("movie_pipeline", C.c_int), ("out_stack", C.c_int * 2), ("align", C.c_int),
        ("in_fm_motion", C.c_int), ("corr_interp", C.c_int), ("frame_fmt", C.c_int),
    ]


class McMovie(C.Structure):
    _fields_ = [("frames", C.c_void_p), ("W", C.c_int), ("H", C.c_int), ("N", C.c_int),
                ("pitch", C.c_longlong), ("frame_stride", C.c_longlong)]


class McResult(C.Structure):
    _fields_ = [
        ("sums", C.c_void_p * 5), ("num_sums", C.c_int), ("out_w", C.c_int),
        ("out_h", C.c_int), ("global_shift", C.POINTER(C.c_float)),
        ("global_iters", C.POINTER(C.c_int)), ("global_errors", C.POINTER(C.c_float)),
        ("patch_shift", C.POINTER(C.c_float)), ("patch_bad", C.POINTER(C.c_ubyte)),
        ("patch_center", C.POINTER(C.c_float)), ("patch_starts", C.POINTER(C.c_int)),
        ("patch_iters", C.POINTER(C.c_int)), ("seconds_align", C.c_double),
        ("seconds_total", C.c_double), ("spectra_fmt_used", C.c_int * 2),
        ("prior_features", C.POINTER(C.c_float)), ("prior_targets", C.POINTER(C.c_float)),
        ("prior_fallbacks", C.c_int), ("stack", C.c_void_p), ("frame_fmt_used", C.c_int),
        ("seconds_stage", C.c_double * 5),
    ]


class McRawStack(C.Structure):
    _fields_ = [("data", C.c_void_p), ("mode", C.c_int), ("sx", C.c_int), ("sy", C.c_int),
                ("n", C.c_int), ("n_int", C.c_int), ("int_start", C.POINTER(C.c_int)),
                ("int_size", C.POINTER(C.c_int))]


class McEerStack(C.Structure):
    _fields_ = [("data", C.c_void_p), ("offsets", C.POINTER(C.c_longlong)), ("n", C.c_int),
                ("cam_w", C.c_int), ("cam_h", C.c_int), ("bits", C.c_int), ("sampling", C.c_int),
                ("n_int", C.c_int), ("int_start", C.POINTER(C.c_int)),
                ("int_size", C.POINTER(C.c_int))]


class McOverride(C.Structure):
    _fields_ = [
        ("replay_global", C.POINTER(C.c_float)), ("replay_patch", C.POINTER(C.c_float)),
        ("replay_centers", C.POINTER(C.c_float)), ("replay_bad", C.POINTER(C.c_ubyte)),
        ("patch_starts", C.POINTER(C.c_int)),
    ]


# Every symbol include/mc_abi.h declares.
ABI_SYMBOLS = (
    "mc_ctx_create", "mc_ctx_destroy", "mc_last_error", "mc_sync", "mc_default_params",
    "mc_output_geometry", "mc_stack_frames", "mc_save_stack", "mc_geometry", "mc_malloc", "mc_free", "mc_memcpy", "mc_align_movie",
    "mc_align_movies", "mc_rfft2", "mc_irfft2",
    "mc_phase_shift_accumulate", "mc_group_correlate_window", "mc_peak_fit",
    "mc_spline_despike_fp64", "mc_patch_stats_fp64", "mc_local_warp", "mc_dose_weight_sum",
    "mc_fourier_crop", "mc_feature_centres", "mc_extract_patches", "mc_finalize_sums",
    "mc_local_warp_noalloc",
    "mc_set_kernel_timing", "mc_get_kernel_timing", "mc_synth_render",
    "mc_prior_create", "mc_prior_destroy", "mc_prior_dims", "mc_prior_params",
    "mc_prior_forward", "mc_prior_loss_grad", "mc_prior_train", "mc_set_prior", "mc_gemm_bf16",
    "mc_autotune", "mc_autotune_schedules", "mc_fft_variant", "mc_fft_variant_name",
    "mc_host_alloc", "mc_host_free", "mc_set_refs", "mc_get_refs", "mc_load_frames", "mc_load_wait",
    "mc_mrc_info", "mc_mrc_read", "mc_mrc_stacks", "mc_tiff_info", "mc_tiff_read", "mc_eer_info", "mc_eer_read",
    "mc_load_eer",
    "mc_save_sums", "mc_save_aln", "mc_load_aln", "mc_save_star", "mc_fm_int_table",
    "mc_detect_bad", "mc_correct_bad",
)


def load():
    """Load libmc3.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmc3.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    lib.mc_last_error.restype = C.c_char_p
    for name in ABI_SYMBOLS:
        if name != "mc_last_error" and name != "mc_default_params":
            getattr(lib, name).restype = C.c_int
    _LIB = lib
    return lib


def default_params(**kw):
    p = McParams()
    load().mc_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("patch", "bft", "group", "sum_range", "out_stack"):
            arr = getattr(p, k)
            arr[0], arr[1] = v[0], v[1]
        elif k in ("log_dir", "movie_name"):
            setattr(p, k, v.encode() if isinstance(v, str) else v)
        elif k in ("fm_int_start", "fm_int_size"):
            a = np.ascontiguousarray(v, np.int32)
            p.__dict__.setdefault("_keep", []).append(a)
            setattr(p, k, a.ctypes.data_as(C.POINTER(C.c_int)))
        else:
            setattr(p, k, v)
    return p


def geometry(W, H, px=0, py=0):
    s = (C.c_int * 4)()
    b = (C.c_float * 2)()
    rc = load().mc_geometry(W, H, px, py, s, b)
    if rc != MC_OK:
        raise ValueError("unsupported geometry")
    return dict(xcf=(s[0], s[1]), patch=(s[2], s[3]), xcf_bin=(b[0], b[1]))


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class DevArray:
    """A C-contiguous device buffer owned by a Context."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(v) for v in shape)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        ctx._check(ctx.lib.mc_malloc(ctx.ptr, C.c_ulonglong(self.nbytes), C.byref(p)), "mc_malloc")
        self.ptr = p.value

    def stride(self, axis):
        """Stride of `axis` in elements."""
        return int(np.prod(self.shape[axis + 1:])) if axis + 1 < len(self.shape) else 1

    def vp(self, offset_elems=0):
        return C.c_void_p(self.ptr + offset_elems * self.dtype.itemsize)

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(self.ctx.lib.mc_memcpy(self.ctx.ptr, out.ctypes.data_as(C.c_void_p),
                                               C.c_void_p(self.ptr), C.c_ulonglong(self.nbytes), 1),
                        "mc_memcpy")
        return out

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        self.ctx._check(self.ctx.lib.mc_memcpy(self.ctx.ptr, C.c_void_p(self.ptr),
                                               a.ctypes.data_as(C.c_void_p),
                                               C.c_ulonglong(self.nbytes), 0), "mc_memcpy")
        return self

    def free(self):
        if self.ptr and self.ctx.ptr:
            self.ctx.lib.mc_free(self.ctx.ptr, C.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One mc_ctx (stream, FFT plans, device workspace) on one GPU."""

    def __init__(self, device=0, stream=None):
        self.lib = load()
        self.ptr = C.c_void_p()
        rc = self.lib.mc_ctx_create(device, C.c_void_p(stream) if stream else None,
                                    C.byref(self.ptr))
        if rc != MC_OK:
            raise RuntimeError("mc_ctx_create failed: %d" % rc)

    def close(self):
        if self.ptr:
            self.lib.mc_ctx_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != MC_OK:
            msg = self.lib.mc_last_error(self.ptr)
            raise RuntimeError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def sync(self):
        self._check(self.lib.mc_sync(self.ptr), "mc_sync")

    def empty(self, shape, dtype=np.float32):
        return DevArray(self, shape, dtype)

    def upload(self, arr, dtype=None):
        arr = np.asarray(arr)
        return DevArray(self, arr.shape, dtype or arr.dtype).upload(arr)

    def copy(self, dst, src, nbytes=None):
        """Device-to-device copy of src into dst (same byte size), or of the
        first nbytes of src."""
        if nbytes is None:
            assert dst.nbytes == src.nbytes
            nbytes = src.nbytes
        assert nbytes <= min(dst.nbytes, src.nbytes)
        self._check(self.lib.mc_memcpy(self.ptr, C.c_void_p(dst.ptr), C.c_void_p(src.ptr),
                                       C.c_ulonglong(nbytes), 2), "mc_memcpy")

    # ---- stage entry points ----
    def rfft2(self, frames, scale=1.0):
        """frames: DevArray float32 [B, H, W] -> complex64 [B, W/2+1, H] ([kx][ky])."""
        B, H, W = frames.shape
        out = self.empty((B, W // 2 + 1, H), np.complex64)
        self._check(self.lib.mc_rfft2(self.ptr, frames.vp(), W, H, C.c_longlong(frames.stride(1)),
                                      C.c_longlong(frames.stride(0)), B, C.c_float(scale),
                                      out.vp()), "mc_rfft2")
        self.sync()
        return out

    def irfft2(self, spec, W, clamp=False):
        """spec: complex64 [B, W/2+1, H] -> float32 [B, H, W]."""
        B, cx, H = spec.shape
        out = self.empty((B, H, W), np.float32)
        self._check(self.lib.mc_irfft2(self.ptr, spec.vp(), W, H, B, out.vp(), C.c_longlong(W),
                                       C.c_longlong(W * H), int(clamp)), "mc_irfft2")
        self.sync()
        return out

    def phase_shift_accumulate(self, spec, shifts, W):
        N, cx, H = spec.shape
        out = self.empty((cx, H), np.complex64)
        sh = np.ascontiguousarray(shifts, dtype=np.float32)
        self._check(self.lib.mc_phase_shift_accumulate(self.ptr, spec.vp(), W, H, N, _fptr(sh),
                                                       out.vp()), "mc_phase_shift_accumulate")
        self.sync()
        return out

    def correlate_window(self, sum_spec, grp_spec, W, bft, phase_only=False, sea=(64, 64)):
        H = sum_spec.shape[-1]
        win = np.zeros((sea[1], sea[0]), np.float32)
        self._check(self.lib.mc_group_correlate_window(
            self.ptr, sum_spec.vp(), grp_spec.vp(), W, H, C.c_float(bft), int(phase_only),
            sea[0], sea[1], _fptr(win)), "mc_group_correlate_window")
        return win

    def peak_fit(self, windows):
        cnt, ny, nx = windows.shape
        out = np.zeros((cnt, 2), np.float32)
        self._check(self.lib.mc_peak_fit(self.ptr, windows.vp(), nx, ny, cnt, _fptr(out)),
                    "mc_peak_fit")
        return out

    def despike(self, sx, sy, quirks=True):
        """sx, sy: one trajectory [n] or several [count, n], despiked in one call."""
        one = np.ndim(sx) == 1
        t = np.ascontiguousarray(np.stack([np.atleast_2d(sx), np.atleast_2d(sy)], 1),
                                 dtype=np.float64).copy()
        self._check(self.lib.mc_spline_despike_fp64(
            self.ptr, t.ctypes.data_as(C.POINTER(C.c_double)), t.shape[-1], t.shape[0],
            int(quirks)), "mc_spline_despike_fp64")
        return (t[0, 0], t[0, 1]) if one else (t[:, 0], t[:, 1])

    def patch_stats(self, shifts, centers, W, H, quirks=True):
        s = np.ascontiguousarray(shifts, dtype=np.float64).copy()
        cen = np.ascontiguousarray(centers, dtype=np.float64)
        N, P = s.shape[0], s.shape[1]
        bad = np.zeros((N, P), np.uint8)
        ref = C.c_int(0)
        self._check(self.lib.mc_patch_stats_fp64(
            self.ptr, s.ctypes.data_as(C.POINTER(C.c_double)),
            cen.ctypes.data_as(C.POINTER(C.c_double)), N, P, W, H, int(quirks),
            bad.ctypes.data_as(C.POINTER(C.c_ubyte)), C.byref(ref)), "mc_patch_stats_fp64")
        return s, bad, ref.value

    def local_warp(self, frames, centers, shifts, bad=None):
        N, H, W = frames.shape
        P = len(centers)
        out = self.empty(frames.shape, np.float32)
        cen = np.ascontiguousarray(centers, dtype=np.float32)
        sh = np.ascontiguousarray(shifts, dtype=np.float32)
        bp = None
        if bad is not None:
            bad = np.ascontiguousarray(bad, dtype=np.uint8)
            bp = bad.ctypes.data_as(C.POINTER(C.c_ubyte))
        self._check(self.lib.mc_local_warp(self.ptr, frames.vp(), out.vp(), W, H, N, P,
                                           _fptr(cen), _fptr(sh), bp), "mc_local_warp")
        self.sync()
        return out

    def dose_weight_sum(self, W, H, N, pix, kv, fm_dose):
        out = self.empty((W // 2 + 1, H), np.float32)
        self._check(self.lib.mc_dose_weight_sum(self.ptr, W, H, N, C.c_float(pix), kv,
                                                C.c_float(fm_dose), out.vp()),
                    "mc_dose_weight_sum")
        self.sync()
        return out

    def fourier_crop(self, spec, in_w, out_w, out_h, out=None, sum=False):
        """spec: complex64 [B, in_w/2+1, in_h] -> [B, out_w/2+1, out_h]
        (GFourierResize2D); a new output starts at zero."""
        B, _, in_h = spec.shape
        if out is None:
            out = self.upload(np.zeros((B, out_w // 2 + 1, out_h), np.complex64))
        self._check(self.lib.mc_fourier_crop(self.ptr, spec.vp(), in_w, in_h, out.vp(), out_w, out_h,
                                             B, int(sum)), "mc_fourier_crop")
        self.sync()
        return out

    def feature_centres(self, xcf_sum_spec, xw, px, py):
        """xcf_sum_spec: complex64 [xw/2+1, xh] -> (128 x 128 feature map,
        patch starts [P, 2] in xcf pixels)."""
        xh = xcf_sum_spec.shape[-1]
        fm = np.zeros((128, 128), np.uint8)
        st = np.zeros((px * py, 2), np.int32)
        self._check(self.lib.mc_feature_centres(
            self.ptr, xcf_sum_spec.vp(), xw, xh, px, py, fm.ctypes.data_as(C.POINTER(C.c_ubyte)),
            st.ctypes.data_as(C.POINTER(C.c_int))), "mc_feature_centres")
        return fm, st

    def extract_patches(self, xcf_frames, starts, pw, ph):
        """xcf_frames: float32 [N, xh, xw]; starts [P, 2] -> complex64
        [P * N, pw/2+1, ph], transform q * N + i (unnormalised R2C)."""
        N, xh, xw = xcf_frames.shape
        st = np.ascontiguousarray(starts, np.int32)
        P = st.shape[0]
        out = self.empty((P * N, pw // 2 + 1, ph), np.complex64)
        self._check(self.lib.mc_extract_patches(self.ptr, xcf_frames.vp(), xw, xh, N,
                                                st.ctypes.data_as(C.POINTER(C.c_int)), P, pw, ph,
                                                out.vp()), "mc_extract_patches")
        self.sync()
        return out

    def finalize_sums(self, sum_specs, W, ft_bin=1.0):
        """sum_specs: complex64 [K, W/2+1, H] -> float32 [K, oh, ow] (crop to
        the -FtBin size, C2R, clamp at 0)."""
        K, _, H = sum_specs.shape
        _, ow, oh = self.output_geometry(default_params(ft_bin=ft_bin), W, H, 1)
        out = self.empty((K, oh, ow), np.float32)
        self._check(self.lib.mc_finalize_sums(self.ptr, sum_specs.vp(), W, H, K, C.c_float(ft_bin),
                                              out.vp()), "mc_finalize_sums")
        self.sync()
        return out

    def synth_render(self, frames, recipe, poisson=True, seed=None):
        """Render a MovieRecipe into a DevArray float32 [N, H, W]."""
        N, H, W = frames.shape
        blobs = np.stack([recipe.bx, recipe.by, recipe.bs, recipe.ba], 1).astype(np.float32)
        disp = np.zeros((N, len(recipe.bx), 2), np.float32)
        for t in range(N):
            dx, dy = recipe.blob_displacement(t)
            disp[t, :, 0] = dx
            disp[t, :, 1] = dy
        self._check(self.lib.mc_synth_render(
            self.ptr, frames.vp(), W, H, N, C.c_longlong(frames.stride(1)),
            C.c_longlong(frames.stride(0)), len(blobs), _fptr(blobs), _fptr(disp),
            C.c_float(recipe.scale), C.c_ulonglong(recipe.seed if seed is None else seed),
            int(poisson)), "mc_synth_render")

    def host_array(self, shape, dtype):
        """Pinned host memory (mc_host_alloc) viewed as a numpy array."""
        return HostArray(self, shape, dtype)

    def set_refs(self, gain=None, dark=None, rot=0, flip=0, inv=0, quirks=True):
        """mc_set_refs: gain / dark references (numpy [ref_h, ref_w]) with
        -RotGain / -FlipGain / -InvGain post-processing."""
        ref = gain if gain is not None else dark
        g = None if gain is None else np.ascontiguousarray(gain, np.float32)
        d = None if dark is None else np.ascontiguousarray(dark, np.float32)
        rh, rw = (ref.shape if ref is not None else (0, 0))
        self._check(self.lib.mc_set_refs(self.ptr, None if g is None else _fptr(g),
                                         None if d is None else _fptr(d), rw, rh, rot, flip, inv,
                                         int(quirks)), "mc_set_refs")

    def get_refs(self):
        w, h = C.c_int(), C.c_int()
        self._check(self.lib.mc_get_refs(self.ptr, None, None, C.byref(w), C.byref(h)), "mc_get_refs")
        g = np.zeros((h.value, w.value), np.float32)
        d = np.zeros((h.value, w.value), np.float32)
        self._check(self.lib.mc_get_refs(self.ptr, _fptr(g), _fptr(d), None, None), "mc_get_refs")
        return g, d

    def load_frames(self, raw, mode, sx, sy, n, frames, width=None, height=None, async_=False,
                    int_start=None, int_size=None):
        """mc_load_frames: `raw` (numpy, HostArray or DevArray) holds n frames
        of an sx x sy stack in MRC `mode`; `frames` is an [N, H, pitch]
        float32 DevArray receiving the centre-cropped corrected frames."""
        N, H, pitch = frames.shape
        W = pitch if width is None else width
        Hh = H if height is None else height
        if isinstance(raw, (DevArray, HostArray)):
            ptr = raw.vp()
        else:
            raw = np.ascontiguousarray(raw)
            ptr = raw.ctypes.data_as(C.c_void_p)
        rs = McRawStack(ptr, mode, sx, sy, n)
        if int_size is not None:
            ist = np.ascontiguousarray(int_start, np.int32)
            isz = np.ascontiguousarray(int_size, np.int32)
            rs.n_int = len(isz)
            rs.int_start = ist.ctypes.data_as(C.POINTER(C.c_int))
            rs.int_size = isz.ctypes.data_as(C.POINTER(C.c_int))
        self._check(self.lib.mc_load_frames(self.ptr, C.byref(rs), frames.vp(), W, Hh,
                                            C.c_longlong(frames.stride(1)),
                                            C.c_longlong(frames.stride(0)), int(async_)),
                    "mc_load_frames")
        if not async_:
            self.sync()

    def load_eer(self, data, offsets, cam_w, cam_h, bits, sampling, frames, width=None,
                 height=None, async_=False, int_start=None, int_size=None):
        """mc_load_eer: `data` (numpy uint8 or HostArray) holds the compressed
        EER fractions, `offsets` their n + 1 byte offsets; `frames` is an
        [N, H, pitch] float32 DevArray receiving the rendered, corrected and
        centre-cropped frames."""
        N, H, pitch = frames.shape
        W = pitch if width is None else width
        Hh = H if height is None else height
        if isinstance(data, HostArray):
            ptr = data.vp()
        else:
            data = np.ascontiguousarray(data, np.uint8)
            ptr = data.ctypes.data_as(C.c_void_p)
        off = np.ascontiguousarray(offsets, np.int64)
        es = McEerStack(ptr, off.ctypes.data_as(C.POINTER(C.c_longlong)), len(off) - 1, cam_w,
                        cam_h, bits, sampling)
        keep = [data, off]
        if int_size is not None:
            ist = np.ascontiguousarray(int_start, np.int32)
            isz = np.ascontiguousarray(int_size, np.int32)
            keep += [ist, isz]
            es.n_int = len(isz)
            es.int_start = ist.ctypes.data_as(C.POINTER(C.c_int))
            es.int_size = isz.ctypes.data_as(C.POINTER(C.c_int))
        self._check(self.lib.mc_load_eer(self.ptr, C.byref(es), frames.vp(), W, Hh,
                                         C.c_longlong(frames.stride(1)),
                                         C.c_longlong(frames.stride(0)), int(async_)),
                    "mc_load_eer")
        if not async_:
            self.sync()
        return keep

    def load_wait(self):
        self._check(self.lib.mc_load_wait(self.ptr), "mc_load_wait")

    def save_sums(self, params, result, out_mrc):
        """mc_save_sums: the result's output sums as float MRC files."""
        r = McResult()
        for i, sm in enumerate(result["sums"]):
            r.sums[i] = sm.vp()
        r.num_sums = result["num_sums"]
        r.out_h, r.out_w = result["sums"][0].shape
        self._check(self.lib.mc_save_sums(self.ptr, C.byref(params), C.byref(r), out_mrc.encode()),
                    "mc_save_sums")

    def save_stack(self, params, result, out_mrc):
        """mc_save_stack: the result's aligned stack as out_mrc + "_Stk.mrc"."""
        r = McResult()
        st = result["stack"]
        r.stack = st.vp()
        r.out_h, r.out_w = st.shape[1], st.shape[2]
        self._check(self.lib.mc_save_stack(self.ptr, C.byref(params), C.byref(r), st.shape[0],
                                           out_mrc.encode()), "mc_save_stack")

    def detect_bad(self, frames, threshold=6.0, defect=(6, 6), defect_file=None, defect_map=None,
                   width=None):
        """mc_detect_bad on an [N, H, pitch] DevArray: (bad map DevArray [H, W]
        uint8, number of bad pixels, (hot threshold, correlation threshold))."""
        N, H, pitch = frames.shape
        W = pitch if width is None else width
        bad = self.empty((H, W), np.uint8)
        nb = C.c_int()
        th = (C.c_float * 2)()
        dm = None if defect_map is None else np.ascontiguousarray(defect_map, np.uint8)
        self._check(self.lib.mc_detect_bad(
            self.ptr, frames.vp(), W, H, N, C.c_longlong(frames.stride(1)),
            C.c_longlong(frames.stride(0)), C.c_float(threshold), defect[0], defect[1],
            defect_file.encode() if defect_file else None,
            None if dm is None else dm.ctypes.data_as(C.c_void_p), bad.vp(), C.byref(nb), th),
            "mc_detect_bad")
        return bad, nb.value, (th[0], th[1])

    def correct_bad(self, frames, bad, win=7, width=None):
        N, H, pitch = frames.shape
        W = pitch if width is None else width
        self._check(self.lib.mc_correct_bad(self.ptr, frames.vp(), W, H, N,
                                            C.c_longlong(frames.stride(1)),
                                            C.c_longlong(frames.stride(0)), bad.vp(), win),
                    "mc_correct_bad")

    def autotune(self, W, H, N, patch=(0, 0), ft_bin=1.0, radix=False):
        """mc_autotune (radix: mc_autotune_schedules with other radix orders
        too): time the FFT variants of this geometry and keep the fastest.
        Returns the report text."""
        buf = C.create_string_buffer(16384)
        self._check(self.lib.mc_autotune_schedules(self.ptr, W, H, N, patch[0], patch[1],
                                                   C.c_float(ft_bin), int(bool(radix)), buf,
                                                   len(buf)), "mc_autotune_schedules")
        return buf.value.decode()

    def fft_variant_name(self, W, H, cols, v):
        """Tile and radix order of variant v of the W x H plan's rows or columns."""
        buf = C.create_string_buffer(256)
        self._check(self.lib.mc_fft_variant_name(self.ptr, W, H, int(cols), v, buf, len(buf)),
                    "mc_fft_variant_name")
        return buf.value.decode()

    def fft_variant(self, W, H, cols, set=-1):
        """(count, current) tile variants of the W x H plan's rows or columns;
        set >= 0 forces one."""
        n, cur = C.c_int(), C.c_int()
        self._check(self.lib.mc_fft_variant(self.ptr, W, H, int(cols), set, C.byref(n),
                                            C.byref(cur)), "mc_fft_variant")
        return n.value, cur.value

    def set_kernel_timing(self, name):
        self._check(self.lib.mc_set_kernel_timing(self.ptr, name.encode() if name else None),
                    "mc_set_kernel_timing")

    def kernel_timing(self):
        ms = C.c_double(0)
        n = C.c_int(0)
        self._check(self.lib.mc_get_kernel_timing(self.ptr, C.byref(ms), C.byref(n)),
                    "mc_get_kernel_timing")
        return ms.value, n.value

    # ---- whole path ----
    def output_geometry(self, params, W, H, N):
        ns, ow, oh = C.c_int(), C.c_int(), C.c_int()
        self._check(self.lib.mc_output_geometry(C.byref(params), W, H, N, C.byref(ns),
                                                C.byref(ow), C.byref(oh)), "mc_output_geometry")
        return ns.value, ow.value, oh.value

    def stack_frames(self, params, N):
        """mc_stack_frames: frames of the -OutStack aligned stack (0 without it)."""
        n = C.c_int()
        self._check(self.lib.mc_stack_frames(C.byref(params), N, C.byref(n)), "mc_stack_frames")
        return n.value

    def _prepare(self, frames, params, sums, width, prior_data=False):
        N, H, pitch = frames.shape
        W = pitch if width is None else width
        ns, ow, oh = self.output_geometry(params, W, H, N)
        P = params.patch[0] * params.patch[1] if (params.patch[0] > 1 and params.patch[1] > 1) else 0
        if sums is None:
            sums = [self.empty((oh, ow), np.float32) for _ in range(ns)]
        mv = McMovie(frames.vp(), W, H, N, frames.stride(1), frames.stride(0))
        res = McResult()
        for i in range(ns):
            res.sums[i] = sums[i].vp()
        out = dict(global_shift=np.zeros((N, 2), np.float32),
                   global_iters=np.zeros(1, np.int32),
                   global_errors=np.zeros(max(params.iter, 1), np.float32))
        res.global_shift = _fptr(out["global_shift"])
        res.global_iters = out["global_iters"].ctypes.data_as(C.POINTER(C.c_int))
        res.global_errors = _fptr(out["global_errors"])
        if P:
            out.update(patch_shift=np.zeros((N, P, 2), np.float32),
                       patch_bad=np.zeros((N, P), np.uint8),
                       patch_center=np.zeros((P, 2), np.float32),
                       patch_starts=np.zeros((P, 2), np.int32),
                       patch_iters=np.zeros(P, np.int32))
            res.patch_shift = _fptr(out["patch_shift"])
            res.patch_bad = out["patch_bad"].ctypes.data_as(C.POINTER(C.c_ubyte))
            res.patch_center = _fptr(out["patch_center"])
            res.patch_starts = out["patch_starts"].ctypes.data_as(C.POINTER(C.c_int))
            res.patch_iters = out["patch_iters"].ctypes.data_as(C.POINTER(C.c_int))
            if prior_data:
                out.update(prior_features=np.zeros((P, 68 * N + 2), np.float32),
                           prior_targets=np.zeros((P, N, 2), np.float32))
                res.prior_features = _fptr(out["prior_features"])
                res.prior_targets = _fptr(out["prior_targets"])
        if params.out_stack[0]:
            stack = self.empty((self.stack_frames(params, N), oh, ow), np.float32)
            res.stack = stack.vp()
            out["stack"] = stack
        out["sums"] = sums
        out["num_sums"] = ns
        return mv, res, out

    @staticmethod
    def _finish(res, out):
        out["global_iters"] = int(out["global_iters"][0])
        out["global_errors"] = out["global_errors"][: out["global_iters"]]
        out["seconds_align"] = res.seconds_align
        out["seconds_total"] = res.seconds_total
        out["spectra_fmt_used"] = (res.spectra_fmt_used[0], res.spectra_fmt_used[1])
        out["prior_fallbacks"] = res.prior_fallbacks
        out["frame_fmt_used"] = res.frame_fmt_used
        out["seconds_stage"] = dict(zip(STAGES, res.seconds_stage))
        return out

    def align_movie(self, frames, params, replay_global=None, replay_patch=None,
                    replay_centers=None, replay_bad=None, patch_starts=None, sums=None,
                    width=None, prior_data=False):
        """Align an [N, H, pitch] float32 DevArray movie whose rows hold `width`
        (default pitch) pixels.  Returns numpy shift tables and DevArray sums;
        with prior_data also each patch's prior features and targets."""
        mv, res, out = self._prepare(frames, params, sums, width, prior_data)
        keep = []

        def fp(a, dt=np.float32, ct=C.c_float):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ct))

        ov = McOverride(fp(replay_global), fp(replay_patch), fp(replay_centers),
                        fp(replay_bad, np.uint8, C.c_ubyte), fp(patch_starts, np.int32, C.c_int))
        self._check(self.lib.mc_align_movie(self.ptr, C.byref(params), C.byref(mv), C.byref(ov),
                                            C.byref(res)), "mc_align_movie")
        return self._finish(res, out)

    def align_movies(self, movies, params, sums=None, prior_data=False):
        """mc_align_movies over a list of [N, H, W] float32 DevArrays with one
        McParams (or a list of them); `sums` optionally gives each movie's
        output DevArrays.  Returns one result dict per movie."""
        n = len(movies)
        plist = params if isinstance(params, (list, tuple)) else [params] * n
        slist = sums if sums is not None else [None] * n
        prep = [self._prepare(f, p, sm, None, prior_data) for f, p, sm in zip(movies, plist, slist)]
        mva = (McMovie * n)(*[q[0] for q in prep])
        ra = (McResult * n)(*[q[1] for q in prep])
        pa = (McParams * n)(*plist)
        self._check(self.lib.mc_align_movies(self.ptr, pa, mva, n, ra), "mc_align_movies")
        return [self._finish(ra[k], prep[k][2]) for k in range(n)]


class Prior:
    """A learned shift prior (mc_prior_*) on a context's device."""

    def __init__(self, ctx, frames, hidden=128, seed=1):
        self.ctx = ctx
        self.ptr = C.c_void_p()
        ctx._check(ctx.lib.mc_prior_create(ctx.ptr, frames, hidden, C.c_ulonglong(seed),
                                           C.byref(self.ptr)), "mc_prior_create")
        i, o, n = C.c_int(), C.c_int(), C.c_longlong()
        ctx.lib.mc_prior_dims(self.ptr, C.byref(i), C.byref(o), C.byref(n))
        self.in_dim, self.out_dim, self.n_params = i.value, o.value, n.value
        self.frames, self.hidden = frames, hidden

    def close(self):
        if self.ptr:
            self.ctx.lib.mc_prior_destroy(self.ctx.ptr, self.ptr)
            self.ptr = C.c_void_p()

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        self.ctx._check(self.ctx.lib.mc_prior_params(self.ctx.ptr, self.ptr, _fptr(out), None),
                        "mc_prior_params")
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.n_params
        self.ctx._check(self.ctx.lib.mc_prior_params(self.ctx.ptr, self.ptr, None, _fptr(w)),
                        "mc_prior_params")

    def layers(self, w=None):
        """Split a parameter vector into [(W0, b0), ... (W3, b3)]."""
        w = self.params() if w is None else w
        dims = [self.in_dim, self.hidden, self.hidden, self.hidden, self.out_dim]
        out, off = [], 0
        for l in range(4):
            W = w[off: off + dims[l] * dims[l + 1]].reshape(dims[l], dims[l + 1])
            off += dims[l] * dims[l + 1]
            b = w[off: off + dims[l + 1]]
            off += dims[l + 1]
            out.append((W, b))
        return out

    def forward(self, x):
        """x: [n, in_dim] DevArray -> [n, out_dim] numpy."""
        n = x.shape[0]
        y = self.ctx.empty((n, self.out_dim), np.float32)
        self.ctx._check(self.ctx.lib.mc_prior_forward(self.ctx.ptr, self.ptr, x.vp(), n, y.vp()),
                        "mc_prior_forward")
        return y.numpy()

    def loss_grad(self, x, t):
        loss = C.c_double()
        g = np.zeros(self.n_params, np.float32)
        self.ctx._check(self.ctx.lib.mc_prior_loss_grad(self.ctx.ptr, self.ptr, x.vp(), t.vp(),
                                                        x.shape[0], C.byref(loss), _fptr(g)),
                        "mc_prior_loss_grad")
        return loss.value, g

    def train(self, x, t, epochs, batch=64, lr=1e-3, seed=7):
        hist = np.zeros(epochs, np.float32)
        self.ctx._check(self.ctx.lib.mc_prior_train(self.ctx.ptr, self.ptr, x.vp(), t.vp(),
                                                    x.shape[0], epochs, batch, C.c_float(lr),
                                                    C.c_ulonglong(seed), _fptr(hist)),
                        "mc_prior_train")
        return hist

    def attach(self, on=True):
        self.ctx._check(self.ctx.lib.mc_set_prior(self.ctx.ptr, self.ptr if on else None),
                        "mc_set_prior")


def gemm_bf16(ctx, a16, sa, b16, sb, c, m, n, k):
    """c[i][j] = sum_k a16[i*sa[0] + k*sa[1]] * b16[k*sb[0] + j*sb[1]] (uint16 BF16
    DevArrays in, [m, n] float32 DevArray out)."""
    ctx._check(ctx.lib.mc_gemm_bf16(ctx.ptr, a16.vp(), C.c_longlong(sa[0]), C.c_longlong(sa[1]),
                                    b16.vp(), C.c_longlong(sb[0]), C.c_longlong(sb[1]), c.vp(),
                                    c.shape[1], m, n, k),
               "mc_gemm_bf16")


class HostArray:
    """Pinned host memory from mc_host_alloc with a numpy view (.a)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(v) for v in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ctx._check(ctx.lib.mc_host_alloc(ctx.ptr, C.c_ulonglong(self.nbytes), C.byref(p)),
                   "mc_host_alloc")
        self.ptr = p.value
        buf = (C.c_char * self.nbytes).from_address(self.ptr)
        self.a = np.frombuffer(buf, dtype=self.dtype).reshape(self.shape)

    def vp(self):
        return C.c_void_p(self.ptr)

    def free(self):
        if self.ptr and self.ctx.ptr:
            self.a = None
            self.ctx.lib.mc_host_free(self.ctx.ptr, C.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def raw_frame_bytes(mode, sx, sy):
    """Bytes of one raw frame in an MRC mode (4-bit: rows of (sx + 1) // 2)."""
    per_row = {0: sx, 5: sx, 1: 2 * sx, 6: 2 * sx, 2: 4 * sx, 101: (sx + 1) // 2}[mode]
    return per_row * sy


def mrc_info(path):
    """mc_mrc_info: dict(sx, sy, n, mode, data_offset)."""
    v = [C.c_int() for _ in range(4)]
    off = C.c_longlong()
    rc = load().mc_mrc_info(path.encode(), *[C.byref(x) for x in v], C.byref(off))
    if rc != MC_OK:
        raise RuntimeError("mc_mrc_info failed for %s" % path)
    return dict(sx=v[0].value, sy=v[1].value, n=v[2].value, mode=v[3].value,
                data_offset=off.value)


def mrc_stacks(path, stack_z=0):
    """mc_mrc_stacks: (frames per movie, movies) of an MRC file under -StackZ."""
    fr, st = C.c_int(), C.c_int()
    if load().mc_mrc_stacks(path.encode(), stack_z, C.byref(fr), C.byref(st)) != MC_OK:
        raise RuntimeError("mc_mrc_stacks failed for %s" % path)
    return fr.value, st.value


def mrc_read(path, first=0, count=None, out=None):
    """mc_mrc_read: frames [first, first+count) as raw bytes into `out`
    (numpy array or HostArray, allocated as uint8 when None)."""
    info = mrc_info(path)
    count = info["n"] - first if count is None else count
    nb = raw_frame_bytes(info["mode"], info["sx"], info["sy"]) * count
    if out is None:
        out = np.empty(nb, np.uint8)
    ptr = out.vp() if isinstance(out, HostArray) else out.ctypes.data_as(C.c_void_p)
    size = out.nbytes
    if size < nb:
        raise ValueError("output buffer too small")
    rc = load().mc_mrc_read(path.encode(), first, count, ptr)
    if rc != MC_OK:
        raise RuntimeError("mc_mrc_read failed for %s" % path)
    return out


def tiff_info(path):
    """mc_tiff_info: dict(sx, sy, n, mode, compression, pix_size)."""
    v = [C.c_int() for _ in range(5)]
    px = C.c_float()
    rc = load().mc_tiff_info(path.encode(), *[C.byref(x) for x in v], C.byref(px))
    if rc != MC_OK:
        raise RuntimeError("mc_tiff_info failed for %s" % path)
    return dict(sx=v[0].value, sy=v[1].value, n=v[2].value, mode=v[3].value,
                compression=v[4].value, pix_size=px.value)


def tiff_read(path, first=0, count=None, out=None):
    """mc_tiff_read: frames [first, first+count) decoded into `out` (numpy
    array or HostArray; uint8 bytes when None), rows bottom-up."""
    info = tiff_info(path)
    if info["mode"] < 0:
        raise ValueError("not a sample TIFF (mode %d)" % info["mode"])
    count = info["n"] - first if count is None else count
    nb = raw_frame_bytes(info["mode"], info["sx"], info["sy"]) * count
    if out is None:
        out = np.empty(nb, np.uint8)
    ptr = out.vp() if isinstance(out, HostArray) else out.ctypes.data_as(C.c_void_p)
    if out.nbytes < nb:
        raise ValueError("output buffer too small")
    rc = load().mc_tiff_read(path.encode(), first, count, ptr)
    if rc != MC_OK:
        raise RuntimeError("mc_tiff_read failed for %s" % path)
    return out


def eer_info(path):
    """mc_eer_info: dict(cam_w, cam_h, n, bits, bytes)."""
    v = [C.c_int() for _ in range(4)]
    nb = C.c_longlong()
    rc = load().mc_eer_info(path.encode(), *[C.byref(x) for x in v], C.byref(nb))
    if rc != MC_OK:
        raise RuntimeError("mc_eer_info failed for %s" % path)
    return dict(cam_w=v[0].value, cam_h=v[1].value, n=v[2].value, bits=v[3].value, bytes=nb.value)


def eer_read(path):
    """mc_eer_read: (data uint8 array, offsets int64 array of n + 1)."""
    info = eer_info(path)
    data = np.empty(max(info["bytes"], 1), np.uint8)
    off = np.empty(info["n"] + 1, np.int64)
    rc = load().mc_eer_read(path.encode(), data.ctypes.data_as(C.POINTER(C.c_ubyte)),
                            C.c_longlong(data.nbytes), off.ctypes.data_as(C.POINTER(C.c_longlong)))
    if rc != MC_OK:
        raise RuntimeError("mc_eer_read failed for %s" % path)
    return data[:info["bytes"]], off


def save_aln(path, params, sx, sy, n_stack, global_shift, patch_shift=None, patch_center=None):
    """mc_save_aln (-OutAln text format)."""
    g = np.ascontiguousarray(global_shift, np.float32)
    ps = None if patch_shift is None else np.ascontiguousarray(patch_shift, np.float32)
    pc = None if patch_center is None else np.ascontiguousarray(patch_center, np.float32)
    rc = load().mc_save_aln(path.encode(), C.byref(params), sx, sy, n_stack, g.shape[0], _fptr(g),
                            None if ps is None else _fptr(ps), None if pc is None else _fptr(pc))
    if rc != MC_OK:
        raise RuntimeError("mc_save_aln failed for %s" % path)


def load_aln(path):
    """mc_load_aln: dict(sx, sy, n, patch, global_shift, patch_shift, patch_center)."""
    v = [C.c_int() for _ in range(5)]
    lib = load()
    if lib.mc_load_aln(path.encode(), *[C.byref(x) for x in v], None, None, None) != MC_OK:
        raise RuntimeError("mc_load_aln failed for %s" % path)
    sx, sy, n, px, py = [x.value for x in v]
    P = px * py if (px > 1 and py > 1) else 0
    g = np.zeros((n, 2), np.float32)
    ps = np.zeros((n, max(P, 1), 2), np.float32)
    pc = np.zeros((max(P, 1), 2), np.float32)
    if lib.mc_load_aln(path.encode(), None, None, None, None, None, _fptr(g), _fptr(ps),
                       _fptr(pc)) != MC_OK:
        raise RuntimeError("mc_load_aln failed for %s" % path)
    out = dict(sx=sx, sy=sy, n=n, patch=(px, py), global_shift=g)
    if P:
        out.update(patch_shift=ps, patch_center=pc)
    return out


def save_star(path, params, sx, sy, n_stack, global_shift, gain_name=None):
    """mc_save_star (Relion global-shift star file)."""
    g = np.ascontiguousarray(global_shift, np.float32)
    rc = load().mc_save_star(path.encode(), C.byref(params), sx, sy, n_stack,
                             gain_name.encode() if gain_name else None, _fptr(g), g.shape[0])
    if rc != MC_OK:
        raise RuntimeError("mc_save_star failed for %s" % path)


def fm_int_table(path, n_raw, fm_dose, throw0=0, throw1=0):
    """mc_fm_int_table: dict(start, size, acc_dose, raw_dose, integrated)."""
    lib = load()
    n = C.c_int()
    pb = path.encode() if path else None
    if lib.mc_fm_int_table(pb, n_raw, C.c_float(fm_dose), throw0, throw1, 0, C.byref(n), None,
                           None, None, None, None) != MC_OK:
        raise RuntimeError("mc_fm_int_table failed")
    st = np.zeros(n.value, np.int32)
    sz = np.zeros(n.value, np.int32)
    acc = np.zeros(n.value, np.float32)
    rd = C.c_float()
    integ = C.c_int()
    if lib.mc_fm_int_table(pb, n_raw, C.c_float(fm_dose), throw0, throw1, n.value, C.byref(n),
                           st.ctypes.data_as(C.POINTER(C.c_int)),
                           sz.ctypes.data_as(C.POINTER(C.c_int)), _fptr(acc), C.byref(rd),
                           C.byref(integ)) != MC_OK:
        raise RuntimeError("mc_fm_int_table failed")
    return dict(start=st, size=sz, acc_dose=acc, raw_dose=rd.value, integrated=bool(integ.value))
