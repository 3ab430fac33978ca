"""ctypes binding of libzola_hip.so (include/zola_hip.h).

Mirrors the reference's host-side objects for the hot path: a context plays
the role of DataPhase_ + PSFmany_float_ + Modelmany_double_
(src/main/java/org/pasteur/imagej/process/gpu/DataPhase_.java,
PSFmany_float_.java, Modelmany_double_.java); errors raise ZolaError with the
library's message, as the reference logs cudaResult failures.  The library
must be built (zola-3d_amd/build.py); there is no CPU fallback.
"""
import ctypes as C
import importlib.util
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzola_hip.so")

FIT_REFERENCE = 0
FIT_LM = 1
LAYOUT_AUTO, LAYOUT_WAVE, LAYOUT_QUAD, LAYOUT_PAIR, LAYOUT_OCT = 0, 1, 2, 3, 4
ST_CONVERGED, ST_MAXITER, ST_NAN, ST_LOWPHOTON = 1, 2, 4, 8

SYMBOLS = [
    "zola_ctx_create", "zola_ctx_destroy", "zola_last_error", "zola_ctx_info",
    "zola_set_coefficients", "zola_set_sigma", "zola_set_stream", "zola_psf_forward",
    "zola_psf_forward_dev", "zola_psf_backward", "zola_psf_backward_dev", "zola_fit_batch",
    "zola_fit_batch_dev", "zola_fit_init_dev", "zola_get_stream", "zola_calib_eval_dev",
    "zola_calib_fit", "zola_detector_create", "zola_detector_destroy", "zola_detector_info",
    "zola_detector_range", "zola_detector_templates", "zola_detect_frames",
    "zola_detector_correlation", "zola_localize_frames", "zola_localize_frames_scmos", "zola_fit_mixed_dev",
    "zola_fit_batch_zoil_dev", "zola_fit_batch_scmos_dev", "zola_fit_batch_ext_dev",
    "zola_search_psf_center", "zola_set_fit_layout", "zola_autotune_fit",
    "zola_java_double_string",
    "zola_calib_json_format", "zola_calib_json_parse", "zola_loc_csv_header", "zola_loc_csv_rows",
    "zola_formats_last_error", "zola_bead_prepare", "zola_bead_prepare_scmos",
    "zola_psf_backward_noalloc_dev", "zola_calib_eval_full_dev", "zola_calib_fit_full",
    "zola_fit_ragged",
]
CAND_FIELDS = 7
LOC_FIELDS = 12


class ZolaError(RuntimeError):
    pass


class Optics(C.Structure):
    _fields_ = [("size_fft", C.c_int), ("xystep", C.c_double), ("zstep", C.c_double),
                ("wavelength", C.c_double), ("noil", C.c_double), ("nwat", C.c_double),
                ("na", C.c_double), ("sigma", C.c_double), ("with_apo", C.c_int),
                ("reserved", C.c_int)]


class FitConfig(C.Structure):
    _fields_ = [("mode", C.c_int), ("max_iter", C.c_int), ("zmin", C.c_double),
                ("zmax", C.c_double), ("zfocus", C.c_double), ("photon_threshold", C.c_double),
                ("tol", C.c_double)]


class BeadDiag(C.Structure):
    _fields_ = [("mean0", C.c_double), ("std0", C.c_double), ("mean", C.c_double), ("std", C.c_double),
                ("coef", C.c_double * 6), ("size_filt", C.c_int), ("miniposition", C.c_int), ("ndata", C.c_int)]


CALIB_FIT_REGISTRATION, CALIB_FIT_SIGMA, CALIB_FIT_BACKGROUND, CALIB_FIT_PHOTONS = 1, 2, 4, 8
CALIB_FIT_ALL = 15


class CalibBeads(C.Structure):
    _fields_ = [("nb", C.c_int), ("S", C.c_int), ("pixels", C.c_void_p), ("dz", C.c_void_p), ("A", C.c_void_p),
                ("offset", C.c_void_p), ("xyz", C.c_void_p), ("poly", C.c_void_p), ("fit_a", C.c_void_p),
                ("fit_a_each", C.c_void_p)]


class AcceptConfig(C.Structure):
    _fields_ = [("photon_threshold", C.c_double), ("min_z", C.c_double), ("max_z", C.c_double),
                ("snr_min", C.c_double), ("model_mask_frac", C.c_double),
                ("patch_mask_frac", C.c_double), ("max_shift_px", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ZolaError("libzola_hip.so is not built: run zola-3d_amd/build.py")
        # A process holds one HIP runtime (soname libamdhip64.so.7): when PyTorch
        # is installed, load it first so that its bundled runtime is the one
        # both use and torch streams/tensors can be passed to the *_dev calls.
        if importlib.util.find_spec("torch") is not None:
            import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.zola_ctx_create.argtypes = [C.POINTER(Optics), C.c_int, vp, C.c_int, C.POINTER(vp)]
        L.zola_ctx_destroy.argtypes = [vp]
        L.zola_last_error.restype = C.c_char_p
        L.zola_last_error.argtypes = [vp]
        L.zola_ctx_info.argtypes = [vp, vp]
        L.zola_set_coefficients.argtypes = [vp, vp, C.c_int]
        L.zola_set_sigma.argtypes = [vp, C.c_double]
        L.zola_set_stream.argtypes = [vp, vp]
        for name in ("zola_psf_forward", "zola_psf_forward_dev"):
            getattr(L, name).argtypes = [vp, C.c_int, vp, vp]
        for name in ("zola_psf_backward", "zola_psf_backward_dev", "zola_psf_backward_noalloc_dev"):
            if hasattr(L, name):
                getattr(L, name).argtypes = [vp, C.c_int, vp, vp, vp, vp]
        for name in ("zola_fit_batch", "zola_fit_batch_dev"):
            if hasattr(L, name):
                getattr(L, name).argtypes = [vp, C.POINTER(FitConfig), C.c_int, vp, vp, vp, vp,
                                             vp, vp, vp]
        if hasattr(L, "zola_calib_fit"):
            L.zola_calib_eval_dev.argtypes = [vp, C.c_int] + [vp] * 9
            L.zola_calib_fit.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.c_int,
                                         C.c_double, vp, vp, vp]
        if hasattr(L, "zola_calib_fit_full"):
            L.zola_calib_eval_full_dev.argtypes = [vp, C.c_int, C.c_int] + [vp] * 11
            L.zola_calib_fit_full.argtypes = [vp, C.POINTER(CalibBeads), vp, vp, vp, C.c_int, C.c_int, C.c_double,
                                              vp, vp, vp]
        if hasattr(L, "zola_fit_init_dev"):
            L.zola_fit_init_dev.argtypes = [vp, C.c_int, vp, C.c_int, C.c_double, C.c_double,
                                            C.c_double, vp]
        if hasattr(L, "zola_search_psf_center"):
            L.zola_search_psf_center.argtypes = [vp, C.c_double, C.c_double, C.POINTER(C.c_double)]
        if hasattr(L, "zola_fit_batch_zoil_dev"):
            L.zola_fit_batch_zoil_dev.argtypes = [vp, C.POINTER(FitConfig), C.c_int] + [vp] * 8
            L.zola_fit_batch_scmos_dev.argtypes = [vp, C.POINTER(FitConfig), C.c_int] + [vp] * 8
            L.zola_fit_batch_ext_dev.argtypes = [vp, C.POINTER(FitConfig), C.c_int] + [vp] * 9
        if hasattr(L, "zola_fit_mixed_dev"):
            L.zola_fit_mixed_dev.argtypes = [vp, C.POINTER(FitConfig), C.c_int] + [vp] * 9
        if hasattr(L, "zola_fit_ragged"):
            L.zola_fit_ragged.argtypes = [vp, C.POINTER(FitConfig), C.c_int, vp, vp, vp, C.c_longlong] + [vp] * 6
        if hasattr(L, "zola_detector_create"):
            L.zola_detector_create.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double,
                                               C.c_double, C.c_double, C.c_double, C.POINTER(vp)]
            L.zola_detector_destroy.argtypes = [vp]
            L.zola_detector_info.argtypes = [vp, vp]
            L.zola_detector_range.argtypes = [vp, vp]
            L.zola_detector_templates.argtypes = [vp, vp]
            L.zola_detect_frames.argtypes = [vp, C.c_int, vp, C.c_int, vp, vp]
            L.zola_detector_correlation.argtypes = [vp, vp, vp]
            L.zola_localize_frames.argtypes = [vp, C.POINTER(FitConfig), C.POINTER(AcceptConfig),
                                               C.c_int, vp, C.c_int, vp, vp, vp, vp]
            L.zola_localize_frames_scmos.argtypes = [vp, C.POINTER(FitConfig), C.POINTER(AcceptConfig),
                                                     C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp]
        if hasattr(L, "zola_bead_prepare"):
            L.zola_bead_prepare.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, C.c_double, C.c_double,
                                            C.c_double, C.c_double, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp,
                                            C.POINTER(BeadDiag)]
            L.zola_bead_prepare_scmos.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, C.c_double,
                                                  C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp,
                                                  C.POINTER(BeadDiag)]
        L.zola_set_fit_layout.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.zola_autotune_fit.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), vp]
        L.zola_java_double_string.argtypes = [C.c_double, C.c_char_p, C.c_int]
        L.zola_calib_json_format.restype = C.c_long
        L.zola_calib_json_format.argtypes = [C.POINTER(Optics), vp, C.c_int, C.c_char_p, C.c_long]
        L.zola_calib_json_parse.argtypes = [C.c_char_p, C.c_int, C.POINTER(Optics),
                                            C.POINTER(C.c_int), vp, C.c_int, C.POINTER(C.c_int)]
        L.zola_loc_csv_header.restype = C.c_long
        L.zola_loc_csv_header.argtypes = [C.c_char_p, C.c_long]
        L.zola_loc_csv_rows.restype = C.c_long
        L.zola_loc_csv_rows.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_long]
        L.zola_formats_last_error.restype = C.c_char_p
        _lib = L
    return _lib


# ---- calibration file and localization table (host code, no GPU) ----------

def _formats_error(rc):
    return ZolaError(f"{rc}: {lib().zola_formats_last_error().decode()}")


def java_double_string(v):
    """The text Java's Double.toString gives for v."""
    buf = C.create_string_buffer(32)
    rc = lib().zola_java_double_string(float(v), buf, 32)
    if rc < 0:
        raise _formats_error(rc)
    return buf.value.decode()


def _text(call):
    n = call(None, 0)
    if n < 0:
        raise _formats_error(n)
    buf = C.create_string_buffer(n + 1)
    call(buf, n + 1)
    return buf.value.decode()


def calib_json_format(opt, coeffs):
    """Calibration file text of DataPhase.saveJSON for optics and Zernike slot
    coefficients."""
    a = np.ascontiguousarray(coeffs, dtype=np.float64)
    return _text(lambda b, c: lib().zola_calib_json_format(C.byref(opt), _ptr(a), a.size, b, c))


def calib_json_parse(text, size_fft):
    """Reads calibration file text as DataPhase.loadJSON does: returns (optics,
    coefficients, the file's own sizeFFT).  nwat is left at 0."""
    if isinstance(text, str):
        text = text.encode()
    opt = Optics()
    fsize, J = C.c_int(0), C.c_int(0)
    L = lib()
    rc = L.zola_calib_json_parse(text, size_fft, C.byref(opt), C.byref(fsize), None, 0, C.byref(J))
    if rc < 0:
        raise _formats_error(rc)
    a = np.zeros(J.value)
    rc = L.zola_calib_json_parse(text, size_fft, C.byref(opt), C.byref(fsize), _ptr(a), a.size,
                                 C.byref(J))
    if rc < 0:
        raise _formats_error(rc)
    return opt, a, fsize.value


def loc_csv(locs, first_id=0, frame_base=0, header=True):
    """Localization table text (StackLocalization.saveCSV) of records as
    zola_localize_frames writes them."""
    r = np.ascontiguousarray(locs, dtype=np.float64).reshape(-1, LOC_FIELDS)
    head = _text(lambda b, c: lib().zola_loc_csv_header(b, c)) if header else ""
    rows = _text(lambda b, c: lib().zola_loc_csv_rows(_ptr(r), r.shape[0], first_id, frame_base, b, c))
    return head + rows


def optics(size_fft=64, xystep=0.11, zstep=0.05, wavelength=0.68, noil=1.518, nwat=1.518,
           na=1.45, sigma=1.0, with_apo=0):
    return Optics(size_fft, xystep, zstep, wavelength, noil, nwat, na, sigma, with_apo, 0)


_DT = {"f4": ("float32", "torch.float32"), "f8": ("float64", "torch.float64"), "i4": ("int32", "torch.int32")}


def _shape_ok(shape, want):
    return len(shape) == len(want) and all(w is None or int(s) == w for s, w in zip(shape, want))


def _host(a, name, want, dt):
    """Host array argument: converted to a C-contiguous array of dtype dt and
    checked against the shape want (None entries are free).  Raises ValueError
    before anything reaches the library, which trusts the sizes it is given."""
    a = np.ascontiguousarray(a, dtype=_DT[dt][0])
    if not _shape_ok(a.shape, want):
        raise ValueError(f"{name}: shape {tuple(a.shape)}, expected {tuple(want)}")
    return a


def _dev(t, name, want, dt, optional=False):
    """Device argument of a *_dev entry point: a torch tensor on the GPU,
    C-contiguous, of dtype dt and shape want (None entries free)."""
    if t is None:
        if optional:
            return None
        raise ValueError(f"{name}: a device tensor is required")
    if not (hasattr(t, "data_ptr") and hasattr(t, "is_cuda")):
        raise ValueError(f"{name}: expected a torch tensor in GPU memory, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name}: tensor is not in GPU memory")
    if str(t.dtype) != _DT[dt][1]:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {_DT[dt][1]}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor is not contiguous")
    if not _shape_ok(tuple(t.shape), want):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(want)}")
    return t


def _ptr(a):
    """Data pointer of a numpy array or a torch tensor."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return C.c_void_p(a.ctypes.data)


class Context:
    """A PSF model (optics + Zernike coefficients) for ROIs of side R."""

    def __init__(self, opt, R, coeffs):
        self.opt = opt
        self.R = int(R)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        self.J = len(c)
        h = C.c_void_p()
        rc = lib().zola_ctx_create(C.byref(opt), self.R, _ptr(c), self.J, C.byref(h))
        self._h = h
        if rc != 0:
            msg = lib().zola_last_error(h).decode() if h.value else "zola_ctx_create failed"
            lib().zola_ctx_destroy(h)
            self._h = None
            raise ZolaError(f"zola_ctx_create: {msg} (rc={rc})")
        info = np.zeros(5, np.int32)
        lib().zola_ctx_info(self._h, _ptr(info))
        self.N, self.D, self.s2 = int(info[0]), int(info[1]), int(info[2])

    def close(self):
        if getattr(self, "_h", None):
            lib().zola_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise ZolaError(f"{what}: {lib().zola_last_error(self._h).decode()} (rc={rc})")

    def set_coefficients(self, coeffs):
        c = _host(coeffs, "coeffs", (self.J,), "f8")
        self._check(lib().zola_set_coefficients(self._h, _ptr(c), len(c)), "set_coefficients")

    def set_sigma(self, sigma):
        self._check(lib().zola_set_sigma(self._h, float(sigma)), "set_sigma")

    def set_stream(self, stream_handle):
        self._check(lib().zola_set_stream(self._h, C.c_void_p(stream_handle)), "set_stream")

    # -- PSF ---------------------------------------------------------------
    def psf_forward(self, params):
        """params: (B, 4) [x, y, z, zoil] um -> (B, R, R) float32 (host arrays)."""
        p = _host(params, "params", (None, 4), "f4")
        B = p.shape[0]
        out = np.zeros((B, self.R, self.R), np.float32)
        self._check(lib().zola_psf_forward(self._h, B, _ptr(p), _ptr(out)), "psf_forward")
        return out

    def psf_forward_dev(self, params, out):
        """Device form: torch float32 tensors already on the GPU."""
        B = _dev(params, "params", (None, 4), "f4").shape[0]
        _dev(out, "out", (B, self.R, self.R), "f4")
        self._check(lib().zola_psf_forward_dev(self._h, params.shape[0], _ptr(params), _ptr(out)),
                    "psf_forward_dev")

    def psf_backward(self, params, w, with_coeffs=True):
        p = _host(params, "params", (None, 4), "f4")
        B = p.shape[0]
        w = _host(w, "w", (B, self.R, self.R), "f4")
        dp = np.zeros((B, 4), np.float32)
        da = np.zeros(self.J, np.float64) if with_coeffs else None
        self._check(lib().zola_psf_backward(self._h, B, _ptr(p), _ptr(w), _ptr(dp), _ptr(da)),
                    "psf_backward")
        return dp, da

    def psf_backward_dev(self, params, w, dparams, da=None):
        B = _dev(params, "params", (None, 4), "f4").shape[0]
        _dev(w, "w", (B, self.R, self.R), "f4")
        _dev(dparams, "dparams", (B, 4), "f4")
        _dev(da, "da", (self.J,), "f8", optional=True)
        self._check(lib().zola_psf_backward_dev(self._h, params.shape[0], _ptr(params), _ptr(w),
                                                _ptr(dparams), _ptr(da)), "psf_backward_dev")

    # -- fits --------------------------------------------------------------
    @staticmethod
    def fit_config(mode=FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, zfocus=0.0,
                   photon_threshold=0.0, tol=1e-4):
        return FitConfig(mode, max_iter, zmin, zmax, zfocus, photon_threshold, tol)

    def fit(self, cfg, pixels, init):
        """Host form. pixels (B, R, R) float32, init (B, 5) float64 [A, B, x, y, z]."""
        px = _host(pixels, "pixels", (None, self.R, self.R), "f4")
        B = px.shape[0]
        ini = _host(init, "init", (B, 5), "f8")
        theta = np.zeros((B, 5)); nll = np.zeros(B); crlb = np.zeros((B, 3))
        iters = np.zeros(B, np.int32); status = np.zeros(B, np.int32)
        self._check(lib().zola_fit_batch(self._h, C.byref(cfg), B, _ptr(px), _ptr(ini),
                                         _ptr(theta), _ptr(nll), _ptr(crlb), _ptr(iters),
                                         _ptr(status)), "fit_batch")
        return dict(theta=theta, nll=nll, crlb=crlb, iters=iters, status=status)

    def fit_ragged(self, cfg, rois, init):
        """Host form for ROIs of several sides: rois is a sequence of square
        float32 arrays (sides the library covers on this grid), init (B, 5)."""
        B = len(rois)
        sizes = np.zeros(B, np.int32)
        offs = np.zeros(B, np.int64)
        parts = []
        n = 0
        for b, r in enumerate(rois):
            a = np.ascontiguousarray(r, dtype=np.float32)
            if a.ndim != 2 or a.shape[0] != a.shape[1]:
                raise ValueError(f"rois[{b}]: expected a square 2-D array, got shape {a.shape}")
            sizes[b] = a.shape[0]
            offs[b] = n
            n += a.size
            parts.append(a.ravel())
        px = np.concatenate(parts) if parts else np.zeros(0, np.float32)
        ini = _host(init, "init", (B, 5), "f8")
        theta = np.zeros((B, 5)); nll = np.zeros(B); crlb = np.zeros((B, 3))
        iters = np.zeros(B, np.int32); status = np.zeros(B, np.int32)
        self._check(lib().zola_fit_ragged(self._h, C.byref(cfg), B, _ptr(offs), _ptr(sizes), _ptr(px), n,
                                          _ptr(ini), _ptr(theta), _ptr(nll), _ptr(crlb), _ptr(iters),
                                          _ptr(status)), "fit_ragged")
        return dict(theta=theta, nll=nll, crlb=crlb, iters=iters, status=status)

    def set_fit_layout(self, R, mode, layout):
        """Work layout of the 64-grid fits of side R: LAYOUT_AUTO / WAVE / QUAD / PAIR
        (PAIR: LM form, R <= 21)."""
        self._check(lib().zola_set_fit_layout(self._h, R, mode, layout), "set_fit_layout")

    def autotune_fit(self, R, mode, B=8192):
        """Times every layout that exists for (R, mode) on a synthetic batch, keeps
        the fastest; returns (layout, ms one wave per ROI, ms four waves per ROI,
        ms two waves per ROI or 0 where that layout does not exist)."""
        lay = C.c_int(0)
        ms = np.zeros(3)
        self._check(lib().zola_autotune_fit(self._h, R, mode, B, C.byref(lay), _ptr(ms)), "autotune_fit")
        return lay.value, float(ms[0]), float(ms[1]), float(ms[2])

    def search_psf_center(self, zfocus, axial_range):
        """Centre of the template z range (SearchPSFcenter.getPosition), um."""
        out = C.c_double()
        self._check(lib().zola_search_psf_center(self._h, zfocus, axial_range, C.byref(out)),
                    "search_psf_center")
        return out.value

    # -- calibration bead preparation ---------------------------------------
    def _bead_out(self, S, W, H, P, xp, yp):
        nb = len(xp)
        return dict(x=np.array(xp, dtype=np.float32), y=np.array(yp, dtype=np.float32),
                    patches=np.zeros((nb, S, P, P)), B=np.zeros(nb), A=np.zeros((nb, S)),
                    background=np.zeros((W, H)))

    @staticmethod
    def _bead_diag(d):
        return dict(mean0=d.mean0, std0=d.std0, mean=d.mean, std=d.std, coef=np.array(d.coef[:]),
                    size_filt=d.size_filt, miniposition=d.miniposition, ndata=d.ndata)

    def bead_prepare(self, adu, xp, yp, P, zstep, adu_factor=1.0, gain=1.0, offset=0.0):
        """InitBackgroundAndPhotonNumber (CCD/EMCCD form).  adu (S, W, H) float32
        raw stack [z][i][ii]; xp, yp float seeds.  Returns refined seeds,
        patches (nb, S, P, P), B (nb), A (nb, S), background (W, H) and diag."""
        a = _host(adu, "adu", (None, None, None), "f4")
        S, W, H = a.shape
        if len(xp) != len(yp):
            raise ValueError(f"xp, yp: {len(xp)} and {len(yp)} seeds")
        o = self._bead_out(S, W, H, P, xp, yp)
        d = BeadDiag()
        self._check(lib().zola_bead_prepare(self._h, S, W, H, _ptr(a), adu_factor, gain, offset, zstep, P,
                                            len(o["x"]), _ptr(o["x"]), _ptr(o["y"]), _ptr(o["patches"]),
                                            _ptr(o["B"]), _ptr(o["A"]), _ptr(o["background"]), C.byref(d)),
                    "bead_prepare")
        o["diag"] = self._bead_diag(d)
        return o

    def bead_prepare_scmos(self, adu, offset, gain, variance, xp, yp, P, zstep):
        """InitBackgroundAndPhotonNumber (sCMOS form) with float camera maps (W, H);
        adds scmos (nb, P, P) as the reference leaves it (last bead only)."""
        a = _host(adu, "adu", (None, None, None), "f4")
        S, W, H = a.shape
        if len(xp) != len(yp):
            raise ValueError(f"xp, yp: {len(xp)} and {len(yp)} seeds")
        maps = [_host(m, n, (W, H), "f4") for m, n in ((offset, "offset"), (gain, "gain"), (variance, "variance"))]
        o = self._bead_out(S, W, H, P, xp, yp)
        o["scmos"] = np.zeros((len(o["x"]), P, P))
        d = BeadDiag()
        self._check(lib().zola_bead_prepare_scmos(self._h, S, W, H, _ptr(a), _ptr(maps[0]), _ptr(maps[1]),
                                                  _ptr(maps[2]), zstep, P, len(o["x"]), _ptr(o["x"]), _ptr(o["y"]),
                                                  _ptr(o["patches"]), _ptr(o["B"]), _ptr(o["A"]),
                                                  _ptr(o["scmos"]), _ptr(o["background"]), C.byref(d)),
                    "bead_prepare_scmos")
        o["diag"] = self._bead_diag(d)
        return o

    def _fit_outputs(self, pixels, init, theta, nll, crlb, iters, status):
        B = _dev(pixels, "pixels", (None, self.R, self.R), "f4").shape[0]
        _dev(init, "init", (B, 5), "f8")
        _dev(theta, "theta", (B, 5), "f8")
        _dev(nll, "nll", (B,), "f8")
        _dev(crlb, "crlb", (B, 3), "f8")
        _dev(iters, "iters", (B,), "i4")
        _dev(status, "status", (B,), "i4")
        return B

    def fit_zoil_dev(self, cfg, pixels, init, zoil, theta, nll, crlb, iters, status):
        """Device form with a per-ROI focus term zoil (B floats)."""
        B = self._fit_outputs(pixels, init, theta, nll, crlb, iters, status)
        _dev(zoil, "zoil", (B,), "f4")
        self._check(lib().zola_fit_batch_zoil_dev(self._h, C.byref(cfg), pixels.shape[0], _ptr(pixels),
                                                  _ptr(init), _ptr(zoil), _ptr(theta), _ptr(nll), _ptr(crlb),
                                                  _ptr(iters), _ptr(status)), "fit_batch_zoil_dev")

    def fit_ext_dev(self, cfg, pixels, init, theta, nll, crlb, iters, status, zoil=None, varg=None):
        """Device form with an optional per-ROI focus term zoil (B floats) and an
        optional sCMOS map varg (B x R x R floats, var/gain^2 added to mu)."""
        B = self._fit_outputs(pixels, init, theta, nll, crlb, iters, status)
        _dev(zoil, "zoil", (B,), "f4", optional=True)
        _dev(varg, "varg", (B, self.R, self.R), "f4", optional=True)
        self._check(lib().zola_fit_batch_ext_dev(self._h, C.byref(cfg), pixels.shape[0], _ptr(pixels), _ptr(init),
                                                 _ptr(zoil) if zoil is not None else None,
                                                 _ptr(varg) if varg is not None else None, _ptr(theta),
                                                 _ptr(nll), _ptr(crlb), _ptr(iters), _ptr(status)),
                    "fit_batch_ext_dev")

    def fit_mixed_dev(self, cfg, groups):
        """groups: list of dicts with keys R, pixels, init, theta, nll, crlb, iters, status
        (device tensors); one graph launch on the context's stream."""
        G = len(groups)
        if not 1 <= G <= 8:
            raise ValueError(f"groups: {G}, expected 1 to 8")
        for g in groups:
            R = int(g["R"])
            B = _dev(g["pixels"], "pixels", (None, R, R), "f4").shape[0]
            _dev(g["init"], "init", (B, 5), "f8")
            _dev(g["theta"], "theta", (B, 5), "f8")
            _dev(g["nll"], "nll", (B,), "f8")
            _dev(g["crlb"], "crlb", (B, 3), "f8")
            _dev(g["iters"], "iters", (B,), "i4")
            _dev(g["status"], "status", (B,), "i4")
        Rs = (C.c_int * G)(*[int(g["R"]) for g in groups])
        Bs = (C.c_int * G)(*[int(g["pixels"].shape[0]) for g in groups])
        arr = {k: (C.c_void_p * G)(*[g[k].data_ptr() for g in groups])
               for k in ("pixels", "init", "theta", "nll", "crlb", "iters", "status")}
        self._check(lib().zola_fit_mixed_dev(self._h, C.byref(cfg), G, Rs, Bs, arr["pixels"], arr["init"],
                                             arr["theta"], arr["nll"], arr["crlb"], arr["iters"],
                                             arr["status"]), "fit_mixed_dev")

    def fit_dev(self, cfg, pixels, init, theta, nll, crlb, iters, status):
        self._fit_outputs(pixels, init, theta, nll, crlb, iters, status)
        self._check(lib().zola_fit_batch_dev(self._h, C.byref(cfg), pixels.shape[0],
                                             _ptr(pixels), _ptr(init), _ptr(theta), _ptr(nll),
                                             _ptr(crlb), _ptr(iters), _ptr(status)),
                    "fit_batch_dev")

    # -- calibration -------------------------------------------------------
    def calib_eval_dev(self, params, pixels, amp, bg, nll, dparams, damp, dbg, da=None):
        B = _dev(params, "params", (None, 4), "f4").shape[0]
        _dev(pixels, "pixels", (B, self.R, self.R), "f4")
        for name, t in (("amp", amp), ("bg", bg), ("nll", nll), ("damp", damp), ("dbg", dbg)):
            _dev(t, name, (B,), "f8")
        _dev(dparams, "dparams", (B, 4), "f4")
        _dev(da, "da", (self.J,), "f8", optional=True)
        self._check(lib().zola_calib_eval_dev(self._h, params.shape[0], _ptr(params), _ptr(pixels),
                                              _ptr(amp), _ptr(bg), _ptr(nll), _ptr(dparams),
                                              _ptr(damp), _ptr(dbg), _ptr(da)), "calib_eval_dev")

    def calib_fit(self, stacks, dz, bead, coeffs, fit_mask=None, max_iter=100, gtol=1e-6):
        """ZernikePhaseRetrieval: stacks (nb, S, R, R), dz (S,), bead (nb, 5) [x, y, z0, A, bg]."""
        px = _host(stacks, "stacks", (None, None, self.R, self.R), "f4")
        nb, S = px.shape[:2]
        bd = _host(bead, "bead", (nb, 5), "f8").copy()
        co = _host(coeffs, "coeffs", (self.J,), "f8").copy()
        dz = _host(dz, "dz", (S,), "f8")
        mask = None if fit_mask is None else _host(fit_mask, "fit_mask", (self.J,), "i4")
        nll = C.c_double()
        it = C.c_int()
        ev = C.c_int()
        self._check(lib().zola_calib_fit(self._h, nb, S, _ptr(px), _ptr(dz),
                                         _ptr(bd), _ptr(co), _ptr(mask), int(max_iter),
                                         float(gtol), C.byref(nll), C.byref(it), C.byref(ev)),
                    "calib_fit")
        return dict(bead=bd, coeffs=co, nll=nll.value, iters=it.value, evals=ev.value)

    def calib_eval_full_dev(self, nb, S, params, pixels, amp, poly, nll, dparams, damp, offset=None, dpoly=None,
                            dsigma=None, da=None):
        """The reference's full calibration model (zola_calib_eval_full_dev), device tensors;
        PSF b*S + s is slice s of bead b."""
        B = nb * S
        _dev(params, "params", (B, 4), "f4")
        _dev(pixels, "pixels", (B, self.R, self.R), "f4")
        for name, t in (("amp", amp), ("nll", nll), ("damp", damp)):
            _dev(t, name, (B,), "f8")
        _dev(poly, "poly", (nb, 6), "f8")
        _dev(dparams, "dparams", (B, 4), "f4")
        _dev(offset, "offset", (nb, self.R, self.R), "f8", optional=True)
        _dev(dpoly, "dpoly", (B, 6), "f8", optional=True)
        _dev(dsigma, "dsigma", (B,), "f8", optional=True)
        _dev(da, "da", (self.J,), "f8", optional=True)
        self._check(lib().zola_calib_eval_full_dev(self._h, nb, S, _ptr(params), _ptr(pixels), _ptr(amp), _ptr(poly),
                                                   _ptr(offset), _ptr(nll), _ptr(dparams), _ptr(damp), _ptr(dpoly),
                                                   _ptr(dsigma), _ptr(da)), "calib_eval_full_dev")

    def calib_fit_full(self, stacks, dz, A, xyz, poly, coeffs, sigma=1.0, fit_a=None, fit_a_each=None,
                       offset=None, fit_mask=None, flags=CALIB_FIT_ALL, max_iter=200, gtol=1e-6):
        """Phase retrieval on the reference's full model (ZernikePhaseRetrieval.run):
        stacks (nb, S, R, R), dz (S,), A (nb, S) photon curves, xyz (nb, 3), poly (nb, 6),
        offset (nb, R, R) sCMOS map or None.  Returns the fitted values."""
        px = _host(stacks, "stacks", (None, None, self.R, self.R), "f4")
        nb, S = px.shape[:2]
        dz = _host(dz, "dz", (S,), "f8")
        A = _host(A, "A", (nb, S), "f8")
        xyz = _host(xyz, "xyz", (nb, 3), "f8").copy()
        poly = _host(poly, "poly", (nb, 6), "f8").copy()
        fa = np.zeros(nb) if fit_a is None else _host(fit_a, "fit_a", (nb,), "f8").copy()
        fae = np.zeros((nb, S)) if fit_a_each is None else _host(fit_a_each, "fit_a_each", (nb, S), "f8").copy()
        off = None if offset is None else _host(offset, "offset", (nb, self.R, self.R), "f8")
        co = _host(coeffs, "coeffs", (self.J,), "f8").copy()
        mask = None if fit_mask is None else _host(fit_mask, "fit_mask", (self.J,), "i4")
        beads = CalibBeads(nb, S, px.ctypes.data, dz.ctypes.data, A.ctypes.data,
                           off.ctypes.data if off is not None else None, xyz.ctypes.data, poly.ctypes.data,
                           fa.ctypes.data, fae.ctypes.data)
        sg = C.c_double(float(sigma))
        nll, it, ev = C.c_double(), C.c_int(), C.c_int()
        self._check(lib().zola_calib_fit_full(self._h, C.byref(beads), C.byref(sg), _ptr(co), _ptr(mask), int(flags),
                                              int(max_iter), float(gtol), C.byref(nll), C.byref(it), C.byref(ev)),
                    "calib_fit_full")
        return dict(coeffs=co, sigma=sg.value, xyz=xyz, poly=poly, fit_a=fa, fit_a_each=fae, nll=nll.value,
                    iters=it.value, evals=ev.value)

    def fit_init_dev(self, pixels, nZ, zmin_t, zstep_t, zfocus, init):
        B = _dev(pixels, "pixels", (None, self.R, self.R), "f4").shape[0]
        _dev(init, "init", (B, 5), "f8")
        self._check(lib().zola_fit_init_dev(self._h, pixels.shape[0], _ptr(pixels), int(nZ),
                                            float(zmin_t), float(zstep_t), float(zfocus),
                                            _ptr(init)), "fit_init_dev")


# -- Zernike index helpers (calibration-file conventions) ------------------
class Detector:
    """Frame detector and frame localization pipeline on a context
    (process/cpu/Predetection.java and the LocalizationThread walk of
    process/cpu/LocalizationPipeline.java)."""

    def __init__(self, ctx, W, H, zmin, zmax, zstep, threshold=0.1, zfocus=0.0):
        self.ctx = ctx
        self.W, self.H = int(W), int(H)
        h = C.c_void_p()
        rc = lib().zola_detector_create(ctx._h, self.W, self.H, zmin, zmax, zstep, threshold, zfocus,
                                        C.byref(h))
        self._h = h if h.value else None
        if rc != 0:
            self.close()
            raise ZolaError("zola_detector_create: %s (%d)" % (lib().zola_last_error(ctx._h).decode(), rc))
        info = np.zeros(5, dtype=np.int32)
        lib().zola_detector_info(self._h, _ptr(info))
        self.nZ, self.FX, self.FY = int(info[0]), int(info[1]), int(info[2])

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.zola_detector_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise ZolaError("%s: %s (%d)" % (what, lib().zola_last_error(self.ctx._h).decode(), rc))

    def range(self):
        r = np.zeros(self.nZ)
        self._check(lib().zola_detector_range(self._h, _ptr(r)), "zola_detector_range")
        return r

    def templates(self):
        t = np.zeros((self.nZ, self.ctx.R, self.ctx.R))
        self._check(lib().zola_detector_templates(self._h, _ptr(t)), "zola_detector_templates")
        return t

    def correlation(self, frame):
        f = np.ascontiguousarray(frame, dtype=np.float32).reshape(self.W, self.H)
        out = np.zeros((self.nZ, self.W, self.H))
        self._check(lib().zola_detector_correlation(self._h, _ptr(f), _ptr(out)), "zola_detector_correlation")
        return out

    def detect(self, frames, cap=100000):
        """Candidate lists (one (n, 7) array per frame) of frames (F, W, H)."""
        fr = np.ascontiguousarray(frames, dtype=np.float32).reshape(-1, self.W, self.H)
        F = fr.shape[0]
        cand = np.zeros((F, cap, CAND_FIELDS))
        counts = np.zeros(F, dtype=np.int32)
        self._check(lib().zola_detect_frames(self._h, F, _ptr(fr), cap, _ptr(cand), _ptr(counts)),
                    "zola_detect_frames")
        return [cand[f, :min(int(counts[f]), cap)].copy() for f in range(F)], counts

    @staticmethod
    def accept_config(photon_threshold=500.0, min_z=-1.0, max_z=1.0, snr_min=1.0,
                      model_mask_frac=0.2, patch_mask_frac=0.5, max_shift_px=10.0):
        return AcceptConfig(photon_threshold, min_z, max_z, snr_min, model_mask_frac,
                            patch_mask_frac, max_shift_px)

    def localize(self, frames, fit_cfg, acc_cfg, cap=1000000):
        """Localizations (n, 12) of frames (F, W, H), per-frame counts and fit rounds."""
        fr = np.ascontiguousarray(frames, dtype=np.float32).reshape(-1, self.W, self.H)
        F = fr.shape[0]
        locs = np.zeros((cap, LOC_FIELDS))
        n = C.c_int()
        per = np.zeros(F, dtype=np.int32)
        rounds = C.c_int()
        self._check(lib().zola_localize_frames(self._h, C.byref(fit_cfg), C.byref(acc_cfg), F, _ptr(fr),
                                               cap, _ptr(locs), C.byref(n), _ptr(per), C.byref(rounds)),
                    "zola_localize_frames")
        return locs[:min(n.value, cap)].copy(), per, rounds.value

    def localize_scmos(self, adu, offset, gain, variance, fit_cfg, acc_cfg, cap=1000000):
        """Localizations of raw sCMOS frames adu (F, W, H) with the camera maps
        offset, gain, variance (W, H)."""
        fr = np.ascontiguousarray(adu, dtype=np.float32).reshape(-1, self.W, self.H)
        maps = [np.ascontiguousarray(x, dtype=np.float32).reshape(self.W, self.H) for x in (offset, gain, variance)]
        F = fr.shape[0]
        locs = np.zeros((cap, LOC_FIELDS))
        n = C.c_int()
        per = np.zeros(F, dtype=np.int32)
        rounds = C.c_int()
        self._check(lib().zola_localize_frames_scmos(self._h, C.byref(fit_cfg), C.byref(acc_cfg), F, _ptr(fr),
                                                     _ptr(maps[0]), _ptr(maps[1]), _ptr(maps[2]), cap, _ptr(locs),
                                                     C.byref(n), _ptr(per), C.byref(rounds)),
                    "zola_localize_frames_scmos")
        return locs[:min(n.value, cap)].copy(), per, rounds.value


def noll_to_nm(j):
    """Noll index j >= 1 -> (n, m), m > 0 for cos and m < 0 for sin terms."""
    n, j1 = 0, j - 1
    while j1 > n:
        n += 1
        j1 -= n
    m = (-1) ** j * ((n % 2) + 2 * int((j1 + ((n + 1) % 2)) / 2))
    return n, m


def nm_to_slot(n, m):
    """(n, m) -> 0-based slot of the reference basis order (utils/Zernike.java:26-79)."""
    return (n * (n + 2) + m) // 2


def coeffs_from_noll(noll_values, J):
    """Dict {noll j: rad} -> coefficient vector in the reference order."""
    a = np.zeros(J)
    for j, v in noll_values.items():
        s = nm_to_slot(*noll_to_nm(j))
        if s >= J:
            raise ValueError(f"Noll {j} does not fit in {J} coefficients")
        a[s] = v
    return a


def astigmatic_coeffs(J=28, seed=0x20A1D3, astig=1.0, sd=0.05):
    """Astigmatic PSF of the isolated-emitter workload: Noll 5 = astig rad,
    Noll 4..J (other than 5) ~ Normal(0, sd)."""
    rng = np.random.default_rng(seed)
    vals = {j: rng.normal(0.0, sd) for j in range(4, J + 1)}
    vals[5] = astig
    return coeffs_from_noll(vals, J)


def tetrapod_coeffs(J=45, seed=0x7E7A, primary=(5, 12, 23), amp=(2.0, 3.0), sd=0.1):
    """Tetrapod-type PSF of the 128-pupil workloads: the astigmatism family of
    orders 2, 4, 6 (Noll 5, 12, 23) ~ Uniform(amp) rad with random sign,
    Noll 4..J otherwise ~ Normal(0, sd)."""
    rng = np.random.default_rng(seed)
    vals = {j: rng.normal(0.0, sd) for j in range(4, J + 1)}
    for j in primary:
        vals[j] = rng.uniform(*amp) * rng.choice([-1.0, 1.0])
    return coeffs_from_noll(vals, J)
