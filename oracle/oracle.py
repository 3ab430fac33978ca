"""ctypes front-end of the FP64 oracle (oracle/zola_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Optics(C.Structure):
    _fields_ = [("size_fft", C.c_int), ("xystep", C.c_double), ("zstep", C.c_double),
                ("wavelength", C.c_double), ("noil", C.c_double), ("nwat", C.c_double),
                ("na", C.c_double), ("sigma", C.c_double), ("with_apo", C.c_int),
                ("reset_kz", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.zo_model_create.restype = vp
        L.zo_model_create.argtypes = [C.POINTER(Optics), C.c_int, C.c_void_p, C.c_int]
        L.zo_model_free.argtypes = [vp]
        L.zo_model_set_coeffs.argtypes = [vp, _dp, C.c_int]
        L.zo_model_set_sigma.argtypes = [vp, C.c_double]
        L.zo_model_set_basis_row.argtypes = [vp, C.c_int, _dp]
        L.zo_model_info.argtypes = [vp, _ip]
        L.zo_model_tables.argtypes = [vp] + [C.c_void_p] * 10
        L.zo_model_basis.argtypes = [vp, _dp]
        L.zo_model_gauss.argtypes = [vp, _dp]
        L.zo_model_radius_px.restype = C.c_double
        L.zo_model_radius_px.argtypes = [vp]
        L.zo_model_psf.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.zo_model_psf_jac.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp]
        L.zo_model_psf_vjp.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, _dp,
                                       _dp, C.c_void_p]
        L.zo_fft2d.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int]
        L.zo_inverse_cofactor.argtypes = [C.c_int, _dp, _dp]
        L.zo_nll.restype = C.c_double
        L.zo_nll.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double]
        L.zo_localize.restype = C.c_int
        L.zo_localize.argtypes = [vp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, _dp, C.POINTER(C.c_double), C.POINTER(C.c_int),
                                  _dp, C.POINTER(C.c_int)]
        L.zo_set_psf_cache.argtypes = [C.c_int]
        L.zo_set_scmos.argtypes = [C.c_void_p]
        L.zo_localize_batch.argtypes = [C.POINTER(Optics), C.c_int, C.c_void_p, C.c_int, C.c_int,
                                        _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                        C.c_double, _dp, _dp, _ip, _ip, C.c_int]
        L.zo_lm_fit.restype = C.c_int
        L.zo_lm_fit.argtypes = [vp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_int), _dp]
        L.zo_fisher_analytic.argtypes = [vp, _dp, C.c_double, _dp]
        L.zo_get_a_and_b.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp]
        L.zo_score.restype = C.c_double
        L.zo_score.argtypes = [C.c_int, _dp, _dp]
        L.zo_detector_create.restype = vp
        L.zo_detector_create.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_double]
        L.zo_detector_free.argtypes = [vp]
        L.zo_detector_nz.restype = C.c_int
        L.zo_detector_nz.argtypes = [vp]
        L.zo_detector_range.argtypes = [vp, _dp]
        L.zo_detector_templates.argtypes = [vp, _dp]
        L.zo_detector_run.restype = C.c_int
        L.zo_detector_run.argtypes = [vp, _dp, _dp, C.c_int, C.c_void_p]
        L.zo_calib_nll.restype = C.c_double
        L.zo_frame_localize.argtypes = [vp, vp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                        C.c_double, _dp, C.c_int]
        L.zo_frame_localize_iters.argtypes = [vp, vp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                              C.c_double, _dp, C.c_int, _ip]
        L.zo_set_psf_perturb.argtypes = [C.c_double]
        L.zo_frame_localize_scmos.argtypes = [vp, vp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                              C.c_int, C.c_double, _dp, C.c_int]
        L.zo_search_psf_center.restype = C.c_double
        L.zo_search_psf_center.argtypes = [vp, C.c_double, C.c_double]
        L.zo_calib_nll.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double]
        L.zo_calib_ref_create.restype = vp
        L.zo_calib_ref_create.argtypes = [C.POINTER(Optics), C.c_int, _dp, C.c_int, C.c_int, C.c_int, _dp,
                                          C.c_double, _dp, C.c_void_p, _dp, C.c_double, C.c_int]
        L.zo_calib_ref_free.argtypes = [vp]
        L.zo_calib_ref_likelihood.restype = C.c_double
        L.zo_calib_ref_likelihood.argtypes = [vp]
        L.zo_calib_ref_iteration.restype = C.c_double
        L.zo_calib_ref_iteration.argtypes = [vp]
        L.zo_calib_ref_state.argtypes = [vp] + [C.c_void_p] * 6
        _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.zo_bead_convert.argtypes = [C.c_long, _dp, C.c_double, C.c_double, C.c_double]
        L.zo_bead_convert_scmos.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.zo_bead_prepare.restype = C.c_int
        L.zo_bead_prepare.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_void_p, C.c_double, C.c_int,
                                      C.c_int, _fp, _fp, _dp, _dp, _dp, C.c_void_p, _dp, _dp, _dp, _ip]
        _lib = L
    return _lib


def optics(size_fft=64, xystep=0.11, zstep=0.05, wavelength=0.68, noil=1.518, nwat=1.518,
           na=1.45, sigma=1.0, with_apo=0, reset_kz=0):
    return Optics(size_fft, xystep, zstep, wavelength, noil, nwat, na, sigma, with_apo, reset_kz)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Model:
    """FP64 PSF model (PhaseParameters + ZernikePhase + PSFPhaseMany)."""

    def __init__(self, opt, R, coeffs):
        self.opt = opt
        self.R = int(R)
        self.coeffs = _d(coeffs)
        self.J = len(self.coeffs)
        self._h = lib().zo_model_create(C.byref(opt), self.R, self.coeffs.ctypes.data, self.J)
        info = np.zeros(5, np.int32)
        lib().zo_model_info(self._h, info)
        self.N, self.D, self.s2 = int(info[0]), int(info[1]), int(info[2])

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.zo_model_free(self._h)
            self._h = None

    def set_coeffs(self, c):
        self.coeffs = _d(c)
        lib().zo_model_set_coeffs(self._h, self.coeffs, len(self.coeffs))

    def set_sigma(self, s):
        lib().zo_model_set_sigma(self._h, float(s))

    def set_basis_row(self, k, row):
        lib().zo_model_set_basis_row(self._h, int(k), _d(row))

    def tables(self):
        D = self.D
        di = np.zeros(D, np.int32); dj = np.zeros(D, np.int32)
        arrs = [np.zeros(D) for _ in range(8)]
        lib().zo_model_tables(self._h, di.ctypes.data, dj.ctypes.data,
                              *[a.ctypes.data for a in arrs])
        names = ["kx", "ky", "kz", "kzoil", "kzflag", "kzoilflag", "pupil", "phase"]
        out = dict(zip(names, arrs))
        out["di"], out["dj"] = di, dj
        return out

    def basis(self):
        Z = np.zeros((self.J, self.D))
        lib().zo_model_basis(self._h, Z)
        return Z

    def gauss(self):
        g = np.zeros((self.s2, self.s2))
        lib().zo_model_gauss(self._h, g)
        return g

    def radius_px(self):
        return lib().zo_model_radius_px(self._h)

    def psf(self, x, y, zoil, zwat):
        out = np.zeros((self.R, self.R))
        lib().zo_model_psf(self._h, x, y, zoil, zwat, out)
        return out

    def psf_jac(self, x, y, zoil, zwat):
        out = np.zeros((self.R, self.R))
        d = np.zeros((3, self.R, self.R))
        lib().zo_model_psf_jac(self._h, x, y, zoil, zwat, out, d)
        return out, d

    def psf_vjp(self, x, y, zoil, zwat, w):
        g = np.zeros(4)
        ga = np.zeros(max(self.J, 1))
        lib().zo_model_psf_vjp(self._h, x, y, zoil, zwat, _d(w), g, ga.ctypes.data)
        return g, ga[: self.J]

    def localize(self, roi, theta, zfocus=0.0, minZ=-1.3, maxZ=1.3, iterMax=20, thr=0.0):
        th = _d(theta).copy()
        crlb = np.zeros(3)
        lik = C.c_double()
        it = C.c_int()
        nanf = C.c_int()
        model = np.zeros((self.R, self.R))
        ok = lib().zo_localize(self._h, _d(roi), th, zfocus, minZ, maxZ, iterMax, thr, crlb,
                               C.byref(lik), C.byref(it), model, C.byref(nanf))
        return dict(ok=ok, theta=th, crlb=crlb, lik=lik.value, iters=it.value, model=model,
                    nan=nanf.value)

    def lm_fit(self, roi, theta, zfocus=0.0, zmin=-1.3, zmax=1.3, maxit=200, tol=1e-9):
        th = _d(theta).copy()
        nll = C.c_double()
        it = C.c_int()
        crlb = np.zeros(3)
        conv = lib().zo_lm_fit(self._h, _d(roi), th, zfocus, zmin, zmax, maxit, tol,
                               C.byref(nll), C.byref(it), crlb)
        return dict(conv=conv, theta=th, nll=nll.value, iters=it.value, crlb=crlb)

    def fisher(self, theta, zfocus=0.0):
        F = np.zeros((5, 5))
        lib().zo_fisher_analytic(self._h, _d(theta), zfocus, F)
        return F


_scmos_keep = None


def set_scmos(varg):
    """sCMOS map (R x R, var/gain^2) added to mu by localize and lm_fit, or None."""
    global _scmos_keep
    if varg is None:
        _scmos_keep = None
        lib().zo_set_scmos(None)
    else:
        _scmos_keep = np.ascontiguousarray(varg, dtype=np.float64)
        lib().zo_set_scmos(_scmos_keep.ctypes.data)


def fft2d(re, im, sign):
    re = _d(re).copy(); im = _d(im).copy()
    lib().zo_fft2d(re.shape[0], re.shape[1], re, im, int(sign))
    return re, im


def inverse_cofactor(a):
    a = _d(a)
    out = np.zeros_like(a)
    lib().zo_inverse_cofactor(a.shape[0], a, out)
    return out


def nll(psf, roi, A, B):
    return lib().zo_nll(psf.size, _d(psf), _d(roi), A, B)


def get_a_and_b(patch, psf, thr=0.1):
    res = np.zeros(3)
    lib().zo_get_a_and_b(patch.size, _d(patch), _d(psf), thr, res)
    return res


def score(patch, model):
    return lib().zo_score(patch.size, _d(patch), _d(model))


def localize_batch(opt, R, coeffs, rois, theta, zfocus=0.0, minZ=-1.3, maxZ=1.3, iterMax=20,
                   thr=0.0, nthreads=1):
    coeffs = _d(coeffs)
    rois = _d(rois)
    B = rois.shape[0]
    th = _d(theta).copy()
    crlb = np.zeros((B, 3)); lik = np.zeros(B)
    iters = np.zeros(B, np.int32); status = np.zeros(B, np.int32)
    lib().zo_localize_batch(C.byref(opt), R, coeffs.ctypes.data, len(coeffs), B, rois, th,
                            zfocus, minZ, maxZ, iterMax, thr, crlb, lik, iters, status,
                            int(nthreads))
    return dict(theta=th, crlb=crlb, lik=lik, iters=iters, status=status)


class Detector:
    def __init__(self, model, W, H, zmin, zmax, zstep, thr=0.1, zfocus=0.0):
        self.model = model
        self.W, self.H = W, H
        self._h = lib().zo_detector_create(model._h, W, H, zmin, zmax, zstep, thr, zfocus)
        self.nZ = lib().zo_detector_nz(self._h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.zo_detector_free(self._h)
            self._h = None

    def range(self):
        r = np.zeros(self.nZ)
        lib().zo_detector_range(self._h, r)
        return r

    def templates(self):
        R = self.model.R
        t = np.zeros((self.nZ, R, R))
        lib().zo_detector_templates(self._h, t)
        return t

    def run(self, image, cap=100000, with_corr=False):
        out = np.zeros((cap, 7))
        corr = np.zeros((self.nZ, self.W, self.H)) if with_corr else None
        n = lib().zo_detector_run(self._h, _d(image), out, cap,
                                  corr.ctypes.data if with_corr else None)
        return (out[: min(n, cap)], corr) if with_corr else out[: min(n, cap)]


def frame_localize(model, det, image, zfocus, minZ, maxZ, iter_max=20, photon_thr=500.0,
                   cap=10000, with_iters=False):
    """LocalizationThread.run on one frame: (n, 12) records (and, with_iters, the
    iteration count of each record's fit)."""
    out = np.zeros((cap, 12))
    if with_iters:
        its = np.zeros(cap, np.int32)
        n = lib().zo_frame_localize_iters(model._h, det._h, _d(image), zfocus, minZ, maxZ, iter_max,
                                          photon_thr, out, cap, its)
        return out[:min(n, cap)], its[:min(n, cap)]
    n = lib().zo_frame_localize(model._h, det._h, _d(image), zfocus, minZ, maxZ, iter_max,
                                photon_thr, out, cap)
    return out[:min(n, cap)]


def set_psf_perturb(eps):
    """Sensitivity probe: deterministic PSF perturbation of up to eps * peak (0 = off)."""
    lib().zo_set_psf_perturb(float(eps))


def frame_localize_scmos(model, det, det_image, image, varg, zfocus, minZ, maxZ, iter_max=20,
                         photon_thr=500.0, cap=10000):
    """The walk of an sCMOS frame: candidates from det_image, patches from image,
    var/gain^2 map varg (W x H)."""
    out = np.zeros((cap, 12))
    n = lib().zo_frame_localize_scmos(model._h, det._h, _d(det_image), _d(image), _d(varg), zfocus, minZ, maxZ,
                                      iter_max, photon_thr, out, cap)
    return out[:min(n, cap)]


def search_psf_center(model, zfocus, axial_range):
    """SearchPSFcenter.getPosition with a model of R = N."""
    return lib().zo_search_psf_center(model._h, zfocus, axial_range)


def calib_nll(model, stack, dz, xs, ys, zs, Aslice, bg, zwat=0.0):
    S = stack.shape[0]
    return lib().zo_calib_nll(model._h, S, _d(stack), _d(dz), _d(xs), _d(ys), _d(zs),
                              _d(Aslice), _d(bg), zwat)


def bead_convert(adu, adu_factor, gain, offset):
    """CCD/EMCCD conversion of a raw bead stack (float ADU) to photons + 1, clamped at 0."""
    img = np.ascontiguousarray(adu, dtype=np.float64).copy()
    lib().zo_bead_convert(img.size, img, adu_factor, gain, offset)
    return img


def bead_convert_scmos(adu, offset, gain):
    S, W, H = adu.shape
    img = np.ascontiguousarray(adu, dtype=np.float64).copy()
    lib().zo_bead_convert_scmos(S, W, H, img, _d(np.asarray(offset, np.float32).astype(np.float64)),
                                _d(np.asarray(gain, np.float32).astype(np.float64)))
    return img


def bead_prepare(image, xp, yp, P, zstep, vargain=None):
    """Bead preparation on a converted stack image (S x W x H FP64) with float
    seeds xp, yp (refined copies are returned)."""
    S, W, H = image.shape
    nb = len(xp)
    xp = np.array(xp, dtype=np.float32)
    yp = np.array(yp, dtype=np.float32)
    patches = np.zeros((nb, S, P, P))
    B = np.zeros(max(nb, 1))
    A = np.zeros((max(nb, 1), S))
    sp = np.zeros((max(nb, 1), P, P)) if vargain is not None else None
    poly = np.zeros((W, H))
    coef = np.zeros(6)
    stats = np.zeros(4)
    info = np.zeros(3, np.int32)
    vg = None if vargain is None else _d(vargain)
    rc = lib().zo_bead_prepare(S, W, H, _d(image), None if vg is None else vg.ctypes.data, zstep, P, nb, xp, yp,
                               patches if nb else np.zeros(1), B, A, None if sp is None else sp.ctypes.data,
                               poly, coef, stats, info)
    return dict(rc=rc, x=xp, y=yp, patches=patches, B=B[:nb], A=A[:nb], scmos=None if sp is None else sp[:nb],
                poly=poly, coef=coef, stats=stats, size_filt=int(info[0]), miniposition=int(info[1]),
                ndata=int(info[2]))



class CalibRef:
    """The reference CPU calibration loop (ZernikePhaseRetrieval.phase_retrieve_zernike)
    restated in FP64 C: stacks (nb, S, R, R), A (nb, S) photon curves, B0 (nb,)
    background levels, scmos (nb, R, R) or None; slices on nthreads threads."""

    def __init__(self, opt, coeffs, stacks, zstep, A, B0, sigma=1.0, scmos=None, nthreads=1):
        self._keep = [_d(coeffs), _d(stacks), _d(A), None if scmos is None else _d(scmos), _d(B0)]
        self.nb, self.S, self.R = stacks.shape[0], stacks.shape[1], stacks.shape[2]
        self.J = len(coeffs)
        self.opt = opt
        co, st, a, sc, b0 = self._keep
        self._h = lib().zo_calib_ref_create(C.byref(opt), self.R, co, self.J, self.nb, self.S, st, zstep, a,
                                            None if sc is None else sc.ctypes.data, b0, sigma, int(nthreads))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.zo_calib_ref_free(self._h)
            self._h = None

    def likelihood(self):
        return lib().zo_calib_ref_likelihood(self._h)

    def iteration(self):
        return lib().zo_calib_ref_iteration(self._h)

    def state(self):
        co = np.zeros(self.J); reg = np.zeros((3, self.nb)); poly = np.zeros((self.nb, 6))
        fa = np.zeros(self.nb); fae = np.zeros((self.nb, self.S)); sg = C.c_double()
        lib().zo_calib_ref_state(self._h, co.ctypes.data, reg.ctypes.data, poly.ctypes.data, fa.ctypes.data,
                                 fae.ctypes.data, C.addressof(sg))
        return dict(coeffs=co, reg=reg, poly=poly, fit_a=fa, fit_a_each=fae, sigma=sg.value)
