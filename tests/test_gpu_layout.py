"""64-grid fits in every work layout (one wave, two waves (LM form, R <= 21) and
four waves per ROI) at every instantiated ROI side, against the FP64 oracle:

* LM form: the FP64 Levenberg-Marquardt optimum (0.5 nm x/y, 1 nm z, 1e-5
  relative NLL);
* reference form: the FP64 restatement of cpu/Localization.localize from the
  same start, on every ROI (tests/refparity.py);
* the autotuner returns a valid layout and the fits that follow use it.
"""
import numpy as np
import pytest

import oracle as O
import refparity
import simdata
import zola_hip as Z

pytestmark = pytest.mark.gpu

B = 12
ZMIN_T, ZSTEP_T, NZ = -1.0, 0.1, 21
_cache = {}


def setup(R):
    if R not in _cache:
        a = Z.astigmatic_coeffs(28)
        m = O.Model(O.optics(), R, a)
        rois, truth = simdata.isolated_rois(m, B, seed=300 + R)
        tmpl = simdata.templates(m, ZMIN_T, ZSTEP_T, NZ)
        init = simdata.init_restated(rois, tmpl, ZMIN_T, ZSTEP_T)
        opt = [m.lm_fit(rois[b].astype(np.float64), init[b], zmin=-1.3, zmax=1.3, maxit=500, tol=1e-10)
               for b in range(B)]
        _cache[R] = (m, a, rois, init, opt)
    return _cache[R]


@pytest.mark.parametrize("layout", [Z.LAYOUT_WAVE, Z.LAYOUT_QUAD, Z.LAYOUT_PAIR])
@pytest.mark.parametrize("R", [17, 21, 31, 41])
def test_lm_layout_matches_fp64_optimum(R, layout):
    if layout == Z.LAYOUT_PAIR and R > 21:
        pytest.skip("the two-wave layout covers R <= 21")
    m, a, rois, init, opt = setup(R)
    print(f"R {R} layout {layout}", flush=True)
    ctx = Z.Context(Z.optics(), R, a)
    ctx.set_fit_layout(R, Z.FIT_LM, layout)
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
    res = ctx.fit(cfg, rois, init)
    ctx.close()
    th = np.array([r["theta"] for r in opt])
    nll = np.array([r["nll"] for r in opt])
    d = (res["theta"][:, 2:] - th[:, 2:]) * 1000.0
    print("max |dx|,|dy|,|dz| nm", np.abs(d).max(axis=0), "iters", res["iters"])
    assert ((res["status"] & Z.ST_CONVERGED) != 0).all()
    assert np.abs(d[:, :2]).max() <= 0.5
    assert np.abs(d[:, 2]).max() <= 1.0
    assert np.all(np.abs(res["nll"] - nll) <= 1e-5 * np.abs(nll))
    # CRLB: the oracle's analytic Fisher matrix at the GPU's own optimum
    crlb_ref = np.array([np.sqrt(np.diag(np.linalg.inv(m.fisher(res["theta"][b])))[2:]) for b in range(B)])
    np.testing.assert_allclose(res["crlb"], crlb_ref, rtol=1e-3)
    np.testing.assert_allclose(res["crlb"], np.array([r["crlb"] for r in opt]), rtol=5e-3)


@pytest.mark.parametrize("R", [31, 41])
def test_reference_form_quad_layout_matches_restatement(R):
    m, a, rois, init, opt = setup(R)
    ctx = Z.Context(Z.optics(), R, a)
    ctx.set_fit_layout(R, Z.FIT_REFERENCE, Z.LAYOUT_QUAD)
    cfg = ctx.fit_config(mode=Z.FIT_REFERENCE, max_iter=20, zmin=-1.3, zmax=1.3, photon_threshold=0.0)
    res = ctx.fit(cfg, rois, init)
    ctx.close()
    O.lib().zo_set_psf_cache(1)
    try:
        ref = [m.localize(rois[b].astype(np.float64), init[b], minZ=-1.3, maxZ=1.3, iterMax=20) for b in range(B)]
    finally:
        O.lib().zo_set_psf_cache(0)
    refparity.compare(res, ref, 20, "R %d quad" % R)


@pytest.mark.parametrize("R", [21, 31])
def test_autotune_picks_a_layout_and_fits_follow_it(R):
    m, a, rois, init, opt = setup(R)
    ctx = Z.Context(Z.optics(), R, a)
    lay, t_wave, t_quad, t_pair = ctx.autotune_fit(R, Z.FIT_LM, B=2048)
    print(f"autotune R {R}:", lay, t_wave, t_quad, t_pair)
    times = {Z.LAYOUT_WAVE: t_wave, Z.LAYOUT_QUAD: t_quad, Z.LAYOUT_PAIR: t_pair}
    assert t_wave > 0 and t_quad > 0
    assert (t_pair > 0) == (R <= 21)
    assert lay == min((t, k) for k, t in times.items() if t > 0)[1]
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
    got = ctx.fit(cfg, rois, init)
    ctx.set_fit_layout(R, Z.FIT_LM, lay)
    again = ctx.fit(cfg, rois, init)
    ctx.close()
    np.testing.assert_array_equal(got["theta"], again["theta"])


def test_pair_layout_is_lm_only_and_small_sides_only():
    m, a, rois, init, opt = setup(31)
    ctx = Z.Context(Z.optics(), 31, a)
    with pytest.raises(Z.ZolaError):
        ctx.set_fit_layout(31, Z.FIT_LM, Z.LAYOUT_PAIR)
    with pytest.raises(Z.ZolaError):
        ctx.set_fit_layout(21, Z.FIT_REFERENCE, Z.LAYOUT_PAIR)
    ctx.close()


def test_eight_wave_layout_is_the_128_grid_lm_form_only():
    ctx = Z.Context(Z.optics(), 21, Z.astigmatic_coeffs(28))
    with pytest.raises(Z.ZolaError):
        ctx.set_fit_layout(21, Z.FIT_LM, Z.LAYOUT_OCT)
    ctx.close()
    opt = dict(size_fft=128, xystep=0.11, wavelength=0.68, noil=1.518, nwat=1.518, na=1.45, sigma=1.0)
    ctx = Z.Context(Z.optics(**opt), 41, Z.tetrapod_coeffs(45))
    ctx.set_fit_layout(41, Z.FIT_LM, Z.LAYOUT_OCT)
    ctx.set_fit_layout(41, Z.FIT_LM, Z.LAYOUT_QUAD)
    with pytest.raises(Z.ZolaError):
        ctx.set_fit_layout(41, Z.FIT_REFERENCE, Z.LAYOUT_OCT)
    with pytest.raises(Z.ZolaError):
        ctx.set_fit_layout(41, Z.FIT_LM, Z.LAYOUT_PAIR)
    ctx.close()
