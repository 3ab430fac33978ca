"""The plugin's own localization shapes on the 128x128 pupil grid against the
FP64 oracle.  ZOLA.java forces an even patch (1090-1092), defaults to 32 (239)
and sets sizeFFT = max(128, 2 patch) (1126-1128), so a patch up to 64 runs on
N = 128 with the blur grid s2 = pow2(min(2R, N)) (process/cpu/PSFPhaseMany.java:
102-108): 64 for R = 32 and 20, 32 for R = 16, i.e. smaller than N.
For R = 16, 20, 32 (astigmatic 28-mode PSF):
* PSF synthesis within 1.5e-6 of the peak;
* LM fits at the FP64 optimum (0.5 nm x/y, 1 nm z, NLL 1e-5) with the CRLB of
  the oracle's analytic Fisher matrix (1e-3): the x, y tangents of the cropped
  blur are exact only through the crop-derivative kernel, so this checks them;
* the reference coordinate descent against its FP64 restatement
  (tests/refparity.py)."""
import concurrent.futures as cf

import numpy as np
import pytest

import oracle as O
import refparity
import zola_hip as Z

pytestmark = pytest.mark.gpu
SIZES = (16, 20, 32)
ZR = (-0.8, 0.8)


def _models(R):
    a = Z.astigmatic_coeffs(28)
    return a, Z.Context(Z.optics(size_fft=128), R, a), O.Model(O.optics(size_fft=128), R, a)


def _rois(m, B, seed):
    R = m.R
    rng = np.random.default_rng(seed)
    x, y = rng.uniform(-0.11, 0.11, B), rng.uniform(-0.11, 0.11, B)
    z = rng.uniform(ZR[0] + 0.1, ZR[1] - 0.1, B)
    A = np.exp(rng.uniform(np.log(2000), np.log(20000), B))
    bg = rng.uniform(5, 50, B)
    rois = np.stack([rng.poisson(A[b] * m.psf(x[b], y[b], 0.0, z[b]) + bg[b]) for b in range(B)]).astype(np.float32)
    truth = np.stack([A, bg, x, y, z], 1)
    return rois, truth


def _pool(R, a, fn, items):
    def work(chunk):
        mm = O.Model(O.optics(size_fft=128), R, a)
        return [fn(mm, it) for it in chunk]
    chunks = [items[i::8] for i in range(8)]
    out = [None] * len(items)
    with cf.ThreadPoolExecutor(8) as ex:
        for i, res in enumerate(ex.map(work, chunks)):
            for k, r in enumerate(res):
                out[i + 8 * k] = r
    return out


@pytest.mark.parametrize("R", SIZES)
def test_plugin_shape_psf(torch_cuda, R):
    print(f"case: N = 128, R = {R}, PSF", flush=True)
    a, ctx, m = _models(R)
    info = np.zeros(5, np.int32)
    Z.lib().zola_ctx_info(ctx._h, info.ctypes.data)
    assert info[0] == 128 and info[2] == m.s2 == (32 if R <= 16 else 64)
    rng = np.random.default_rng(R)
    B = 24
    par = np.stack([rng.uniform(-0.11, 0.11, B), rng.uniform(-0.11, 0.11, B), rng.uniform(-0.8, 0.8, B),
                    rng.uniform(-0.2, 0.2, B)], 1)
    got = ctx.psf_forward(par)
    ctx.close()
    err = 0.0
    for b in range(B):
        x, y, z, zo = par[b].astype(np.float32).astype(np.float64)
        ref = m.psf(x, y, zo, z)
        err = max(err, np.abs(got[b] - ref).max() / ref.max())
    print("max |psf - fp64| / peak", err)
    assert err < 1.5e-6


@pytest.mark.parametrize("layout", [Z.LAYOUT_QUAD, Z.LAYOUT_OCT])
@pytest.mark.parametrize("R", SIZES)
def test_plugin_shape_lm_fit(torch_cuda, R, layout):
    print(f"case: N = 128, R = {R}, LM fit, layout {layout}", flush=True)
    a, ctx, m = _models(R)
    ctx.set_fit_layout(R, Z.FIT_LM, layout)
    B = 16
    rois, truth = _rois(m, B, seed=500 + R)
    init = truth.copy()
    init[:, 0] *= 0.8
    init[:, 2:4] = 0.0
    init[:, 4] += 0.15
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=ZR[0] - 0.3, zmax=ZR[1] + 0.3, tol=1e-4)
    res = ctx.fit(cfg, rois, init)
    ctx.close()
    ref = _pool(R, a, lambda mm, b: mm.lm_fit(rois[b].astype(np.float64), init[b], zmin=ZR[0] - 0.3,
                                              zmax=ZR[1] + 0.3, maxit=300, tol=1e-10), list(range(B)))
    th = np.array([r["theta"] for r in ref])
    nll = np.array([r["nll"] for r in ref])
    d = np.abs(res["theta"][:, 2:] - th[:, 2:]) * 1000
    crlb_ref = np.array([np.sqrt(np.diag(np.linalg.inv(m.fisher(res["theta"][b])))[2:]) for b in range(B)])
    crel = np.abs(res["crlb"] - crlb_ref) / crlb_ref
    print("|d| nm max", d.max(axis=0), "NLL rel", (np.abs(res["nll"] - nll) / np.abs(nll)).max(), "CRLB rel",
          crel.max(), "iters", res["iters"].tolist())
    assert ((res["status"] & Z.ST_CONVERGED) != 0).all()
    assert d[:, :2].max() <= 0.5 and d[:, 2].max() <= 1.0
    assert np.all(np.abs(res["nll"] - nll) <= 1e-5 * np.abs(nll))
    assert crel.max() <= 1e-3


@pytest.mark.parametrize("R", SIZES)
def test_plugin_shape_reference_form(torch_cuda, R):
    print(f"case: N = 128, R = {R}, reference form", flush=True)
    a, ctx, m = _models(R)
    B = 8
    rois, truth = _rois(m, B, seed=600 + R)
    init = truth.copy()
    init[:, 4] += 0.1
    cfg = ctx.fit_config(mode=Z.FIT_REFERENCE, max_iter=20, zmin=ZR[0] - 0.3, zmax=ZR[1] + 0.3, photon_threshold=0.0)
    res = ctx.fit(cfg, rois, init)
    ctx.close()
    O.lib().zo_set_psf_cache(1)
    try:
        refs = _pool(R, a, lambda mm, b: mm.localize(rois[b].astype(np.float64), init[b], minZ=ZR[0] - 0.3,
                                                     maxZ=ZR[1] + 0.3, iterMax=20), list(range(B)))
    finally:
        O.lib().zo_set_psf_cache(0)
    refparity.compare(res, refs, 20, "N 128 R %d" % R)


def test_plugin_default_frame_walk(torch_cuda):
    """Detection and the acceptance walk with the plugin's default patch (32 on the
    128 grid) on three 128x128 frames of isolated emitters: the same candidates
    accepted in the same order as the FP64 oracle's walk, positions to 0.5 nm (x,
    y) and 1 nm (z) where the oracle's fit converged (tests/refparity.py's bar
    where it stopped at its cap), photons 1e-3 relative."""
    import simdata
    print("case: frame walk, N = 128, R = 32, 3 frames 128x128", flush=True)
    R, W = 32, 128
    a, ctx, m = _models(R)
    frames = []
    for f in range(3):
        em = simdata.random_emitters(W, W, 7, 31 + 97 * f, margin=18, zrange=(-0.5, 0.5), photons=(1500, 6000),
                                     min_sep=20.0)
        frames.append(simdata.render_frame(m, W, W, em, 12.0, seed=77 + f))
    frames = np.stack(frames)
    zlim = (-0.7, 0.7)
    det = Z.Detector(ctx, W, W, zlim[0], zlim[1], 0.1)
    fit = ctx.fit_config(mode=Z.FIT_REFERENCE, max_iter=20, zmin=zlim[0] - 0.3, zmax=zlim[1] + 0.3,
                         photon_threshold=0.0)
    acc = det.accept_config(photon_threshold=500.0, min_z=zlim[0], max_z=zlim[1])
    locs, per, rounds = det.localize(frames, fit, acc)
    det.close()
    ctx.close()

    def run(f):
        mm = O.Model(O.optics(size_fft=128), R, a)
        od = O.Detector(mm, W, W, zlim[0], zlim[1], 0.1, thr=0.1)
        r, its = O.frame_localize(mm, od, frames[f].astype(np.float64), 0.0, zlim[0], zlim[1], with_iters=True)
        r[:, 0] = f
        return r, its

    O.lib().zo_set_psf_cache(1)
    try:
        with cf.ThreadPoolExecutor(3) as ex:
            out = list(ex.map(run, range(3)))
    finally:
        O.lib().zo_set_psf_cache(0)
    ref = np.concatenate([o[0] for o in out])
    its = np.concatenate([o[1] for o in out])
    print("records gpu", len(locs), "oracle", len(ref), "per frame", per.tolist())
    assert len(locs) == len(ref) and len(ref) >= 12
    np.testing.assert_array_equal(locs[:, [0, 11]], ref[:, [0, 11]])
    d = np.abs(locs[:, 1:4] - ref[:, 1:4])
    conv = its < 20
    print("max |d| nm x, y, z (converged):", d[conv].max(axis=0), "capped", int((~conv).sum()))
    assert d[conv][:, :2].max() <= 0.5 and d[conv][:, 2].max() <= 1.0
    if (~conv).any():
        assert d[~conv].max() <= refparity.CAPPED_NM
    np.testing.assert_allclose(locs[conv, 4], ref[conv, 4], rtol=1e-3)
