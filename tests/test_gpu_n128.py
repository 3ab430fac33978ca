"""128x128 pupil grid (ROI 41 x 41, tetrapod PSF from 45 modes) on the GPU
against the FP64 oracle: PSF synthesis, LM fits against the FP64 optimum
(0.5 nm x/y, 1 nm z, 1e-5 relative NLL) and the reference coordinate descent
against its FP64 restatement."""
import concurrent.futures as cf

import numpy as np
import pytest

import oracle as O
import refparity
import zola_hip as Z

pytestmark = pytest.mark.gpu
R = 41
ZR = (-1.5, 1.5)


def _opt():
    return dict(size_fft=128, xystep=0.11, wavelength=0.68, noil=1.518, nwat=1.518, na=1.45, sigma=1.0)


@pytest.fixture(scope="module")
def setup(torch_cuda):
    coef = Z.tetrapod_coeffs(45)
    ctx = Z.Context(Z.optics(**_opt()), R, coef)
    m = O.Model(O.optics(**_opt()), R, coef)
    yield ctx, m, coef
    ctx.close()


def _rois(m, B, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.11, 0.11, B)
    y = rng.uniform(-0.11, 0.11, B)
    z = rng.uniform(ZR[0] + 0.2, ZR[1] - 0.2, B)
    A = np.exp(rng.uniform(np.log(2000), np.log(20000), B))
    bg = rng.uniform(20, 100, B)
    rois = np.zeros((B, R, R), np.float32)
    for b in range(B):
        rois[b] = rng.poisson(A[b] * m.psf(x[b], y[b], 0.0, z[b]) + bg[b])
    return rois, np.stack([A, bg, x, y, z], axis=1)


def _pool(fn, items, coef):
    """FP64 oracle work on up to 8 threads, one oracle model per thread."""
    def work(chunk):
        mm = O.Model(O.optics(**_opt()), R, coef)
        return [fn(mm, it) for it in chunk]
    chunks = [items[i::8] for i in range(8)]
    out = [None] * len(items)
    with cf.ThreadPoolExecutor(8) as ex:
        for i, res in enumerate(ex.map(work, chunks)):
            for k, r in enumerate(res):
                out[i + 8 * k] = r
    return out


def test_psf128_matches_fp64(setup):
    ctx, m, _ = setup
    print("case: N=128 R=41 PSF parity", flush=True)
    params = np.array([[0.0, 0.0, 0.0, 0.0], [0.05, -0.03, 0.8, 0.0], [-0.1, 0.07, -1.2, 0.0],
                       [0.02, 0.01, 2.5, 0.0], [0.0, 0.0, -2.8, 0.4]], np.float32)
    got = ctx.psf_forward(params)
    err = 0.0
    for b, p in enumerate(params.astype(np.float64)):
        ref = m.psf(p[0], p[1], p[3], p[2])
        e = np.abs(got[b] - ref).max() / ref.max()
        err = max(err, e)
    print("max |psf - fp64| / peak =", err)
    assert err <= 1.5e-6


@pytest.mark.parametrize("layout", [Z.LAYOUT_QUAD, Z.LAYOUT_OCT])
def test_lm128_matches_fp64_optimum(setup, layout):
    ctx, m, coef = setup
    print(f"case: N=128 LM vs FP64 optimum, B=16, layout {layout}", flush=True)
    ctx.set_fit_layout(R, Z.FIT_LM, layout)
    B = 16
    rois, truth = _rois(m, B, seed=128)
    init = truth.copy()
    rng = np.random.default_rng(3)
    init[:, 0] *= rng.uniform(0.8, 1.2, B)
    init[:, 2:4] += rng.uniform(-0.05, 0.05, (B, 2))
    init[:, 4] += rng.uniform(-0.15, 0.15, B)
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=ZR[0] - 0.3, zmax=ZR[1] + 0.3, tol=1e-4)
    try:
        res = ctx.fit(cfg, rois, init)
    finally:
        ctx.set_fit_layout(R, Z.FIT_LM, Z.LAYOUT_AUTO)
    refs = _pool(lambda mm, b: mm.lm_fit(rois[b].astype(np.float64), init[b], zmin=ZR[0] - 0.3,
                                         zmax=ZR[1] + 0.3, maxit=200, tol=1e-10), list(range(B)), coef)
    th_ref = np.array([r["theta"] for r in refs])
    nll_ref = np.array([r["nll"] for r in refs])
    d = (res["theta"][:, 2:] - th_ref[:, 2:]) * 1000.0
    print("|d| nm max:", np.abs(d).max(axis=0), "iters", res["iters"])
    conv = (res["status"] & Z.ST_CONVERGED) != 0
    assert conv.all()
    assert np.abs(d[:, :2]).max() <= 0.5 and np.abs(d[:, 2]).max() <= 1.0
    assert np.all(np.abs(res["nll"] - nll_ref) <= 1e-5 * np.abs(nll_ref))
    # CRLB: the oracle's analytic Fisher matrix at the GPU's own optimum
    crlb_ref = np.array([np.sqrt(np.diag(np.linalg.inv(m.fisher(res["theta"][b])))[2:]) for b in range(B)])
    np.testing.assert_allclose(res["crlb"], crlb_ref, rtol=2e-3)
    np.testing.assert_allclose(res["crlb"], np.array([r["crlb"] for r in refs]), rtol=5e-3)


def test_ref128_matches_restatement(setup):
    ctx, m, coef = setup
    print("case: N=128 reference coordinate descent, B=8", flush=True)
    B = 8
    rois, truth = _rois(m, B, seed=129)
    init = truth.copy()
    init[:, 4] += 0.1
    cfg = ctx.fit_config(mode=Z.FIT_REFERENCE, max_iter=20, zmin=ZR[0] - 0.3, zmax=ZR[1] + 0.3,
                         photon_threshold=0.0)
    res = ctx.fit(cfg, rois, init)
    O.lib().zo_set_psf_cache(1)
    try:
        refs = _pool(lambda mm, b: mm.localize(rois[b].astype(np.float64), init[b], minZ=ZR[0] - 0.3,
                                               maxZ=ZR[1] + 0.3, iterMax=20), list(range(B)), coef)
    finally:
        O.lib().zo_set_psf_cache(0)
    refparity.compare(res, refs, 20, "N 128")


def test_lm128_per_roi_focus_index_mismatch(torch_cuda):
    torch = torch_cuda
    print("case: N=128 LM, n_sample 1.33, per-ROI focus U[0,4], B=8", flush=True)
    opt = dict(_opt(), nwat=1.33)
    coef = Z.tetrapod_coeffs(45)
    ctx = Z.Context(Z.optics(**opt), R, coef)
    m = O.Model(O.optics(**opt), R, coef)
    B = 8
    rng = np.random.default_rng(130)
    zf = rng.uniform(0.0, 4.0, B)
    truth = np.stack([np.exp(rng.uniform(np.log(3000), np.log(30000), B)), rng.uniform(20, 80, B),
                      rng.uniform(-0.1, 0.1, B), rng.uniform(-0.1, 0.1, B), rng.uniform(-1.0, 1.0, B)], 1)
    rois = np.stack([rng.poisson(t[0] * m.psf(t[2], t[3], zf[b], t[4]) + t[1]) for b, t in enumerate(truth)]).astype(np.float32)
    init = truth.copy()
    init[:, 4] += 0.1
    init[:, 2] += 0.02
    d = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")
    th = torch.zeros(B, 5, dtype=torch.float64, device="cuda")
    nl = torch.zeros(B, dtype=torch.float64, device="cuda")
    cr = torch.zeros(B, 3, dtype=torch.float64, device="cuda")
    it = torch.zeros(B, dtype=torch.int32, device="cuda")
    st = torch.zeros(B, dtype=torch.int32, device="cuda")
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.5, zmax=1.5, tol=1e-4)
    ctx.fit_zoil_dev(cfg, d(rois, torch.float32), d(init, torch.float64), d(zf, torch.float32), th, nl, cr, it, st)
    torch.cuda.synchronize()
    mm = O.Model(O.optics(**opt), R, coef)
    ref = np.array([mm.lm_fit(rois[b].astype(np.float64), init[b], zfocus=float(np.float32(zf[b])), zmin=-1.5,
                              zmax=1.5, maxit=200, tol=1e-10)["theta"] for b in range(B)])
    dd = (th.cpu().numpy()[:, 2:] - ref[:, 2:]) * 1000
    print("per-ROI focus |d| nm max:", np.abs(dd).max(axis=0))
    assert np.abs(dd[:, :2]).max() <= 0.5 and np.abs(dd[:, 2]).max() <= 1.0
    ctx.close()
