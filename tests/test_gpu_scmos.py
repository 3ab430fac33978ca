"""Fits with an sCMOS map (var/gain^2 per pixel added to mu) against the FP64
oracle with the same map: LM form to the FP64 optimum (0.5 nm x/y, 1 nm z,
1e-5 relative NLL), reference form to the restatement of
cpu/Localization.localize (0.5 nm x/y, 1 nm z where both converge)."""
import numpy as np
import pytest

import oracle as O
import simdata
import refparity
import zola_hip as Z

pytestmark = pytest.mark.gpu

R = 21
B = 12


def _data():
    a = Z.astigmatic_coeffs(28)
    m = O.Model(O.optics(), R, a)
    rng = np.random.default_rng(77)
    truth = np.stack([np.exp(rng.uniform(np.log(800), np.log(5000), B)), rng.uniform(5, 40, B),
                      rng.uniform(-0.1, 0.1, B), rng.uniform(-0.1, 0.1, B), rng.uniform(-0.9, 0.9, B)], 1)
    varg = rng.uniform(0.5, 12.0, (B, R, R))
    rois = np.stack([rng.poisson(t[0] * m.psf(t[2], t[3], 0.0, t[4]) + t[1] + varg[b])
                     for b, t in enumerate(truth)]).astype(np.float32)
    tmpl = simdata.templates(m, -1.0, 0.1, 21)
    init = simdata.init_restated(rois - varg.astype(np.float32), tmpl, -1.0, 0.1)
    return a, m, rois, varg, init


def _fit(torch, ctx, cfg, rois, varg, init):
    d = lambda x, dt: torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device="cuda")
    th = torch.zeros(B, 5, dtype=torch.float64, device="cuda")
    nl = torch.zeros(B, dtype=torch.float64, device="cuda")
    cr = torch.zeros(B, 3, dtype=torch.float64, device="cuda")
    it = torch.zeros(B, dtype=torch.int32, device="cuda")
    st = torch.zeros(B, dtype=torch.int32, device="cuda")
    ctx.fit_ext_dev(cfg, d(rois, torch.float32), d(init, torch.float64), th, nl, cr, it, st,
                    varg=d(varg, torch.float32))
    torch.cuda.synchronize()
    return th.cpu().numpy(), nl.cpu().numpy(), it.cpu().numpy(), st.cpu().numpy()


@pytest.mark.parametrize("layout", [Z.LAYOUT_WAVE, Z.LAYOUT_QUAD, Z.LAYOUT_PAIR])
def test_lm_scmos_matches_fp64_optimum(torch_cuda, layout):
    print(f"case: sCMOS LM R {R} layout {layout}", flush=True)
    a, m, rois, varg, init = _data()
    ctx = Z.Context(Z.optics(), R, a)
    ctx.set_stream(torch_cuda.cuda.current_stream().cuda_stream)
    ctx.set_fit_layout(R, Z.FIT_LM, layout)
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
    th, nl, it, st = _fit(torch_cuda, ctx, cfg, rois, varg, init)
    ctx.close()
    ref, nref = [], []
    for b in range(B):
        O.set_scmos(varg[b].astype(np.float32).astype(np.float64))
        try:
            r = m.lm_fit(rois[b].astype(np.float64), init[b], zmin=-1.3, zmax=1.3, maxit=500, tol=1e-10)
        finally:
            O.set_scmos(None)
        ref.append(r["theta"])
        nref.append(r["nll"])
    ref, nref = np.array(ref), np.array(nref)
    d = (th[:, 2:] - ref[:, 2:]) * 1000
    print("sCMOS LM |d| nm max:", np.abs(d).max(axis=0), "iters", it)
    assert ((st & Z.ST_CONVERGED) != 0).all()
    assert np.abs(d[:, :2]).max() <= 0.5 and np.abs(d[:, 2]).max() <= 1.0
    assert np.all(np.abs(nl - nref) <= 1e-5 * np.abs(nref))


def test_reference_form_scmos_matches_restatement(torch_cuda):
    print(f"case: sCMOS reference form R {R}", flush=True)
    a, m, rois, varg, init = _data()
    ctx = Z.Context(Z.optics(), R, a)
    ctx.set_stream(torch_cuda.cuda.current_stream().cuda_stream)
    cfg = ctx.fit_config(mode=Z.FIT_REFERENCE, max_iter=20, zmin=-1.3, zmax=1.3, photon_threshold=0.0)
    th, nl, it, st = _fit(torch_cuda, ctx, cfg, rois, varg, init)
    ctx.close()
    O.lib().zo_set_psf_cache(1)
    ref = []
    try:
        for b in range(B):
            O.set_scmos(varg[b].astype(np.float32).astype(np.float64))
            ref.append(m.localize(rois[b].astype(np.float64), init[b], minZ=-1.3, maxZ=1.3, iterMax=20))
    finally:
        O.set_scmos(None)
        O.lib().zo_set_psf_cache(0)
    refparity.compare(dict(theta=th, nll=nl, iters=it), ref, 20, "sCMOS")
