"""The 256 pupil grid: the plugin's default calibration geometry (sizeFFT 140,
rounded up to N = 256 by PhaseParameters.java:278-296; 64 px bead patches, so
s2 = 128 by PSFPhaseMany.java:102-108), in FP64 on the GPU, against the FP64
oracle:

* PSF synthesis (float output: 2e-7 of the peak);
* the adjoint for a random dL/dpsf (float dparams: 1e-6; FP64 dL/da: 1e-9);
* the full calibration likelihood and its gradient (FP64: 1e-9 relative NLL,
  1e-6 of each block's largest gradient component; the float dparams limit the
  registration block), dNLL/dsigma against central differences of the FP64 NLL;
* a calibration fit on the 256 grid recovering the coefficients;
* localization entry points refuse N = 256 contexts.
"""
import numpy as np
import pytest

import calibdata
import oracle as O
import zola_hip as Z

pytestmark = pytest.mark.gpu

SIZE_FFT, R = 140, 64


def _coeffs(J, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(0, 0.15, J)
    a[:3] = 0.0
    a[Z.nm_to_slot(2, -2)] = 0.8
    return a


def test_grid256_geometry_and_psf_match_oracle():
    print("case: N=256 PSF, R 64", flush=True)
    a = _coeffs(28, 1)
    ctx = Z.Context(Z.optics(size_fft=SIZE_FFT), R, a)
    m = O.Model(O.optics(size_fft=SIZE_FFT), R, a)
    assert (ctx.N, ctx.s2) == (256, 128) and (m.N, m.s2) == (256, 128)
    assert ctx.D == m.D
    rng = np.random.default_rng(4)
    B = 6
    par = np.stack([rng.uniform(-0.1, 0.1, B), rng.uniform(-0.1, 0.1, B), rng.uniform(-0.8, 0.8, B),
                    rng.uniform(-0.5, 0.5, B)], 1).astype(np.float32)
    got = ctx.psf_forward(par)
    ctx.close()
    for b in range(B):
        x, y, z, zo = par[b].astype(np.float64)
        ref = m.psf(x, y, zo, z)
        err = np.abs(got[b] - ref).max() / ref.max()
        print("psf", b, "err / peak", err)
        assert err < 2e-7


def test_grid256_adjoint_matches_oracle():
    print("case: N=256 adjoint, R 64, J 45", flush=True)
    J = 45
    a = _coeffs(J, 2)
    ctx = Z.Context(Z.optics(size_fft=SIZE_FFT), R, a)
    m = O.Model(O.optics(size_fft=SIZE_FFT), R, a)
    rng = np.random.default_rng(5)
    B = 4
    par = np.stack([rng.uniform(-0.1, 0.1, B), rng.uniform(-0.1, 0.1, B), rng.uniform(-0.8, 0.8, B),
                    rng.uniform(-0.3, 0.3, B)], 1).astype(np.float32)
    w = rng.normal(size=(B, R, R)).astype(np.float32)
    dp, da = ctx.psf_backward(par, w)
    ctx.close()
    ref_dp = np.zeros((B, 4))
    ref_da = np.zeros(J)
    for b in range(B):
        x, y, z, zo = par[b].astype(np.float64)
        g, ga = m.psf_vjp(x, y, zo, z, w[b].astype(np.float64))
        ref_dp[b] = g
        ref_da += ga
    err_p = np.abs(dp - ref_dp).max(axis=0) / np.abs(ref_dp).max(axis=0)
    err_a = np.abs(da - ref_da).max() / np.abs(ref_da).max()
    print("rel err dparams", err_p, "da", err_a)
    assert np.all(err_p < 1e-6)
    assert err_a < 1e-9


def test_grid256_full_calibration_model_matches_fp64(torch_cuda):
    torch = torch_cuda
    print("case: N=256 full calibration model, 2 beads x 5 slices x 64^2, J 45", flush=True)
    J, NB, S = 45, 2, 5
    a = _coeffs(J, 3)
    m = O.Model(O.optics(size_fft=SIZE_FFT), R, a)
    st, dz, tr = calibdata.full_stacks(m, NB, S, 0.1875, seed=13)
    rng = np.random.default_rng(22)
    a1 = a + rng.normal(0, 0.02, J)
    a1[:3] = 0.0
    # positions on a 2^-10 um grid and a z step of 3/16 um: the float32 parameters
    # of the GPU call then hold exactly the FP64 restatement's values
    xyz = np.round((tr["xyz"] + 0.01) * 1024) / 1024
    poly = tr["poly"] * 1.02 + np.array([0.3, 0.01, 1e-4, -0.01, 1e-4, -1e-4])
    fa = np.full(NB, 400.0)
    fae = rng.normal(0, 300.0, (NB, S))
    off = rng.uniform(0.5, 3.0, (NB, R, R))
    sig = 1.05
    m.set_coeffs(a1)
    m.set_sigma(sig)
    f_ref, g_ref = calibdata.full_nll_grad(m, st, dz, tr["A"], xyz, poly, fa, fae, off)
    m.set_sigma(sig + 1e-3)
    fp, _ = calibdata.full_nll_grad(m, st, dz, tr["A"], xyz, poly, fa, fae, off)
    m.set_sigma(sig - 1e-3)
    fm, _ = calibdata.full_nll_grad(m, st, dz, tr["A"], xyz, poly, fa, fae, off)
    dsig_fd = (fp - fm) / 2e-3

    ctx = Z.Context(Z.optics(size_fft=SIZE_FFT), R, a1)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.set_sigma(sig)
    B = NB * S
    par = np.zeros((B, 4), np.float32)
    amp = np.zeros(B)
    for b in range(NB):
        for s in range(S):
            par[b * S + s] = [xyz[b, 0], xyz[b, 1], 0.0, dz[s] + xyz[b, 2]]
            amp[b * S + s] = tr["A"][b, s] + fa[b] + fae[b, s]
    assert all(par[b * S + s, 3] == dz[s] + xyz[b, 2] for b in range(NB) for s in range(S))
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    nll = torch.zeros(B, dtype=torch.float64, device="cuda")
    dpar = torch.zeros(B, 4, device="cuda")
    damp = torch.zeros_like(nll)
    dpoly = torch.zeros(B, 6, dtype=torch.float64, device="cuda")
    dsig = torch.zeros_like(nll)
    da = torch.zeros(J, dtype=torch.float64, device="cuda")
    ctx.calib_eval_full_dev(NB, S, dev(par), dev(st.reshape(B, R, R)), dev(amp), dev(poly), nll, dpar, damp,
                            offset=dev(off), dpoly=dpoly, dsigma=dsig, da=da)
    torch.cuda.synchronize()
    ctx.close()
    f = nll.sum().item()
    dp = dpar.cpu().numpy().reshape(NB, S, 4)
    dA = damp.cpu().numpy().reshape(NB, S)
    got = dict(a=da.cpu().numpy(), xyz=dp[:, :, [0, 1, 3]].sum(1), poly=dpoly.cpu().numpy().reshape(NB, S, 6).sum(1),
               fa=dA.sum(1), fae=dA)
    ds = dsig.cpu().numpy().sum()
    rel = {k: float(np.abs(got[k] - g_ref[k]).max() / np.abs(g_ref[k]).max()) for k in got}
    rel["a"] = float(np.abs(got["a"][3:] - g_ref["a"][3:]).max() / np.abs(g_ref["a"][3:]).max())
    rel["poly"] = float(max(np.abs(got["poly"][:, c] - g_ref["poly"][:, c]).max() / np.abs(g_ref["poly"][:, c]).max()
                            for c in range(6)))
    print("nll rel", abs(f - f_ref) / abs(f_ref), "gradient rel", rel, "dsigma", ds, "fd", dsig_fd,
          "rel", abs(ds - dsig_fd) / abs(dsig_fd))
    assert abs(f - f_ref) <= 1e-9 * abs(f_ref)
    for k, v in rel.items():
        assert v <= 1e-6, (k, v)
    assert abs(ds - dsig_fd) <= 1e-3 * abs(dsig_fd)


def test_grid256_calibration_fit_recovers_coefficients():
    """zola_calib_fit_full on the 256 grid, 3 beads x 9 slices x 64^2, J = 21:
    every coefficient but the defocus slot 4 (degenerate with the beads' common
    z) to 0.03 rad, sigma to 2e-2 px."""
    print("case: N=256 calibration fit, 3 beads x 9 slices x 64^2, J 21", flush=True)
    J, NB, S = 21, 3, 9
    a = _coeffs(J, 7)
    m = O.Model(O.optics(size_fft=SIZE_FFT), R, a)
    st, dz, tr = calibdata.full_stacks(m, NB, S, 0.2, seed=17, sigma_true=1.1)
    ctx = Z.Context(Z.optics(size_fft=SIZE_FFT), R, np.zeros(J))
    c0 = np.array([np.median(np.concatenate([st[b, :, 0, :].ravel(), st[b, :, -1, :].ravel()])) for b in range(NB)])
    A0 = st.reshape(NB, S, -1).sum(2) - R * R * c0[:, None]
    poly0 = np.zeros((NB, 6))
    poly0[:, 0] = c0
    mask = np.ones(J, np.int32)
    mask[:3] = 0
    res = ctx.calib_fit_full(st, dz, A0, np.zeros((NB, 3)), poly0, np.zeros(J), sigma=1.0, fit_mask=mask,
                             max_iter=1000, gtol=1e-4)
    ctx.close()
    err = np.abs(res["coeffs"] - a)
    err[:3] = 0.0
    print("iters", res["iters"], "evals", res["evals"], "coef err max (slot 4 aside)", np.delete(err, 4).max(),
          "sigma", res["sigma"])
    assert np.delete(err, 4).max() < 0.03
    assert abs(res["sigma"] - 1.1) < 2e-2


def test_grid256_refuses_localization():
    ctx = Z.Context(Z.optics(size_fft=SIZE_FFT), R, _coeffs(15, 1))
    cfg = ctx.fit_config(mode=Z.FIT_LM)
    with pytest.raises(Z.ZolaError, match="N=256"):
        ctx.fit(cfg, np.zeros((2, R, R), np.float32), np.zeros((2, 5)))
    with pytest.raises(Z.ZolaError, match="N=256"):
        ctx.search_psf_center(0.0, 1.0)
    ctx.close()
