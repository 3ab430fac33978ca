"""GPU calibration: one forward+adjoint pass (zola_calib_eval_dev) against the FP64
restatement of ZernikePhaseRetrieval.Model's likelihood (process/cpu/
ZernikePhaseRetrieval.java:1842-1879) and central finite differences of it
(h = 1e-4 rad), then a small L-BFGS calibration against an FP64 optimum."""
import numpy as np
import pytest

import calibdata
import oracle as O
import zola_hip as Z

pytestmark = pytest.mark.gpu

R = 31


def truth_coeffs(J, seed=3):
    rng = np.random.default_rng(seed)
    a = rng.normal(0, 0.15, J)
    a[:3] = 0.0                    # piston and tilts: registration terms
    a[Z.nm_to_slot(2, -2)] = 0.8   # Noll 5
    return a


def test_calib_gradient_matches_oracle_and_fd(torch_cuda):
    torch = torch_cuda
    J = 15
    a = truth_coeffs(J)
    m = O.Model(O.optics(), R, a)
    stacks, dz, beads = calibdata.bead_stacks(m, 2, 9, 0.1, seed=5)
    # evaluate at a perturbed point
    a1 = a + np.random.default_rng(9).normal(0, 0.02, J)
    a1[:3] = 0
    b1 = beads.copy()
    b1[:, :3] += 0.01
    b1[:, 3] *= 0.97
    b1[:, 4] += 0.5
    m.set_coeffs(a1)
    f_ref, ga_ref, gb_ref = calibdata.calib_nll_grad(m, stacks, dz, b1)
    # central FD of the FP64 likelihood in a few coefficients, h = 1e-4 rad
    for jz in [3, 5, 8, 14]:
        ap = a1.copy(); ap[jz] += 1e-4
        am = a1.copy(); am[jz] -= 1e-4
        m.set_coeffs(ap); fp, _, _ = calibdata.calib_nll_grad(m, stacks, dz, b1)
        m.set_coeffs(am); fm, _, _ = calibdata.calib_nll_grad(m, stacks, dz, b1)
        fd = (fp - fm) / 2e-4
        assert abs(fd - ga_ref[jz]) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(ga_ref).max())
    m.set_coeffs(a1)
    ctx = Z.Context(Z.optics(), R, a1)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    nb, S = stacks.shape[:2]
    par = np.zeros((nb * S, 4), np.float32)
    amp = np.zeros(nb * S); bg = np.zeros(nb * S)
    for b in range(nb):
        for s in range(S):
            par[b * S + s] = [b1[b, 0], b1[b, 1], 0.0, dz[s] + b1[b, 2]]
            amp[b * S + s] = b1[b, 3]
            bg[b * S + s] = b1[b, 4]
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    P, X, A_, B_ = dev(par), dev(stacks.reshape(-1, R, R)), dev(amp), dev(bg)
    nll = torch.zeros(nb * S, dtype=torch.float64, device="cuda")
    dpar = torch.zeros(nb * S, 4, device="cuda")
    dA = torch.zeros_like(nll); dB = torch.zeros_like(nll)
    da = torch.zeros(J, dtype=torch.float64, device="cuda")
    ctx.calib_eval_dev(P, X, A_, B_, nll, dpar, dA, dB, da)
    torch.cuda.synchronize()
    f = nll.sum().item()
    ga = da.cpu().numpy()
    dp = dpar.cpu().numpy().reshape(nb, S, 4).sum(1)
    gb = np.stack([dp[:, 0], dp[:, 1], dp[:, 3], dA.cpu().numpy().reshape(nb, S).sum(1),
                   dB.cpu().numpy().reshape(nb, S).sum(1)], 1)
    print("nll rel", abs(f - f_ref) / abs(f_ref), "ga rel", np.abs(ga - ga_ref).max() / np.abs(ga_ref).max())
    assert abs(f - f_ref) <= 1e-6 * abs(f_ref)
    assert np.abs(ga[3:] - ga_ref[3:]).max() <= 1e-5 * np.abs(ga_ref[3:]).max()
    np.testing.assert_allclose(gb, gb_ref, rtol=2e-3, atol=1e-3 * np.abs(gb_ref).max())


def test_calibration_recovers_coefficients():
    J = 15
    a = truth_coeffs(J)
    m = O.Model(O.optics(), R, a)
    stacks, dz, beads = calibdata.bead_stacks(m, 3, 21, 0.1, seed=8)
    ctx = Z.Context(Z.optics(), R, np.zeros(J))
    start = beads.copy()
    start[:, :3] = 0.0
    start[:, 3] = stacks.reshape(3, 21, -1).sum(2).mean(1) - R * R * 10.0
    start[:, 4] = 10.0
    mask = np.ones(J, np.int32); mask[:3] = 0
    res = ctx.calib_fit(stacks, dz, start, np.zeros(J), fit_mask=mask, max_iter=300, gtol=1e-3)
    print("iters", res["iters"], "evals", res["evals"], "coef err", np.abs(res["coeffs"] - a)[3:].max())
    # noise-limited recovery (2e5 photons per slice)
    assert np.abs(res["coeffs"] - a)[3:].max() < 0.02
    assert np.abs(res["bead"][:, :3] - beads[:, :3]).max() < 0.005
    # the GPU optimum is a stationary point of the FP64 likelihood
    m.set_coeffs(res["coeffs"])
    f_ref, ga_ref, gb_ref = calibdata.calib_nll_grad(m, stacks, dz, res["bead"])
    m.set_coeffs(a)
    f_true, _, _ = calibdata.calib_nll_grad(m, stacks, dz, beads)
    assert f_ref <= f_true + 1e-7 * abs(f_true)


# ---- the reference's full model at the calibration workload's size ----------
NB, S3, J3 = 5, 41, 45


def _config3(seed=3, sigma_true=1.0):
    a = truth_coeffs(J3, seed)
    m = O.Model(O.optics(), R, a)
    st, dz, tr = calibdata.full_stacks(m, NB, S3, 0.05, seed=seed + 10, sigma_true=sigma_true)
    return a, m, st, dz, tr


def test_full_model_config3_matches_fp64(torch_cuda):
    """One evaluation of zola_calib_eval_full_dev at a point off the truth (per-slice
    photon offsets, order-2 backgrounds, sCMOS maps, sigma = 1.05) against the FP64
    restatement: NLL to 1e-6 relative; the gradient in the coefficients, bead
    registration, background coefficients and photon offsets to 1e-4 of each
    block's largest component (the FP64 adjoint is pinned by central differences
    in tests/test_oracle_pins.py); dNLL/dsigma to 1e-3 of central differences of
    the FP64 NLL (h = 1e-3 px; the float-rounded Gaussian of the reference limits
    the difference quotient)."""
    torch = torch_cuda
    print("case: full calibration model, 5 beads x 41 slices x 31^2, J = 45", flush=True)
    a, m, st, dz, tr = _config3()
    rng = np.random.default_rng(21)
    a1 = a + rng.normal(0, 0.02, J3)
    a1[:3] = 0.0
    xyz = tr["xyz"] + 0.01
    poly = tr["poly"] * 1.02 + np.array([0.3, 0.01, 1e-4, -0.01, 1e-4, -1e-4])
    fa = np.full(NB, 400.0)
    fae = rng.normal(0, 300.0, (NB, S3))
    off = rng.uniform(0.5, 3.0, (NB, R, R))
    sig = 1.05
    m.set_coeffs(a1)
    m.set_sigma(sig)
    f_ref, g_ref = calibdata.full_nll_grad(m, st, dz, tr["A"], xyz, poly, fa, fae, off)
    m.set_sigma(sig + 1e-3)
    fp, _ = calibdata.full_nll_grad(m, st, dz, tr["A"], xyz, poly, fa, fae, off)
    m.set_sigma(sig - 1e-3)
    fm, _ = calibdata.full_nll_grad(m, st, dz, tr["A"], xyz, poly, fa, fae, off)
    m.set_sigma(1.0)
    dsig_fd = (fp - fm) / 2e-3

    ctx = Z.Context(Z.optics(), R, a1)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    ctx.set_sigma(sig)
    B = NB * S3
    par = np.zeros((B, 4), np.float32)
    amp = np.zeros(B)
    for b in range(NB):
        for s in range(S3):
            par[b * S3 + s] = [xyz[b, 0], xyz[b, 1], 0.0, dz[s] + xyz[b, 2]]
            amp[b * S3 + s] = tr["A"][b, s] + fa[b] + fae[b, s]
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    nll = torch.zeros(B, dtype=torch.float64, device="cuda")
    dpar = torch.zeros(B, 4, device="cuda")
    damp = torch.zeros_like(nll)
    dpoly = torch.zeros(B, 6, dtype=torch.float64, device="cuda")
    dsig = torch.zeros_like(nll)
    da = torch.zeros(J3, dtype=torch.float64, device="cuda")
    ctx.calib_eval_full_dev(NB, S3, dev(par), dev(st.reshape(B, R, R)), dev(amp), dev(poly), nll, dpar, damp,
                            offset=dev(off), dpoly=dpoly, dsigma=dsig, da=da)
    torch.cuda.synchronize()
    ctx.close()
    f = nll.sum().item()
    dp = dpar.cpu().numpy().reshape(NB, S3, 4)
    dA = damp.cpu().numpy().reshape(NB, S3)
    got = dict(a=da.cpu().numpy(), xyz=dp[:, :, [0, 1, 3]].sum(1), poly=dpoly.cpu().numpy().reshape(NB, S3, 6).sum(1),
               fa=dA.sum(1), fae=dA)
    ds = dsig.cpu().numpy().sum()
    rel = {k: float(np.abs(got[k] - g_ref[k]).max() / np.abs(g_ref[k]).max()) for k in got}
    rel["a"] = float(np.abs(got["a"][3:] - g_ref["a"][3:]).max() / np.abs(g_ref["a"][3:]).max())
    # background coefficients: each monomial on its own scale (i^2 reaches 900)
    rel["poly"] = float(max(np.abs(got["poly"][:, c] - g_ref["poly"][:, c]).max() / np.abs(g_ref["poly"][:, c]).max()
                            for c in range(6)))
    print("nll rel", abs(f - f_ref) / abs(f_ref), "gradient rel", rel, "dsigma", ds, "fd", dsig_fd,
          "rel", abs(ds - dsig_fd) / abs(dsig_fd))
    assert abs(f - f_ref) <= 1e-6 * abs(f_ref)
    for k, v in rel.items():
        assert v <= 1e-4, (k, v)
    assert abs(ds - dsig_fd) <= 1e-3 * abs(dsig_fd)


def test_full_calibration_recovers_config3():
    """zola_calib_fit_full on the calibration workload (blur sigma 1.1 px in the
    data): from zero coefficients, sigma 1, the beads at the patch centre, the
    photon curves as plain sums over the patch minus the background level and a
    flat background, every parameter of the reference's model fitted.

    A global defocus a_4 and a common z shift of all beads give nearly the same
    PSFs, so the data fix only their combination (the reference fits both too:
    its registration z and slot 4).  Checked: every coefficient but a_4 to the
    noise level (0.02 rad at 2e5 photons per slice), sigma to 1e-2 px, bead x, y
    and z0 relative to the beads' mean z0 to 5 nm, the fitted PSFs of every bead
    slice against the true ones to 5e-3 of the peak (a peak pixel holds ~1e4
    photons, so its Poisson noise is ~1e-2 of the peak), and the FP64 likelihood
    at the result not above its value at the truth."""
    print("case: full calibration fit, 5 beads x 41 slices x 31^2, J = 45", flush=True)
    a, m, st, dz, tr = _config3(seed=5, sigma_true=1.1)
    ctx = Z.Context(Z.optics(), R, np.zeros(J3))
    c0 = np.array([np.median(np.concatenate([st[b, :, 0, :].ravel(), st[b, :, -1, :].ravel()])) for b in range(NB)])
    A0 = st.reshape(NB, S3, -1).sum(2) - R * R * c0[:, None]
    poly0 = np.zeros((NB, 6))
    poly0[:, 0] = c0
    mask = np.ones(J3, np.int32)
    mask[:3] = 0
    res = ctx.calib_fit_full(st, dz, A0, np.zeros((NB, 3)), poly0, np.zeros(J3), sigma=1.0, fit_mask=mask,
                             max_iter=2000, gtol=1e-4)
    ctx.close()
    err = np.abs(res["coeffs"] - a)
    err[:3] = 0.0
    d = res["xyz"] - tr["xyz"]
    dz0 = d[:, 2] - d[:, 2].mean()
    m.set_coeffs(res["coeffs"])
    m.set_sigma(res["sigma"])
    fit_psf = [[m.psf(res["xyz"][b, 0], res["xyz"][b, 1], dz[s] + res["xyz"][b, 2], 0.0) for s in range(S3)]
               for b in range(NB)]
    f_fit, _ = calibdata.full_nll_grad(m, st, dz, A0, res["xyz"], res["poly"], res["fit_a"], res["fit_a_each"])
    m.set_coeffs(a)
    m.set_sigma(1.1)
    psf_err = max(np.abs(fit_psf[b][s] - (p := m.psf(tr["xyz"][b, 0], tr["xyz"][b, 1], dz[s] + tr["xyz"][b, 2], 0.0)))
                  .max() / p.max() for b in range(NB) for s in range(S3))
    f_true, _ = calibdata.full_nll_grad(m, st, dz, tr["A"], tr["xyz"], tr["poly"], np.zeros(NB), np.zeros((NB, S3)))
    m.set_sigma(1.0)
    print("iters", res["iters"], "evals", res["evals"], "coef err max (slot 4 aside)", np.delete(err, 4).max(),
          "slot 4", err[4], "sigma", res["sigma"], "xy err nm", np.abs(d[:, :2]).max() * 1000,
          "z0 err nm (common", d[:, 2].mean() * 1000, ") relative", np.abs(dz0).max() * 1000,
          "psf err / peak", psf_err, "nll fit - truth", f_fit - f_true)
    assert np.delete(err, 4).max() < 0.02
    assert abs(res["sigma"] - 1.1) < 1e-2
    assert np.abs(d[:, :2]).max() < 0.005 and np.abs(dz0).max() < 0.005
    assert psf_err < 5e-3
    assert f_fit <= f_true + 1e-7 * abs(f_true)
