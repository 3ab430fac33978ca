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
    assert np.abs(ga[3:] - ga_ref[3:]).max() <= 1e-3 * np.abs(ga_ref[3:]).max()
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
