"""Edge cases of the GPU entry points: empty and single-ROI batches, ROIs
without signal, very bright emitters, starts at the z clamp, an empty group in
a mixed-size fit, empty frame batches.  Results are checked against the FP64
oracle where one exists, otherwise against what the model fixes (status
flags, clamps, no NaN in a converged fit)."""
import numpy as np
import pytest

import oracle as O
import simdata
import zola_hip as Z

pytestmark = pytest.mark.gpu

R = 21


@pytest.fixture(scope="module")
def env():
    a = Z.astigmatic_coeffs(28)
    ctx = Z.Context(Z.optics(), R, a)
    m = O.Model(O.optics(), R, a)
    yield ctx, m, a
    ctx.close()


def test_empty_batch_is_a_no_op(env):
    ctx, m, a = env
    print("case: fit with B = 0", flush=True)
    for mode in (Z.FIT_LM, Z.FIT_REFERENCE):
        cfg = ctx.fit_config(mode=mode)
        res = ctx.fit(cfg, np.zeros((0, R, R), np.float32), np.zeros((0, 5)))
        assert res["theta"].shape == (0, 5)


@pytest.mark.parametrize("layout", [Z.LAYOUT_WAVE, Z.LAYOUT_QUAD, Z.LAYOUT_PAIR])
def test_single_roi_matches_oracle(env, layout):
    ctx, m, a = env
    print(f"case: one ROI, layout {layout}", flush=True)
    rois, truth = simdata.isolated_rois(m, 1, seed=501)
    init = truth.copy()
    init[:, 2:4] += 0.03
    init[:, 4] += 0.08
    ctx.set_fit_layout(R, Z.FIT_LM, layout)
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
    res = ctx.fit(cfg, rois, init)
    ctx.set_fit_layout(R, Z.FIT_LM, Z.LAYOUT_AUTO)
    ref = m.lm_fit(rois[0].astype(np.float64), init[0], zmin=-1.3, zmax=1.3, maxit=500, tol=1e-10)
    d = (res["theta"][0, 2:] - ref["theta"][2:]) * 1000
    print("one ROI |d| nm:", d)
    assert (res["status"][0] & Z.ST_CONVERGED) and np.abs(d[:2]).max() <= 0.5 and abs(d[2]) <= 1.0


def test_roi_without_signal_does_not_corrupt_its_neighbours(env):
    """An all-zero ROI between two normal ones: its fit ends flagged or with A at
    the clamp, never with NaN in theta; the neighbours' fits are unaffected."""
    ctx, m, a = env
    print("case: zero ROI between two normal ROIs", flush=True)
    rois, truth = simdata.isolated_rois(m, 3, seed=502)
    init = truth.copy()
    init[:, 4] += 0.05
    alone = ctx.fit(ctx.fit_config(mode=Z.FIT_LM), rois[[0, 2]], init[[0, 2]])
    rois[1] = 0.0
    init[1] = [1000.0, 10.0, 0.0, 0.0, 0.0]
    for mode in (Z.FIT_LM, Z.FIT_REFERENCE):
        res = ctx.fit(ctx.fit_config(mode=mode, max_iter=30 if mode == Z.FIT_LM else 20), rois, init)
        print("mode", mode, "zero-ROI theta", res["theta"][1], "status", res["status"][1])
        th = res["theta"][1]
        assert np.all(np.isfinite(th[2:])) or (res["status"][1] & Z.ST_NAN)
        if mode == Z.FIT_LM:
            # no photons: the likelihood sum(mu) falls to the clamps A = 1, B = 0.01
            assert th[0] < 2.0 and th[1] < 0.02
            np.testing.assert_array_equal(res["theta"][[0, 2]], alone["theta"])


def test_bright_emitter_converges(env):
    ctx, m, a = env
    print("case: 1e6-photon emitter", flush=True)
    rng = np.random.default_rng(503)
    truth = np.array([1e6, 50.0, 0.04, -0.03, 0.3])
    roi = rng.poisson(truth[0] * m.psf(truth[2], truth[3], 0.0, truth[4]) + truth[1]).astype(np.float32)[None]
    init = truth.copy()[None]
    init[0, 0] *= 0.7
    init[0, 2:] += [0.02, 0.02, 0.06]
    res = ctx.fit(ctx.fit_config(mode=Z.FIT_LM, max_iter=30, tol=1e-4), roi, init)
    ref = m.lm_fit(roi[0].astype(np.float64), init[0], zmin=-1.3, zmax=1.3, maxit=500, tol=1e-10)
    d = (res["theta"][0, 2:] - ref["theta"][2:]) * 1000
    print("bright |d| nm:", d, "iters", res["iters"][0])
    assert res["status"][0] & Z.ST_CONVERGED
    assert np.abs(d[:2]).max() <= 0.5 and abs(d[2]) <= 1.0
    assert abs(res["theta"][0, 0] / truth[0] - 1) < 0.01


def test_start_at_the_z_clamp_stays_inside(env):
    ctx, m, a = env
    print("case: starts at zmin / zmax", flush=True)
    rois, truth = simdata.isolated_rois(m, 4, seed=504, zrange=(-0.95, 0.95))
    init = truth.copy()
    init[:2, 4] = -1.0
    init[2:, 4] = 1.0
    for mode in (Z.FIT_LM, Z.FIT_REFERENCE):
        res = ctx.fit(ctx.fit_config(mode=mode, max_iter=30 if mode == Z.FIT_LM else 20, zmin=-1.0, zmax=1.0),
                      rois, init)
        print("mode", mode, "z", res["theta"][:, 4], "truth", truth[:, 4])
        assert np.all(res["theta"][:, 4] >= -1.0 - 1e-12) and np.all(res["theta"][:, 4] <= 1.0 + 1e-12)


def test_mixed_fit_with_an_empty_group(env, torch_cuda):
    torch = torch_cuda
    ctx, m, a = env
    print("case: mixed fit, groups of 0 and 8 ROIs", flush=True)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    rois, truth = simdata.isolated_rois(m, 8, seed=505)
    init = truth.copy()
    init[:, 4] += 0.05

    def group(Rg, px, ini):
        B = px.shape[0]
        return dict(R=Rg, pixels=torch.as_tensor(px, device="cuda").contiguous(),
                    init=torch.as_tensor(ini, dtype=torch.float64, device="cuda").contiguous(),
                    theta=torch.zeros(B, 5, dtype=torch.float64, device="cuda"),
                    nll=torch.zeros(B, dtype=torch.float64, device="cuda"),
                    crlb=torch.zeros(B, 3, dtype=torch.float64, device="cuda"),
                    iters=torch.zeros(B, dtype=torch.int32, device="cuda"),
                    status=torch.zeros(B, dtype=torch.int32, device="cuda"))

    groups = [group(17, np.zeros((0, 17, 17), np.float32), np.zeros((0, 5))), group(R, rois, init)]
    cfg = ctx.fit_config(mode=Z.FIT_LM)
    ctx.fit_mixed_dev(cfg, groups)
    torch.cuda.synchronize()
    sep = ctx.fit(cfg, rois, init)
    np.testing.assert_array_equal(groups[1]["theta"].cpu().numpy(), sep["theta"])


def test_zero_frames(env):
    ctx, m, a = env
    print("case: localize and detect with F = 0", flush=True)
    det = Z.Detector(ctx, 64, 64, -0.5, 0.5, 0.1)
    locs, per, rounds = det.localize(np.zeros((0, 64, 64), np.float32), ctx.fit_config(mode=Z.FIT_LM),
                                     det.accept_config())
    assert len(locs) == 0 and rounds == 0
    lists, counts = det.detect(np.zeros((0, 64, 64), np.float32), cap=16)
    assert len(lists) == 0
    det.close()
