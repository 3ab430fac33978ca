"""Mixed ROI sizes fitted in one graph launch (zola_fit_mixed_dev) give, group by
group, exactly the results of separate zola_fit_batch_dev launches (same kernels,
same model: the 64 grid's tables and blur do not depend on the ROI side)."""
import numpy as np
import pytest

import oracle as O
import simdata
import zola_hip as Z

pytestmark = pytest.mark.gpu
SIZES = (17, 21, 31, 41)
COUNTS = (24, 18, 12, 6)


@pytest.mark.parametrize("mode", [Z.FIT_LM, Z.FIT_REFERENCE])
def test_mixed_groups_equal_separate_launches(torch_cuda, mode):
    torch = torch_cuda
    print("case: mixed sizes", SIZES, "mode", mode, flush=True)
    coef = Z.astigmatic_coeffs(28)
    opt = Z.optics(size_fft=64)
    ctx = Z.Context(opt, 21, coef)
    groups, ref = [], []
    for k, (R, B) in enumerate(zip(SIZES, COUNTS)):
        m = O.Model(O.optics(size_fft=64), R, coef)
        rois, truth = simdata.isolated_rois(m, B, seed=700 + k, photons=(200, 20000))
        init = truth.copy()
        init[:, 4] += 0.05
        d = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")
        g = dict(R=R, pixels=d(rois, torch.float32), init=d(init, torch.float64),
                 theta=torch.zeros(B, 5, dtype=torch.float64, device="cuda"),
                 nll=torch.zeros(B, dtype=torch.float64, device="cuda"),
                 crlb=torch.zeros(B, 3, dtype=torch.float64, device="cuda"),
                 iters=torch.zeros(B, dtype=torch.int32, device="cuda"),
                 status=torch.zeros(B, dtype=torch.int32, device="cuda"))
        groups.append(g)
        cR = Z.Context(opt, R, coef)
        cfg = cR.fit_config(mode=mode, max_iter=30 if mode == Z.FIT_LM else 20, zmin=-1.3, zmax=1.3,
                            tol=1e-4, photon_threshold=0.0)
        ref.append(cR.fit(cfg, rois, init))
        cR.close()
    cfg = ctx.fit_config(mode=mode, max_iter=30 if mode == Z.FIT_LM else 20, zmin=-1.3, zmax=1.3, tol=1e-4,
                         photon_threshold=0.0)
    for rep in range(2):   # second call reuses the executable graph
        ctx.fit_mixed_dev(cfg, groups)
        torch.cuda.synchronize()
        for g, r in zip(groups, ref):
            np.testing.assert_array_equal(g["theta"].cpu().numpy(), r["theta"])
            np.testing.assert_array_equal(g["nll"].cpu().numpy(), r["nll"])
            np.testing.assert_array_equal(g["status"].cpu().numpy(), r["status"])
            g["theta"].zero_()
    ctx.close()


def test_ragged_host_batch_equals_per_side_fits(torch_cuda):
    """zola_fit_ragged on interleaved ROIs of four sides: each ROI's results
    equal those of zola_fit_batch on its side's ROIs (same order within a side);
    an offset past the pixel array is refused before any device work."""
    print("case: ragged host batch", SIZES, flush=True)
    coef = Z.astigmatic_coeffs(28)
    opt = Z.optics(size_fft=64)
    ctx = Z.Context(opt, 21, coef)
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
    per = {}
    for k, (R, B) in enumerate(zip(SIZES, COUNTS)):
        m = O.Model(O.optics(size_fft=64), R, coef)
        rois, truth = simdata.isolated_rois(m, B, seed=900 + k, photons=(200, 20000))
        init = truth.copy()
        init[:, 4] += 0.05
        cR = Z.Context(opt, R, coef)
        per[R] = (rois, init, cR.fit(cR.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4),
                                      rois, init))
        cR.close()
    # interleave the sides: ROI b takes the next ROI of side SIZES[b % 4] while any is left
    order, nxt = [], {R: 0 for R in SIZES}
    while any(nxt[R] < B for R, B in zip(SIZES, COUNTS)):
        for R, B in zip(SIZES, COUNTS):
            if nxt[R] < B:
                order.append((R, nxt[R]))
                nxt[R] += 1
    rois = [per[R][0][i] for R, i in order]
    init = np.stack([per[R][1][i] for R, i in order])
    got = ctx.fit_ragged(cfg, rois, init)
    for b, (R, i) in enumerate(order):
        ref = per[R][2]
        assert np.array_equal(got["theta"][b], ref["theta"][i]), (b, R, i)
        assert got["nll"][b] == ref["nll"][i] and got["iters"][b] == ref["iters"][i]
        assert got["status"][b] == ref["status"][i]
        assert np.array_equal(got["crlb"][b], ref["crlb"][i])
    # a ROI reaching past the pixel array is an argument error
    import ctypes as C
    lib = Z.lib()
    offs = np.array([0, 441], np.int64)
    sizes = np.array([21, 21], np.int32)
    px = np.zeros(2 * 21 * 21 - 1, np.float32)
    out = [np.zeros((2, 5)), np.zeros(2), np.zeros((2, 3)), np.zeros(2, np.int32), np.zeros(2, np.int32)]
    rc = lib.zola_fit_ragged(ctx._h, C.byref(cfg), 2, Z._ptr(offs), Z._ptr(sizes), Z._ptr(px), px.size,
                             Z._ptr(init[:2].copy()), *[Z._ptr(a) for a in out])
    assert rc == -1 and b"outside" in lib.zola_last_error(ctx._h)
    ctx.close()
