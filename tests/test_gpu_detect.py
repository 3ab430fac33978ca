"""Frame detector and frame localization on the GPU against the FP64 oracle.

Detector (process/cpu/Predetection.java): templates and correlation maps to FP64
rounding, candidate lists identical -- same count, same (r, i, j) in the same
order -- with scores to 1e-12 and parabola offsets to 1e-6 relative.
Frame walk (process/cpu/LocalizationPipeline.java:721-1073): with the reference
fit form, the same candidates are accepted in the same order, positions within
0.5 nm (x, y) and 1 nm (z) of the oracle's."""
import numpy as np
import pytest

import oracle as O
import simdata
import zola_hip as Z

pytestmark = pytest.mark.gpu
R = 21


@pytest.fixture(scope="module")
def models(torch_cuda):
    coef = Z.astigmatic_coeffs(28)
    ctx = Z.Context(Z.optics(size_fft=64), R, coef)
    om = O.Model(O.optics(size_fft=64), R, coef)
    yield ctx, om
    ctx.close()


def _frames(W, H, F, seed, n=5, bg=15.0):
    m = O.Model(O.optics(size_fft=64), R, Z.astigmatic_coeffs(28))
    out = []
    for f in range(F):
        em = simdata.random_emitters(W, H, n, seed + 97 * f, margin=12, zrange=(-0.5, 0.5),
                                     photons=(1500, 5000), min_sep=14.0)
        out.append(simdata.render_frame(m, W, H, em, bg, seed=seed + 31 * f))
    return np.stack(out)


def test_templates_and_range_match_fp64(models):
    ctx, om = models
    print("case: templates 96x80, z -0.7..0.7 step 0.1", flush=True)
    det = Z.Detector(ctx, 96, 80, -0.7, 0.7, 0.1)
    odet = O.Detector(om, 96, 80, -0.7, 0.7, 0.1)
    np.testing.assert_array_equal(det.range(), odet.range())
    t, ot = det.templates(), odet.templates()
    assert t.shape == ot.shape
    assert np.abs(t - ot).max() <= 1e-13 * ot.max()
    det.close()


def test_correlation_maps_match_fp64(models):
    ctx, om = models
    print("case: correlation maps 96x80 seed 3", flush=True)
    fr = _frames(96, 80, 1, seed=3)[0]
    det = Z.Detector(ctx, 96, 80, -0.5, 0.5, 0.1)
    odet = O.Detector(om, 96, 80, -0.5, 0.5, 0.1)
    corr = det.correlation(fr)
    _, ocorr = odet.run(fr.astype(np.float64), with_corr=True)
    err = np.abs(corr - ocorr).max()
    print("max |corr - oracle| =", err)
    assert err <= 1e-11
    det.close()


@pytest.mark.parametrize("W,H,seed", [(96, 96, 1), (96, 80, 2), (64, 128, 4)])
def test_candidate_lists_identical(models, W, H, seed):
    ctx, om = models
    print("case: candidates %dx%d seed %d" % (W, H, seed), flush=True)
    frames = _frames(W, H, 3, seed)
    det = Z.Detector(ctx, W, H, -0.6, 0.6, 0.1)
    odet = O.Detector(om, W, H, -0.6, 0.6, 0.1)
    lists, counts = det.detect(frames)
    for f in range(frames.shape[0]):
        ref = odet.run(frames[f].astype(np.float64))
        got = lists[f]
        assert counts[f] == len(ref) and len(ref) > 0
        np.testing.assert_array_equal(got[:, 1:4], ref[:, 1:4])
        np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=1e-12)
        # parabola offsets: -b/2a amplifies 1e-16 map differences where the
        # curvature is small, so relative (1e-6 of up to a few um: picometres)
        np.testing.assert_allclose(got[:, 4:], ref[:, 4:], rtol=1e-6, atol=1e-12)
    det.close()


def test_flat_frame_has_no_candidates(models):
    ctx, om = models
    print("case: flat frame", flush=True)
    det = Z.Detector(ctx, 64, 64, -0.3, 0.3, 0.1)
    lists, counts = det.detect(np.full((2, 64, 64), 7.0, np.float32))
    assert list(counts) == [0, 0]
    det.close()


def _score_and_nll(locs, ref):
    """Column 6: scoreCompute's chi2 of the accepted model (LocalizationPipeline.java:
    1099-1115); column 10: the fit's NLL.  Both depend on the fitted theta, which
    agrees to a fraction of a nanometre."""
    ds = np.abs(locs[:, 6] - ref[:, 6]) / np.abs(ref[:, 6])
    dn = np.abs(locs[:, 10] - ref[:, 10]) / np.abs(ref[:, 10])
    print("score rel max", ds.max(), "NLL rel max", dn.max())
    assert ds.max() <= 1e-3
    assert dn.max() <= 1e-5


def _oracle_walk(om, frames, zlim):
    recs = []
    for f in range(frames.shape[0]):
        det = O.Detector(om, frames.shape[1], frames.shape[2], zlim[0], zlim[1], 0.1)
        r = O.frame_localize(om, det, frames[f].astype(np.float64), 0.0, zlim[0], zlim[1])
        r[:, 0] = f
        recs.append(r)
    return np.concatenate(recs)


def test_frame_walk_reference_form_matches_oracle(models):
    ctx, om = models
    print("case: frame walk, reference fit form, 3 frames 96x96", flush=True)
    frames = _frames(96, 96, 3, seed=21, n=6)
    zlim = (-0.7, 0.7)
    det = Z.Detector(ctx, 96, 96, zlim[0], zlim[1], 0.1)
    fit = ctx.fit_config(mode=Z.FIT_REFERENCE, max_iter=20, zmin=zlim[0] - 0.3, zmax=zlim[1] + 0.3,
                         photon_threshold=0.0)
    acc = det.accept_config(photon_threshold=500.0, min_z=zlim[0], max_z=zlim[1])
    locs, per, rounds = det.localize(frames, fit, acc)
    ref = _oracle_walk(om, frames, zlim)
    print("records gpu", len(locs), "oracle", len(ref), "fit rounds", rounds)
    assert len(locs) == len(ref) and len(ref) >= 12
    np.testing.assert_array_equal(locs[:, [0, 11]], ref[:, [0, 11]])
    d = np.abs(locs[:, 1:4] - ref[:, 1:4])
    print("max |d| nm x, y, z:", d.max(axis=0))
    assert d[:, :2].max() <= 0.5 and d[:, 2].max() <= 1.0
    np.testing.assert_allclose(locs[:, 4], ref[:, 4], rtol=1e-3)
    np.testing.assert_allclose(locs[:, 7:10], ref[:, 7:10], rtol=2e-2)
    _score_and_nll(locs, ref)
    det.close()


def test_frame_walk_lm_form_finds_same_emitters(models):
    ctx, om = models
    print("case: frame walk, LM fit form, 3 frames 96x96", flush=True)
    frames = _frames(96, 96, 3, seed=21, n=6)
    zlim = (-0.7, 0.7)
    det = Z.Detector(ctx, 96, 96, zlim[0], zlim[1], 0.1)
    fit = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=zlim[0] - 0.3, zmax=zlim[1] + 0.3, tol=1e-4)
    acc = det.accept_config(photon_threshold=500.0, min_z=zlim[0], max_z=zlim[1])
    locs, per, rounds = det.localize(frames, fit, acc)
    ref = _oracle_walk(om, frames, zlim)
    assert len(locs) == len(ref)
    np.testing.assert_array_equal(locs[:, [0, 11]], ref[:, [0, 11]])
    d = np.abs(locs[:, 1:4] - ref[:, 1:4])
    print("LM vs reference form, max |d| nm x, y, z:", d.max(axis=0))
    assert d[:, :2].max() <= 5.0 and d[:, 2].max() <= 10.0
    det.close()


@pytest.mark.parametrize("N,Rc,cases", [(64, R, ((0.0, 1.0), (1.5, 0.8))), (128, 32, ((0.0, 1.0), (0.6, 0.8)))])
def test_search_psf_center_matches_oracle(torch_cuda, N, Rc, cases):
    """SearchPSFcenter: the FP64 device Fisher matrices of every scan, inverted
    on the host, pick the same depth as the oracle's host search (an argmin over
    hundreds of CRLB values, so the whole chain must agree)."""
    print(f"case: focus centre, {N} grid, astigmatic", flush=True)
    coef = Z.astigmatic_coeffs(28)
    ctx = Z.Context(Z.optics(size_fft=N, nwat=1.33), Rc, coef)
    m = O.Model(O.optics(size_fft=N, nwat=1.33), N, coef)
    for zf, ax in cases:
        c = ctx.search_psf_center(zf, ax)
        ref = O.search_psf_center(m, zf, ax)
        print("centre", zf, ax, c, ref)
        assert abs(c - ref) <= 1e-9
    ctx.close()


def _scmos_camera(frames, seed):
    """Raw frames of an sCMOS camera and its maps (offset, gain ADU/e-, read
    variance ADU^2), and the frames the pipeline derives from them (the
    library's conversion: FP64, then float32)."""
    rng = np.random.default_rng(seed)
    F, W, H = frames.shape
    off = rng.uniform(90, 110, (W, H)).astype(np.float32)
    gain = rng.uniform(1.8, 2.4, (W, H)).astype(np.float32)
    var = rng.uniform(4.0, 60.0, (W, H)).astype(np.float32)
    read = rng.normal(0.0, 1.0, frames.shape) * np.sqrt(var.astype(np.float64))
    adu = np.round(off + gain * frames.astype(np.float64) + read).astype(np.float32)
    e = (adu.astype(np.float64) - off.astype(np.float64)) / gain.astype(np.float64)
    vmap = (var.astype(np.float64) / (gain.astype(np.float64) ** 2)).astype(np.float32)
    v0 = float(var[0, 0]) / (float(gain[0, 0]) ** 2)
    det = np.maximum(e + v0, 0.0).astype(np.float32)
    img = (e + vmap.astype(np.float64)).astype(np.float32)
    return adu, off, gain, var, det, img, vmap


def test_frame_walk_scmos_matches_oracle(models):
    ctx, om = models
    print("case: sCMOS frame walk, reference fit form, 3 frames 96x96", flush=True)
    frames = _frames(96, 96, 3, seed=41, n=6)
    adu, off, gain, var, det_f, img_f, vmap = _scmos_camera(frames, 42)
    zlim = (-0.7, 0.7)
    det = Z.Detector(ctx, 96, 96, zlim[0], zlim[1], 0.1)
    fit = ctx.fit_config(mode=Z.FIT_REFERENCE, max_iter=20, zmin=zlim[0] - 0.3, zmax=zlim[1] + 0.3,
                         photon_threshold=0.0)
    acc = det.accept_config(photon_threshold=500.0, min_z=zlim[0], max_z=zlim[1])
    locs, per, rounds = det.localize_scmos(adu, off, gain, var, fit, acc)
    recs = []
    for f in range(3):
        od = O.Detector(om, 96, 96, zlim[0], zlim[1], 0.1)
        r = O.frame_localize_scmos(om, od, det_f[f].astype(np.float64), img_f[f].astype(np.float64),
                                   vmap.astype(np.float64), 0.0, zlim[0], zlim[1])
        r[:, 0] = f
        recs.append(r)
    ref = np.concatenate(recs)
    print("records gpu", len(locs), "oracle", len(ref), "fit rounds", rounds)
    assert len(locs) == len(ref) and len(ref) >= 12
    np.testing.assert_array_equal(locs[:, [0, 11]], ref[:, [0, 11]])
    d = np.abs(locs[:, 1:4] - ref[:, 1:4])
    print("max |d| nm x, y, z:", d.max(axis=0))
    assert d[:, :2].max() <= 0.5 and d[:, 2].max() <= 1.0
    np.testing.assert_allclose(locs[:, 4:6], ref[:, 4:6], rtol=1e-3)
    _score_and_nll(locs, ref)
    det.close()
