"""Detection and the acceptance walk at the frame workload's optics, against the
FP64 oracle: 45-mode tetrapod PSF on the 128x128 pupil, 41x41 patches, 61
templates at 100 nm on [-3, 3.05) um, emitters at 0.5 per um^2 (patches overlap,
so the masks decide most outcomes), photons log-uniform 1e3..2e4, background
U[20, 100] per frame, Poisson noise.

* Detector (process/cpu/Predetection.java:63-593): candidate lists of 256x256
  frames identical to the oracle's -- count, (r, i, j) and order -- scores to
  1e-12, parabola offsets to 1e-6 relative.
* Walk (process/cpu/LocalizationPipeline.java:721-1073, reference fit form):
  on eight 128x128 crops the same candidates are accepted in the same order, x, y
  within 0.5 nm and z within 1 nm where the oracle's fit converged (the looser
  bar of tests/refparity.py where it stopped at the iteration cap), photons 1e-3,
  CRLB 2e-2 and score 1e-3 relative (5e-2 for capped fits), NLL 1e-5.  The GPU
  builds the accepted-model masks from its FP32 model at its own theta; the
  oracle from the FP64 model.  The test reports the
  smallest distance of any oracle mask decision from its threshold (model - B
  against 0.2 max), relative to the maximum, next to the FP32 model error
  (~1e-6 of the peak).
"""
import concurrent.futures as cf

import numpy as np
import pytest

import oracle as O
import refparity
import simdata
import zola_hip as Z

pytestmark = pytest.mark.gpu

R = 41
DX = simdata.DX
ZLIM = (-3.0, 3.05)


def _frame(m, W, H, seed):
    rng = np.random.default_rng(seed)
    n = rng.poisson(0.5 * (W * DX) * (H * DX))
    em = [(rng.uniform(0, W), rng.uniform(0, H), rng.uniform(-3, 3),
           float(np.exp(rng.uniform(np.log(1e3), np.log(2e4))))) for _ in range(n)]
    return simdata.render_frame(m, W, H, em, rng.uniform(20, 100), seed=seed + 1), n


@pytest.fixture(scope="module")
def setup(torch_cuda):
    coef = Z.tetrapod_coeffs(45)
    ctx = Z.Context(Z.optics(size_fft=128), R, coef)
    om = O.Model(O.optics(size_fft=128), R, coef)
    yield ctx, om, coef
    ctx.close()


def test_detector_candidates_identical_at_frame_optics(setup):
    ctx, om, coef = setup
    frames, ns = zip(*[_frame(om, 256, 256, 0xD1 + 7 * f) for f in range(2)])
    frames = np.stack(frames)
    print("case: detector 2 frames 256x256, emitters", ns, flush=True)
    det = Z.Detector(ctx, 256, 256, ZLIM[0], ZLIM[1], 0.1, 0.1, 0.0)
    assert det.nZ == 61
    lists, counts = det.detect(frames, cap=16384)

    def run(f):
        od = O.Detector(O.Model(O.optics(size_fft=128), R, coef), 256, 256, ZLIM[0], ZLIM[1], 0.1, thr=0.1)
        return od.run(frames[f].astype(np.float64), cap=1 << 16)

    with cf.ThreadPoolExecutor(2) as ex:
        refs = list(ex.map(run, range(2)))
    for f, ref in enumerate(refs):
        got = lists[f]
        print("frame", f, "candidates gpu", counts[f], "oracle", len(ref), flush=True)
        assert counts[f] == len(ref) and len(ref) > 1000
        np.testing.assert_array_equal(got[:, 1:4], ref[:, 1:4])
        np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=1e-12)
        np.testing.assert_allclose(got[:, 4:], ref[:, 4:], rtol=1e-6, atol=1e-12)
    det.close()


def _mask_margin(om, cand, recs):
    """Smallest |psf - 0.2 max psf| / max psf over the patches of the accepted
    records, the oracle's model at the oracle's theta (model - B = A psf, so the
    accepted-model mask is psf > 0.2 max psf)."""
    out = np.inf
    for r in recs:
        k = int(r[11])
        ci, cj = int(round(cand[k, 2])), int(round(cand[k, 3]))
        x, y = ci * DX - r[1] / 1000.0, cj * DX - r[2] / 1000.0
        p = om.psf(x, y, 0.0, r[3] / 1000.0)
        out = min(out, float(np.abs(p - 0.2 * p.max()).min() / p.max()))
    return out


def test_dense_walk_matches_oracle_at_frame_optics(setup):
    ctx, om, coef = setup
    W = H = 128
    F = 8
    frames, ns = zip(*[_frame(om, W, H, 0xE1 + 11 * f) for f in range(F)])
    frames = np.stack(frames)
    print("case: dense walk, reference form, %d frames 128x128, emitters" % F, ns, flush=True)
    det = Z.Detector(ctx, W, H, ZLIM[0], ZLIM[1], 0.1, 0.1, 0.0)
    fit = ctx.fit_config(mode=Z.FIT_REFERENCE, max_iter=20, zmin=ZLIM[0] - 0.3, zmax=ZLIM[1] + 0.3,
                         photon_threshold=0.0)
    acc = det.accept_config(photon_threshold=500.0, min_z=ZLIM[0], max_z=ZLIM[1])
    locs, per, rounds = det.localize(frames, fit, acc)
    det.close()

    def run(f):
        m = O.Model(O.optics(size_fft=128), R, coef)
        od = O.Detector(m, W, H, ZLIM[0], ZLIM[1], 0.1, thr=0.1)
        r, its = O.frame_localize(m, od, frames[f].astype(np.float64), 0.0, ZLIM[0], ZLIM[1], with_iters=True)
        r[:, 0] = f
        return r, od.run(frames[f].astype(np.float64), cap=1 << 16), its

    O.lib().zo_set_psf_cache(1)
    try:
        with cf.ThreadPoolExecutor(4) as ex:
            out = list(ex.map(run, range(F)))
        ref = np.concatenate([o[0] for o in out])
        its = np.concatenate([o[2] for o in out])
    finally:
        O.lib().zo_set_psf_cache(0)
    print("records gpu", len(locs), "oracle", len(ref), "per frame", per.tolist(), "fit rounds", rounds, flush=True)
    assert len(locs) == len(ref) and len(ref) >= 20
    np.testing.assert_array_equal(locs[:, [0, 11]], ref[:, [0, 11]])
    d = np.abs(locs[:, 1:4] - ref[:, 1:4])
    capped = its >= 20
    conv = ~capped
    print("max |d| nm x, y, z (converged fits):", d[conv].max(axis=0), "capped fits", int(capped.sum()),
          "max |d| nm capped:", d[capped].max(axis=0) if capped.any() else None)
    rel = lambda c: np.abs(locs[:, c] - ref[:, c]) / np.abs(ref[:, c])
    dA, dc, ds, dn = rel(4), np.abs(locs[:, 7:10] - ref[:, 7:10]) / ref[:, 7:10], rel(6), rel(10)
    print("converged: A rel", dA[conv].max(), "CRLB rel", dc[conv].max(), "score rel", ds[conv].max())
    if capped.any():
        print("capped: A rel", dA[capped].max(), "CRLB rel", dc[capped].max(), "score rel", ds[capped].max())
    print("NLL rel max", dn.max())
    assert d[conv][:, :2].max() <= 0.5 and d[conv][:, 2].max() <= 1.0
    assert dA[conv].max() <= 1e-3 and dc[conv].max() <= 2e-2 and ds[conv].max() <= 1e-3
    if capped.any():
        assert d[capped].max() <= refparity.CAPPED_NM
        assert max(dA[capped].max(), dc[capped].max(), ds[capped].max()) <= refparity.CAPPED_REL
    assert dn.max() <= 1e-5
    margins = [_mask_margin(om, out[f][1], ref[ref[:, 0] == f]) for f in range(F)]
    print("smallest mask-decision margin / max(model - B):", min(margins))
