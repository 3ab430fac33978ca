"""Calibration bead preparation on the GPU (zola_bead_prepare and
zola_bead_prepare_scmos) against the FP64 oracle restating
process/InitBackgroundAndPhotonNumber.java on the same raw stacks.

Bit-exact: the refined seeds, the centre slice, the number of background
pixels, and every patch pixel inside the frame (the converted stack).  The
stack moments, per-slice means and the 6x6 normal equations are summed as
fixed-shape trees on the GPU and as one running sum in the reference, so the
polynomial background, B and A are compared to 1e-9 relative (the rounding of
the sums, amplified by the conditioning of the raw-coordinate normal
equations, stays orders of magnitude below that).
"""
import numpy as np
import pytest

import oracle as O
import zola_hip as Z

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Z.Context(Z.optics(), 21, Z.astigmatic_coeffs(28))
    yield c
    c.close()


def _stack(seed, S, W, H, beads, adu_factor=1.0, gain=1.0, offset=100.0):
    """Raw ADU stack: quadratic background with a slice-dependent level,
    Gaussian beads whose width grows away from focus, Poisson noise."""
    rng = np.random.default_rng(seed)
    x = np.arange(W)[:, None]
    y = np.arange(H)[None, :]
    bg = 30 + 0.05 * x + 0.03 * y + 2e-4 * x * x - 1e-4 * x * y + 1.5e-4 * y * y
    lam = np.empty((S, W, H))
    for z in range(S):
        lam[z] = bg + 0.8 * abs(z - S // 2)
        for (bx, by, n) in beads:
            s = 1.2 + 0.15 * abs(z - S // 2)
            lam[z] += n / (2 * np.pi * s * s) * np.exp(-((x - bx) ** 2 + (y - by) ** 2) / (2 * s * s))
    ph = rng.poisson(lam).astype(np.float64)
    return (ph * gain / adu_factor + offset).astype(np.float32)


def _compare(got, ref, nb, S, P, W, H, x0s, y0s):
    np.testing.assert_array_equal(got["x"], ref["x"])
    np.testing.assert_array_equal(got["y"], ref["y"])
    d = got["diag"]
    assert d["size_filt"] == ref["size_filt"]
    assert d["miniposition"] == ref["miniposition"]
    assert d["ndata"] == ref["ndata"]
    np.testing.assert_allclose([d["mean0"], d["std0"], d["mean"], d["std"]], ref["stats"], rtol=1e-12)
    np.testing.assert_allclose(got["background"], ref["poly"], rtol=1e-9)
    np.testing.assert_allclose(got["B"], ref["B"], rtol=1e-9)
    scale = np.abs(ref["A"]).max() if nb else 1.0
    np.testing.assert_allclose(got["A"], ref["A"], rtol=1e-9, atol=1e-9 * scale)
    for p in range(nb):
        x0 = int(ref["x"][p]) - P // 2
        y0 = int(ref["y"][p]) - P // 2
        u = np.arange(P)
        inside = ((x0 + u >= 0) & (x0 + u < W))[:, None] & ((y0 + u >= 0) & (y0 + u < H))[None, :]
        g, r = got["patches"][p], ref["patches"][p]
        np.testing.assert_array_equal(g[:, inside], r[:, inside])
        np.testing.assert_allclose(g[:, ~inside], r[:, ~inside], rtol=1e-9)


@pytest.mark.parametrize("S,W,H,P,zstep", [(41, 128, 96, 32, 0.05), (21, 64, 80, 16, 0.1)])
def test_ccd_stack_matches_oracle(ctx, S, W, H, P, zstep):
    print(f"case: bead prep CCD S {S} W {W} H {H} P {P}", flush=True)
    beads = [(30.4, 40.2, 4e4), (90.7, 20.5, 2e4), (60.0, 70.3, 3e4)] if W > 100 else [(20.3, 30.6, 3e4),
                                                                                       (45.2, 50.1, 2e4)]
    beads = [b for b in beads if b[0] < W and b[1] < H]
    adu = _stack(11 + S, S, W, H, beads, adu_factor=2.0, gain=3.0, offset=100.0)
    seeds = [(b[0] + 1.7, b[1] - 1.1) for b in beads]
    xp = [s[0] for s in seeds]
    yp = [s[1] for s in seeds]
    got = ctx.bead_prepare(adu, xp, yp, P, zstep, adu_factor=2.0, gain=3.0, offset=100.0)
    img = O.bead_convert(adu, 2.0, 3.0, 100.0)
    ref = O.bead_prepare(img, xp, yp, P, zstep)
    assert ref["rc"] == 0
    print("seeds", got["x"], got["y"], "mp", got["diag"]["miniposition"], "ndata", got["diag"]["ndata"])
    _compare(got, ref, len(beads), S, P, W, H, None, None)


def test_scmos_stack_matches_oracle(ctx):
    S, W, H, P, zstep = 25, 72, 64, 16, 0.08
    print("case: bead prep sCMOS", flush=True)
    rng = np.random.default_rng(5)
    beads = [(20.2, 25.7, 3e4), (50.6, 40.1, 2.5e4)]
    off = rng.uniform(95, 105, (W, H)).astype(np.float32)
    gain = rng.uniform(1.8, 2.4, (W, H)).astype(np.float32)
    var = rng.uniform(2, 30, (W, H)).astype(np.float32)
    ph = _stack(21, S, W, H, beads, offset=0.0).astype(np.float64)
    adu = (ph * gain + off).astype(np.float32)
    xp, yp = [21.1, 49.9], [24.8, 41.3]
    got = ctx.bead_prepare_scmos(adu, off, gain, var, xp, yp, P, zstep)
    img = O.bead_convert_scmos(adu, off, gain)
    vg = var.astype(np.float64) / (gain.astype(np.float64) * gain.astype(np.float64))
    ref = O.bead_prepare(img, xp, yp, P, zstep, vargain=vg)
    _compare(got, ref, 2, S, P, W, H, None, None)
    np.testing.assert_array_equal(got["scmos"], ref["scmos"])


def test_bead_at_the_frame_edge_and_short_stack(ctx):
    # S = 4 < 2 sizeFilt + 1: the centre slice is clamped below zero as the reference does
    S, W, H, P, zstep = 4, 48, 40, 16, 0.05
    print("case: bead prep edge bead, short stack", flush=True)
    beads = [(2.5, 3.1, 2e4), (44.0, 37.5, 2e4)]
    adu = _stack(31, S, W, H, beads)
    xp, yp = [2.0, 45.0], [3.0, 38.0]
    got = ctx.bead_prepare(adu, xp, yp, P, zstep, offset=100.0)
    ref = O.bead_prepare(O.bead_convert(adu, 1.0, 1.0, 100.0), xp, yp, P, zstep)
    assert ref["rc"] == 0
    _compare(got, ref, 2, S, P, W, H, None, None)


def test_no_beads(ctx):
    S, W, H = 9, 40, 40
    print("case: bead prep without beads", flush=True)
    adu = _stack(41, S, W, H, [])
    got = ctx.bead_prepare(adu, [], [], 16, 0.1, offset=100.0)
    ref = O.bead_prepare(O.bead_convert(adu, 1.0, 1.0, 100.0), [], [], 16, 0.1)
    np.testing.assert_allclose(got["background"], ref["poly"], rtol=1e-9)
    assert got["diag"]["ndata"] == ref["ndata"]


def test_singular_background_fit_raises(ctx):
    print("case: bead prep singular fit", flush=True)
    adu = np.full((3, 2, 2), 110.0, np.float32)
    with pytest.raises(Z.ZolaError, match="singular"):
        ctx.bead_prepare(adu, [], [], 2, 0.1, offset=100.0)
