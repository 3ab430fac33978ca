"""Pins of the oracle's calibration bead preparation (oracle/zola_beads.c,
restating process/InitBackgroundAndPhotonNumber.java with
utils/PolynomialFit.java and utils/GaussianElimination.java).  No reference
test holds vectors for this class, so each part is checked against what the
construction fixes in closed form:

* a stack whose background is an exact quadratic surface returns that surface
  and its coefficients in the reference's monomial order (1, x, x^2, y, xy,
  y^2 with x the first image index): pins the order, the elimination and the
  evaluation;
* a V-shaped per-slice background puts the centre slice at the V's vertex;
  bright voxels and their 5x5x5 neighbourhood leave the background fit;
* a constant background with a box-shaped bead gives B = background and the
  per-slice photon sums of the box, averaged over +-sizeFilt slices;
* a symmetric blob moves a nearby seed onto its centre; a seed beyond P/4
  stays;
* the camera conversions, and the sCMOS patches as the reference leaves them.
"""
import numpy as np
import pytest

import oracle as O


def _quad(W, H, c):
    x = np.arange(W, dtype=np.float64)[:, None]
    y = np.arange(H, dtype=np.float64)[None, :]
    return c[0] + c[1] * x + c[2] * x * x + c[3] * y + c[4] * x * y + c[5] * y * y


def test_quadratic_background_is_recovered():
    S, W, H = 9, 48, 40
    c = np.array([100.0, 0.5, 0.01, 0.3, -0.005, 0.008])
    bg = _quad(W, H, c)
    img = np.broadcast_to(bg, (S, W, H)).copy()
    r = O.bead_prepare(img, [], [], 16, 0.1)
    assert r["rc"] == 0
    assert r["ndata"] == W * H
    np.testing.assert_allclose(r["poly"], bg, rtol=1e-11)
    np.testing.assert_allclose(r["coef"], c, rtol=1e-6, atol=1e-9)
    # the stack moments of a stack of identical slices are those of one slice
    np.testing.assert_allclose(r["stats"][:2], [bg.mean(), bg.std()], rtol=1e-12)


def test_centre_slice_and_bright_voxels():
    S, W, H = 15, 40, 36
    zstep = 0.1  # sizeFilt = (int)(0.5/0.1)/2 = 2
    lev = 100.0 + 2.0 * np.abs(np.arange(S) - 7.0)
    img = np.repeat(lev[:, None, None], W * H, axis=1).reshape(S, W, H).copy()
    img[7, 20, 18] = 1e5  # one bright voxel in the centre slice
    r = O.bead_prepare(img, [], [], 16, zstep)
    assert r["size_filt"] == 2 and r["miniposition"] == 7
    # is3 drops the slices 5..9 of every pixel within 2 px of the bright voxel
    # (its 5x5x5 neighbourhood reaches slices 5..9 only): 5 x 5 pixels leave the fit
    assert r["ndata"] == W * H - 25
    # the kept per-pixel means are the mean level of slices 5..9
    np.testing.assert_allclose(r["poly"], lev[5:10].mean(), rtol=1e-10)


def test_centre_slice_is_clamped_to_the_filter_window():
    S, W, H = 12, 24, 24
    lev = 100.0 + 3.0 * np.arange(S)  # minimum at slice 0 -> clamped to sizeFilt
    img = np.repeat(lev[:, None, None], W * H, axis=1).reshape(S, W, H).copy()
    r = O.bead_prepare(img, [], [], 8, 0.05)  # sizeFilt = (int)(10.0)/2 = 5
    assert r["size_filt"] == 5 and r["miniposition"] == 5
    lev2 = lev[::-1].copy()  # minimum at the last slice -> S - sizeFilt - 1
    img2 = np.repeat(lev2[:, None, None], W * H, axis=1).reshape(S, W, H).copy()
    r2 = O.bead_prepare(img2, [], [], 8, 0.05)
    assert r2["miniposition"] == S - 5 - 1


def test_b_and_photon_curve_of_a_box_bead():
    S, W, H, P = 11, 64, 64, 16
    zstep = 0.1
    c = 40.0
    img = np.full((S, W, H), c)
    sig = np.array([50.0 + 30.0 * z for z in range(S)])  # per-slice signal per voxel
    # a 3x3 box bead at (30..32, 20..22); bright enough to leave is2
    for z in range(S):
        img[z, 30:33, 20:23] += sig[z] * 10
    r = O.bead_prepare(img, [31.0], [21.0], P, zstep)
    assert r["rc"] == 0
    np.testing.assert_allclose(r["B"], [c], rtol=1e-10)
    raw = 9 * sig * 10  # photon sum above the background, per slice
    sF = 2
    want = np.array([raw[max(0, z - sF):min(S, z + sF + 1)].mean() for z in range(S)])
    np.testing.assert_allclose(r["A"][0], want, rtol=1e-9)
    # patches: the stack inside the frame
    np.testing.assert_array_equal(r["patches"][0], img[:, 31 - 8:31 + 8, 21 - 8:21 + 8])


def test_patch_outside_the_frame_is_filled_with_b():
    S, W, H, P = 5, 32, 32, 16
    img = np.full((S, W, H), 25.0)
    img[:, 2, 3] += 500.0
    r = O.bead_prepare(img, [2.0], [3.0], P, 0.1)
    pt = r["patches"][0]
    # x0 = 2 - 8 = -6: rows u < 6 lie outside the frame
    np.testing.assert_allclose(pt[:, :6, :], r["B"][0], rtol=0)
    np.testing.assert_array_equal(pt[:, 6:, 5:], img[:, 0:10, 0:11])


def _blob(S, W, H, cx, cy, amp=400.0, s=1.3, bg=30.0):
    x = np.arange(W)[:, None] - cx
    y = np.arange(H)[None, :] - cy
    g = amp * np.exp(-(x * x + y * y) / (2 * s * s))
    return np.broadcast_to(bg + g, (S, W, H)).copy()


def test_seed_moves_onto_a_symmetric_blob():
    img = _blob(7, 64, 64, 30, 33)
    r = O.bead_prepare(img, [31.3], [32.2], 16, 0.1)
    assert abs(r["x"][0] - 30.0) < 1e-5 and abs(r["y"][0] - 33.0) < 1e-5


def test_seed_far_from_the_blob_stays():
    img = _blob(7, 64, 64, 30, 33)
    r = O.bead_prepare(img, [37.5], [33.0], 16, 0.1)
    # |x - centroid| >= P/4 = 4 on x (or no pixel above the threshold in the
    # inner window): x keeps its seed; y is refined onto the blob row
    assert r["x"][0] == np.float32(37.5)


def test_camera_conversions():
    rng = np.random.default_rng(3)
    adu = rng.uniform(0, 400, (3, 8, 6)).astype(np.float32)
    got = O.bead_convert(adu, 2.0, 4.0, 100.0)
    want = np.maximum(adu.astype(np.float64) * 0.5 - 50.0 + 1.0, 0.0)
    np.testing.assert_allclose(got, want, rtol=1e-15, atol=1e-13)
    off = rng.uniform(90, 110, (8, 6)).astype(np.float32)
    gain = rng.uniform(1.5, 2.5, (8, 6)).astype(np.float32)
    got = O.bead_convert_scmos(adu, off, gain)
    want = np.maximum((adu.astype(np.float64) - off) / gain.astype(np.float64), 0.0)
    np.testing.assert_array_equal(got, want)


def test_scmos_patches_hold_the_map_for_the_last_bead_only():
    S, W, H, P = 5, 48, 48, 8
    img = np.full((S, W, H), 20.0)
    img[:, 10, 10] += 300.0
    img[:, 35, 30] += 300.0
    vg = np.linspace(0.5, 3.0, W * H).reshape(W, H)
    r = O.bead_prepare(img, [10.0, 35.0], [10.0, 30.0], P, 0.1, vargain=vg)
    assert np.all(r["scmos"][0] == 0.0)
    np.testing.assert_array_equal(r["scmos"][1], vg[35 - 4:35 + 4, 30 - 4:30 + 4])
    # the image patches carry + var/gain^2 for every bead
    np.testing.assert_array_equal(r["patches"][0], img[:, 6:14, 6:14] + vg[6:14, 6:14])


def test_singular_background_fit_is_reported():
    img = np.full((3, 2, 2), 10.0)  # 4 pixels cannot fix 6 coefficients
    r = O.bead_prepare(img, [], [], 2, 0.1)
    assert r["rc"] == -1
