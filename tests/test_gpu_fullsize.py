"""Parity at the benchmark's full sizes, through properties that do not depend
on the batch size (the FP64 oracle cannot fit a whole batch in test time):

* isolated 21x21 ROIs, 64x64 pupil, 65,536 per launch (the headline workload):
  - noise-free ROIs (n = mu exactly) return the parameters they were made from,
    for every ROI of the batch (0.5 nm x/y, 1 nm z);
  - each ROI's result does not depend on the batch around it: a 1,024-ROI
    subset fitted alone is bit-identical to the same ROIs inside the full batch;
  - a random sample of the batch matches the oracle's FP64 optimum (0.5 nm x/y,
    1 nm z, 1e-5 relative NLL);
  - on Poisson data the errors divided by the returned CRLB, over the emitters
    of 1,000 photons or more, have a robust (MAD) width within 5% of 1; for
    |z| < 0.6 um their mean is within 0.05 of 0 and their standard deviation
    within 5% of 1 (CRLB and fit checked together; the MLE is near-efficient
    there); fewer than 3% of the fits, mostly at |z| > 0.6 um, end in the
    other lobe of the astigmatic z response, where their template start lay;
* the same batch with a per-pixel sCMOS map: noise-free recovery for every ROI
  and batch independence;
* 41x41 ROIs on the 128x128 pupil with a per-ROI focus term, 65,536 per launch:
  noise-free recovery for every ROI and a sample against the oracle;
* 200,000 ROIs of mixed sides in one graph launch: every group equals its own
  launch, bit for bit, and noise-free ROIs return their parameters;
* 2,000 frames of 256x256 through detection and the acceptance walk: a few
  frames processed alone give the same candidate lists and records, bit for bit,
  as inside the whole run;
* the calibration bead stack of 101 slices of 512x512 against the oracle
  directly (it finishes in about a second there).
"""
import numpy as np
import pytest

import oracle as O
import zola_hip as Z

pytestmark = pytest.mark.gpu
DX = 0.11
ZMIN_T, ZSTEP_T, NZ = -1.0, 0.1, 21


def _outs(torch, B):
    f64 = lambda *s: torch.zeros(*s, dtype=torch.float64, device="cuda")
    return dict(theta=f64(B, 5), nll=f64(B), crlb=f64(B, 3),
                iters=torch.zeros(B, dtype=torch.int32, device="cuda"),
                status=torch.zeros(B, dtype=torch.int32, device="cuda"))


def _synth(torch, ctx, R, B, seed, noise, alo=500.0, ahi=5000.0):
    """Isolated ROIs as the benchmark draws them: x, y U(+-1 px), z U[-1, 1] um,
    photons log-uniform, background U[5, 50]; mu exactly when noise is False."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    u = lambda: torch.rand(B, device="cuda", generator=g, dtype=torch.float64)
    x, y, z = (2 * u() - 1) * DX, (2 * u() - 1) * DX, 2 * u() - 1
    A = torch.exp(np.log(alo) + (np.log(ahi) - np.log(alo)) * u())
    bg = 5 + 45 * u()
    par = torch.stack([x, y, z, torch.zeros_like(x)], 1).float().contiguous()
    psf = torch.empty(B, R, R, device="cuda")
    ctx.psf_forward_dev(par, psf)
    mu = A.float()[:, None, None] * psf + bg.float()[:, None, None]
    px = (torch.poisson(mu, generator=g) if noise else mu).float().contiguous()
    # the parameters the pixels were made from, at the precision they entered
    truth = torch.stack([A.float().double(), bg.float().double(), par[:, 0].double(), par[:, 1].double(),
                         par[:, 2].double()], 1)
    return px, truth


def _fit(torch, ctx, cfg, px, init):
    o = _outs(torch, px.shape[0])
    ctx.fit_dev(cfg, px, init, o["theta"], o["nll"], o["crlb"], o["iters"], o["status"])
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in o.items()}


@pytest.fixture(scope="module")
def ctx21(torch_cuda):
    c = Z.Context(Z.optics(), 21, Z.astigmatic_coeffs(28))
    c.set_stream(torch_cuda.cuda.current_stream().cuda_stream)
    yield c
    c.close()


def _init(torch, ctx, px):
    init = torch.empty(px.shape[0], 5, dtype=torch.float64, device="cuda")
    ctx.fit_init_dev(px, NZ, ZMIN_T, ZSTEP_T, 0.0, init)
    return init


def test_headline_batch_noise_free_recovers_every_roi(torch_cuda, ctx21):
    torch = torch_cuda
    B = 65536
    print(f"case: {B} noise-free 21x21 ROIs", flush=True)
    px, truth = _synth(torch, ctx21, 21, B, 0x51, noise=False)
    cfg = ctx21.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
    res = _fit(torch, ctx21, cfg, px, _init(torch, ctx21, px))
    t = truth.cpu().numpy()
    d = np.abs(res["theta"][:, 2:] - t[:, 2:]) * 1000
    print("max |d| nm", d.max(axis=0), "converged", np.mean((res["status"] & Z.ST_CONVERGED) != 0))
    assert np.all((res["status"] & Z.ST_CONVERGED) != 0)
    assert d[:, :2].max() <= 0.5 and d[:, 2].max() <= 1.0
    # A stops within 1% of its statistical precision (at least sqrt(A)) per step
    assert np.all(np.abs(res["theta"][:, 0] - t[:, 0]) <= 0.05 * np.sqrt(t[:, 0]))


def test_headline_batch_results_do_not_depend_on_the_batch(torch_cuda, ctx21):
    torch = torch_cuda
    B = 65536
    print(f"case: subset of {B} fitted alone", flush=True)
    px, _ = _synth(torch, ctx21, 21, B, 0x52, noise=True)
    init = _init(torch, ctx21, px)
    cfg = ctx21.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
    full = _fit(torch, ctx21, cfg, px, init)
    idx = torch.from_numpy(np.sort(np.random.default_rng(3).choice(B, 1024, replace=False))).cuda()
    sub = _fit(torch, ctx21, cfg, px[idx].contiguous(), init[idx].contiguous())
    ii = idx.cpu().numpy()
    for k in ("theta", "nll", "crlb", "iters", "status"):
        np.testing.assert_array_equal(sub[k], full[k][ii])


def test_headline_batch_sample_matches_oracle_and_pulls_match_crlb(torch_cuda, ctx21):
    torch = torch_cuda
    B = 65536
    print(f"case: {B} Poisson 21x21 ROIs, oracle sample and CRLB pulls", flush=True)
    px, truth = _synth(torch, ctx21, 21, B, 0x53, noise=True)
    init = _init(torch, ctx21, px)
    cfg = ctx21.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
    res = _fit(torch, ctx21, cfg, px, init)
    t = truth.cpu().numpy()
    ok = (res["status"] & Z.ST_CONVERGED) != 0
    assert ok.mean() > 0.999
    ok &= t[:, 0] >= 1000.0  # efficiency is asymptotic: keep emitters of 1,000 photons or more
    pull = (res["theta"][ok, 2:] - t[ok, 2:]) / res["crlb"][ok]
    zt = t[ok, 4]
    # fits whose template start lay in the other lobe of the astigmatic z
    # response converge to that local optimum (as the FP64 fit from the same
    # start does): about 1% of the batch, mostly at |z| > 0.6 um
    core = np.all(np.abs(pull) < 5, axis=1)
    width = 1.4826 * np.median(np.abs(pull - np.median(pull, axis=0)), axis=0)
    mid = core & (np.abs(zt) < 0.6)
    print("pull robust width", width, "outside 5 sigma", 1 - core.mean(), "|z|<0.6 um mean", pull[mid].mean(axis=0),
          "std", pull[mid].std(axis=0))
    assert core.mean() > 0.97
    assert np.all(np.abs(width - 1.0) < 0.05)
    assert np.all(np.abs(pull[mid].mean(axis=0)) < 0.05)
    assert np.all(np.abs(pull[mid].std(axis=0) - 1.0) < 0.05)
    m = O.Model(O.optics(), 21, Z.astigmatic_coeffs(28))
    sample = np.random.default_rng(4).choice(B, 16, replace=False)
    hp, hi = px[sample].cpu().numpy(), init[sample].cpu().numpy()
    for k, b in enumerate(sample):
        r = m.lm_fit(hp[k].astype(np.float64), hi[k], zmin=-1.3, zmax=1.3, maxit=500, tol=1e-10)
        d = (res["theta"][b, 2:] - r["theta"][2:]) * 1000
        assert np.abs(d[:2]).max() <= 0.5 and abs(d[2]) <= 1.0, (b, d)
        assert abs(res["nll"][b] - r["nll"]) <= 1e-5 * abs(r["nll"])


def test_depth128_batch_noise_free_and_oracle_sample(torch_cuda):
    torch = torch_cuda
    B, R = 65536, 41
    print(f"case: {B} 41x41 ROIs on the 128 pupil with per-ROI focus", flush=True)
    opt = dict(size_fft=128, nwat=1.33)
    coef = Z.tetrapod_coeffs(45)
    ctx = Z.Context(Z.optics(**opt), R, coef)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    try:
        g = torch.Generator(device="cuda")
        g.manual_seed(0x128)
        u = lambda: torch.rand(B, device="cuda", generator=g, dtype=torch.float64)
        x, y = (2 * u() - 1) * DX, (2 * u() - 1) * DX
        z, zf = 6 * u() - 3, 4 * u()
        A = torch.exp(np.log(1e3) + (np.log(5e4) - np.log(1e3)) * u())
        bg = 10 + 90 * u()
        par = torch.stack([x, y, z, zf], 1).float().contiguous()
        psf = torch.empty(B, R, R, device="cuda")
        ctx.psf_forward_dev(par, psf)
        mu = (A.float()[:, None, None] * psf + bg.float()[:, None, None]).contiguous()
        noisy = torch.poisson(mu, generator=g).contiguous()
        del psf
        init = torch.stack([A * 1.1, bg * 0.9, x + 0.03, y - 0.03, z + 0.1], 1).contiguous()
        zoil = zf.float().contiguous()
        cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-3.3, zmax=3.3, tol=1e-4)
        o = _outs(torch, B)
        ctx.fit_zoil_dev(cfg, mu, init, zoil, o["theta"], o["nll"], o["crlb"], o["iters"], o["status"])
        torch.cuda.synchronize()
        th, st = o["theta"].cpu().numpy(), o["status"].cpu().numpy()
        t = torch.stack([x, y, z], 1).float().double().cpu().numpy()
        d = np.abs(th[:, 2:] - t) * 1000
        print("noise-free max |d| nm", d.max(axis=0))
        assert np.all((st & Z.ST_CONVERGED) != 0)
        assert d[:, :2].max() <= 0.5 and d[:, 2].max() <= 1.0
        ctx.fit_zoil_dev(cfg, noisy, init, zoil, o["theta"], o["nll"], o["crlb"], o["iters"], o["status"])
        torch.cuda.synchronize()
        th, nl = o["theta"].cpu().numpy(), o["nll"].cpu().numpy()
        m = O.Model(O.optics(**opt), R, coef)
        sample = np.random.default_rng(5).choice(B, 6, replace=False)
        hp, hi, hz = noisy[sample].cpu().numpy(), init[sample].cpu().numpy(), zoil[sample].cpu().numpy()
        for k, b in enumerate(sample):
            r = m.lm_fit(hp[k].astype(np.float64), hi[k], zfocus=float(hz[k]), zmin=-3.3, zmax=3.3, maxit=500,
                         tol=1e-10)
            dd = (th[b, 2:] - r["theta"][2:]) * 1000
            assert np.abs(dd[:2]).max() <= 0.5 and abs(dd[2]) <= 1.0, (b, dd)
            assert abs(nl[b] - r["nll"]) <= 1e-5 * abs(r["nll"])
    finally:
        ctx.close()


def test_mixed_200k_groups_equal_their_own_launch_and_recover_noise_free(torch_cuda):
    torch = torch_cuda
    sizes, frac, total = (17, 21, 31, 41), (0.4, 0.3, 0.2, 0.1), 200000
    print(f"case: {total} mixed ROIs in one graph launch", flush=True)
    coef = Z.astigmatic_coeffs(28)
    ctxs, groups, truths = [], [], []
    try:
        for k, (Rg, fr) in enumerate(zip(sizes, frac)):
            Bg = int(round(total * fr))
            c = Z.Context(Z.optics(), Rg, coef)
            c.set_stream(torch.cuda.current_stream().cuda_stream)
            ctxs.append(c)
            # half of each group noise-free, half Poisson
            pn, tn = _synth(torch, c, Rg, Bg // 2, 0x3000 + k, noise=False, alo=200.0, ahi=2e4)
            pp, tp = _synth(torch, c, Rg, Bg - Bg // 2, 0x3100 + k, noise=True, alo=200.0, ahi=2e4)
            px = torch.cat([pn, pp]).contiguous()
            gr = dict(R=Rg, pixels=px, init=_init(torch, c, px), **_outs(torch, Bg))
            groups.append(gr)
            truths.append(tn.cpu().numpy())
        ctx = ctxs[1]
        cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
        ctx.fit_mixed_dev(cfg, groups)
        torch.cuda.synchronize()
        for gr, c, tn in zip(groups, ctxs, truths):
            alone = _fit(torch, c, cfg, gr["pixels"], gr["init"])
            for k in ("theta", "nll", "crlb", "iters", "status"):
                np.testing.assert_array_equal(gr[k].cpu().numpy(), alone[k])
            nf = tn.shape[0]
            d = np.abs(alone["theta"][:nf, 2:] - tn[:, 2:]) * 1000
            print("R", gr["R"], "noise-free max |d| nm", d.max(axis=0))
            assert np.all((alone["status"][:nf] & Z.ST_CONVERGED) != 0)
            assert d[:, :2].max() <= 0.5 and d[:, 2].max() <= 1.0
    finally:
        for c in ctxs:
            c.close()


def _frames_2000(torch, ctx, F, W=256, H=256):
    """The frame workload of the benchmark: emitters Poisson at 0.5 per um^2 per
    frame, tetrapod PSF on the 128 pupil, z U[-3, 3] um, photons log-uniform
    1e3..2e4, background U[20, 100] per frame, Poisson noise."""
    Rn, h = 41, 20
    rng = np.random.default_rng(0xF1)
    counts = rng.poisson(0.5 * (W * DX) * (H * DX), F)
    n = int(counts.sum())
    fidx = np.repeat(np.arange(F), counts)
    pi, pj = rng.uniform(0, W, n), rng.uniform(0, H, n)
    ci, cj = np.rint(pi).astype(np.int64), np.rint(pj).astype(np.int64)
    z = rng.uniform(-3, 3, n)
    A = np.exp(rng.uniform(np.log(1e3), np.log(2e4), n))
    par = np.stack([(ci - pi) * DX, (cj - pj) * DX, z, np.zeros(n)], 1).astype(np.float32)
    Wp, Hp = W + 2 * Rn, H + 2 * Rn
    frames = torch.zeros(F * Wp * Hp, device="cuda")
    a = torch.arange(Rn, device="cuda")
    off = (a[:, None] * Hp + a[None, :]).reshape(-1)
    base_all = fidx * Wp * Hp + (ci - h + Rn) * Hp + (cj - h + Rn)
    chunk = 65536
    psf = torch.empty(chunk, Rn, Rn, device="cuda")
    for e0 in range(0, n, chunk):
        e1 = min(n, e0 + chunk)
        ctx.psf_forward_dev(torch.from_numpy(par[e0:e1]).cuda(), psf[:e1 - e0])
        base = torch.from_numpy(base_all[e0:e1]).cuda()
        frames.index_add_(0, (base[:, None] + off[None, :]).reshape(-1),
                          (torch.from_numpy(A[e0:e1]).float().cuda()[:, None, None] * psf[:e1 - e0]).reshape(-1))
    frames = frames.view(F, Wp, Hp)[:, Rn:Rn + W, Rn:Rn + H]
    bg = torch.from_numpy(rng.uniform(20, 100, F)).float().cuda()
    g = torch.Generator(device="cuda")
    g.manual_seed(0xF1)
    return torch.poisson(frames + bg[:, None, None], generator=g).cpu().numpy().astype(np.float32)


def test_frames_2000_records_do_not_depend_on_the_batch(torch_cuda):
    """2,000 frames of 256x256 (the benchmark's frame workload) through detection
    and the acceptance walk: the candidate lists and the records of a few frames
    processed alone equal, bit for bit, those frames' lists and records inside
    the 2,000-frame run; the records obey the acceptance rules."""
    torch = torch_cuda
    F = 2000
    print(f"case: {F} frames 256x256, detection and LM walk", flush=True)
    ctx = Z.Context(Z.optics(size_fft=128), 41, Z.tetrapod_coeffs(45))
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    det = None
    try:
        host = _frames_2000(torch, ctx, F)
        det = Z.Detector(ctx, 256, 256, -3.0, 3.05, 0.1, 0.1, 0.0)
        acc = det.accept_config(photon_threshold=500.0, min_z=-3.0, max_z=3.0)
        fit = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-3.3, zmax=3.3, tol=1e-4, photon_threshold=0.0)
        lists, cnt = det.detect(host, cap=8192)
        locs, per, rounds = det.localize(host, fit, acc)
        print("records", len(locs), "per frame", per.mean(), "rounds", rounds, flush=True)
        assert per.sum() == len(locs) and np.all(np.diff(locs[:, 0]) >= 0)
        assert np.all(locs[:, 4] > 500.0) and np.all(np.abs(locs[:, 3]) < 3000.0)
        assert np.all((locs[:, 1] > 0) & (locs[:, 1] < 256 * DX * 1000) & (locs[:, 2] > 0) &
                      (locs[:, 2] < 256 * DX * 1000))
        for a, b in ((0, 8), (996, 1004), (1992, 2000)):
            sub_lists, sub_cnt = det.detect(host[a:b], cap=8192)
            np.testing.assert_array_equal(sub_cnt, cnt[a:b])
            for f in range(b - a):
                np.testing.assert_array_equal(sub_lists[f], lists[a + f])
            sub, sub_per, _ = det.localize(host[a:b], fit, acc)
            ref = locs[(locs[:, 0] >= a) & (locs[:, 0] < b)].copy()
            ref[:, 0] -= a
            np.testing.assert_array_equal(sub_per, per[a:b])
            np.testing.assert_array_equal(sub, ref)
    finally:
        if det is not None:
            det.close()
        ctx.close()


def test_bead_preparation_full_stack_matches_oracle(torch_cuda):
    """The benchmark's bead stack (101 slices of 512x512, 12 beads, 32 px
    patches, 20 nm steps) against the FP64 oracle: refined seeds, centre slice
    and background pixel count identical, in-frame patch pixels bit-identical,
    background, B and A to 1e-9 relative."""
    print("case: bead preparation 101 x 512 x 512", flush=True)
    S, W, H, P, nb, zstep = 101, 512, 512, 32, 12, 0.02
    rng = np.random.default_rng(0x20A1D3 + 5)
    lam = np.full((S, W, H), 30.0)
    beads = rng.uniform(40, 470, (nb, 2))
    u = np.arange(-16, 17)
    for z in range(S):
        s = 1.2 + 0.02 * abs(z - S // 2)
        for bx, by in beads:
            ix, iy = int(bx) + u, int(by) + u
            lam[z][np.ix_(ix, iy)] += 3e4 / (2 * np.pi * s * s) * np.exp(
                -((ix[:, None] - bx) ** 2 + (iy[None, :] - by) ** 2) / (2 * s * s))
    adu = (rng.poisson(lam) + 100).astype(np.float32)
    del lam
    xp, yp = beads[:, 0] + 0.7, beads[:, 1] - 0.6
    ctx = Z.Context(Z.optics(), 21, Z.astigmatic_coeffs(28))
    try:
        got = ctx.bead_prepare(adu, xp, yp, P, zstep, offset=100.0)
    finally:
        ctx.close()
    ref = O.bead_prepare(O.bead_convert(adu, 1.0, 1.0, 100.0), xp, yp, P, zstep)
    assert ref["rc"] == 0
    np.testing.assert_array_equal(got["x"], ref["x"])
    np.testing.assert_array_equal(got["y"], ref["y"])
    d = got["diag"]
    assert (d["size_filt"], d["miniposition"], d["ndata"]) == (ref["size_filt"], ref["miniposition"], ref["ndata"])
    np.testing.assert_allclose(got["background"], ref["poly"], rtol=1e-9)
    np.testing.assert_allclose(got["B"], ref["B"], rtol=1e-9)
    np.testing.assert_allclose(got["A"], ref["A"], rtol=1e-9, atol=1e-9 * np.abs(ref["A"]).max())
    for p in range(nb):
        x0, y0 = int(ref["x"][p]) - P // 2, int(ref["y"][p]) - P // 2
        np.testing.assert_array_equal(got["patches"][p], ref["patches"][p])  # beads lie inside the frame
        assert 0 <= x0 and x0 + P <= W and 0 <= y0 and y0 + P <= H


def test_headline_batch_with_scmos_map_noise_free(torch_cuda, ctx21):
    """65,536 noise-free 21x21 ROIs with a per-pixel sCMOS map v (mu = A psf +
    B + v): every ROI returns the parameters it was made from (0.5 nm x/y,
    1 nm z), and a subset fitted alone is bit-identical to the full batch."""
    torch = torch_cuda
    B = 65536
    print(f"case: {B} noise-free 21x21 ROIs with an sCMOS map", flush=True)
    px, truth = _synth(torch, ctx21, 21, B, 0x54, noise=False)
    g = torch.Generator(device="cuda")
    g.manual_seed(0x55)
    varg = (0.5 + 11.5 * torch.rand(B, 21, 21, device="cuda", generator=g)).contiguous()
    px = (px + varg).contiguous()
    init = _init(torch, ctx21, (px - varg).contiguous())
    cfg = ctx21.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)

    def fit(p, i, v):
        o = _outs(torch, p.shape[0])
        ctx21.fit_ext_dev(cfg, p, i, o["theta"], o["nll"], o["crlb"], o["iters"], o["status"], varg=v)
        torch.cuda.synchronize()
        return {k: x.cpu().numpy() for k, x in o.items()}

    res = fit(px, init, varg)
    t = truth.cpu().numpy()
    d = np.abs(res["theta"][:, 2:] - t[:, 2:]) * 1000
    print("max |d| nm", d.max(axis=0))
    assert np.all((res["status"] & Z.ST_CONVERGED) != 0)
    assert d[:, :2].max() <= 0.5 and d[:, 2].max() <= 1.0
    idx = torch.from_numpy(np.sort(np.random.default_rng(6).choice(B, 512, replace=False))).cuda()
    sub = fit(px[idx].contiguous(), init[idx].contiguous(), varg[idx].contiguous())
    ii = idx.cpu().numpy()
    for k in ("theta", "nll", "crlb", "iters", "status"):
        np.testing.assert_array_equal(sub[k], res[k][ii])
