#!/usr/bin/env python3
"""Throughput of the ZOLA-3D localization hot path on one MI355X.

Workload: isolated-emitter MLE localization, 21x21-pixel ROIs, 64x64 pupil,
astigmatic PSF from 28 Zernike modes (Noll 5 = 1 rad, Noll 4..28 ~ N(0, 0.05)),
NA 1.45, lambda 680 nm, n 1.518, 110 nm pixels; emitters x, y ~ U(+-1 px),
z ~ U[-1, 1] um, photons log-uniform 500..5000, background U[5, 50], Poisson
noise, float32 pixels.  Synthetic data are generated on the GPU with the
library's own PSF synthesis.

One step = per-ROI initial values (template correlation + getAandB) and the
batched Levenberg-Marquardt fit with analytic tangent Jacobians, for BATCH ROIs
resident in HBM.  The plain per-ROI GPU form (the reference's coordinate
descent) is timed on the same ROIs, and the reference CPU algorithm (the FP64
restatement in oracle/) on a bounded sample on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "zola-3d_amd"))

F64 = 5 * 64 * 64 * np.log2(64 * 64)     # flops of one 64x64 complex FFT (5 N log2 N)
PEAK_FP32 = 157.3e12                      # MI355X FP32 vector peak (AMD spec)
DX = 0.11
R = 21
ZMIN_T, ZSTEP_T, NZ = -1.0, 0.1, 21       # initial-z templates at 100 nm


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--batch", type=int, default=65536)
    p.add_argument("--plain-batch", type=int, default=16384)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--no-calib", action="store_true")
    p.add_argument("--no-n128", action="store_true")
    p.add_argument("--no-mixed", action="store_true")
    p.add_argument("--no-frames", action="store_true")
    p.add_argument("--no-formats", action="store_true")
    p.add_argument("--no-beads", action="store_true")
    p.add_argument("--no-host", action="store_true")
    p.add_argument("--n128-batch", type=int, default=1_000_000)
    p.add_argument("--n128-layout", choices=["auto", "quad", "oct"], default="auto",
                   help="LM kernel layout of the 128-grid leg: four or eight waves per ROI")
    p.add_argument("--mixed-total", type=int, default=200000)
    p.add_argument("--frames", type=int, default=2000)
    p.add_argument("--cpu-sample", type=int, default=0, help="ROIs for the CPU baseline (0: auto)")
    p.add_argument("--copy-probe-mb", type=int, default=0,
                   help="also run one device copy of this size (reference for HBM counters)")
    return p.parse_args()


def synth(ctx, torch, B, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    u = lambda: torch.rand(B, device="cuda", generator=g, dtype=torch.float64)
    x = (2 * u() - 1) * DX
    y = (2 * u() - 1) * DX
    z = 2 * u() - 1
    A = torch.exp(np.log(500) + (np.log(5000) - np.log(500)) * u())
    bg = 5 + 45 * u()
    par = torch.stack([x, y, z, torch.zeros_like(x)], 1).float().contiguous()
    psf = torch.empty(B, R, R, device="cuda")
    ctx.psf_forward_dev(par, psf)
    mu = A.float()[:, None, None] * psf + bg.float()[:, None, None]
    px = torch.poisson(mu, generator=g).float().contiguous()
    truth = torch.stack([A, bg, x, y, z], 1)
    return px, truth


def calibration_bench(torch, Z, stream, cpu=True):
    """PSF calibration on the reference's full model (ZernikePhaseRetrieval.Model):
    5 bead z-stacks of 41 slices (50 nm, -1..+1 um), 31x31 px, 2e5 photons per
    slice on a per-slice curve, order-2 background around 10 photons/px, 45 Zernike
    modes ~ N(0, 0.15 rad) with Noll 5 = 0.8 rad.  Fitted by L-BFGS on exact
    adjoint gradients from a zero start: the 42 non-registration coefficients,
    sigma, and per bead x, y, z0, the six background coefficients, fit_a and the
    41 fit_a_each."""
    Rc, J, nb, S = 31, 45, 5, 41
    rng = np.random.default_rng(0x20A1D3 + 3)
    truth = rng.normal(0.0, 0.15, J)
    truth[:3] = 0.0
    truth[Z.nm_to_slot(*Z.noll_to_nm(5))] = 0.8
    dz = (np.arange(S) - S // 2) * 0.05
    xyz = np.stack([rng.uniform(-0.05, 0.05, nb), rng.uniform(-0.05, 0.05, nb), rng.uniform(-0.05, 0.05, nb)], 1)
    A = 2e5 * (1.0 + 0.15 * np.sin(np.arange(S)[None, :] / 5.0 + rng.uniform(0, 6, (nb, 1))))
    poly = np.zeros((nb, 6))
    poly[:, 0] = 10.0 + rng.uniform(-2, 2, nb)
    poly[:, 1] = rng.uniform(-0.05, 0.05, nb)
    poly[:, 3] = rng.uniform(-0.05, 0.05, nb)
    ii = np.arange(Rc, dtype=np.float64)
    mono = np.stack([np.ones((Rc, Rc)), ii[:, None] * np.ones(Rc), (ii * ii)[:, None] * np.ones(Rc),
                     np.ones(Rc)[:, None] * ii, ii[:, None] * ii, np.ones(Rc)[:, None] * ii * ii])
    gen = Z.Context(Z.optics(), Rc, truth)
    gen.set_stream(stream.cuda_stream)
    par = np.zeros((nb * S, 4), np.float32)
    for b in range(nb):
        for s_ in range(S):
            par[b * S + s_] = [xyz[b, 0], xyz[b, 1], 0.0, dz[s_] + xyz[b, 2]]
    P = torch.from_numpy(par).cuda()
    psf = torch.empty(nb * S, Rc, Rc, device="cuda")
    gen.psf_forward_dev(P, psf)
    bgmap = torch.from_numpy(np.repeat(np.tensordot(poly, mono, 1), S, axis=0)).float().cuda()
    amp_t = torch.from_numpy(A.reshape(-1)).float().cuda()
    g = torch.Generator(device="cuda")
    g.manual_seed(0x20A1D3 + 3)
    stacks = torch.poisson(amp_t[:, None, None] * psf + bgmap, generator=g).float().contiguous()
    ctx = Z.Context(Z.optics(), Rc, np.zeros(J))
    ctx.set_stream(stream.cuda_stream)
    # one forward+adjoint evaluation of the full model (device form), HIP events on the context stream
    B = nb * S
    amp = torch.from_numpy(A.reshape(-1)).cuda()
    pol = torch.from_numpy(poly).cuda()
    nll = torch.empty(B, dtype=torch.float64, device="cuda")
    dpar = torch.empty(B, 4, device="cuda")
    dA = torch.empty_like(nll)
    dpoly = torch.empty(B, 6, dtype=torch.float64, device="cuda")
    dsig = torch.empty_like(nll)
    da = torch.zeros(J, dtype=torch.float64, device="cuda")
    args = (nb, S, P, stacks, amp, pol, nll, dpar, dA)
    kw = dict(dpoly=dpoly, dsigma=dsig, da=da)
    ctx.calib_eval_full_dev(*args, **kw)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    reps = 20
    for _ in range(reps):
        ctx.calib_eval_full_dev(*args, **kw)
    e1.record(stream)
    torch.cuda.synchronize()
    eval_ms = e0.elapsed_time(e1) / reps
    # starting point as the plugin has it after the bead preparation: photon curves as
    # patch sums minus a flat background level from the patch border
    host = stacks.cpu().numpy().reshape(nb, S, Rc, Rc)
    c0 = np.array([np.median(np.concatenate([host[b, :, 0, :].ravel(), host[b, :, -1, :].ravel()]))
                   for b in range(nb)])
    A0 = host.reshape(nb, S, -1).sum(2) - Rc * Rc * c0[:, None]
    poly0 = np.zeros((nb, 6))
    poly0[:, 0] = c0
    mask = np.ones(J, np.int32)
    mask[:3] = 0
    t0 = time.perf_counter()
    res = ctx.calib_fit_full(host, dz, A0, np.zeros((nb, 3)), poly0, np.zeros(J), sigma=1.0, fit_mask=mask,
                             max_iter=2000, gtol=1e-4)
    dt = time.perf_counter() - t0
    # algorithmic work of one evaluation: a forward and an adjoint pupil
    # transform per PSF plus the Zernike projection, 2 D J FLOP per PSF
    D = ctx.D
    fft = 5 * 64 * 64 * np.log2(64 * 64)
    W = nb * S * (2 * fft + 2 * D * J)
    out = {
        "workload": "full calibration model: 5 beads x 41 slices x 31x31 px, 45 Zernike modes, per-slice photon "
                    "curves, order-2 backgrounds, sigma; L-BFGS on adjoint gradients (forward+adjoint per "
                    "evaluation), zero start",
        "iterations": res["iters"], "evaluations": res["evals"],
        "iters_per_s": round(res["iters"] / dt, 2),
        "evals_per_s": round(res["evals"] / dt, 2),
        "eval_kernel_ms": round(eval_ms, 4),
        "max_coeff_err_rad": round(float(np.abs(res["coeffs"] - truth)[3:].max()), 5),
        "sigma": round(res["sigma"], 5),
        "max_xyz_err_nm": round(float(np.abs(res["xyz"] - xyz).max() * 1000), 3),
        "roofline": {"bound": "valu", "achieved": round(W / (eval_ms * 1e-3) / 1e12, 4), "peak": 157.3,
                     "unit": "TFLOP/s", "frac": round(W / (eval_ms * 1e-3) / 1e12 / 157.3, 5),
                     "algorithmic_flops_per_eval": float(W)},
    }
    if cpu:
        # one iteration of the reference calibration loop (ZernikePhaseRetrieval.
        # phase_retrieve_zernike: registration, sigma, background polynomial, photon
        # offsets and the 42 Zernike coordinates by finite differences), FP64 C
        # restatement, the slices of a bead on all host cores, from the same start
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads = host_cores()
        ref = O.CalibRef(O.optics(), np.zeros(J), host.astype(np.float64), 0.05, A0, c0, sigma=1.0,
                         nthreads=threads)
        t0 = time.perf_counter()
        ref.iteration()
        t_it = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": round(1.0 / t_it, 5), "unit": "iterations/s", "cores": threads, "kind": "port",
            "sample": f"one iteration of the reference calibration loop (FP64 C restatement of "
                      f"ZernikePhaseRetrieval.phase_retrieve_zernike, same stacks and start), {t_it:.1f} s "
                      f"on {threads} threads"}
    gen.close()
    ctx.close()
    return out


def calibration256_bench(torch, Z, stream, max_iter=400):
    """The plugin's default calibration geometry: sizeFFT 140 (pupil grid 256),
    64 px bead patches (blur grid 128), 5 beads x 41 slices, 45 modes, the full
    model, FP64 throughout.  One forward+adjoint evaluation timed with HIP events,
    then L-BFGS from a zero start, bounded at max_iter iterations."""
    Rc, J, nb, S = 64, 45, 5, 41
    rng = np.random.default_rng(0x20A1D3 + 30)
    truth = rng.normal(0.0, 0.15, J)
    truth[:3] = 0.0
    truth[Z.nm_to_slot(*Z.noll_to_nm(5))] = 0.8
    dz = (np.arange(S) - S // 2) * 0.05
    xyz = np.stack([rng.uniform(-0.05, 0.05, nb), rng.uniform(-0.05, 0.05, nb), rng.uniform(-0.05, 0.05, nb)], 1)
    A = 2e5 * (1.0 + 0.15 * np.sin(np.arange(S)[None, :] / 5.0 + rng.uniform(0, 6, (nb, 1))))
    opt = Z.optics(size_fft=140)
    gen = Z.Context(opt, Rc, truth)
    gen.set_stream(stream.cuda_stream)
    par = np.zeros((nb * S, 4), np.float32)
    for b in range(nb):
        for s_ in range(S):
            par[b * S + s_] = [xyz[b, 0], xyz[b, 1], 0.0, dz[s_] + xyz[b, 2]]
    P = torch.from_numpy(par).cuda()
    psf = torch.empty(nb * S, Rc, Rc, device="cuda")
    gen.psf_forward_dev(P, psf)
    amp_t = torch.from_numpy(A.reshape(-1)).float().cuda()
    g = torch.Generator(device="cuda")
    g.manual_seed(0x20A1D3 + 30)
    stacks = torch.poisson(amp_t[:, None, None] * psf + 10.0, generator=g).float().contiguous()
    ctx = Z.Context(opt, Rc, np.zeros(J))
    ctx.set_stream(stream.cuda_stream)
    B = nb * S
    amp = torch.from_numpy(A.reshape(-1)).cuda()
    pol = torch.zeros(nb, 6, dtype=torch.float64, device="cuda")
    pol[:, 0] = 10.0
    nll = torch.empty(B, dtype=torch.float64, device="cuda")
    dpar = torch.empty(B, 4, device="cuda")
    dA = torch.empty_like(nll)
    dpoly = torch.empty(B, 6, dtype=torch.float64, device="cuda")
    dsig = torch.empty_like(nll)
    da = torch.zeros(J, dtype=torch.float64, device="cuda")
    args = (nb, S, P, stacks, amp, pol, nll, dpar, dA)
    kw = dict(dpoly=dpoly, dsigma=dsig, da=da)
    ctx.calib_eval_full_dev(*args, **kw)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record(stream)
    for _ in range(reps):
        ctx.calib_eval_full_dev(*args, **kw)
    e1.record(stream)
    torch.cuda.synchronize()
    eval_ms = e0.elapsed_time(e1) / reps
    host = stacks.cpu().numpy().reshape(nb, S, Rc, Rc)
    c0 = np.array([np.median(np.concatenate([host[b, :, 0, :].ravel(), host[b, :, -1, :].ravel()]))
                   for b in range(nb)])
    A0 = host.reshape(nb, S, -1).sum(2) - Rc * Rc * c0[:, None]
    poly0 = np.zeros((nb, 6))
    poly0[:, 0] = c0
    mask = np.ones(J, np.int32)
    mask[:3] = 0
    t0 = time.perf_counter()
    res = ctx.calib_fit_full(host, dz, A0, np.zeros((nb, 3)), poly0, np.zeros(J), sigma=1.0, fit_mask=mask,
                             max_iter=max_iter, gtol=1e-4)
    dt = time.perf_counter() - t0
    # algorithmic work of one evaluation: a forward and an adjoint 256 x 256 pupil
    # transform per PSF plus the Zernike projection (2 D J)
    fft = 5 * 256 * 256 * np.log2(256 * 256)
    W = B * (2 * fft + 2 * ctx.D * J)
    out = {
        "workload": "plugin default calibration grid: sizeFFT 140 (pupil 256), 64 px patches (blur grid 128), "
                    "5 beads x 41 slices, 45 Zernike modes, full model, FP64; L-BFGS from a zero start, "
                    f"at most {max_iter} iterations",
        "iterations": res["iters"], "evaluations": res["evals"],
        "iters_per_s": round(res["iters"] / dt, 2), "evals_per_s": round(res["evals"] / dt, 2),
        "eval_ms": round(eval_ms, 4),
        "max_coeff_err_rad": round(float(np.abs(res["coeffs"] - truth)[3:].max()), 5),
        "max_coeff_err_rad_but_defocus": round(float(np.delete(np.abs(res["coeffs"] - truth), [0, 1, 2, 4]).max()), 5),
        "sigma": round(res["sigma"], 5),
        "nll": res["nll"],
        "roofline": {"bound": "valu-fp64", "achieved": round(W / (eval_ms * 1e-3) / 1e12, 4), "peak": 78.6,
                     "unit": "TFLOP/s", "frac": round(W / (eval_ms * 1e-3) / 1e12 / 78.6, 5),
                     "algorithmic_flops_per_eval": float(W)},
    }
    gen.close()
    ctx.close()
    return out


def _cgroup_cpus():
    """CPUs granted by the cgroup v2 quota (cpu.max), or None without a quota."""
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        return None if q == "max" else max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        return None


def host_cores():
    """Host cores this process may use: the CPU affinity set, capped by the
    cgroup quota (the CPU legs run one thread per core)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    q = _cgroup_cpus()
    return min(n, q) if q else n


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = None
    return {"cores_used": host_cores(), "affinity_cpus": aff, "cgroup_cpus": _cgroup_cpus(),
            "cpu_count": os.cpu_count(), "cpu_model": model}


def _events(torch, n=2):
    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


def depth128_bench(torch, Z, stream, B, layout="auto"):
    """Large-range depth-dependent run on the 128x128 pupil: 45-mode tetrapod PSF
    (Noll 5, 12, 23 at 2-3 rad, others N(0, 0.1)), n_sample 1.33 / n_imm 1.518,
    focal depth U[0, 4] um per ROI, z U[-3, 3] um, x, y U(+-1 px), photons
    log-uniform 1e3..5e4, background U[10, 100], 41x41 ROIs, Poisson noise.
    Batched LM with per-ROI focus and CRLB; the plain reference form on B/8."""
    Rn = 41
    ctx = Z.Context(Z.optics(size_fft=128, nwat=1.33), Rn, Z.tetrapod_coeffs(45))
    ctx.set_stream(stream.cuda_stream)
    lay = {"auto": Z.LAYOUT_AUTO, "quad": Z.LAYOUT_QUAD, "oct": Z.LAYOUT_OCT}[layout]
    ctx.set_fit_layout(Rn, Z.FIT_LM, lay)
    g = torch.Generator(device="cuda")
    g.manual_seed(0x128)
    u = lambda: torch.rand(B, device="cuda", generator=g, dtype=torch.float64)
    x, y = (2 * u() - 1) * DX, (2 * u() - 1) * DX
    z, zf = 6 * u() - 3, 4 * u()
    A = torch.exp(np.log(1e3) + (np.log(5e4) - np.log(1e3)) * u())
    bg = 10 + 90 * u()
    par = torch.stack([x, y, z, zf], 1).float().contiguous()
    psf = torch.empty(B, Rn, Rn, device="cuda")
    ctx.psf_forward_dev(par, psf)
    px = torch.poisson(A.float()[:, None, None] * psf + bg.float()[:, None, None], generator=g).contiguous()
    del psf
    init = torch.stack([A * 1.1, bg * 0.9, x + 0.03, y - 0.03, z + 0.1], 1).contiguous()
    zoil = zf.float().contiguous()
    th = torch.empty(B, 5, dtype=torch.float64, device="cuda")
    nl = torch.empty(B, dtype=torch.float64, device="cuda")
    cr = torch.empty(B, 3, dtype=torch.float64, device="cuda")
    it = torch.empty(B, dtype=torch.int32, device="cuda")
    st = torch.empty(B, dtype=torch.int32, device="cuda")
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-3.3, zmax=3.3, tol=1e-4)
    # warm-up on a slice (kernel load), then the whole batch in one persistent pass
    Bw = min(B, 8192)
    ctx.fit_zoil_dev(cfg, px[:Bw], init[:Bw], zoil[:Bw], th[:Bw], nl[:Bw], cr[:Bw], it[:Bw], st[:Bw])
    e = _events(torch)
    e[0].record(stream)
    ctx.fit_zoil_dev(cfg, px, init, zoil, th, nl, cr, it, st)
    e[1].record(stream)
    torch.cuda.synchronize()
    ms = e[0].elapsed_time(e[1])
    sn = st.cpu().numpy()
    err = (th[:, 2:] - torch.stack([x, y, z], 1)).abs().cpu().numpy() * 1000
    out = {
        "workload": f"{B} ROIs 41x41, 128x128 pupil, tetrapod 45 modes, n 1.33/1.518, focal depth "
                    "U[0,4] um per ROI, z U[-3,3] um, LM max 30 it, tol 1e-4 um, CRLB",
        "rois": B, "kernel_ms": round(ms, 3), "value": round(B / (ms / 1e3), 1), "unit": "localizations/s",
        "fields_per_roi": round(float((sn >> 8).mean()), 2),
        "iters_mean": round(float(it.float().mean().item()), 3),
        "converged_frac": round(float(np.mean((sn & Z.ST_CONVERGED) != 0)), 5),
        "median_abs_err_vs_truth_nm": [round(float(v), 3) for v in np.median(err, axis=0)],
        "median_crlb_nm": [round(float(v), 3) for v in np.median(cr.cpu().numpy() * 1000, axis=0)],
        "lm_layout": {"auto": "default (eight waves per ROI)", "quad": "four waves per ROI",
                      "oct": "eight waves per ROI"}[layout],
    }
    # algorithmic FLOPs: 5 n log2 n per 128x128 pupil field (n = 128^2), as for the headline's 64^2 fields
    f128 = 5.0 * 128 * 128 * np.log2(128 * 128)
    ach = float((sn >> 8).astype(np.float64).sum()) * f128 / (ms / 1e3) / 1e12
    out["roofline"] = {"bound": "valu", "kernel": "k_fit_lm128o<41>" if layout != "quad" else "k_fit_lm128<41,4>",
                       "achieved": round(ach, 4), "peak": 157.3, "unit": "TFLOP/s", "frac": round(ach / 157.3, 5),
                       "algorithmic_flops_per_field": f128}
    Bp = max(1, min(B // 8, 8192))
    cfgp = ctx.fit_config(mode=Z.FIT_REFERENCE, max_iter=20, zmin=-3.3, zmax=3.3, photon_threshold=0.0)
    args = (px[:Bp].contiguous(), init[:Bp].contiguous(), zoil[:Bp].contiguous(), th[:Bp], nl[:Bp], cr[:Bp],
            it[:Bp], st[:Bp])
    ctx.fit_zoil_dev(cfgp, *args)
    e[0].record(stream)
    ctx.fit_zoil_dev(cfgp, *args)
    e[1].record(stream)
    torch.cuda.synchronize()
    pms = e[0].elapsed_time(e[1])
    out["plain_gpu"] = {"rois": Bp, "kernel_ms": round(pms, 3), "value": round(Bp / (pms / 1e3), 1),
                        "fields_per_roi": round(float((st[:Bp].cpu().numpy() >> 8).mean()), 2)}
    out["batched_vs_plain"] = round((B / ms) / (Bp / pms), 3)
    ctx.close()
    return out


def mixed_bench(torch, Z, stream, total):
    """Uneven work: 40% 17x17, 30% 21x21, 20% 31x31, 10% 41x41 ROIs, astigmatic
    28-mode PSF (64x64 pupil), photons log-uniform 200..20000 (so LM iteration
    counts spread), background U[5, 50], z U[-1, 1]; initial values from the
    template correlation; all groups fitted by one graph launch."""
    sizes, frac = (17, 21, 31, 41), (0.4, 0.3, 0.2, 0.1)
    coeffs = Z.astigmatic_coeffs(28)
    groups, ctxs = [], []
    for k, (Rg, fr) in enumerate(zip(sizes, frac)):
        Bg = int(round(total * fr))
        c = Z.Context(Z.optics(), Rg, coeffs)
        c.set_stream(stream.cuda_stream)
        g = torch.Generator(device="cuda")
        g.manual_seed(0x2000 + k)
        u = lambda: torch.rand(Bg, device="cuda", generator=g, dtype=torch.float64)
        x, y, z = (2 * u() - 1) * DX, (2 * u() - 1) * DX, 2 * u() - 1
        A = torch.exp(np.log(200) + (np.log(2e4) - np.log(200)) * u())
        bg = 5 + 45 * u()
        psf = torch.empty(Bg, Rg, Rg, device="cuda")
        c.psf_forward_dev(torch.stack([x, y, z, torch.zeros_like(x)], 1).float().contiguous(), psf)
        px = torch.poisson(A.float()[:, None, None] * psf + bg.float()[:, None, None], generator=g).contiguous()
        init = torch.empty(Bg, 5, dtype=torch.float64, device="cuda")
        c.fit_init_dev(px, NZ, ZMIN_T, ZSTEP_T, 0.0, init)
        groups.append(dict(R=Rg, pixels=px, init=init,
                           theta=torch.empty(Bg, 5, dtype=torch.float64, device="cuda"),
                           nll=torch.empty(Bg, dtype=torch.float64, device="cuda"),
                           crlb=torch.empty(Bg, 3, dtype=torch.float64, device="cuda"),
                           iters=torch.empty(Bg, dtype=torch.int32, device="cuda"),
                           status=torch.empty(Bg, dtype=torch.int32, device="cuda")))
        ctxs.append(c)
    ctx = ctxs[1]
    # autotuner: per ROI side, the fastest of one, two (R <= 21) and four waves per ROI
    layouts = {}
    names = {Z.LAYOUT_WAVE: "wave", Z.LAYOUT_QUAD: "quad", Z.LAYOUT_PAIR: "pair"}
    for Rg in sizes:
        lay, t_wave, t_quad, t_pair = ctx.autotune_fit(Rg, Z.FIT_LM, B=8192)
        layouts[str(Rg)] = {"layout": names.get(lay, str(lay)), "ms_wave": round(t_wave, 3),
                            "ms_quad": round(t_quad, 3), "ms_pair": round(t_pair, 3) if t_pair else None}
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
    ctx.fit_mixed_dev(cfg, groups)
    e = _events(torch)
    e[0].record(stream)
    ctx.fit_mixed_dev(cfg, groups)
    e[1].record(stream)
    torch.cuda.synchronize()
    ms = e[0].elapsed_time(e[1])
    B = sum(int(gr["pixels"].shape[0]) for gr in groups)
    per = {}
    for gr in groups:
        itn = gr["iters"].cpu().numpy()
        sn = gr["status"].cpu().numpy()
        per[str(gr["R"])] = {"rois": int(itn.size), "iters_mean": round(float(itn.mean()), 2),
                             "iters_min": int(itn.min()), "iters_max": int(itn.max()),
                             "converged_frac": round(float(np.mean((sn & Z.ST_CONVERGED) != 0)), 4)}
    for c in ctxs:
        c.close()
    return {"workload": f"{B} ROIs mixed 40/30/20/10% of 17/21/31/41 px, 64x64 pupil, photons 200..2e4, "
                        "one graph launch (one kernel node per size)",
            "rois": B, "graph_ms": round(ms, 3), "value": round(B / (ms / 1e3), 1), "unit": "localizations/s",
            "groups": per, "autotune_8192_rois": layouts}


def frames_bench(torch, Z, stream, F):
    """Detection + localization on full frames: F frames of 256x256 px, emitters
    Poisson at 0.5 per um^2 per frame, tetrapod 45-mode PSF on the 128x128 pupil,
    z U[-3, 3] um, photons log-uniform 1e3..2e4, background U[20, 100] per frame;
    61 templates at 100 nm on [-3, 3] um, 41x41 patches, reference acceptance
    (photons > 500, z in (-3, 3))."""
    W = H = 256
    Rn = 41
    h = Rn // 2
    ctx = Z.Context(Z.optics(size_fft=128), Rn, Z.tetrapod_coeffs(45))
    ctx.set_stream(stream.cuda_stream)
    rng = np.random.default_rng(0xF1)
    counts = rng.poisson(0.5 * (W * DX) * (H * DX), F)
    n = int(counts.sum())
    fidx = np.repeat(np.arange(F), counts)
    pi, pj = rng.uniform(0, W, n), rng.uniform(0, H, n)
    ci, cj = np.rint(pi).astype(np.int64), np.rint(pj).astype(np.int64)
    z = rng.uniform(-3, 3, n)
    A = np.exp(rng.uniform(np.log(1e3), np.log(2e4), n))
    par_all = np.stack([(ci - pi) * DX, (cj - pj) * DX, z, np.zeros(n)], 1).astype(np.float32)
    Wp, Hp = W + 2 * Rn, H + 2 * Rn
    frames = torch.zeros(F * Wp * Hp, device="cuda")
    a = torch.arange(Rn, device="cuda")
    off = (a[:, None] * Hp + a[None, :]).reshape(-1)
    base_all = fidx * Wp * Hp + (ci - h + Rn) * Hp + (cj - h + Rn)
    chunk = 65536
    psf = torch.empty(chunk, Rn, Rn, device="cuda")
    for e0 in range(0, n, chunk):
        e1 = min(n, e0 + chunk)
        ctx.psf_forward_dev(torch.from_numpy(par_all[e0:e1]).cuda(), psf[:e1 - e0])
        base = torch.from_numpy(base_all[e0:e1]).cuda()
        frames.index_add_(0, (base[:, None] + off[None, :]).reshape(-1),
                          (torch.from_numpy(A[e0:e1]).float().cuda()[:, None, None] * psf[:e1 - e0]).reshape(-1))
    frames = frames.view(F, Wp, Hp)[:, Rn:Rn + W, Rn:Rn + H]
    bg = torch.from_numpy(rng.uniform(20, 100, F)).float().cuda()
    g = torch.Generator(device="cuda")
    g.manual_seed(0xF1)
    host = torch.poisson(frames + bg[:, None, None], generator=g).cpu().numpy().astype(np.float32)
    det = Z.Detector(ctx, W, H, -3.0, 3.05, 0.1, 0.1, 0.0)
    det.detect(host, cap=8192)  # allocates the batch buffers outside the timed call
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lists, cnt = det.detect(host, cap=8192)
    t_det = time.perf_counter() - t0
    # detection roofline: per frame one forward transform of the padded frame and
    # one inverse per template, FX x FY complex FP64, 5 N log2 N flops each
    # (windowed statistics, products and local maxima excluded)
    fl_det = (1 + det.nZ) * 5.0 * det.FX * det.FY * np.log2(det.FX * det.FY)
    ach = fl_det * F / t_det / 1e12
    out = {"workload": f"{F} frames 256x256, 0.5 emitters/um^2 ({n} emitters), tetrapod 45 modes, 128x128 "
                       f"pupil, {det.nZ} templates at 100 nm, 41x41 patches",
           "frames": F, "emitters": n, "templates": det.nZ,
           "detect_ms_per_frame": round(t_det / F * 1e3, 3),
           "candidates_per_frame": round(float(np.mean(cnt)), 1),
           "detect_roofline": {"bound": "valu-fp64", "achieved": round(ach, 4), "peak": 78.6, "unit": "TFLOP/s",
                               "frac": round(ach / 78.6, 5), "algorithmic_flops_per_frame": float(fl_det),
                               "note": "host-array call, frame upload and candidate download included"}}
    acc = det.accept_config(photon_threshold=500.0, min_z=-3.0, max_z=3.0)
    for mode, name in ((Z.FIT_LM, "lm"), (Z.FIT_REFERENCE, "plain")):
        fit = ctx.fit_config(mode=mode, max_iter=30 if mode == Z.FIT_LM else 20, zmin=-3.3, zmax=3.3, tol=1e-4,
                             photon_threshold=0.0)
        t0 = time.perf_counter()
        locs, per, rounds = det.localize(host, fit, acc)
        dt = time.perf_counter() - t0
        out[name] = {"s": round(dt, 3), "frames_per_s": round(F / dt, 2),
                     "localizations": int(len(locs)), "localizations_per_s": round(len(locs) / dt, 1),
                     "fit_rounds": rounds}
    det.close()
    ctx.close()
    return out


def formats_bench(Z, n=1_000_000):
    """Localization table (ThunderSTORM CSV rows as the plugin writes them) for
    n synthetic records, and calibration-file write + read of a 45-mode model."""
    rng = np.random.default_rng(0xC5)
    locs = np.zeros((n, Z.LOC_FIELDS))
    locs[:, 0] = np.sort(rng.integers(0, 10000, n))
    locs[:, 1:3] = rng.uniform(0, 50000, (n, 2))
    locs[:, 3] = rng.uniform(-3000, 3000, n)
    locs[:, 4] = rng.uniform(500, 2e4, n)
    locs[:, 5] = rng.uniform(20, 100, n)
    locs[:, 6] = rng.uniform(0, 3, n)
    locs[:, 7:10] = rng.uniform(5, 40, (n, 3))
    t0 = time.perf_counter()
    text = Z.loc_csv(locs)
    dt = time.perf_counter() - t0
    opt = Z.optics(size_fft=128)
    a = Z.tetrapod_coeffs(45)
    reps = 2000
    t0 = time.perf_counter()
    for _ in range(reps):
        js = Z.calib_json_format(opt, a)
    t_w = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        Z.calib_json_parse(js, 128)
    t_r = (time.perf_counter() - t0) / reps
    return {"workload": f"{n} localization records -> CSV rows; 45-mode calibration file write/read",
            "csv_rows_per_s": round(n / dt, 1), "csv_mb": round(len(text) / 1e6, 2),
            "calib_write_us": round(t_w * 1e6, 2), "calib_read_us": round(t_r * 1e6, 2)}


def beads_bench(Z, stream, cpu=True):
    """Calibration bead preparation (InitBackgroundAndPhotonNumber): a raw
    101-slice 512x512 stack (20 nm steps), 12 beads of 3e4 photons on a
    background of 30, Poisson noise, 32 px patches; host arrays in and out."""
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
    ctx = Z.Context(Z.optics(), R, Z.astigmatic_coeffs(28))
    ctx.set_stream(stream.cuda_stream)
    ctx.bead_prepare(adu, xp, yp, P, zstep, offset=100.0)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        res = ctx.bead_prepare(adu, xp, yp, P, zstep, offset=100.0)
        ts.append(time.perf_counter() - t0)
    dt = float(np.median(ts))
    ctx.close()
    out = {"workload": f"{S} slices x {W}x{H} px raw stack, {nb} beads, {P} px patches, z step {zstep} um "
                       "(background fit window +-12 slices)",
           "call_ms": round(dt * 1e3, 3), "stacks_per_s": round(1.0 / dt, 3),
           "voxels_per_s": round(S * W * H / dt, 1), "centre_slice": res["diag"]["miniposition"],
           "background_pixels": res["diag"]["ndata"],
           "note": "host arrays in and out: includes the 106 MB host-to-device copy of the stack"}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        img = O.bead_convert(adu, 1.0, 1.0, 100.0)
        t0 = time.perf_counter()
        O.bead_prepare(img, xp, yp, P, zstep)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(1.0 / tc, 4), "unit": "stacks/s", "cores": 1, "kind": "port",
                               "sample": f"the same stack through the FP64 C restatement on one core, "
                                         f"{tc:.2f} s (conversion excluded)"}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the path shards ROIs across GPUs with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    import zola_hip as Z

    torch.cuda.init()
    stream = torch.cuda.current_stream()
    coeffs = Z.astigmatic_coeffs(28)
    ctx = Z.Context(Z.optics(), R, coeffs)
    ctx.set_stream(stream.cuda_stream)
    B = args.batch
    px, truth = synth(ctx, torch, B, seed=0x20A1D3)
    init = torch.empty(B, 5, dtype=torch.float64, device="cuda")
    theta = torch.empty_like(init)
    nll = torch.empty(B, dtype=torch.float64, device="cuda")
    crlb = torch.empty(B, 3, dtype=torch.float64, device="cuda")
    iters = torch.empty(B, dtype=torch.int32, device="cuda")
    status = torch.empty(B, dtype=torch.int32, device="cuda")
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
    # the default layout of the LM form at 21 px: two waves (one block) per ROI
    ctx.set_fit_layout(R, Z.FIT_LM, Z.LAYOUT_PAIR)
    kernel_name = "k_fit_lm64p<21>"

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fit_ms = []

    def step(timed):
        if timed:
            ev[0].record(stream)
        ctx.fit_init_dev(px, NZ, ZMIN_T, ZSTEP_T, 0.0, init)
        if timed:
            ev[1].record(stream)
        ctx.fit_dev(cfg, px, init, theta, nll, crlb, iters, status)
        if timed:
            ev[2].record(stream)

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    init_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
        torch.cuda.synchronize()
        init_ms.append(ev[0].elapsed_time(ev[1]))
        fit_ms.append(ev[1].elapsed_time(ev[2]))
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = B / (ms_per_step / 1e3)

    st = status.cpu().numpy()
    nfield = (st >> 8).astype(np.int64)
    it_np = iters.cpu().numpy()
    th = theta.cpu().numpy()
    tr = truth.cpu().numpy()
    fit_kernel_ms = float(np.mean(fit_ms))
    flops = float(nfield.sum()) * F64
    achieved = flops / (fit_kernel_ms / 1e3)
    conv = float(np.mean((st & Z.ST_CONVERGED) != 0))
    err_nm = np.abs(th[:, 2:] - tr[:, 2:]) * 1000.0

    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_fit_lm64_r21.json")
    if os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f)
        if t.get("batch") == B and t.get("kernel") == kernel_name:
            traffic = t.get("bytes_per_launch")

    out = {
        "metric": "MLE localizations/s (21^2 ROI, 64^2 pupil), batched LM with analytic "
                  "tangents vs plain GPU; %roofline",
        "value": round(value, 1),
        "unit": "localizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 PSF synthesis, f64 likelihood/normal equations",
        "data": "synthetic (GPU-generated Poisson ROIs from the library's PSF model, seed 0x20A1D3)",
        "config": {
            "workload": f"isolated-emitter localization: {B} ROIs of 21x21 px per step, 64x64 "
                        "pupil, 28 Zernike modes (astigmatic), init + LM fit (max 30 it, "
                        "tol 1e-4 um)",
            "batch_rois": B, "roi": R, "pupil": 64, "zernike_modes": 28,
        },
        "roofline": {
            "bound": "valu",
            "kernel": kernel_name,
            "achieved": round(achieved / 1e12, 4),
            "peak": PEAK_FP32 / 1e12,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32, 5),
            "traffic": traffic,
            "kernel_ms": round(fit_kernel_ms, 4),
            "algorithmic_flops_per_launch": flops,
            "fields_per_roi": round(float(nfield.mean()), 3),
        },
        "fit": {
            "init_kernel_ms": round(float(np.mean(init_ms)), 4),
            "iters_mean": round(float(it_np.mean()), 3),
            "converged_frac": round(conv, 5),
            "median_abs_err_vs_truth_nm": [round(float(v), 3) for v in np.median(err_nm, axis=0)],
        },
    }

    if not args.no_plain:
        Bp = min(args.plain_batch, B)
        cfgp = ctx.fit_config(mode=Z.FIT_REFERENCE, max_iter=20, zmin=-1.3, zmax=1.3,
                              photon_threshold=0.0)
        thp = torch.empty(Bp, 5, dtype=torch.float64, device="cuda")
        nlp = torch.empty(Bp, dtype=torch.float64, device="cuda")
        crp = torch.empty(Bp, 3, dtype=torch.float64, device="cuda")
        itp = torch.empty(Bp, dtype=torch.int32, device="cuda")
        stp = torch.empty(Bp, dtype=torch.int32, device="cuda")
        pxp, inp = px[:Bp].contiguous(), init[:Bp].contiguous()
        ctx.fit_dev(cfgp, pxp, inp, thp, nlp, crp, itp, stp)   # warm-up
        torch.cuda.synchronize()
        ev[0].record(stream)
        ctx.fit_dev(cfgp, pxp, inp, thp, nlp, crp, itp, stp)
        ev[1].record(stream)
        torch.cuda.synchronize()
        pms = ev[0].elapsed_time(ev[1])
        nfp = (stp.cpu().numpy() >> 8).astype(np.int64)
        pflops = float(nfp.sum()) * F64
        plain_rate = Bp / (pms / 1e3)
        out["plain_gpu"] = {
            "form": "reference coordinate descent (k_fit_ref64<21,4>), FD steps of "
                    "Localization.java, one PSF synthesis per likelihood",
            "rois": Bp, "kernel_ms": round(pms, 3),
            "value": round(plain_rate, 1),
            "achieved_tflops": round(pflops / (pms / 1e3) / 1e12, 4),
            "frac": round(pflops / (pms / 1e3) / PEAK_FP32, 5),
            "fields_per_roi": round(float(nfp.mean()), 2),
            "iters_mean": round(float(itp.float().mean().item()), 3),
        }
        # both forms fit the same ROIs from the same start: compare on common ROIs
        lm_rate_kernel = B / ((np.mean(fit_ms) + np.mean(init_ms)) / 1e3)
        out["batched_vs_plain"] = round(lm_rate_kernel / (Bp / ((pms + np.mean(init_ms) * Bp / B) / 1e3)), 3)

    if not args.no_host:
        # The same ROIs through the host-array entry point (zola_fit_batch):
        # pixels and starts copied in from pageable host memory, results
        # copied out, inside the timed call.  Starts come from the device
        # initialiser above.
        px_h = px.cpu().numpy()
        init_h = init.cpu().numpy()
        ctx.fit(cfg, px_h, init_h)
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            res_h = ctx.fit(cfg, px_h, init_h)
        th_ms = (time.perf_counter() - t0) * 1e3 / reps
        out["host_arrays"] = {
            "rois": B, "call_ms": round(th_ms, 3), "value": round(B / (th_ms / 1e3), 1),
            "unit": "localizations/s",
            "in_bytes": int(px_h.nbytes + init_h.nbytes),
            "same_theta_as_device_form": bool(np.array_equal(res_h["theta"], theta.cpu().numpy())),
            "note": "zola_fit_batch on host arrays: PCIe copies in and out included, init not included",
        }
    if not args.no_calib:
        out["calibration"] = calibration_bench(torch, Z, stream, cpu=not args.no_cpu)
        out["calibration_n256"] = calibration256_bench(torch, Z, stream)
    if not args.no_n128:
        out["depth_n128"] = depth128_bench(torch, Z, stream, args.n128_batch, args.n128_layout)
    if not args.no_mixed:
        out["mixed_sizes"] = mixed_bench(torch, Z, stream, args.mixed_total)
    if not args.no_frames:
        out["frames_n128"] = frames_bench(torch, Z, stream, args.frames)
    if not args.no_formats:
        out["formats"] = formats_bench(Z)
    if not args.no_beads:
        out["bead_prep"] = beads_bench(Z, stream, cpu=not args.no_cpu)

    if args.copy_probe_mb:
        n = args.copy_probe_mb * (1 << 20) // 4
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = 2 * n * 4

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads = host_cores()
        S = args.cpu_sample or threads * 6
        S = min(S, B)
        rois = px[:S].double().cpu().numpy()
        ini = init[:S].cpu().numpy()
        t0 = time.perf_counter()
        cpu = O.localize_batch(O.optics(), R, coeffs, rois, ini, minZ=-1.3, maxZ=1.3,
                               iterMax=20, thr=0.0, nthreads=threads)
        dt = time.perf_counter() - t0
        # Agreement of the batched LM fit with the CPU path on the same ROIs
        # and starts: the CPU path is the reference's coordinate descent
        # (cpu/Localization.java:491-653), capped at iterMax = 20.
        d_nm = np.abs(th[:S, 2:] - cpu["theta"][:, 2:]) * 1000.0
        close = (d_nm[:, 0] <= 0.5) & (d_nm[:, 1] <= 0.5) & (d_nm[:, 2] <= 1.0)
        capped = cpu["iters"] >= 20
        dnll = nll[:S].cpu().numpy() - cpu["lik"]
        out["lm_vs_cpu_path"] = {
            "rois": S,
            "within_0.5_0.5_1nm": round(float(close.mean()), 4),
            "within_among_cpu_converged": round(float(close[~capped].mean()), 4)
            if (~capped).any() else None,
            "cpu_hit_iter_max": round(float(capped.mean()), 4),
            "median_abs_diff_nm": [round(float(v), 4) for v in np.median(d_nm, axis=0)],
            "nll_lm_minus_cpu": {"median": round(float(np.median(dnll)), 6),
                                 "min": round(float(dnll.min()), 6),
                                 "max": round(float(dnll.max()), 6)},
            "note": "LM results vs the FP64 restatement of the CPU path (iterMax 20); a "
                    "negative NLL difference means the LM fit reached a lower Poisson NLL",
        }
        out["cpu_baseline"] = {
            "value": round(S / dt, 3),
            "unit": "localizations/s",
            "cores": threads,
            "kind": "port",
            "sample": f"first {S} ROIs of the same batch, same initial values; FP64 C "
                      f"restatement of cpu/Localization.localize + PSFPhaseMany (no PSF cache), "
                      f"{dt:.1f} s on {threads} threads",
            "host": host_info(),
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
