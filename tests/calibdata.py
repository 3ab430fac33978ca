"""Bead z-stacks for the calibration tests (ground truth from the FP64 oracle)
and an FP64 restatement of the calibration likelihood with its gradient."""
import numpy as np

import oracle as O


def bead_stacks(model, nb, S, zstep, seed, photons=2e5, bg=10.0):
    rng = np.random.default_rng(seed)
    R = model.R
    dz = (np.arange(S) - S // 2) * zstep          # ZernikePhaseRetrieval.java:131-140
    beads = np.stack([rng.uniform(-0.05, 0.05, nb), rng.uniform(-0.05, 0.05, nb),
                      rng.uniform(-0.05, 0.05, nb), np.full(nb, photons), np.full(nb, bg)], 1)
    stacks = np.zeros((nb, S, R, R), np.float32)
    for b in range(nb):
        x, y, z0, A, B = beads[b]
        for s in range(S):
            mu = A * model.psf(x, y, dz[s] + z0, 0.0) + B
            stacks[b, s] = rng.poisson(mu)
    return stacks, dz, beads


def calib_nll_grad(model, stacks, dz, beads):
    """NLL and gradient (coefficients, then per bead x, y, z0, A, bg) in FP64."""
    nb, S, R, _ = stacks.shape
    J = model.J
    f = 0.0
    ga = np.zeros(J)
    gb = np.zeros((nb, 5))
    for b in range(nb):
        x, y, z0, A, B = beads[b]
        for s in range(S):
            psf = model.psf(x, y, dz[s] + z0, 0.0)
            n = stacks[b, s].astype(np.float64)
            mu = np.maximum(A * psf + B, 1e-8)
            f += np.sum(mu - n * np.log(mu))
            r = 1.0 - n / mu
            g, gj = model.psf_vjp(x, y, dz[s] + z0, 0.0, A * r)
            ga += gj
            gb[b] += [g[0], g[1], g[3], np.sum(psf * r), np.sum(r)]
    return f, ga, gb
