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


def monomials(R):
    """The six background monomials of PolynomialBackground (order 2) over an R x R
    patch, in its order: 1, i, i^2, ii, i ii, ii^2 (i the first index)."""
    i = np.arange(R, dtype=np.float64)[:, None] * np.ones((1, R))
    ii = np.ones((R, 1)) * np.arange(R, dtype=np.float64)[None, :]
    return np.stack([np.ones((R, R)), i, i * i, ii, i * ii, ii * ii])


def full_stacks(model, nb, S, zstep, seed, photons=2e5, sigma_true=None):
    """Bead z-stacks of the reference's full calibration model: a per-slice photon
    curve, an order-2 background per bead and, when sigma_true is given, that blur."""
    rng = np.random.default_rng(seed)
    R = model.R
    dz = (np.arange(S) - S // 2) * zstep
    xyz = np.stack([rng.uniform(-0.05, 0.05, nb), rng.uniform(-0.05, 0.05, nb), rng.uniform(-0.05, 0.05, nb)], 1)
    A = photons * (1.0 + 0.15 * np.sin(np.arange(S)[None, :] / 5.0 + rng.uniform(0, 6, (nb, 1))))
    poly = np.zeros((nb, 6))
    poly[:, 0] = rng.uniform(8, 14, nb)
    poly[:, 1] = rng.uniform(-0.1, 0.1, nb)
    poly[:, 3] = rng.uniform(-0.1, 0.1, nb)
    poly[:, 2] = rng.uniform(-2e-3, 2e-3, nb)
    poly[:, 4] = rng.uniform(-2e-3, 2e-3, nb)
    poly[:, 5] = rng.uniform(-2e-3, 2e-3, nb)
    M = monomials(R)
    if sigma_true is not None:
        model.set_sigma(sigma_true)
    stacks = np.zeros((nb, S, R, R), np.float32)
    for b in range(nb):
        bg = np.tensordot(poly[b], M, 1)
        for s in range(S):
            mu = A[b, s] * model.psf(xyz[b, 0], xyz[b, 1], dz[s] + xyz[b, 2], 0.0) + bg
            stacks[b, s] = rng.poisson(np.maximum(mu, 0.0))
    if sigma_true is not None:
        model.set_sigma(1.0)
    return stacks, dz, dict(xyz=xyz, A=A, poly=poly)


def full_nll_grad(model, stacks, dz, A, xyz, poly, fa, fae, offset=None):
    """NLL of the full model (ZernikePhaseRetrieval.Model.computeLikelihood) and its
    gradient in FP64: coefficients, per-bead x, y, z0, background coefficients,
    fit_a, and per-slice fit_a_each.  Clamped pixels (mu <= 0 -> 1e-8) add a
    constant, so no gradient."""
    nb, S, R, _ = stacks.shape
    M = monomials(R)
    f = 0.0
    ga = np.zeros(model.J)
    gx = np.zeros((nb, 3))
    gp = np.zeros((nb, 6))
    gfa = np.zeros(nb)
    gfae = np.zeros((nb, S))
    for b in range(nb):
        bg = np.tensordot(poly[b], M, 1) + (0.0 if offset is None else offset[b])
        x, y, z0 = xyz[b]
        for s in range(S):
            amp = A[b, s] + fa[b] + fae[b, s]
            psf = model.psf(x, y, dz[s] + z0, 0.0)
            n = stacks[b, s].astype(np.float64)
            mu = amp * psf + bg
            cl = mu <= 0
            mu = np.where(cl, 1e-8, mu)
            f += float(np.sum(mu - n * np.log(mu)))
            r = np.where(cl, 0.0, 1.0 - n / mu)
            g, gj = model.psf_vjp(x, y, dz[s] + z0, 0.0, amp * r)
            ga += gj
            gx[b] += [g[0], g[1], g[3]]
            gp[b] += np.tensordot(M, r, axes=([1, 2], [0, 1]))
            d = float(np.sum(psf * r))
            gfa[b] += d
            gfae[b, s] = d
    return f, dict(a=ga, xyz=gx, poly=gp, fa=gfa, fae=gfae)
