"""Seeded synthetic ROIs for the parity tests (ground truth from the FP64 oracle)."""
import numpy as np

import oracle as O
import zola_hip as Z

DX = 0.11


def isolated_rois(model, B, seed, zrange=(-1.0, 1.0), photons=(500, 5000), bg=(5, 50)):
    """Isolated emitters, one per ROI: x, y ~ U(+-1 px), z ~ U(zrange), photons
    log-uniform, background uniform, Poisson noise, float32 pixels."""
    rng = np.random.default_rng(seed)
    R = model.R
    x = rng.uniform(-DX, DX, B)
    y = rng.uniform(-DX, DX, B)
    z = rng.uniform(*zrange, B)
    A = np.exp(rng.uniform(np.log(photons[0]), np.log(photons[1]), B))
    Bg = rng.uniform(*bg, B)
    rois = np.zeros((B, R, R), np.float32)
    for b in range(B):
        mu = A[b] * model.psf(x[b], y[b], 0.0, z[b]) + Bg[b]
        rois[b] = rng.poisson(mu).astype(np.float32)
    truth = np.stack([A, Bg, x, y, z], axis=1)
    return rois, truth


def templates(model, zmin, zstep, nZ):
    """PSF(0, 0, z_r) stack used by the isolated-ROI initialisation."""
    return np.stack([model.psf(0.0, 0.0, 0.0, zmin + r * zstep) for r in range(nZ)])


def init_restated(rois, tmpl, zmin, zstep):
    """CPU restatement of zola_fit_init_dev: best zero-shift Pearson correlation
    against the templates, z refined by the parabola of Predetection.java:589-593,
    A/B by getAandB (LocalizationPipeline.java:1119-1162)."""
    B = rois.shape[0]
    nZ = tmpl.shape[0]
    out = np.zeros((B, 5))
    for b in range(B):
        n = rois[b].astype(np.float64).ravel()
        c = np.array([np.corrcoef(t.ravel(), n)[0, 1] for t in tmpl])
        r = int(np.argmax(c))
        z0 = zmin + r * zstep
        if 1 < r < nZ - 1:
            ya, yb, yc = c[r - 1], c[r], c[r + 1]
            xa, xb, xc = -zstep, 0.0, zstep
            a = (yc - ya) / ((xc - xa) * (xc - xb)) - (yb - ya) / ((xb - xa) * (xc - xb))
            bq = ((yb - ya) / (xb - xa)) - a * (xb + xa)
            z0 += -bq / (2 * a)
        AB = O.get_a_and_b(rois[b].astype(np.float64), tmpl[r], 0.1)
        out[b] = [AB[0], AB[1], 0.0, 0.0, z0]
    return out


def render_frame(model, W, H, emitters, bg, seed=None):
    """Frame W x H ([i][ii], i the x column) with emitters (pi, pj, z, A): each
    rendered as A * PSF on the R x R patch around round(pi), round(pj), clipped
    at the frame edge (the
    model's emitter sits at pixel R//2 - x/dx, so x = (ci - pi) dx), plus a
    uniform background, Poisson noise when seed is given; float32."""
    R = model.R
    mu = np.full((W, H), float(bg))
    for (pi, pj, z, A) in emitters:
        ci, cj = int(round(pi)), int(round(pj))
        p = A * model.psf((ci - pi) * DX, (cj - pj) * DX, 0.0, z)
        i0, j0 = ci - R // 2, cj - R // 2
        a0, b0 = max(0, -i0), max(0, -j0)
        a1, b1 = min(R, W - i0), min(R, H - j0)
        mu[i0 + a0:i0 + a1, j0 + b0:j0 + b1] += p[a0:a1, b0:b1]
    if seed is None:
        return mu.astype(np.float32)
    return np.random.default_rng(seed).poisson(mu).astype(np.float32)


def random_emitters(W, H, n, seed, margin=12, zrange=(-0.6, 0.6), photons=(1500, 5000), min_sep=0.0):
    """n emitters uniform in the frame interior, z uniform, photons log-uniform;
    with min_sep > 0, positions closer than min_sep px to an earlier one are redrawn."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        pi = rng.uniform(margin, W - margin)
        pj = rng.uniform(margin, H - margin)
        if min_sep > 0 and any((pi - a) ** 2 + (pj - b) ** 2 < min_sep ** 2 for a, b, _, _ in out):
            continue
        z = rng.uniform(*zrange)
        A = float(np.exp(rng.uniform(np.log(photons[0]), np.log(photons[1]))))
        out.append((pi, pj, z, A))
    return out
