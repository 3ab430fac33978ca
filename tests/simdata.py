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
