"""GPU PSF synthesis (zola_psf_forward) against the FP64 oracle restatement of
PSFPhaseMany.computePSF (process/cpu/PSFPhaseMany.java:553-622)."""
import numpy as np
import pytest

import oracle as O
import zola_hip as Z

pytestmark = pytest.mark.gpu


def astig_coeffs(J, seed=5):
    rng = np.random.default_rng(seed)
    a = np.zeros(J)
    a[4:] = rng.normal(0, 0.05, J - 4)
    a[3] = 1.0
    return a


@pytest.mark.parametrize("R", [17, 21, 31, 41])
def test_psf_forward_matches_oracle(R):
    J = 28
    a = astig_coeffs(J)
    opt_o = O.optics(size_fft=64)
    opt_z = Z.optics(size_fft=64)
    m = O.Model(opt_o, R, a)
    ctx = Z.Context(opt_z, R, a)
    rng = np.random.default_rng(R)
    B = 37
    params = np.stack([rng.uniform(-0.11, 0.11, B), rng.uniform(-0.11, 0.11, B),
                       rng.uniform(-1, 1, B), np.zeros(B)], axis=1)
    got = ctx.psf_forward(params)
    err = 0.0
    for b in range(B):
        x, y, z, zoil = params[b].astype(np.float32).astype(np.float64)
        ref = m.psf(x, y, zoil, z)
        err = max(err, np.abs(got[b] - ref).max() / ref.max())
    print("max err / peak", err)
    assert err < 1.5e-6
