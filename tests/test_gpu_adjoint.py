"""GPU adjoint (zola_psf_backward) against the oracle's FP64 adjoint, which is
itself pinned by central finite differences (tests/test_oracle_pins.py)."""
import numpy as np
import pytest

import oracle as O
import zola_hip as Z

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("R,J", [(21, 28), (31, 45)])
def test_backward_matches_oracle_adjoint(R, J):
    rng = np.random.default_rng(R + J)
    a = np.zeros(J)
    a[3] = 0.8
    a[4:] = rng.normal(0, 0.15, J - 4)
    m = O.Model(O.optics(), R, a)
    ctx = Z.Context(Z.optics(), R, a)
    B = 9
    params = np.stack([rng.uniform(-0.1, 0.1, B), rng.uniform(-0.1, 0.1, B),
                       rng.uniform(-0.9, 0.9, B), rng.uniform(-0.2, 0.2, B)], 1).astype(np.float32)
    w = rng.normal(size=(B, R, R)).astype(np.float32)
    dp, da = ctx.psf_backward(params, w)
    ref_dp = np.zeros((B, 4))
    ref_da = np.zeros(J)
    for b in range(B):
        x, y, z, zo = params[b].astype(np.float64)
        g, ga = m.psf_vjp(x, y, zo, z, w[b].astype(np.float64))
        ref_dp[b] = g
        ref_da += ga
    err_p = np.abs(dp - ref_dp).max(axis=0) / np.abs(ref_dp).max(axis=0)
    err_a = np.abs(da - ref_da).max() / np.abs(ref_da).max()
    print("rel err dparams", err_p, "da", err_a)
    assert np.all(err_p < 1e-6)
    assert err_a < 1e-6
