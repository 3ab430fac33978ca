"""The FP64 restatement of the reference calibration loop (oracle/zola_calib_ref.c),
which times the CPU baseline of the calibration benchmark.  Its likelihood is the
full calibration model's (tests/calibdata.full_nll_grad, pinned by central
differences and by the clamped Poisson sum of the oracle), before and after an
iteration; the registration coefficients of basis slots kx, ky, kz_oil act as
(x, y, -z0); one iteration lowers the likelihood and leaves a result that does not
depend on the thread count.  The step rules of its individual updates are pinned
only by restatement."""
import numpy as np

import calibdata as cd
import oracle as O


def _setup():
    a = np.random.default_rng(3).normal(0, 0.15, 15)
    a[:3] = 0.0
    a[3] = 0.8
    m = O.Model(O.optics(), 21, a)
    st, dz, tr = cd.full_stacks(m, 2, 9, 0.1, seed=7, photons=5e4)
    c0 = tr["poly"][:, 0]
    A0 = st.reshape(2, 9, -1).sum(2) - 21 * 21 * c0[:, None]
    start = a.copy()
    start[3:] += np.random.default_rng(8).normal(0, 0.05, 12)
    return m, st, dz, A0, c0, start


def _full(m, st, dz, A0, state):
    m.set_coeffs(np.r_[np.zeros(3), state["coeffs"][3:]])
    m.set_sigma(state["sigma"])
    xyz = np.stack([state["reg"][0], state["reg"][1], -state["reg"][2]], 1)
    f, _ = cd.full_nll_grad(m, st, dz, A0, xyz, state["poly"], state["fit_a"], state["fit_a_each"])
    m.set_sigma(1.0)
    return f


def test_calibration_restatement_likelihood_and_one_iteration():
    m, st, dz, A0, c0, start = _setup()
    ref = O.CalibRef(O.optics(), start, st, 0.1, A0, c0, sigma=1.0, nthreads=3)
    f0 = ref.likelihood()
    s0 = ref.state()
    assert abs(f0 - _full(m, st, dz, A0, s0)) <= 1e-10 * abs(f0)
    f1 = ref.iteration()
    s1 = ref.state()
    assert f1 < f0
    assert np.abs(s1["reg"]).max() > 0 and s1["sigma"] != 1.0 and np.abs(s1["poly"][:, 1:]).max() > 0
    assert abs(f1 - _full(m, st, dz, A0, s1)) <= 1e-10 * abs(f1)
    ref1 = O.CalibRef(O.optics(), start, st, 0.1, A0, c0, sigma=1.0, nthreads=1)
    assert ref1.iteration() == f1
    for k in s1:
        np.testing.assert_array_equal(ref1.state()[k], s1[k])
