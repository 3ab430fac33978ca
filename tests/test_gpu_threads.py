"""Host threads sharing one context.

The reference drives one model object from up to 100 Java threads per part
(process/gpu/Modelmany_double_.java:865-940).  Here every entry point locks its
context, so threads may share a handle: concurrent calls must give exactly the
results of the same calls made one after another, and an error raised in one
thread must not leak into another thread's message."""
import threading

import numpy as np
import pytest

import zola_hip as Z

pytestmark = pytest.mark.gpu

R, B, THREADS, CALLS = 21, 64, 4, 3


def _batches(ctx):
    rng = np.random.default_rng(0x7E4D)
    out = []
    for t in range(THREADS * CALLS):
        par = np.stack([rng.uniform(-0.1, 0.1, B), rng.uniform(-0.1, 0.1, B),
                        rng.uniform(-0.8, 0.8, B), np.zeros(B)], 1).astype(np.float32)
        psf = ctx.psf_forward(par)
        A = rng.uniform(800, 4000, B)
        bg = rng.uniform(5, 40, B)
        rois = rng.poisson(A[:, None, None] * psf + bg[:, None, None]).astype(np.float32)
        init = np.stack([0.8 * A, bg, np.zeros(B), np.zeros(B), par[:, 2] + 0.1], 1)
        out.append((par, rois, init))
    return out


def test_threads_sharing_a_context_match_sequential_calls(torch_cuda):
    ctx = Z.Context(Z.optics(), R, Z.astigmatic_coeffs(28))
    cfg = ctx.fit_config(mode=Z.FIT_LM, max_iter=30, zmin=-1.3, zmax=1.3, tol=1e-4)
    work = _batches(ctx)
    seq = [(ctx.psf_forward(p), ctx.fit(cfg, r, i)) for (p, r, i) in work]

    got = [None] * len(work)
    errors = []

    def run(t):
        try:
            for k in range(CALLS):
                n = t * CALLS + k
                p, r, i = work[n]
                got[n] = (ctx.psf_forward(p), ctx.fit(cfg, r, i))
        except Exception as e:  # reported below, with the thread
            errors.append((t, repr(e)))

    th = [threading.Thread(target=run, args=(t,)) for t in range(THREADS)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in th), "a thread did not finish"
    assert not errors, errors
    for n, ((p0, f0), (p1, f1)) in enumerate(zip(seq, got)):
        assert np.array_equal(p0, p1), n
        for key in ("theta", "nll", "crlb", "iters", "status"):
            assert np.array_equal(f0[key], f1[key]), (n, key)


def test_error_messages_are_per_thread(torch_cuda):
    """A thread whose call fails reads its own message; a thread whose calls
    succeed is not disturbed by it."""
    import ctypes as C
    ctx = Z.Context(Z.optics(), R, Z.astigmatic_coeffs(28))
    lib = Z.lib()
    msgs = {}

    def bad():
        rc = lib.zola_set_sigma(ctx._h, C.c_double(-1.0))
        msgs["bad"] = (rc, lib.zola_last_error(ctx._h))

    def good():
        par = np.zeros((8, 4), np.float32)
        for _ in range(20):
            ctx.psf_forward(par)
        msgs["good"] = True

    th = [threading.Thread(target=bad), threading.Thread(target=good)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=60)
    assert msgs.get("good") is True
    rc, m = msgs["bad"]
    assert rc == -1 and b"sigma" in m, (rc, m)   # ZOLA_EINVAL
