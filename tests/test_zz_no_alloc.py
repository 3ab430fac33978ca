"""zola_psf_backward_noalloc_dev, the allocation-free form of the PSF adjoint,
called through the C ABI: on two calls in a row, the second with a smaller
batch, its outputs equal those of zola_psf_backward_dev (same kernels, same
inputs: bit for bit); a batch above ZOLA_NOALLOC_MAX_PSF is rejected with a
clear error."""
import ctypes as C

import numpy as np
import pytest

import zola_hip as Z

pytestmark = pytest.mark.gpu
R, J = 21, 28
MAX_PSF = 2048


def _call(name, ctx, params, w, dparams, da):
    fn = getattr(Z.lib(), name)
    return fn(ctx._h, int(params.shape[0]), C.c_void_p(params.data_ptr()), C.c_void_p(w.data_ptr()),
              C.c_void_p(dparams.data_ptr()), C.c_void_p(da.data_ptr()) if da is not None else None)


def test_noalloc_backward_equals_backward_dev(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(77)
    a = Z.astigmatic_coeffs(J)
    ctx = Z.Context(Z.optics(), R, a)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    try:
        for B in (300, 37):
            print(f"case: allocation-free adjoint, B = {B}", flush=True)
            par = np.stack([rng.uniform(-0.1, 0.1, B), rng.uniform(-0.1, 0.1, B), rng.uniform(-0.9, 0.9, B),
                            rng.uniform(-0.2, 0.2, B)], 1).astype(np.float32)
            P = torch.from_numpy(par).cuda()
            W = torch.from_numpy(rng.normal(size=(B, R, R)).astype(np.float32)).cuda()
            outs = []
            for name in ("zola_psf_backward_dev", "zola_psf_backward_noalloc_dev"):
                dp = torch.full((B, 4), float("nan"), device="cuda")
                da = torch.zeros(J, dtype=torch.float64, device="cuda")
                assert _call(name, ctx, P, W, dp, da) == 0, Z.lib().zola_last_error(ctx._h)
                torch.cuda.synchronize()
                outs.append((dp.cpu().numpy(), da.cpu().numpy()))
            np.testing.assert_array_equal(outs[1][0], outs[0][0])
            np.testing.assert_array_equal(outs[1][1], outs[0][1])
            assert np.isfinite(outs[1][0]).all() and np.abs(outs[1][1]).max() > 0
        # without da no scratch is touched; dparams still equal
        dp0 = torch.zeros((37, 4), device="cuda")
        assert _call("zola_psf_backward_noalloc_dev", ctx, P, W, dp0, None) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dp0.cpu().numpy(), outs[0][0])
        # above the limit: rejected before any launch
        big = torch.zeros((MAX_PSF + 1, 4), device="cuda")
        wbig = torch.zeros((MAX_PSF + 1, R, R), device="cuda")
        rc = _call("zola_psf_backward_noalloc_dev", ctx, big, wbig, big.clone(),
                   torch.zeros(J, dtype=torch.float64, device="cuda"))
        assert rc == -1
        assert b"allocation-free limit" in Z.lib().zola_last_error(ctx._h)
    finally:
        ctx.close()
