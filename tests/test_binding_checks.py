"""The Python binding checks every array's shape, dtype, contiguity and device
before it calls the library, which trusts the sizes it is given (a short theta
or a float64 pixel tensor would otherwise be read or written out of bounds).
Runs without a GPU: the library entry point is replaced by one that fails the
test if it is reached."""
import numpy as np
import pytest

import zola_hip as Z

R, J = 21, 28


@pytest.fixture
def ctx(monkeypatch):
    def tripwire():
        raise AssertionError("a native call was reached")
    monkeypatch.setattr(Z, "lib", tripwire)
    c = object.__new__(Z.Context)
    c._h, c.R, c.J, c.N, c.D, c.s2 = None, R, J, 64, 716, 64
    c.opt = Z.optics()
    return c


def _t(shape, dtype="float32"):
    import torch
    return torch.zeros(shape, dtype=getattr(torch, dtype))


def test_host_forms_reject_mismatched_shapes(ctx):
    cfg = ctx.fit_config()
    with pytest.raises(ValueError, match="pixels"):
        ctx.fit(cfg, np.zeros((4, R, R + 1)), np.zeros((4, 5)))
    with pytest.raises(ValueError, match="init"):
        ctx.fit(cfg, np.zeros((4, R, R)), np.zeros((3, 5)))
    with pytest.raises(ValueError, match="params"):
        ctx.psf_forward(np.zeros((4, 3)))
    with pytest.raises(ValueError, match="w"):
        ctx.psf_backward(np.zeros((4, 4)), np.zeros((5, R, R)))
    with pytest.raises(ValueError, match="coeffs"):
        ctx.set_coefficients(np.zeros(J - 1))
    st = np.zeros((2, 9, R, R))
    with pytest.raises(ValueError, match="bead"):
        ctx.calib_fit(st, np.zeros(9), np.zeros((3, 5)), np.zeros(J))
    with pytest.raises(ValueError, match="dz"):
        ctx.calib_fit(st, np.zeros(8), np.zeros((2, 5)), np.zeros(J))
    with pytest.raises(ValueError, match="fit_mask"):
        ctx.calib_fit(st, np.zeros(9), np.zeros((2, 5)), np.zeros(J), fit_mask=np.ones(J - 3))
    with pytest.raises(ValueError, match="xp, yp"):
        ctx.bead_prepare(np.zeros((5, 16, 16)), [1.0, 2.0], [1.0], 8, 0.05)
    with pytest.raises(ValueError, match="gain"):
        ctx.bead_prepare_scmos(np.zeros((5, 16, 16)), np.zeros((16, 16)), np.zeros((16, 15)),
                               np.zeros((16, 16)), [1.0], [1.0], 8, 0.05)


def test_device_forms_reject_host_wrong_dtype_and_short_tensors(ctx):
    pytest.importorskip("torch")
    cfg = ctx.fit_config()
    B = 4
    good = dict(pixels=_t((B, R, R)), init=_t((B, 5), "float64"), theta=_t((B, 5), "float64"),
                nll=_t((B,), "float64"), crlb=_t((B, 3), "float64"), iters=_t((B,), "int32"),
                status=_t((B,), "int32"))
    # CPU tensors are not device memory
    with pytest.raises(ValueError, match="not in GPU memory"):
        ctx.fit_dev(cfg, **good)
    with pytest.raises(ValueError, match="expected a torch tensor"):
        ctx.fit_dev(cfg, **dict(good, pixels=np.zeros((B, R, R), np.float32)))
    with pytest.raises(ValueError, match="not in GPU memory"):
        ctx.psf_forward_dev(_t((B, 4)), _t((B, R, R)))
    with pytest.raises(ValueError, match="not in GPU memory"):
        ctx.calib_eval_dev(_t((B, 4)), _t((B, R, R)), *[_t((B,), "float64")] * 3, _t((B, 4)),
                           *[_t((B,), "float64")] * 2)


def test_device_checks_on_gpu_shaped_tensors(ctx, monkeypatch):
    """The same checks past the device test: a stand-in tensor type that reports
    itself in GPU memory carries wrong dtypes, shapes and strides."""
    torch = pytest.importorskip("torch")

    class Dev:
        def __init__(self, shape, dtype=torch.float32, contiguous=True):
            self.shape, self.dtype, self._c = tuple(shape), dtype, contiguous
            self.is_cuda = True

        def is_contiguous(self):
            return self._c

        def data_ptr(self):
            raise AssertionError("a pointer was taken")

    cfg = ctx.fit_config()
    B = 4
    good = dict(pixels=Dev((B, R, R)), init=Dev((B, 5), torch.float64), theta=Dev((B, 5), torch.float64),
                nll=Dev((B,), torch.float64), crlb=Dev((B, 3), torch.float64), iters=Dev((B,), torch.int32),
                status=Dev((B,), torch.int32))
    for key, bad, msg in (("pixels", Dev((B, R, R), torch.float64), "dtype"),
                          ("theta", Dev((B - 1, 5), torch.float64), "shape"),
                          ("status", Dev((B,), torch.int64), "dtype"),
                          ("crlb", Dev((B, 3), torch.float64, contiguous=False), "contiguous")):
        with pytest.raises(ValueError, match=msg):
            ctx.fit_dev(cfg, **dict(good, **{key: bad}))
    with pytest.raises(ValueError, match="varg"):
        ctx.fit_ext_dev(cfg, **good, varg=Dev((B, R, R - 1)))
    with pytest.raises(ValueError, match="zoil"):
        ctx.fit_zoil_dev(cfg, good["pixels"], good["init"], Dev((B + 1,)), good["theta"], good["nll"],
                         good["crlb"], good["iters"], good["status"])
    with pytest.raises(ValueError, match="da"):
        ctx.psf_backward_dev(Dev((B, 4)), Dev((B, R, R)), Dev((B, 4)), Dev((J + 1,), torch.float64))
    with pytest.raises(ValueError, match="groups"):
        ctx.fit_mixed_dev(cfg, [])
