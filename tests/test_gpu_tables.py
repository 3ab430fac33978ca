"""Coefficient and sigma updates on a live context (the calibration loops'
path): on the 64 and 256 grids zola_set_coefficients forms the phase tables on
the device and zola_set_sigma (256 grid) the Gaussian spectrum; the host's FP64
phase is recomputed only when a host routine needs it.  A context updated this
way must synthesise the same PSFs, bit for bit, as a context created with the
final coefficients, and its host-side routines (the focus-centre search) must
see the new phase.  The phase is bit-exact (the device forms the host's
products and sums in the host's order); the 256 grid's Gaussian uses the
device exp, so there PSFs after a sigma update agree to 1e-7 of the peak."""
import numpy as np
import pytest

import zola_hip as Z

pytestmark = pytest.mark.gpu


def _coeffs(J, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(0, 0.2, J)
    a[:3] = 0.0
    return a


@pytest.mark.parametrize("size_fft,R,J", [(64, 21, 28), (140, 64, 45)])
def test_updated_context_equals_fresh_context(size_fft, R, J):
    print(f"case: table update, size_fft {size_fft}, R {R}, J {J}", flush=True)
    a0, a1 = _coeffs(J, 1), _coeffs(J, 2)
    rng = np.random.default_rng(3)
    par = np.stack([rng.uniform(-0.1, 0.1, 6), rng.uniform(-0.1, 0.1, 6), rng.uniform(-0.8, 0.8, 6),
                    np.zeros(6)], 1).astype(np.float32)
    upd = Z.Context(Z.optics(size_fft=size_fft), R, a0)
    upd.set_coefficients(a1)
    upd.set_sigma(1.2)
    fresh = Z.Context(Z.optics(size_fft=size_fft, sigma=1.2), R, a1)
    got, ref = upd.psf_forward(par), fresh.psf_forward(par)
    if size_fft == 64:
        zu, zr = upd.search_psf_center(0.0, 1.0), fresh.search_psf_center(0.0, 1.0)
    upd.close()
    fresh.close()
    err = np.abs(got - ref).max() / ref.max()
    print("max |updated - fresh| / peak", err)
    if size_fft == 64:
        np.testing.assert_array_equal(got, ref)
        assert zu == zr
    else:
        assert err <= 1e-7


@pytest.mark.parametrize("size_fft,R,J", [(64, 21, 28), (140, 64, 45)])
def test_phase_update_is_bit_exact(size_fft, R, J):
    print(f"case: phase update, size_fft {size_fft}, R {R}, J {J}", flush=True)
    a0, a1 = _coeffs(J, 4), _coeffs(J, 5)
    rng = np.random.default_rng(6)
    par = np.stack([rng.uniform(-0.1, 0.1, 4), rng.uniform(-0.1, 0.1, 4), rng.uniform(-0.8, 0.8, 4),
                    np.zeros(4)], 1).astype(np.float32)
    upd = Z.Context(Z.optics(size_fft=size_fft), R, a0)
    upd.set_coefficients(a1)
    fresh = Z.Context(Z.optics(size_fft=size_fft), R, a1)
    got, ref = upd.psf_forward(par), fresh.psf_forward(par)
    upd.close()
    fresh.close()
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("size_fft,R,J", [(64, 31, 45), (140, 64, 45)])
def test_backward_dev_accumulates_da(torch_cuda, size_fft, R, J):
    """zola_psf_backward_dev adds dL/da to da (the calibration loops use an
    internal overwriting form): two calls on the same da give twice one call."""
    torch = torch_cuda
    print(f"case: dL/da accumulation, size_fft {size_fft}, R {R}", flush=True)
    ctx = Z.Context(Z.optics(size_fft=size_fft), R, _coeffs(J, 7))
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(8)
    B = 5
    par = torch.from_numpy(np.stack([rng.uniform(-0.1, 0.1, B), rng.uniform(-0.1, 0.1, B),
                                     rng.uniform(-0.8, 0.8, B), np.zeros(B)], 1).astype(np.float32)).cuda()
    w = torch.from_numpy(rng.normal(size=(B, R, R)).astype(np.float32)).cuda()
    dp = torch.zeros(B, 4, device="cuda")
    da1 = torch.zeros(J, dtype=torch.float64, device="cuda")
    ctx.psf_backward_dev(par, w, dp, da1)
    da2 = torch.full((J,), 0.5, dtype=torch.float64, device="cuda")
    ctx.psf_backward_dev(par, w, dp, da2)
    ctx.psf_backward_dev(par, w, dp, da2)
    torch.cuda.synchronize()
    ctx.close()
    one, two = da1.cpu().numpy(), da2.cpu().numpy()
    np.testing.assert_allclose(two, 0.5 + 2.0 * one, rtol=1e-12, atol=1e-12 * np.abs(one).max())
