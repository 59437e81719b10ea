"""Autograd through the HIP transforms: the backward pass is the adjoint
transform, checked by the dot-product identity, by torch's gradcheck and by
FP64 central finite differences of a soft-threshold loss (the differentiable
use of a 3-D separable CDF 9/7 transform, four levels per axis)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CDF9_7 = 100907


def _torch():
    import torch
    return torch


@pytest.mark.parametrize("transform", [0, 1, 2, 3])
@pytest.mark.parametrize("level", [(3, 2, 4), 2])
def test_dot_product(gpu, transform, level):
    """<W x, y> == <x, W^T y> with W^T from the backward pass (fp64)."""
    torch = _torch()
    import lifted_amd.autograd as AG
    g = torch.Generator(device="cuda").manual_seed(7 + transform)
    x = torch.randn(96, 80, 72, dtype=torch.float64, device="cuda", generator=g, requires_grad=True)
    y = torch.randn(96, 80, 72, dtype=torch.float64, device="cuda", generator=g)
    for w, bc in ((CDF9_7, 3), (2, 0), (10404, 1), (10202, 4)):
        wx = AG.lwt(x, w, bc, transform, axes=(0, 1, 2), level=level)
        (gx,) = torch.autograd.grad((wx * y).sum(), x)
        lhs = float((wx.detach() * y).sum())
        rhs = float((x.detach() * gx).sum())
        assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs)) * 100, (w, bc, lhs, rhs)


def test_gradcheck(gpu):
    torch = _torch()
    import lifted_amd.autograd as AG
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(12, 10, dtype=torch.float64, device="cuda", generator=g, requires_grad=True)
    for tf in (0, 1):
        assert torch.autograd.gradcheck(lambda t: AG.lwt(t, CDF9_7, 3, tf, axes=(0, 1), level=(2, 2)), (x,),
                                        eps=1e-6, atol=1e-8)


def soft(c, t):
    torch = _torch()
    return torch.sign(c) * torch.clamp(c.abs() - t, min=0.0)


def test_finite_differences_3d_crop(gpu, oracle):
    """64^3 fp64, CDF 9/7 with four levels per axis (past max_level = 3 at 64,
    well defined for Zero boundaries); loss = sum |soft(W x, 0.1)|; the
    directional derivative along a random v from the backward pass matches
    central differences at h = 1e-6, and the forward matches the oracle."""
    torch = _torch()
    import lifted_amd.autograd as AG
    rng = np.random.default_rng(20261201 + 2)
    xh = rng.standard_normal((64, 64, 64))
    vh = rng.standard_normal((64, 64, 64))
    ref = oracle.lwt(xh, CDF9_7, oracle.FORWARD, oracle.ZERO, level=[4, 4, 4])
    x = torch.tensor(xh, device="cuda", requires_grad=True)
    c = AG.forward(x, CDF9_7, 0, axes=(0, 1, 2), level=(4, 4, 4))
    from helpers import assert_parity, band_slices
    assert_parity(c.detach().cpu().numpy(), ref, "64^3 forward", bands=band_slices(ref.shape, [0, 1, 2], [4, 4, 4]))
    loss = soft(c, 0.1).abs().sum()
    (gx,) = torch.autograd.grad(loss, x)
    v = torch.tensor(vh, device="cuda")
    h = 1e-6

    def L(t):
        with torch.no_grad():
            return float(soft(AG.forward(t, CDF9_7, 0, axes=(0, 1, 2), level=(4, 4, 4)), 0.1).abs().sum())

    fd = (L(x.detach() + h * v) - L(x.detach() - h * v)) / (2 * h)
    an = float((gx * v).sum())
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_grad_512_cube_float32(gpu):
    """The full-size 512^3 float32 separable transform: forward + backward run
    and the backward equals the forward-adjoint applied directly."""
    torch = _torch()
    import lifted_amd as la
    import lifted_amd.autograd as AG
    g = torch.Generator(device="cuda").manual_seed(11)
    x = torch.randn(512, 512, 512, device="cuda", generator=g, requires_grad=True)
    c = AG.forward(x, CDF9_7, 0, axes=(0, 1, 2), level=(4, 4, 4))
    loss = soft(c, 0.1).abs().sum()
    loss.backward()
    with torch.no_grad():
        gy = torch.sign(c) * (c.abs() > 0.1).to(c.dtype)
        direct = la.lwt_forward_adjoint(gy, CDF9_7, 0, axes=(0, 1, 2), level=[4, 4, 4])
    assert torch.equal(x.grad, direct)
