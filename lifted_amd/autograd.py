"""Differentiable lifting transforms for torch tensors on the GPU.

Every transform of the library is linear, y = W x, so its gradient is the
adjoint, dL/dx = W^T dL/dy, and the adjoint of each transform type is another
transform type of the same wavelet, boundary, axes and levels:

    Forward        <-> ForwardAdjoint
    Inverse        <-> InverseAdjoint

(the reference's ForwardAdjoint / InverseAdjoint, src/lifted-common.hpp:966-996,
are exactly these transposes; its tests check <W x, y> = <x, W^T y>,
tests/test_dwt.cpp:376-679). Nothing is saved for the backward pass: it
recomputes nothing and keeps no activations, only the call's arguments.
"""
import torch

from . import Transform, _lwt

_ADJOINT = {
    int(Transform.Forward): int(Transform.ForwardAdjoint),
    int(Transform.ForwardAdjoint): int(Transform.Forward),
    int(Transform.Inverse): int(Transform.InverseAdjoint),
    int(Transform.InverseAdjoint): int(Transform.Inverse),
}


class LwtFunction(torch.autograd.Function):
    """y = lwt(x) with transform `transform`; backward applies the adjoint
    transform to the incoming gradient."""

    @staticmethod
    def forward(ctx, x, wavelet, mode, transform, axes, level):
        ctx.args = (wavelet, mode, int(transform), axes, level)
        return _lwt(x, wavelet, mode, int(transform), axes, level)

    @staticmethod
    def backward(ctx, grad):
        wavelet, mode, transform, axes, level = ctx.args
        g = _lwt(grad.contiguous(), wavelet, mode, _ADJOINT[transform], axes, level)
        return g, None, None, None, None, None


def lwt(x, wavelet, mode=0, transform=Transform.Forward, axes=None, level=1):
    """Differentiable lifted wavelet transform of a device tensor (arguments as
    lifted_amd.lwt_forward and friends: a scalar `level` selects the
    level-by-level path, a sequence the separable one)."""
    if isinstance(level, list):
        level = tuple(level)
    if isinstance(axes, list):
        axes = tuple(axes)
    return LwtFunction.apply(x, wavelet, mode, int(transform), axes, level)


def forward(x, wavelet, mode=0, axes=None, level=1):
    return lwt(x, wavelet, mode, Transform.Forward, axes, level)


def inverse(x, wavelet, mode=0, axes=None, level=1):
    return lwt(x, wavelet, mode, Transform.Inverse, axes, level)


def forward_adjoint(x, wavelet, mode=0, axes=None, level=1):
    return lwt(x, wavelet, mode, Transform.ForwardAdjoint, axes, level)


def inverse_adjoint(x, wavelet, mode=0, axes=None, level=1):
    return lwt(x, wavelet, mode, Transform.InverseAdjoint, axes, level)
