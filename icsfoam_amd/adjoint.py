"""torch binding of the residual adjoint.

`ResidualFunction` makes the inviscid AUSM+up residual b(W) of a
`dbnsCase` an autograd node: the forward pass calls
`icsf_residual_device`, the backward pass `icsf_residual_vjp_device`
(hand-written reverse kernels, reconstruction recomputed), both on device
tensors owned by torch.  The reference has no adjoint; the differentiated
chain is updateFields.H:7-104, ausmPlusUpFluxScheme.C:73-299 and
residualsUpdate.H:1-14.

torch is used for device memory only.  Each call synchronises torch's
current stream before handing pointers to the library and returns after the
library's stream has finished, so the tensors are safe to use afterwards.
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch

from ._lib import check, lib


def _dev_f64(t: torch.Tensor, shape, name: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise TypeError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != torch.float64:
        raise TypeError(f"{name} must be float64")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    return t.contiguous()


def residual_device(case, W: Optional[torch.Tensor]) -> torch.Tensor:
    """b = R*V ([5, nCells], float64, same device) at state W (None: the
    case's current state); W becomes the case state."""
    n = case.nCells
    Wc = None if W is None else _dev_f64(W, (5, n), "W")
    dev = Wc.device if Wc is not None else torch.device("cuda", torch.cuda.current_device())
    b = torch.empty((5, n), dtype=torch.float64, device=dev)
    torch.cuda.current_stream(dev).synchronize()
    check(lib().icsf_residual_device(case._h, None if Wc is None else Wc.data_ptr(), b.data_ptr()))
    return b


def residual_vjp_device(case, W: Optional[torch.Tensor], v: torch.Tensor,
                        geometry: bool = True) -> Tuple[torch.Tensor, Optional[torch.Tensor],
                                                         Optional[torch.Tensor]]:
    """(gW [5, nC], gSf [nF, 3], gCf [nF, 3]) = J^T v on device tensors."""
    n, nF = case.nCells, case.mesh.nFaces
    vc = _dev_f64(v, (5, n), "v")
    Wc = None if W is None else _dev_f64(W, (5, n), "W")
    gW = torch.empty((5, n), dtype=torch.float64, device=vc.device)
    gS = torch.empty((nF, 3), dtype=torch.float64, device=vc.device) if geometry else None
    gC = torch.empty((nF, 3), dtype=torch.float64, device=vc.device) if geometry else None
    torch.cuda.current_stream(vc.device).synchronize()
    check(lib().icsf_residual_vjp_device(case._h, None if Wc is None else Wc.data_ptr(), vc.data_ptr(),
                                         gW.data_ptr(), None if gS is None else gS.data_ptr(),
                                         None if gC is None else gC.data_ptr()))
    return gW, gS, gC


def residual_records_device(case, W: Optional[torch.Tensor]) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Records of the residual's non-differentiable points at state W
    (icsf_residual_records_device): (AUSM+up branch bits [nF], limiter
    max/min element [nC, 8], limiter face [nC, 8]), int32 device tensors.
    Equal records at two states: one smooth piece of the residual."""
    n, nF = case.nCells, case.mesh.nFaces
    Wc = None if W is None else _dev_f64(W, (5, n), "W")
    dev = Wc.device if Wc is not None else torch.device("cuda", torch.cuda.current_device())
    fb = torch.empty((nF,), dtype=torch.int32, device=dev)
    te = torch.empty((n, 8), dtype=torch.int32, device=dev)
    tf = torch.empty((n, 8), dtype=torch.int32, device=dev)
    torch.cuda.current_stream(dev).synchronize()
    check(lib().icsf_residual_records_device(case._h, None if Wc is None else Wc.data_ptr(), fb.data_ptr(),
                                             te.data_ptr(), tf.data_ptr()))
    return fb, te, tf


class ResidualFunction(torch.autograd.Function):
    """b = residual(W) with d/dW by the library's reverse pass."""

    @staticmethod
    def forward(ctx, W: torch.Tensor, case) -> torch.Tensor:
        ctx.case = case
        ctx.save_for_backward(W)
        return residual_device(case, W.detach())

    @staticmethod
    def backward(ctx, gb: torch.Tensor):
        (W,) = ctx.saved_tensors
        gW, _, _ = residual_vjp_device(ctx.case, W.detach(), gb.to(torch.float64).contiguous(), geometry=False)
        return gW, None


def residual(case, W: torch.Tensor) -> torch.Tensor:
    """Differentiable b(W) of the inviscid AUSM+up residual."""
    return ResidualFunction.apply(W, case)
