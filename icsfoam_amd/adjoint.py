"""torch binding of the residual adjoint.

`ResidualFunction` makes the residual b(W) of a `dbnsCase` (AUSM+up, Roe
or HLLC, inviscid or viscous) an autograd node: the forward pass calls
`icsf_residual_device`, the backward pass `icsf_residual_vjp_device`
(hand-written reverse kernels, reconstruction recomputed), both on device
tensors owned by torch.  The reference has no adjoint; the differentiated
chain is updateFields.H:7-104, ausmPlusUpFluxScheme.C:73-299 or
roeFluxScheme.C:41-241,276-409, and residualsUpdate.H:1-43.

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
    """(gW [5, nC], gSf [nF, 3], gCf [nF, 3]) = J^T v on device tensors
    (gSf, gCf: geometry=True, inviscid residuals only; else None)."""
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


def fd_gate(case, W: torch.Tensor, X: torch.Tensor, n_dirs: int = 256, h: float = 1e-7, seed: int = 1234,
            tol: float = 1e-6, tol_one_sided: float = 1e-5) -> dict:
    """Finite-difference check of the VJP at state W (device [5, nC]) along
    n_dirs seeded smooth directions d (sums of plane waves over the cell
    centres X [nC, 3], scaled by 1 + |W|) with seeds v: v.(b(W+hd) -
    b(W-hd))/2h against (J^T v).d, relative to sum |v fd|.  A direction off
    by more than tol is explained by the residual's switch records
    (residual_records_device): records equal at W-hd, W and W+hd (one
    smooth piece) -> the Richardson extrapolation of the central
    differences of steps h and h/2 must match to tol; records switching on
    one side -> the one-sided difference on the other side must match to
    tol_one_sided (truncation O(h)); switching on both sides -> the same at
    the first of h/10, h/100, h/1000 where one side keeps W's records.
    Returns the error statistics and the unexplained directions."""
    import time

    import numpy as np

    dev = W.device
    scale = 1.0 + W.abs()
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)

    def smooth_dir():
        kv = torch.tensor(rng.standard_normal((5, 4, 3)) * 3.0, device=dev)
        ph = torch.tensor(rng.uniform(0, 2 * np.pi, (5, 4)), device=dev)
        return torch.stack([sum(torch.sin(X @ kv[q, j] + ph[q, j]) for j in range(4)) / 4 for q in range(5)])

    def same(ra, rb):
        return all(torch.equal(x, y) for x, y in zip(ra, rb))

    def rel(v, fd, rhs):
        return abs(float((v * fd).sum()) - rhs) / float((v * fd).abs().sum())

    rec0 = residual_records_device(case, W)
    b0 = residual_device(case, W)
    errs, onesided, unexplained = [], [], []
    kinks = both = smooth_r = 0
    t_vjp = 0.0
    for i in range(n_dirs):
        v = torch.randn(W.shape, generator=gen, device=dev, dtype=torch.float64)
        d = smooth_dir() * scale
        t0 = time.perf_counter()
        gW, _, _ = residual_vjp_device(case, W, v, geometry=False)
        t_vjp += time.perf_counter() - t0
        bp = residual_device(case, W + h * d)
        bm = residual_device(case, W - h * d)
        fd = (bp - bm) / (2 * h)
        rhs = float((gW * d).sum())
        err = rel(v, fd, rhs)
        errs.append(err)
        if err <= tol:
            continue
        kinks += 1
        sp = same(rec0, residual_records_device(case, W + h * d))
        sm = same(rec0, residual_records_device(case, W - h * d))
        if sp and sm:
            fd2 = (residual_device(case, W + 0.5 * h * d) - residual_device(case, W - 0.5 * h * d)) / h
            er = rel(v, (4.0 * fd2 - fd) / 3.0, rhs)
            if er <= tol:
                smooth_r += 1
                continue
            unexplained.append((i, err, f"no switch between W-hd and W+hd (Richardson {er:.2e})"))
            continue
        hs = h
        while not (sp or sm) and hs > h / 1000:
            hs /= 10
            sp = same(rec0, residual_records_device(case, W + hs * d))
            sm = same(rec0, residual_records_device(case, W - hs * d))
        if not (sp or sm):
            unexplained.append((i, err, "switches on both sides down to a step of h/1000"))
            continue
        sgn = 1.0 if sp else -1.0
        b1 = (bp if sp else bm) if hs == h else residual_device(case, W + sgn * hs * d)
        e1 = rel(v, (b1 - b0) / (sgn * hs), rhs)
        onesided.append(e1)
        if hs < h:
            both += 1
        if e1 > tol_one_sided:
            unexplained.append((i, err, f"one-sided difference (step {hs:.0e}) on the smooth side off by {e1:.2e}"))
    errs = np.array(errs)
    return {"directions": n_dirs, "median_rel_err": float(np.median(errs)), "max_rel_err": float(errs.max()),
            "above_tol": kinks, "smooth_richardson": smooth_r, "switch_both_sides_at_h": both,
            "one_sided_max": float(max(onesided)) if onesided else 0.0, "vjp_ms": 1e3 * t_vjp / n_dirs,
            "unexplained": unexplained}
