"""icsf_residual_vjp_device_noalloc (the residual VJP whose adjoint work
buffers live in the library's fixed scratch array, no device allocation)
against icsf_residual_vjp_device (its buffers allocated by the library):
the same bits for gW, gSf and gCf, on two calls in a row on cases of
decreasing size (the second call finds the first one's contents in the
scratch array), with the Sf/Cf outputs requested and not; a case needing
more than the fixed array is refused."""
import ctypes as C

import numpy as np
import pytest

from icsfoam_amd import IcsfError, cases, dbnsCase
from icsfoam_amd._lib import check, lib
from icsfoam_amd.mesh import prepare_case

from helpers import ausm_inviscid, generic_state

pytestmark = pytest.mark.gpu

_vp = C.c_void_p


def _bind():
    L = lib()
    for name in ("icsf_residual_vjp_device", "icsf_residual_vjp_device_noalloc"):
        f = getattr(L, name)
        f.restype = C.c_int
        f.argtypes = [_vp] * 6
    return L


def _case(n, seed):
    c = ausm_inviscid(prepare_case(cases.case_box(n=n, jitter=0.1, flux="AUSMPlusUp", mu=0.0, noise=0.05),
                                   permute_seed=seed))
    return c


def _run(fn, g, W, v, want_geo):
    import torch
    dev = torch.device("cuda")
    nC, nF = g.nCells, g.mesh.nFaces
    Wt = torch.from_numpy(np.ascontiguousarray(W)).to(dev)
    vt = torch.from_numpy(np.ascontiguousarray(v)).to(dev)
    gW = torch.full((5 * nC,), np.nan, dtype=torch.float64, device=dev)
    gS = torch.full((3 * nF,), np.nan, dtype=torch.float64, device=dev) if want_geo else None
    gC = torch.full((3 * nF,), np.nan, dtype=torch.float64, device=dev) if want_geo else None
    p = lambda t: None if t is None else _vp(t.data_ptr())  # noqa: E731
    check(fn(g._h, p(Wt), p(vt), p(gW), p(gS), p(gC)))
    torch.cuda.synchronize()
    out = [gW.cpu().numpy()]
    if want_geo:
        out += [gS.cpu().numpy(), gC.cpu().numpy()]
    return out


def test_noalloc_vjp_equals_allocating_vjp_on_two_calls():
    L = _bind()
    for n, seed in (((12, 10, 8), 3), ((6, 5, 4), 4)):
        print(f"[noalloc vjp {n}]", flush=True)
        c = _case(n, seed)
        g = dbnsCase.from_case(c)
        W = generic_state(g.fields()["W"], seed)
        v = np.random.default_rng(seed).standard_normal(W.shape)
        for geo in (True, False):
            ref = _run(L.icsf_residual_vjp_device, g, W, v, geo)
            got = _run(L.icsf_residual_vjp_device_noalloc, g, W, v, geo)
            for a, b in zip(got, ref):
                assert np.isfinite(b).all()
                np.testing.assert_array_equal(a, b)


def test_noalloc_vjp_rejects_cases_beyond_the_scratch():
    L = _bind()
    c = _case((64, 60, 60), 5)
    g = dbnsCase.from_case(c)
    W = g.fields()["W"]
    with pytest.raises(IcsfError, match="scratch"):
        _run(L.icsf_residual_vjp_device_noalloc, g, W, np.ones_like(W), False)
