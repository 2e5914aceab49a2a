"""Rotational cyclic patches on the GPU against the oracle.

Annular sectors whose theta planes are a rotational cyclic pair (seeded
rotation-periodic swirling state, permuted + RCM with the coupled cells +
colour-major numbering), AUSM+up / Roe second order (inviscid and
viscous: the viscous terms and their LF Jacobian across the cyclic faces)
and HLLC first order:

* face fluxes, residual, rPseudoDeltaT, Jacobian blocks (diagonal,
  upper/lower and the cyclic interface blocks), block SpMV and LU-SGS:
  bitwise (same operation order on both sides, no FMA contraction);
* GMRES increment: relative 1e-10; 5 pseudo-steps: first-step residuals
  1e-12, histories and fields 1e-8;
* the fused Arnoldi step equals the unfused one to 1e-12 and a replayed
  hipGraph step equals the eager step bit for bit;
* mixed precision (FGMRES + FP32 inner GMRES/LU-SGS) stays within 1e-3 of
  the FP64 oracle after 5 steps.
"""
import dataclasses

import numpy as np
import pytest

from icsfoam_amd import cases, dbnsCase
from icsfoam_amd.mesh import prepare_case
from oracle.oracle import OracleCase

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _sector(flux, recon="linearUpwind limitedGrad", mu=0.0, **sol):
    c = cases.case_sector(8, 6, 5, np.pi / 8, flux=flux, recon=recon, swirl=True, co=20.0, mu=mu, **sol)
    return prepare_case(c, permute_seed=19)


CASES = {
    "sector_roe": lambda: _sector("ROE"),
    "sector_ausm": lambda: _sector("AUSMPlusUp"),
    "sector_hllc_upwind": lambda: _sector("HLLC", recon="upwind"),
    "sector_roe_viscous": lambda: _sector("ROE", mu=2.0e-3),
    "sector_ausm_viscous": lambda: _sector("AUSMPlusUp", mu=2.0e-3),
}


@pytest.fixture(params=sorted(CASES))
def pair(request):
    print(f"[cyclic case {request.param}]", flush=True)
    c = CASES[request.param]()
    return c, dbnsCase.from_case(c), OracleCase(c)


def test_cyclic_fluxes_residual_bitwise(pair):
    c, g, o = pair
    for a, b in zip(g.calcFlux(), o.calcFlux()):
        np.testing.assert_array_equal(a, b)
    bg, rg = g.residual()
    bo, ro = o.residual()
    np.testing.assert_array_equal(bg, bo)
    np.testing.assert_array_equal(rg, ro)


def test_cyclic_jacobian_bitwise(pair):
    c, g, o = pair
    g.residual(); o.residual()
    g.createJacobian()
    dg, ug, lg = g.jacobian()
    do, uo, lo = o.createJacobian()
    np.testing.assert_array_equal(dg, do)
    np.testing.assert_array_equal(ug, uo)
    np.testing.assert_array_equal(lg, lo)
    ig, io = g.interfaces(), o.interfaces()
    assert np.count_nonzero(np.abs(io).sum(axis=(1, 2))) > 0
    np.testing.assert_array_equal(ig, io)


def test_cyclic_spmv_lusgs_bitwise(pair):
    c, g, o = pair
    g.residual(); o.residual()
    g.createJacobian(); o.createJacobian()
    x = np.random.default_rng(2).standard_normal((5, g.nCells))
    np.testing.assert_array_equal(g.matrixMul(x), o.matrixMul(x))
    np.testing.assert_array_equal(g.precondition(x), o.precondition(x))


def test_cyclic_gmres_and_pseudo_steps(pair):
    c, g, o = pair
    bg, _ = g.residual(); bo, _ = o.residual()
    g.createJacobian(); o.createJacobian()
    W = g.fields()["W"]
    dWg, rg, _ = g.solveForIncr()
    dWo, ro, _ = o.solveForIncr(W, bo)
    assert rel_err(dWg, dWo) < 1e-10
    assert rel_err(rg, ro) < 1e-10
    g2, o2 = dbnsCase.from_case(c), OracleCase(c)
    hg = g2.pseudo_steps(5)
    ho = o2.pseudo_steps(5)
    assert rel_err(hg[0], ho[0]) < 1e-12
    assert rel_err(hg, ho) < 1e-8
    fg, fo = g2.fields(), o2.fields()
    for k in ("rho", "U", "T", "p", "W"):
        assert rel_err(fg[k], fo[k]) < 1e-8, k


def _with(c, **kw):
    base = dict(co=20.0)
    base.update(kw)
    return dataclasses.replace(c, fvSolution=cases.solution_dict(**base))


def test_cyclic_fused_graph_and_mixed():
    c = _sector("ROE")
    print("[cyclic fused/graph/mixed]", flush=True)
    gf = dbnsCase.from_case(_with(c, fusedSweeps=True))
    gu = dbnsCase.from_case(_with(c, fusedSweeps=False))
    assert gf.mesh.nFwdLevels == 2 and gf.mesh.nBwdLevels == 2
    for g in (gf, gu):
        g.residual()
        g.createJacobian()
    dWf, rf, _ = gf.solveForIncr()
    dWu, ru, _ = gu.solveForIncr()
    assert rel_err(dWf, dWu) < 1e-12
    ge = dbnsCase.from_case(_with(c, graphCapture=False))
    gg = dbnsCase.from_case(_with(c, graphCapture=True))
    np.testing.assert_array_equal(gg.pseudo_steps(6), ge.pseudo_steps(6))
    np.testing.assert_array_equal(gg.fields()["W"], ge.fields()["W"])
    gm = dbnsCase.from_case(_with(c, solver="FGMRES", nDirections=1, innerDirections=30, precision="mixed",
                                  orthogonalisation="CGS"))
    o = OracleCase(c)
    gm.pseudo_steps(5)
    o.pseudo_steps(5)
    fm, fo = gm.fields(), o.fields()
    errs = {k: rel_err(fm[k], fo[k]) for k in ("rho", "U", "T", "p")}
    print(f"cyclic sector, mixed vs FP64 oracle after 5 steps: {errs}")
    assert max(errs.values()) < 1e-3
