"""Rotating zones (MRF) on the GPU against the oracle.

A jittered box with half of its cells rotating about an oblique axis
(viscous Roe and AUSM+up second order, HLLC first order; slip walls that
do not rotate, or no-slip walls that rotate with the zone), the same half
translating (alone, and beside the other half rotating), the rotating half
split into two zones of a zone list (other axis, origin and speed in the
second), and a rotational cyclic sector rotating about its own axis (Roe
second order):

* face fluxes, residual (with the Coriolis source), rPseudoDeltaT,
  Jacobian blocks (diagonal with the frame divergence and the Coriolis
  block, upper/lower, cyclic interface blocks): bitwise;
* 5 pseudo-steps: first-step residuals 1e-12, histories and fields 1e-8
  (FP64 form); the mixed form within 1e-3 of the FP64 oracle; a replayed
  hipGraph step equals the eager step bit for bit;
* a case with generator storage (the frame divergence stored per slot
  beside the ten generators) gives the same bits as full-block storage;
* set-up errors: a rotating patch without a fixedValue U condition, the
  residual VJP.
"""
import dataclasses

import numpy as np
import pytest

from icsfoam_amd import cases, dbnsCase
from icsfoam_amd._lib import IcsfError
from icsfoam_amd.mesh import prepare_case
from oracle.oracle import OracleCase, OracleMesh

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _sol(**kw):
    base = dict(solver="GMRES", nDirections=30, precision="double", co=20.0, jacobianStorage="blocks")
    base.update(kw)
    return cases.solution_dict(**base)


def _box(flux, rotating_walls, recon="linearUpwind limitedGrad", mu=1.8e-5):
    c = cases.case_box(n=(8, 6, 5), jitter=0.15, flux=flux, mu=mu, noise=0.1)
    if recon != "linearUpwind limitedGrad":
        c.fvSchemes = c.fvSchemes.replace("linearUpwind limitedGrad", recon)
    m = c.mesh
    C = OracleMesh(m).geometry()["C"]
    zone = C[:, 0] < np.median(C[:, 0])
    walls = [i for i, p in enumerate(m.patches) if p.name == "walls"]
    bcs = list(c.bcs)
    if rotating_walls:
        for i in walls:
            bcs[i] = cases.BoundarySpec(cases.BC_ZERO_GRADIENT, cases.BC_FIXED_VALUE, cases.BC_ZERO_GRADIENT,
                                        U_val=np.zeros((m.patches[i].size, 3)))
    non_rot = [i for i in range(len(m.patches)) if not (rotating_walls and i in walls)]
    z = cases.MRFZone(cells=zone, origin=np.array([0.1, 0.4, -0.2]), axis=np.array([0.3, -0.2, 1.0]),
                      omega=400.0, non_rotating=non_rot)
    c = dataclasses.replace(c, bcs=bcs, mrf=z, fvSolution=_sol())
    return prepare_case(c, permute_seed=23)


def _sector():
    c = cases.case_sector(8, 6, 5, np.pi / 8, flux="ROE", swirl=True, co=20.0)
    m = c.mesh
    non_rot = [i for i, p in enumerate(m.patches) if p.kind != cases.PATCH_CYCLIC]
    z = cases.MRFZone(cells=np.ones(m.n_cells, bool), origin=np.zeros(3), axis=np.array([0.0, 0.0, 1.0]),
                      omega=-900.0, non_rotating=non_rot)
    c = dataclasses.replace(c, mrf=z, fvSolution=_sol())
    return prepare_case(c, permute_seed=19)


def _translating(both):
    """The box of _box with moving no-slip walls, its zone translating (and,
    with `both`, the other half rotating as well: the face velocities add)."""
    c = _box("ROE", rotating_walls=True)
    z = c.mrf
    t = cases.TranslatingZone(cells=z.cells.copy(), velocity=np.array([40.0, -25.0, 15.0]),
                              non_moving=list(z.non_rotating))
    rot = dataclasses.replace(z, cells=~z.cells, non_rotating=list(range(len(c.mesh.patches)))) if both else None
    return dataclasses.replace(c, mrf=rot, translating=t)


def _two_zones():
    """The box of _box with rotating walls, its zone split into two zones
    (the zone list) that meet inside it: the second with another axis,
    origin and speed."""
    c = _box("ROE", rotating_walls=True)
    z = c.mrf
    C = OracleMesh(c.mesh).geometry()["C"]
    half = C[:, 1] < np.median(C[:, 1])
    z1 = dataclasses.replace(z, cells=z.cells & half)
    z2 = dataclasses.replace(z, cells=z.cells & ~half, axis=np.array([-0.4, 0.1, 0.8]), omega=-250.0,
                             origin=np.array([0.0, 0.2, 0.3]))
    return dataclasses.replace(c, mrf=[z1, z2])


CASES = {
    "box_two_rotating_zones": _two_zones,
    "box_translating": lambda: _translating(False),
    "box_translating_and_rotating": lambda: _translating(True),
    "box_roe_viscous": lambda: _box("ROE", rotating_walls=False),
    "box_roe_viscous_generators": lambda: dataclasses.replace(_box("ROE", rotating_walls=False),
                                                              fvSolution=_sol(jacobianStorage="generators")),
    "box_ausm_rotating_walls": lambda: _box("AUSMPlusUp", rotating_walls=True),
    "box_hllc_upwind": lambda: _box("HLLC", rotating_walls=True, recon="upwind", mu=0.0),
    "sector_roe": _sector,
}


@pytest.fixture(params=sorted(CASES))
def pair(request):
    print(f"[mrf case {request.param}]", flush=True)
    c = CASES[request.param]()
    return c, dbnsCase.from_case(c), OracleCase(c)


def test_mrf_fluxes_residual_and_jacobian_bitwise(pair):
    c, g, o = pair
    for a, b in zip(g.calcFlux(), o.calcFlux()):
        np.testing.assert_array_equal(a, b)
    bg, rg = g.residual()
    bo, ro = o.residual()
    np.testing.assert_array_equal(bg, bo)
    np.testing.assert_array_equal(rg, ro)
    g.createJacobian()
    for a, b in zip(g.jacobian(), o.createJacobian()):
        np.testing.assert_array_equal(a, b)
    if c.mesh.has_cyclic:
        np.testing.assert_array_equal(g.interfaces(), o.interfaces())
    # the frame terms are present: the same case without the zone differs
    o0 = OracleCase(dataclasses.replace(c, mrf=None, translating=None))
    b0, _ = o0.residual()
    assert rel_err(bo, b0) > 1e-3


def test_mrf_pseudo_steps(pair):
    c, g, o = pair
    hg = g.pseudo_steps(5)
    ho = o.pseudo_steps(5)
    assert rel_err(hg[0], ho[0]) < 1e-12
    assert rel_err(hg, ho) < 1e-8
    fg, fo = g.fields(), o.fields()
    for k in ("rho", "U", "T", "p", "W"):
        assert rel_err(fg[k], fo[k]) < 1e-8, k


def test_mrf_graph_replay_and_mixed_form():
    c = _box("ROE", rotating_walls=True)
    print("[mrf graph/mixed]", flush=True)
    ge = dbnsCase.from_case(dataclasses.replace(c, fvSolution=_sol(graphCapture=False)))
    gg = dbnsCase.from_case(dataclasses.replace(c, fvSolution=_sol(graphCapture=True)))
    np.testing.assert_array_equal(gg.pseudo_steps(6), ge.pseudo_steps(6))
    np.testing.assert_array_equal(gg.fields()["W"], ge.fields()["W"])
    gm = dbnsCase.from_case(dataclasses.replace(c, fvSolution=_sol(solver="FGMRES", nDirections=1,
                                                                   precision="mixed", innerDirections=30,
                                                                   orthogonalisation="CGS")))
    o = OracleCase(c)
    gm.pseudo_steps(5)
    o.pseudo_steps(5)
    fm, fo = gm.fields(), o.fields()
    errs = {k: rel_err(fm[k], fo[k]) for k in ("rho", "U", "T", "p")}
    print(f"mixed MRF vs FP64 oracle after 5 steps: {errs}")
    assert max(errs.values()) < 1e-3


def test_mrf_set_up_errors():
    c = _box("ROE", rotating_walls=False)
    print("[mrf errors]", flush=True)
    # generator storage carries the frame divergence per slot: the same
    # results as full blocks
    gg = dbnsCase.from_case(dataclasses.replace(c, fvSolution=_sol(jacobianStorage="generators")))
    gb = dbnsCase.from_case(c)
    np.testing.assert_array_equal(gg.pseudo_steps(3), gb.pseudo_steps(3))
    np.testing.assert_array_equal(gg.fields()["W"], gb.fields()["W"])
    # slip walls that would rotate
    with pytest.raises(IcsfError):
        dbnsCase.from_case(dataclasses.replace(c, mrf=dataclasses.replace(c.mrf, non_rotating=[])))
    # the residual VJP refuses rotating zones
    import torch

    from icsfoam_amd.adjoint import residual_vjp_device

    g = dbnsCase.from_case(c)
    W = torch.tensor(g.fields()["W"], device="cuda")
    with pytest.raises(IcsfError):
        residual_vjp_device(g, W, torch.ones_like(W), geometry=False)


def test_mrf_harmonic_balance():
    """Three instants of the cyclic HB sector, all cells rotating about the
    sector axis (every instance shares the zone): residual with the HB and
    Coriolis sources, coupled SpMV and LU-SGS bitwise; 3 pseudo-steps
    1e-8."""
    from icsfoam_amd.solver import hbCase
    from oracle.oracle import OracleHB

    print("[mrf hb]", flush=True)
    insts, D, _ = cases.case_hb(n=(8, 6, 5), n_harmonics=1, flux="ROE", jitter=0.1)
    m = insts[0].mesh
    non_rot = [i for i, p in enumerate(m.patches) if p.kind != cases.PATCH_CYCLIC]
    z = cases.MRFZone(cells=np.ones(m.n_cells, bool), origin=np.zeros(3), axis=np.array([0.0, 0.0, 1.0]),
                      omega=600.0, non_rotating=non_rot)
    sol = None
    out = []
    for c in insts:
        sol = c.fvSolution.replace("jacobianStorage generators", "jacobianStorage blocks")
        out.append(prepare_case(dataclasses.replace(c, mrf=z, fvSolution=sol), permute_seed=17))
    assert "jacobianStorage blocks" in sol
    g, o = hbCase.from_cases(out, D), OracleHB(out, D)
    np.testing.assert_array_equal(g.residual(), o.residual())
    g.createJacobian(); o.createJacobian()
    x = np.random.default_rng(4).standard_normal((g.nO, 5, g.nCells))
    np.testing.assert_array_equal(g.matrixMul(x), o.matrixMul(x))
    np.testing.assert_array_equal(g.precondition(x), o.precondition(x))
    g2, o2 = hbCase.from_cases(out, D), OracleHB(out, D)
    hg = g2.pseudo_steps(3)
    ho = o2.pseudo_steps(3)
    assert rel_err(hg[0], ho[0]) < 1e-12
    assert rel_err(hg, ho) < 1e-8
