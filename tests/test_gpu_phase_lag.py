"""Phase-lagged cyclic patches of harmonic balance on the GPU against the
oracle: an annular sector whose rotational cyclic pair is phase-lagged
(inter-blade phase angle 2 pi/9), 3 and 5 instants, Roe, AUSM+up and HLLC,
velocity combined unrotated or in cylindrical components (cylCoords), and a
viscous Roe sector (cylCoords; the unrotated form is refused for viscous
cases).

* residual with the lagged far-side values and the HB source, coupled SpMV
  and LU-SGS (the Jacobian's face states lagged, the interfaces plain):
  bitwise;
* pseudo-steps with the bounded local time step (its far density lagged):
  first-step residuals 1e-12, histories and fields 1e-8;
* a replayed hipGraph step equals the eager step bit for bit; the mixed
  form within 1e-3 of the FP64 oracle.
"""
import dataclasses

import numpy as np
import pytest

from icsfoam_amd import cases
from icsfoam_amd.mesh import prepare_case
from icsfoam_amd.solver import hbCase
from oracle.oracle import OracleHB

from helpers import rel_err

pytestmark = pytest.mark.gpu

FREQ = 100.0


def _hb(K, flux, cyl, edit=None, mu=0.0):
    insts, D, t = cases.case_hb(n=(8, 6, 5), n_harmonics=K, frequency_hz=FREQ, flux=flux, jitter=0.1, mu=mu)
    m = insts[0].mesh
    lo = next(i for i, p in enumerate(m.patches) if p.kind == cases.PATCH_CYCLIC)
    pair = cases.PhaseLagPair(owner=lo, ibpa=2 * np.pi / 9, frequencies=np.array([2 * np.pi * FREQ]),
                              harmonics=np.array([K]), instants=t, cyl_coords=cyl)
    out = []
    for c in insts:
        sol = edit(c.fvSolution) if edit else c.fvSolution
        out.append(prepare_case(dataclasses.replace(c, phase_lag=pair, fvSolution=sol), permute_seed=17))
    return out, D


def _oracle(insts, D, g):
    o = OracleHB(insts, D)
    pl = insts[0].phase_lag
    for patch, Dl in g.phase_lag_sides:
        o.set_phase_lag(patch, Dl, cyl=pl.cyl_coords, axis=pl.axis, origin=pl.centre)
    return o


CASES = {
    "hb3_roe": lambda: _hb(1, "ROE", False),
    "hb3_roe_cyl": lambda: _hb(1, "ROE", True),
    "hb5_ausm": lambda: _hb(2, "AUSMPlusUp", False),
    "hb3_hllc_cyl": lambda: _hb(1, "HLLC", True),
    "hb3_roe_viscous_cyl": lambda: _hb(1, "ROE", True, mu=2.0e-3),
}


@pytest.fixture(params=sorted(CASES))
def pair(request):
    print(f"[phase-lag case {request.param}]", flush=True)
    insts, D = CASES[request.param]()
    g = hbCase.from_cases(insts, D)
    return g, _oracle(insts, D, g)


def test_phase_lag_residual_and_operators_bitwise(pair):
    g, o = pair
    np.testing.assert_array_equal(g.residual(), o.residual())
    g.createJacobian(); o.createJacobian()
    x = np.random.default_rng(4).standard_normal((g.nO, 5, g.nCells))
    np.testing.assert_array_equal(g.matrixMul(x), o.matrixMul(x))
    np.testing.assert_array_equal(g.precondition(x), o.precondition(x))


def test_phase_lag_pseudo_steps(pair):
    g, o = pair
    hg = g.pseudo_steps(4)
    ho = o.pseudo_steps(4)
    assert rel_err(hg[0], ho[0]) < 1e-12
    assert rel_err(hg, ho) < 1e-8
    for J in range(g.nO):
        fg, fo = g.fields(J), o.fields(J)
        for k in ("rho", "U", "T", "p", "W"):
            assert rel_err(fg[k], fo[k]) < 1e-8, (J, k)


def test_phase_lag_viscous_needs_cyl_coords():
    """The viscous terms take grad(U) whole: the unrotated lagged velocity
    components are refused for viscous cases."""
    from icsfoam_amd._lib import IcsfError
    insts, D = _hb(1, "ROE", False, mu=2.0e-3)
    with pytest.raises(IcsfError, match="cylCoords"):
        hbCase.from_cases(insts, D)


def test_phase_lag_graph_replay_and_mixed_form():
    print("[phase-lag graph/mixed]", flush=True)
    ie, D = _hb(1, "ROE", True, lambda s: s.replace("graphCapture true", "graphCapture false"))
    ig, _ = _hb(1, "ROE", True, lambda s: s.replace("graphCapture false", "graphCapture true"))
    assert "graphCapture true" in ig[0].fvSolution and "graphCapture false" in ie[0].fvSolution
    ge, gg = hbCase.from_cases(ie, D), hbCase.from_cases(ig, D)
    np.testing.assert_array_equal(gg.pseudo_steps(6), ge.pseudo_steps(6))
    for J in range(ge.nO):
        np.testing.assert_array_equal(gg.fields(J)["W"], ge.fields(J)["W"])
    mixed = cases.solution_dict(solver="FGMRES", nDirections=1, precision="mixed", innerDirections=30,
                                orthogonalisation="CGS", co=20.0)
    im, _ = _hb(1, "ROE", True, lambda s: mixed)
    gm = hbCase.from_cases(im, D)
    io, _ = _hb(1, "ROE", True)
    go = hbCase.from_cases(io, D)
    o = _oracle(io, D, go)
    gm.pseudo_steps(3)
    o.pseudo_steps(3)
    errs = [rel_err(gm.fields(J)[k], o.fields(J)[k]) for J in range(gm.nO) for k in ("rho", "U", "T", "p")]
    print(f"mixed phase-lag HB vs FP64 oracle after 3 steps: max {max(errs):.2e}")
    assert max(errs) < 1e-3
