"""The derived inflow conditions on the GPU against the oracle: the
tutorials' inflow patch (totalPressure, pressureInletOutletVelocity,
totalTemperature) on the bump channel and on a viscous jittered box.

* Initial state (boundary values from the `value` entries / p0): fluxes,
  residual, local time step and Jacobian bit for bit.
* After a state update (the conditions corrected from lagged values; the
  total relations use pow on both sides, whose last bit may differ between
  the device and the host libm): boundary values within 1e-14, residual
  and Jacobian within 1e-12 relative, the inflow switch of
  pressureInletOutletVelocity set identically.
* Five pseudo-steps: residual histories and fields within 1e-8, the first
  step's residuals within 1e-12, boundary values within 1e-8.
"""
import numpy as np
import pytest

from icsfoam_amd import cases, dbnsCase
from icsfoam_amd.mesh import prepare_case
from oracle.oracle import OracleCase

from helpers import rel_err, total_inlet_box

pytestmark = pytest.mark.gpu

CASES = {
    "bump_tutorial_bcs": lambda: prepare_case(cases.case_bump(40, 10, bcs="tutorial"), permute_seed=11),
    "box_total_inlet_viscous": lambda: total_inlet_box(),
    "box_total_inlet_hllc": lambda: total_inlet_box(flux="HLLC", mu=0.0),
}


@pytest.fixture(params=sorted(CASES))
def pair(request):
    print(f"[case {request.param}]", flush=True)
    c = CASES[request.param]()
    return c, dbnsCase.from_case(c), OracleCase(c)


def _inlet(c):
    m = c.mesh
    pt = [p for p in m.patches if p.name == "inlet"][0]
    return slice(pt.start - m.n_internal, pt.start - m.n_internal + pt.size)


def test_initial_state_bitwise(pair):
    c, g, o = pair
    bg, bo = g.boundary(), o.boundary()
    for k in ("p", "U", "T"):
        np.testing.assert_array_equal(bg[k], bo[k])
    for a, b in zip(g.calcFlux(), o.calcFlux()):
        np.testing.assert_array_equal(a, b)
    rg, dtg = g.residual()
    ro, dto = o.residual()
    np.testing.assert_array_equal(rg, ro)
    np.testing.assert_array_equal(dtg, dto)
    g.createJacobian()
    for a, b in zip(g.jacobian(), o.createJacobian()):
        np.testing.assert_array_equal(a, b)


def test_after_state_update(pair):
    c, g, o = pair
    g.residual(); o.residual()      # phi of the boundary faces for the correction
    W = o.fields()["W"]
    g.set_state(W)
    o.set_state_w(W)
    bg, bo = g.boundary(), o.boundary()
    sl = _inlet(c)
    assert (bo["inflow"][sl] == 1).all()
    for k in ("p", "U", "T"):
        assert rel_err(bg[k], bo[k]) < 1e-14, k
    assert not np.array_equal(bo["p"][sl], c.bcs[[p.name for p in c.mesh.patches].index("inlet")].p_val)
    rg, dtg = g.residual()
    ro, dto = o.residual()
    assert rel_err(rg, ro) < 1e-12
    assert rel_err(dtg, dto) < 1e-12
    g.createJacobian()
    for a, b in zip(g.jacobian(), o.createJacobian()):
        assert rel_err(a, b) < 1e-12


def test_pseudo_steps(pair):
    c, g, o = pair
    hg = g.pseudo_steps(5)
    ho = o.pseudo_steps(5)
    assert rel_err(hg[0], ho[0]) < 1e-12
    assert rel_err(hg, ho) < 1e-8
    fg, fo = g.fields(), o.fields()
    for k in ("rho", "U", "T", "p", "W"):
        assert rel_err(fg[k], fo[k]) < 1e-8, k
    bg, bo = g.boundary(), o.boundary()
    for k in ("p", "U", "T"):
        assert rel_err(bg[k], bo[k]) < 1e-8, k


def test_unsupported_condition_kinds_rejected():
    from icsfoam_amd import IcsfError
    c = total_inlet_box(n=(4, 3, 3))
    g = dbnsCase.from_case(c)
    with pytest.raises(IcsfError, match="unsupported boundary-condition type"):
        g.set_boundary(0, cases.BC_SLIP, cases.BC_ZERO_GRADIENT, cases.BC_ZERO_GRADIENT)
    with pytest.raises(IcsfError, match="unsupported boundary-condition type"):
        g.set_boundary(0, cases.BC_ZERO_GRADIENT, cases.BC_TOTAL_TEMPERATURE, cases.BC_ZERO_GRADIENT)
    with pytest.raises(IcsfError, match="without values"):
        g.set_boundary(0, cases.BC_TOTAL_PRESSURE, cases.BC_ZERO_GRADIENT, cases.BC_ZERO_GRADIENT)
