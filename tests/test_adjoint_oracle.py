"""Pins of the oracle's tangent-linear (dual-number) residual, the checker of
the residual adjoint (AUSM+up, Roe and HLLC, inviscid and viscous).  CPU
only.

* its FP64 primal is bit-identical to the oracle residual of the reference
  path (same state obtained from W through updateFields);
* its Jacobian-vector products agree with central finite differences of
  that primal along smooth directions in W, Sf and Cf (h = 1e-7 relative;
  tolerance 1e-6 relative to the largest entry - FD truncation and
  rounding are ~1e-8 here; smooth directions keep the +-h interval clear of
  the limiter's and AUSM+up's switch surfaces, which random per-cell
  directions do cross).
"""
import numpy as np
import pytest

from oracle.oracle import OracleCase

from helpers import ausm_inviscid, box_small, bump_small, generic_state, hllc, roe, smooth_field, with_mu

VENK = "cellLimited<Venkatakrishnan> Gauss linear 1"
MINMOD = "cellLimited<minmod> Gauss linear 1"

# viscous cases take a viscosity at which the viscous terms are of the
# order of the convective ones (a mistake in them must show at 1e-6)
CASES = {
    "bump_venk": lambda: ausm_inviscid(bump_small()),
    "box_venk": lambda: ausm_inviscid(box_small()),
    "box_minmod": lambda: ausm_inviscid(box_small(), limiter=MINMOD),
    "box_unlimited": lambda: ausm_inviscid(box_small(), limiter="Gauss linear"),
    "box_upwind": lambda: ausm_inviscid(box_small(), recon="upwind"),
    "box_roe_venk": lambda: roe(box_small(), mu=0.0),
    "box_roe_minmod": lambda: roe(box_small(), mu=0.0, limiter=MINMOD),
    "box_roe_viscous": lambda: roe(box_small(), mu=20.0),
    "bump_roe_viscous": lambda: roe(bump_small(), mu=0.5),
    "box_ausm_viscous": lambda: with_mu(ausm_inviscid(box_small()), 20.0),
    "box_hllc_venk": lambda: hllc(box_small(), mu=0.0),
    "box_hllc_minmod": lambda: hllc(box_small(), mu=0.0, limiter=MINMOD),
    "box_hllc_viscous": lambda: hllc(box_small(), mu=20.0),
}


@pytest.fixture(params=sorted(CASES))
def oc(request):
    o = OracleCase(CASES[request.param]())
    W = generic_state(o.fields()["W"], 5)
    o.viscous = "viscous" in request.param
    return o, W


def test_tangent_primal_matches_oracle_residual(oc):
    o, W = oc
    o.set_state_w(W)
    b_ref, _ = o.residual()
    np.testing.assert_array_equal(o.residual_w(W), b_ref)
    b, _ = o.residual_jvp(W, np.zeros_like(W))
    np.testing.assert_array_equal(b, b_ref)


def test_tangent_vs_fd_state(oc):
    o, W = oc
    C = o.mesh.geometry()["C"]
    rng = np.random.default_rng(3)
    h = 1e-7
    for _ in range(3):
        d = smooth_field(C, rng, 5) * (1 + np.abs(W))
        _, db = o.residual_jvp(W, d)
        fd = (o.residual_w(W + h * d) - o.residual_w(W - h * d)) / (2 * h)
        assert np.abs(fd - db).max() <= 1e-6 * np.abs(db).max()


def test_tangent_vs_fd_geometry(oc):
    o, W = oc
    if o.viscous:
        pytest.skip("the geometry derivatives cover the inviscid residual")
    g = o.mesh.geometry()
    Sf, Cf = g["Sf"], g["Cf"]
    rng = np.random.default_rng(4)
    h = 1e-7
    dS = smooth_field(Cf, rng, 3).T * np.abs(Sf).max()
    # Boundary faces stay fixed: on the planar slip walls the tangential
    # boundary velocity equals the cell's exactly, and tilting the wall
    # normal moves it off that tie in the direction of h's sign (a kink of
    # the limiter's max/min at this very point).
    dS[o.nInternal:] = 0.0
    _, db = o.residual_jvp(W, dSf=dS)
    fd = (o.residual_w(W, Sf=Sf + h * dS) - o.residual_w(W, Sf=Sf - h * dS)) / (2 * h)
    assert np.abs(fd - db).max() <= 1e-6 * np.abs(db).max()
    # b depends on Cf only through the O(cell size) extrapolations: a
    # direction of the cells' own scale keeps FD rounding below 1e-7
    dC = smooth_field(Cf, rng, 3).T * 0.1
    _, db = o.residual_jvp(W, dCf=dC)
    if np.abs(db).max() == 0:  # first order: no dependence on Cf
        return
    hc = 1e-6
    fd = (o.residual_w(W, Cf=Cf + hc * dC) - o.residual_w(W, Cf=Cf - hc * dC)) / (2 * hc)
    assert np.abs(fd - db).max() <= 1e-6 * np.abs(db).max()


def test_dense_jacobian_columns(oc):
    """The dense Jacobian's columns are the JVPs of unit tangents (the
    dense routine and the single-direction routine agree), checked on a
    few columns of a small case."""
    o, W = oc
    if o.nCells > 320:
        pytest.skip("dense Jacobian only on the small cases")
    JW, JS, JC = o.residual_jac_dense(W)
    rng = np.random.default_rng(6)
    d = rng.standard_normal(W.shape)
    _, db = o.residual_jvp(W, d)
    np.testing.assert_allclose(JW @ d.ravel(), db.ravel(), rtol=0, atol=1e-9 * np.abs(db).max())
