"""Residual adjoint (vector-Jacobian product) on the GPU against the oracle.

The product's reverse pass (icsf_residual_vjp*) is compared with the
oracle's forward-mode dual-number restatement of the same residual:

* tiny meshes: gW, gSf, gCf equal J^T v with J assembled column by column
  from unit tangents (relative 1e-11 of the largest entry);
* small meshes: dot-product test <v, J d> (oracle JVP) = <J^T v, d> (GPU)
  for random d in W, Sf and Cf (relative 1e-11);
* the torch autograd binding gives the same gradient, and its forward
  residual equals the oracle's bit for bit;
* 100^3 jittered box (1 M cells): 256 seeded v, each checked by a central
  finite difference of the GPU residual along a smooth direction d with
  per-entry step 1e-7 (1 + |W|), every mismatch explained by the residual's
  switch records (test_vjp_fd_100cube).

States are made generic (small uncorrelated noise on W) so that no limiter
max/min ties and no velocity component is exactly zero.
"""

import numpy as np
import pytest

from icsfoam_amd import cases, dbnsCase
from icsfoam_amd.mesh import prepare_case
from oracle.oracle import OracleCase

from helpers import ausm_inviscid, box_small, bump_small, generic_state, hllc, roe, smooth_field, with_mu

pytestmark = pytest.mark.gpu

VENK = "cellLimited<Venkatakrishnan> Gauss linear 1"
MINMOD = "cellLimited<minmod> Gauss linear 1"

# Viscous cases take a viscosity at which the viscous terms are a tenth or
# more of the convective ones, so that a mistake in their adjoint shows at
# the 1e-11 tolerance; their geometry adjoint is not built (gW only).
TINY = {
    "bump_venk": lambda: ausm_inviscid(bump_small(nx=12, ny=4)),
    "box_venk": lambda: ausm_inviscid(box_small(n=(5, 4, 3))),
    "box_minmod": lambda: ausm_inviscid(box_small(n=(5, 4, 3)), limiter=MINMOD),
    "box_unlimited": lambda: ausm_inviscid(box_small(n=(5, 4, 3)), limiter="Gauss linear"),
    "box_upwind": lambda: ausm_inviscid(box_small(n=(5, 4, 3)), recon="upwind"),
    "box_roe_venk": lambda: roe(box_small(n=(5, 4, 3)), mu=0.0),
    "bump_roe_minmod": lambda: roe(bump_small(nx=12, ny=4), mu=0.0, limiter=MINMOD),
    "box_roe_viscous": lambda: roe(box_small(n=(5, 4, 3)), mu=20.0),
    "box_roe_viscous_upwind": lambda: roe(box_small(n=(5, 4, 3)), mu=20.0, recon="upwind"),
    "bump_roe_viscous": lambda: roe(bump_small(nx=12, ny=4), mu=0.5),
    "box_ausm_viscous": lambda: with_mu(ausm_inviscid(box_small(n=(5, 4, 3))), 20.0),
    "box_hllc_venk": lambda: hllc(box_small(n=(5, 4, 3)), mu=0.0),
    "bump_hllc_minmod": lambda: hllc(bump_small(nx=12, ny=4), mu=0.0, limiter=MINMOD),
    "box_hllc_viscous": lambda: hllc(box_small(n=(5, 4, 3)), mu=20.0),
}
SMALL = {
    "bump_venk": lambda: ausm_inviscid(bump_small()),
    "box_venk": lambda: ausm_inviscid(box_small()),
    "box_minmod": lambda: ausm_inviscid(box_small(), limiter=MINMOD),
    "box_roe_venk": lambda: roe(box_small(), mu=0.0),
    "box_roe_viscous": lambda: roe(box_small(), mu=20.0),
    "bump_roe_viscous": lambda: roe(bump_small(), mu=0.5),
    "box_hllc_venk": lambda: hllc(box_small(), mu=0.0),
    "box_hllc_viscous": lambda: hllc(box_small(), mu=20.0),
}


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("name", sorted(TINY))
def test_vjp_equals_dense_transpose(name):
    print(f"[tiny {name}]", flush=True)
    c = TINY[name]()
    g, o = dbnsCase.from_case(c), OracleCase(c)
    W = generic_state(o.fields()["W"], 5)
    JW, JS, JC = o.residual_jac_dense(W)
    v = np.random.default_rng(2).standard_normal((5, g.nCells))
    viscous = "viscous" in name
    gW, gS, gC = g.residual_vjp(v, W, geometry=not viscous)
    assert _rel(gW.ravel(), JW.T @ v.ravel()) < 1e-11
    if viscous:
        return
    assert _rel(gS.ravel(), JS.T @ v.ravel()) < 1e-11
    refC = JC.T @ v.ravel()
    if np.abs(refC).max() == 0:
        assert np.abs(gC).max() == 0
    else:
        assert _rel(gC.ravel(), refC) < 1e-11


@pytest.mark.parametrize("name", sorted(SMALL))
def test_vjp_dot_product_with_oracle_jvp(name):
    print(f"[small {name}]", flush=True)
    c = SMALL[name]()
    g, o = dbnsCase.from_case(c), OracleCase(c)
    W = generic_state(o.fields()["W"], 6)
    rng = np.random.default_rng(8)
    nF = o.nFaces
    viscous = "viscous" in name
    for _ in range(4):
        v = rng.standard_normal((5, g.nCells))
        gW, gS, gC = g.residual_vjp(v, W, geometry=not viscous)
        dW = rng.standard_normal(W.shape) * (1 + np.abs(W))
        dS = rng.standard_normal((nF, 3))
        dC = rng.standard_normal((nF, 3))
        checks = [((dW, None, None), gW, dW)]
        if not viscous:
            checks += [((None, dS, None), gS, dS), ((None, None, dC), gC, dC)]
        for args, grad, d in checks:
            _, jd = o.residual_jvp(W, *args)
            lhs = float(np.sum(v * jd))
            rhs = float(np.sum(grad * d))
            scale = float(np.sum(np.abs(v * jd))) + 1e-300
            assert abs(lhs - rhs) <= 1e-11 * scale, (lhs, rhs)


def test_torch_autograd_binding():
    import torch

    from icsfoam_amd.adjoint import residual, residual_vjp_device

    print("[torch binding box_venk]", flush=True)
    c = ausm_inviscid(box_small())
    g, o = dbnsCase.from_case(c), OracleCase(c)
    W = generic_state(o.fields()["W"], 9)
    Wt = torch.tensor(W, device="cuda", requires_grad=True)
    b = residual(g, Wt)
    o.set_state_w(W)
    b_ref, _ = o.residual()
    np.testing.assert_array_equal(b.detach().cpu().numpy(), b_ref)
    v = np.random.default_rng(3).standard_normal(W.shape)
    (b * torch.tensor(v, device="cuda")).sum().backward()
    gW_np, gS_np, _ = g.residual_vjp(v, W)
    np.testing.assert_array_equal(Wt.grad.cpu().numpy(), gW_np)
    gW_d, gS_d, _ = residual_vjp_device(g, torch.tensor(W, device="cuda"), torch.tensor(v, device="cuda"))
    np.testing.assert_array_equal(gW_d.cpu().numpy(), gW_np)
    np.testing.assert_array_equal(gS_d.cpu().numpy(), gS_np)


def test_vjp_fd_100cube():
    """The adjoint configuration: 100^3 jittered box, inviscid AUSM+up,
    second order (linearUpwind + cellLimited<Venkatakrishnan>), FP64; 256
    seeded directions d (smooth fields scaled by 1 + |W|) and adjoint seeds
    v.  Each direction either matches the central difference of step
    h = 1e-7 to 1e-6 relative, v.(b(W+hd) - b(W-hd))/2h against gW.d, or
    crosses a point where the residual is not differentiable: its records
    (AUSM+up switches of every face, limiter max/min element and active face
    of every cell and field; icsf_residual_records_device) differ between W
    and W+hd or W-hd.  Such a direction must match the one-sided difference
    on the side whose records equal W's, to 1e-5 (one-sided truncation
    O(h)); where both sides switch at step h, at the first of h/10, h/100,
    h/1000 where one side does not.  A direction above 1e-6 whose records
    do not switch on either side lies in one smooth piece and must match
    the Richardson extrapolation of the central differences of steps h and
    h/2 to 1e-6.  No direction may be left unexplained."""
    import torch

    from icsfoam_amd.adjoint import fd_gate

    n = 100
    print(f"[fd {n}^3 AUSM+up 2nd order]", flush=True)
    c = cases.case_box(n=(n, n, n), lengths=(1.0, 1.0, 1.0), jitter=0.1, seed=7, flux="AUSMPlusUp", mu=0.0)
    c = prepare_case(c, permute_seed=2024)
    g = dbnsCase.from_case(c)
    dev = torch.device("cuda")
    W = torch.tensor(generic_state(g.fields()["W"], 11), device=dev)
    X = torch.tensor(cases.approx_cell_centres(c.mesh), device=dev)
    r = fd_gate(g, W, X, n_dirs=256)
    print(f"[fd {n}^3] {r['directions']} directions: median rel err {r['median_rel_err']:.2e}, max "
          f"{r['max_rel_err']:.2e}; {r['above_tol']} above 1e-6 ({r['smooth_richardson']} without a switch, "
          f"Richardson; {r['switch_both_sides_at_h']} switching on both sides at step h; one-sided on the smooth "
          f"side: max {r['one_sided_max']:.2e}); VJP {r['vjp_ms']:.2f} ms/call", flush=True)
    assert not r["unexplained"], r["unexplained"]


def test_vjp_fd_viscous_roe_48cube():
    """The discretisation of the 160^3 headline workload (viscous Roe,
    linearUpwind + cellLimited<Venkatakrishnan>, jittered box, slip walls)
    at 48^3, on the state after 300 mixed-precision pseudo-steps from the
    workload's initial field made generic (1e-3 per-cell noise, as in
    test_vjp_fd_100cube): the finite-difference gate over 64 directions
    (Roe switches in the records: |u|, |u +- c|, the entropy fix, the max of
    the two acoustic speeds).  Without the noise the state sits on switches
    itself - the flow relaxes towards the uniform stream, where neighbouring
    values tie in every limiter and the cross-stream velocity is at
    rounding level - and central differences there average the two
    one-sided slopes at every step size."""
    import torch

    from icsfoam_amd.adjoint import fd_gate

    n = 48
    print(f"[fd {n}^3 viscous Roe]", flush=True)
    c = cases.case_box(n=(n, n, n), lengths=(1.0, 1.0, 1.0), jitter=0.1, seed=7, flux="ROE", mu=1.8e-5,
                       mach=(0.5, 0.9), p_jump=2.0, noise=0.02, co=50.0, solver="FGMRES", nDirections=1,
                       precision="mixed", innerDirections=30, orthogonalisation="CGS")
    c = prepare_case(c, permute_seed=2024)
    g = dbnsCase.from_case(c)
    g.pseudo_steps(300, residuals=False)
    dev = torch.device("cuda")
    W = torch.tensor(generic_state(g.fields()["W"], 11), device=dev)
    X = torch.tensor(cases.approx_cell_centres(c.mesh), device=dev)
    r = fd_gate(g, W, X, n_dirs=64)
    print(f"[fd {n}^3 viscous Roe] {r['directions']} directions: median rel err {r['median_rel_err']:.2e}, max "
          f"{r['max_rel_err']:.2e}; {r['above_tol']} above 1e-6 ({r['smooth_richardson']} without a switch, "
          f"Richardson; {r['switch_both_sides_at_h']} switching on both sides at step h; one-sided max "
          f"{r['one_sided_max']:.2e}); VJP {r['vjp_ms']:.2f} ms/call", flush=True)
    assert not r["unexplained"], r["unexplained"]


def test_vjp_on_converged_viscous_roe_state_matches_oracle_tangent():
    """The headline workload's discretisation at 32^3 run to a 1e-8 drop of
    ||R V||_2 (the uniform stream: limiter ties in every cell, cross-stream
    velocity at rounding level, so the residual has no derivative there and
    finite differences do not apply): the VJP equals the oracle's
    dual-number Jacobian-vector product, which resolves the same ties the
    same way, in the dot-product test <v, J d> = <J^T v, d> (1e-11 of
    sum |v J d|) for four random v and smooth d."""
    import torch

    from icsfoam_amd.adjoint import residual_vjp_device

    n = 32
    print(f"[converged {n}^3 viscous Roe vs oracle tangent]", flush=True)
    c = cases.case_box(n=(n, n, n), lengths=(1.0, 1.0, 1.0), jitter=0.1, seed=7, flux="ROE", mu=1.8e-5,
                       mach=(0.5, 0.9), p_jump=2.0, noise=0.02, co=50.0, solver="FGMRES", nDirections=1,
                       precision="mixed", innerDirections=30, orthogonalisation="CGS")
    c = prepare_case(c, permute_seed=2024)
    g = dbnsCase.from_case(c)
    b, _ = g.residual()
    r0 = float(np.sqrt(np.sum(b * b)))
    rel, steps = 1.0, 0
    while rel >= 1e-8 and steps < 4000:
        g.pseudo_steps(50, residuals=False)
        steps += 50
        b, _ = g.residual()
        rel = float(np.sqrt(np.sum(b * b))) / r0
    print(f"[converged {n}^3] {steps} steps, drop {rel:.2e}", flush=True)
    assert rel < 1e-8
    W = g.fields()["W"]
    o = OracleCase(c)
    X = cases.approx_cell_centres(c.mesh)
    rng = np.random.default_rng(21)
    dev = torch.device("cuda")
    for _ in range(4):
        v = rng.standard_normal(W.shape)
        d = smooth_field(X, rng, 5) * (1 + np.abs(W))
        gW, _, _ = residual_vjp_device(g, torch.tensor(W, device=dev), torch.tensor(v, device=dev), geometry=False)
        _, jd = o.residual_jvp(W, d)
        lhs, rhs = float(np.sum(v * jd)), float(np.sum(gW.cpu().numpy() * d))
        assert abs(lhs - rhs) <= 1e-11 * float(np.sum(np.abs(v * jd))), (lhs, rhs)
