"""Failure handling through the C ABI and the solver options that only
change the preconditioner's precision or depth.

Each error code is forced from the host with valid kernels only (no GPU
fault, no bounded wait shortened): unknown keywords are refused at case
creation (newConvectiveFluxScheme.C:51-60, coupledMatrixSolver.C:53-61,
coupledMatrixPreconditioner.C:41-64), a diagonal whose LU-SGS coefficient
falls below VSMALL is refused (lusgs.C:119-123, which tests
1/max|diag| < VSMALL: an oversized diagonal, not an all-zero one), a
non-finite boundary value is refused, and a finite inflow velocity whose
square overflows makes the conserved state non-finite during a pseudo-step,
which the state update flags on the device.
"""
import dataclasses

import numpy as np
import pytest

from icsfoam_amd import cases, dbnsCase
from icsfoam_amd._lib import ERR_NONFINITE, IcsfError, KeywordError, ZeroDiagonalError
from icsfoam_amd.solver import fvMesh, hbCase
from oracle.oracle import OracleCase

from helpers import box_small, rel_err

pytestmark = pytest.mark.gpu


def _with(case, **kw):
    base = dict(solver="GMRES", nDirections=30, precision="double", co=50.0)
    base.update(kw)
    return dataclasses.replace(case, fvSolution=cases.solution_dict(**base))


@pytest.mark.parametrize("what", ["fluxScheme", "solver", "preconditioner", "precision"])
def test_unknown_keyword_is_refused(what):
    print(f"[keyword {what}]", flush=True)
    c = box_small(n=(4, 3, 3))
    if what == "fluxScheme":
        c = dataclasses.replace(c, fvSchemes=cases.schemes_dict("Rusanov"))
    elif what == "solver":
        c = dataclasses.replace(c, fvSolution=cases.solution_dict(solver="BiCGStab"))
    elif what == "preconditioner":
        c = dataclasses.replace(c, fvSolution=cases.solution_dict(precond="ILU0"))
    else:
        c = dataclasses.replace(c, fvSolution=cases.solution_dict(solver="FGMRES", precision="half"))
    with pytest.raises(KeywordError):
        dbnsCase.from_case(c)


def test_oversized_diagonal_is_refused_and_zero_diagonal_passes():
    """The reference's check (lusgs.C:119-123) fires when 1/max|diag| falls
    below VSMALL; an all-zero diagonal gives 1/SMALL and passes."""
    print("[diag]", flush=True)
    c = _with(box_small(n=(4, 3, 3)), jacobianStorage="blocks")
    g = dbnsCase.from_case(c)
    g.residual()
    g.createJacobian()
    d, u, l = g.jacobian()
    d0 = d.copy()
    d0[3] = 0.0
    g.set_jacobian(d0, u, l)          # accepted: rDiag = 1/SMALL on cell 3
    r = g.precondition(np.ones((5, g.nCells)))
    assert np.all(np.isfinite(r))
    dh = d.copy()
    dh[5, 1, 1] = 1e301
    with pytest.raises(ZeroDiagonalError):
        g.set_jacobian(dh, u, l)


def test_nonfinite_boundary_value_is_refused():
    print("[bc nan]", flush=True)
    c = box_small(n=(4, 3, 3))
    g = dbnsCase.from_case(c)
    i = [p.name for p in c.mesh.patches].index("outlet")
    size = c.mesh.patches[i].size
    with pytest.raises(IcsfError) as e:
        g.set_boundary(i, cases.BC_FIXED_VALUE, cases.BC_ZERO_GRADIENT, cases.BC_ZERO_GRADIENT,
                       np.full(size, np.nan))
    assert e.value.code == ERR_NONFINITE


def _overflowing_inlet(c):
    i = [p.name for p in c.mesh.patches].index("inlet")
    bcs = list(c.bcs)
    U = np.zeros((c.mesh.patches[i].size, 3))
    U[:, 0] = 1e200  # finite, but |U|^2 overflows: the flux and the state turn non-finite
    bcs[i] = dataclasses.replace(bcs[i], U_val=U)
    return dataclasses.replace(c, bcs=bcs)


@pytest.mark.parametrize("graph", [False, True])
def test_nonfinite_state_during_pseudo_steps(graph):
    print(f"[nan steps graph={graph}]", flush=True)
    c = _with(_overflowing_inlet(box_small(n=(6, 4, 4))), graphCapture=graph)
    g = dbnsCase.from_case(c)
    with pytest.raises(IcsfError) as e:
        g.pseudo_steps(4, residuals=False)
    assert e.value.code == ERR_NONFINITE
    # the flag is cleared: a healthy case on the same device steps normally
    ok = dbnsCase.from_case(box_small(n=(6, 4, 4)))
    h = ok.pseudo_steps(2)
    assert np.all(np.isfinite(h))


def _fixed_inlet(c):
    """The harmonic-balance case's inflow patch as a fixed velocity and
    temperature (so that _overflowing_inlet acts in the first step)."""
    i = [p.name for p in c.mesh.patches].index("inlet")
    bcs = list(c.bcs)
    n = c.mesh.patches[i].size
    bcs[i] = cases.BoundarySpec(cases.BC_ZERO_GRADIENT, cases.BC_FIXED_VALUE, cases.BC_FIXED_VALUE, None,
                                np.tile([0.0, 0.0, 170.0], (n, 1)), np.full(n, 300.0))
    return dataclasses.replace(c, bcs=bcs)


def test_nonfinite_state_during_hb_pseudo_steps():
    print("[nan hb]", flush=True)
    inst, D, _ = cases.case_hb(n=(8, 6, 5), n_harmonics=1, noise=0.02, jitter=0.1)
    inst[1] = _overflowing_inlet(_fixed_inlet(inst[1]))
    h = hbCase.from_cases(inst, D, fvMesh(inst[0].mesh))
    with pytest.raises(IcsfError) as e:
        h.pseudo_steps(2, residuals=False)
    assert e.value.code == ERR_NONFINITE


def test_fp32_block_jacobi_is_the_rounded_fp32_inverse():
    """Mixed precision stores the block-Jacobi inverse in FP32: the device
    inverts the FP32 diagonal in FP64 (LU with partial pivoting) and rounds
    the inverse to FP32; applied to an FP64 vector the product is summed in
    FP64.  Checked against numpy's inverse of the same FP32 blocks, rounded
    the same way, to FP32 resolution, and shown to differ from the FP64
    form by the FP32 rounding."""
    print("[jacobi fp32]", flush=True)
    c = box_small()
    g32 = dbnsCase.from_case(_with(c, solver="FGMRES", precision="mixed", precond="Jacobi"))
    g64 = dbnsCase.from_case(_with(c, precond="Jacobi"))
    for g in (g32, g64):
        g.residual()
        g.createJacobian()
    d32, _, _ = g32.jacobian()          # FP32-stored diagonal blocks, widened
    assert np.array_equal(d32.astype(np.float32).astype(np.float64), d32)
    inv32 = np.linalg.inv(d32).astype(np.float32).astype(np.float64)
    x = np.random.default_rng(9).standard_normal((5, g32.nCells))
    ref = np.einsum("cij,jc->ic", inv32, x)
    z32 = g32.precondition(x)
    z64 = g64.precondition(x)
    e32 = rel_err(z32, ref)
    e64 = rel_err(z64, ref)
    print(f"FP32 Jacobi vs rounded FP32 inverse {e32:.2e}; FP64 Jacobi vs it {e64:.2e}")
    assert e32 < 1e-6
    assert e64 > 1e-9  # the FP64 form is not the FP32 one


@pytest.mark.parametrize("precond", ["Jacobi", "LUSGS"])
@pytest.mark.parametrize("inner", [0, 30])
def test_mixed_preconditioner_forms_vs_fp64_oracle(precond, inner):
    """Mixed FGMRES with an FP32 inner GMRES(30), or with the FP32
    preconditioner alone (innerDirections 0), follows the FP64 oracle over
    5 pseudo-steps within 1e-3 relative (the error is printed)."""
    print(f"[mixed {precond} inner={inner}]", flush=True)
    c = box_small()
    gm = dbnsCase.from_case(_with(c, solver="FGMRES", nDirections=30 if inner == 0 else 1, innerDirections=inner,
                                  precision="mixed", precond=precond, orthogonalisation="CGS"))
    o = OracleCase(_with(c, precond=precond))
    hm = gm.pseudo_steps(5)
    ho = o.pseudo_steps(5)
    fm, fo = gm.fields(), o.fields()
    errs = {k: rel_err(fm[k], fo[k]) for k in ("rho", "U", "T", "p")}
    print(f"mixed {precond} inner={inner} vs FP64 oracle after 5 steps: {errs}")
    assert max(errs.values()) < 1e-3
    assert hm[-1][0] < 1.5 * ho[-1][0]
