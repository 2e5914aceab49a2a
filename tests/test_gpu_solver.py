"""Solver variants on the GPU against the oracle and dense linear algebra.

* CGS orthogonalisation vs the oracle's MGS GMRES: same Krylov space, so the
  increments agree to rounding (1e-8 relative).
* Block-Jacobi preconditioner vs the oracle: same 5x5 LU with partial
  pivoting in the same order (bit-exact).
* FGMRES (FP64, full Krylov space) equals the dense direct solve.
* Mixed precision (FP64 outer FGMRES, FP32 inner GMRES + FP32 LU-SGS on the
  FP32 Jacobian): converges to the FP64 direct solution when given enough
  outer iterations; after 5 pseudo-steps its state stays within 1e-3
  relative of the FP64-everywhere oracle (the error is printed per case).
"""
import dataclasses

import numpy as np
import pytest

from icsfoam_amd import cases, dbnsCase
from icsfoam_amd._lib import IcsfError
from icsfoam_amd.mesh import prepare_case
from oracle.oracle import OracleCase

from helpers import box_small, bump_small, rel_err

pytestmark = pytest.mark.gpu


def _dense(d, u, l, m):
    n = m.n_cells
    A = np.zeros((5 * n, 5 * n))
    for c in range(n):
        A[5 * c:5 * c + 5, 5 * c:5 * c + 5] = d[c]
    for f in range(m.n_internal):
        P, N = m.owner[f], m.neighbour[f]
        A[5 * P:5 * P + 5, 5 * N:5 * N + 5] = u[f]
        A[5 * N:5 * N + 5, 5 * P:5 * P + 5] = l[f]
    return A


def _with(case, **kw):
    base = dict(solver="GMRES", nDirections=30, precision="double", co=50.0)
    base.update(kw)
    return dataclasses.replace(case, fvSolution=cases.solution_dict(**base))


@pytest.mark.parametrize("which", ["bump", "box"])
def test_cgs_gmres_matches_oracle_mgs(which):
    c = bump_small() if which == "bump" else box_small()
    print(f"[cgs {which}]", flush=True)
    g = dbnsCase.from_case(_with(c, orthogonalisation="CGS"))
    o = OracleCase(c)
    g.residual(); bo, _ = o.residual()
    g.createJacobian(); o.createJacobian()
    dWg, rg, _ = g.solveForIncr()
    dWo, ro, _ = o.solveForIncr(g.fields()["W"], bo)
    assert rel_err(dWg, dWo) < 1e-8
    assert rel_err(rg, ro) < 1e-8


@pytest.mark.parametrize("which", ["bump", "box"])
def test_fused_spmv_lusgs_matches_unfused(which):
    """The Arnoldi SpMV fused with the LU-SGS sweeps (two-colour numbering)
    is the same operator: GMRES increments agree with the unfused path to
    1e-12 relative (only the association of the forward-sweep sums
    differs), in FP64 and in the mixed-precision inner solve."""
    c = bump_small() if which == "bump" else box_small()
    print(f"[fused {which}]", flush=True)
    for kw in (dict(), dict(solver="FGMRES", nDirections=2, precision="mixed", innerDirections=20)):
        gf = dbnsCase.from_case(_with(c, fusedSweeps=True, **kw))
        gu = dbnsCase.from_case(_with(c, fusedSweeps=False, **kw))
        assert gf.mesh.nFwdLevels == 2 and gf.mesh.nBwdLevels == 2  # the fused path applies
        for g in (gf, gu):
            g.residual()
            g.createJacobian()
        dWf, rf, _ = gf.solveForIncr()
        dWu, ru, _ = gu.solveForIncr()
        tol = 1e-12 if not kw else 1e-4
        assert rel_err(dWf, dWu) < tol, kw
        assert rel_err(rf, ru) < tol, kw


@pytest.mark.parametrize("which", ["bump", "box"])
def test_generator_storage_equals_block_storage(which):
    """Off-diagonal blocks stored as their 10 generators and rebuilt inside
    the SpMV/LU-SGS kernels are the stored blocks bit for bit in FP64: same
    Jacobian read-back, same 4-step history and state.  In the mixed form
    the FP32 generators rebuild FP32 blocks in FP32 arithmetic (instead of
    rounding FP64 blocks): the state after 3 steps agrees to 1e-6."""
    c = bump_small() if which == "bump" else box_small()
    print(f"[generators {which}]", flush=True)
    gg = dbnsCase.from_case(_with(c, jacobianStorage="generators"))
    gb = dbnsCase.from_case(_with(c, jacobianStorage="blocks"))
    for g in (gg, gb):
        g.residual()
        g.createJacobian()
    for a, b in zip(gg.jacobian(), gb.jacobian()):
        np.testing.assert_array_equal(a, b)
    gg = dbnsCase.from_case(_with(c, jacobianStorage="generators"))
    gb = dbnsCase.from_case(_with(c, jacobianStorage="blocks"))
    np.testing.assert_array_equal(gg.pseudo_steps(4), gb.pseudo_steps(4))
    np.testing.assert_array_equal(gg.fields()["W"], gb.fields()["W"])
    kw = dict(solver="FGMRES", nDirections=1, precision="mixed", innerDirections=20)
    mg = dbnsCase.from_case(_with(c, jacobianStorage="generators", **kw))
    mb = dbnsCase.from_case(_with(c, jacobianStorage="blocks", **kw))
    mg.pseudo_steps(3)
    mb.pseudo_steps(3)
    assert rel_err(mg.fields()["W"], mb.fields()["W"]) < 1e-6


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("which", ["bump", "box"])
def test_fused_cell_sums_give_the_same_bits(which, form):
    """The pseudo-step forms the residual's cell sums (b, rPseudoDeltaT) in
    the Jacobian kernel's face walk: same residual history and state bit
    for bit as the separate cell-residual and Jacobian kernels."""
    print(f"[fused cell sums {which} {form}]", flush=True)
    kw = dict()
    if form == "mixed":
        kw.update(solver="FGMRES", nDirections=1, precision="mixed", innerDirections=20)
    c = bump_small() if which == "bump" else box_small()
    gf = dbnsCase.from_case(_with(c, fusedCellSums=True, **kw))
    gu = dbnsCase.from_case(_with(c, fusedCellSums=False, **kw))
    np.testing.assert_array_equal(gf.pseudo_steps(4), gu.pseudo_steps(4))
    np.testing.assert_array_equal(gf.fields()["W"], gu.fields()["W"])


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_autotuned_launches_give_the_same_bits(form):
    """The autotuner only picks workgroup sizes and projection chunks: a
    tuned and an untuned case give the same residual history and state bit
    for bit, and the chosen parameters are among the candidates."""
    print(f"[autotune {form}]", flush=True)
    kw = dict(orthogonalisation="CGS")
    if form == "mixed":
        kw.update(solver="FGMRES", nDirections=1, precision="mixed", innerDirections=20)
    c = box_small()
    ga = dbnsCase.from_case(_with(c, autotune=True, **kw))
    gn = dbnsCase.from_case(_with(c, autotune=False, **kw))
    ha = ga.pseudo_steps(4)
    hn = gn.pseudo_steps(4)
    np.testing.assert_array_equal(ha, hn)
    np.testing.assert_array_equal(ga.fields()["W"], gn.fields()["W"])
    t, u = ga.tuning(), gn.tuning()
    print(f"tuned: {t}")
    assert t["tuned"] and not u["tuned"]
    for k in ("op_workgroup", "face_workgroup", "cell_workgroup", "jacobian_workgroup"):
        assert t[k] in (64, 128, 256) and u[k] == 256
    assert t["projection_chunk"] in (8, 16, 32) and u["projection_chunk"] == 16


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("which", ["bump", "box"])
def test_every_launch_variant_gives_the_same_bits(which, form):
    """Every launch-parameter candidate the autotuner may pick, pinned with
    icsf_set_tuning - workgroup sizes, the 128- and 168-VGPR budgets,
    8/16/32 vectors per projection pass - gives the untuned case's residual
    history and state bit for bit (AUSM+up bump: 8 reconstructed fields;
    viscous Roe box: 7); parameters outside the candidates are refused."""
    print(f"[launch variants {which} {form}]", flush=True)
    kw = dict(orthogonalisation="CGS")
    if form == "mixed":
        kw.update(solver="FGMRES", nDirections=1, precision="mixed", innerDirections=20)
    c = bump_small() if which == "bump" else box_small()
    g0 = dbnsCase.from_case(_with(c, autotune=False, **kw))
    h0 = g0.pseudo_steps(3)
    W0 = g0.fields()["W"]
    # [op, face, cell, jacobian workgroups, chunk, -, op, face, cell, jacobian budgets]
    for p in ([128, 64, 128, 64, 8, 0, 1, 1, 1, 1], [64, 128, 64, 128, 32, 0, 0, 2, 2, 2],
              [256, 256, 256, 256, 16, 0, 1, 0, 0, 0], [256, 64, 64, 256, 32, 0, 0, 1, 2, 2]):
        g = dbnsCase.from_case(_with(c, autotune=False, **kw))
        g.set_tuning(p)
        np.testing.assert_array_equal(g.tuning_params()[[0, 1, 2, 3, 4, 6, 7, 8, 9]],
                                      np.array(p)[[0, 1, 2, 3, 4, 6, 7, 8, 9]])
        np.testing.assert_array_equal(g.pseudo_steps(3), h0)
        np.testing.assert_array_equal(g.fields()["W"], W0)
    for bad in ([96, 64, 64, 64, 16, 0, 0, 0, 0, 0], [64, 64, 64, 64, 12, 0, 0, 0, 0, 0],
                [64, 64, 64, 64, 16, 0, 2, 0, 0, 0], [64, 64, 64, 64, 16, 0, 0, 0, 3, 0]):
        with pytest.raises(IcsfError):
            g0.set_tuning(bad)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_graph_replay_equals_eager(form):
    """From the third step on a pseudo-step is replayed from a captured HIP
    graph: 7 steps (5 replays, each followed by a host read of the
    residuals) give the eager path's residual history and state bit for
    bit."""
    print(f"[graph {form}]", flush=True)
    kw = dict() if form == "fp64" else dict(solver="FGMRES", nDirections=1, precision="mixed", innerDirections=20)
    c = box_small()
    ge = dbnsCase.from_case(_with(c, graphCapture=False, **kw))
    gg = dbnsCase.from_case(_with(c, graphCapture=True, **kw))
    he = ge.pseudo_steps(7)
    hg = gg.pseudo_steps(7)
    np.testing.assert_array_equal(hg, he)
    for k in ("rho", "U", "T", "p", "W"):
        np.testing.assert_array_equal(gg.fields()[k], ge.fields()[k])
    np.testing.assert_array_equal(gg.pseudoCo(), ge.pseudoCo())


def test_block_jacobi_bitwise():
    c = _with(box_small(), precond="Jacobi")
    print("[jacobi]", flush=True)
    g = dbnsCase.from_case(c)
    o = OracleCase(c)
    g.residual(); o.residual()
    g.createJacobian(); o.createJacobian()
    x = np.random.default_rng(5).standard_normal((5, g.nCells))
    np.testing.assert_array_equal(g.precondition(x), o.precondition(x))


def _tiny():
    c = cases.case_box(n=(3, 2, 2), jitter=0.1, flux="ROE", mu=1.8e-5, noise=0.1)
    return prepare_case(c, permute_seed=3)


@pytest.mark.parametrize("orth", ["MGS", "CGS"])
def test_fgmres_fp64_full_space_is_direct_solve(orth):
    c = _with(_tiny(), solver="FGMRES", nDirections=60, innerDirections=1, tol=1e-300, relTol=0.0,
              orthogonalisation=orth)
    print(f"[fgmres fp64 {orth}]", flush=True)
    g = dbnsCase.from_case(c)
    b, _ = g.residual()
    g.createJacobian()
    d, u, l = g.jacobian()
    A = _dense(d, u, l, c.mesh)
    dW, res, it = g.solveForIncr()
    x = np.linalg.solve(A, b.T.ravel())
    assert rel_err(dW.T.ravel(), x) < 1e-8


def test_mixed_fgmres_converges_to_fp64_solution():
    c = _with(_tiny(), solver="FGMRES", nDirections=6, innerDirections=60, precision="mixed", tol=1e-300,
              relTol=0.0, maxIter=2, orthogonalisation="CGS")
    print("[fgmres mixed]", flush=True)
    g = dbnsCase.from_case(c)
    b, _ = g.residual()
    g.createJacobian()
    d, u, l = g.jacobian()           # FP32-stored blocks, widened
    A = _dense(d, u, l, c.mesh)
    dW, res, it = g.solveForIncr()
    x = np.linalg.solve(A, b.T.ravel())
    err = rel_err(dW.T.ravel(), x)
    print(f"mixed FGMRES vs direct solve of the FP32 Jacobian: {err:.2e}")
    assert err < 1e-9


@pytest.mark.parametrize("which", ["bump", "box"])
def test_mixed_pseudo_steps_vs_fp64_oracle(which):
    c = bump_small() if which == "bump" else box_small()
    print(f"[mixed steps {which}]", flush=True)
    gm = dbnsCase.from_case(_with(c, solver="FGMRES", nDirections=1, innerDirections=30, precision="mixed",
                                  orthogonalisation="CGS"))
    o = OracleCase(c)
    hm = gm.pseudo_steps(5)
    ho = o.pseudo_steps(5)
    fm, fo = gm.fields(), o.fields()
    errs = {k: rel_err(fm[k], fo[k]) for k in ("rho", "U", "T", "p")}
    print(f"mixed vs FP64 oracle after 5 steps ({which}): {errs}; init res mixed {hm[-1][:2]} fp64 {ho[-1][:2]}")
    assert max(errs.values()) < 1e-3
    # the nonlinear residual keeps falling at the FP64 form's pace
    assert hm[-1][0] < 1.5 * ho[-1][0] and hm[-1][1] < 1.5 * ho[-1][1]


def test_matrix_calls_after_pseudo_steps_need_a_new_assembly():
    """After pseudo-steps the state is the updated one; the residual and the
    matrix of the last step belong to the previous state, so the matrix
    calls refuse until the residual and the Jacobian are formed again."""
    from icsfoam_amd._lib import IcsfError

    g = dbnsCase.from_case(box_small())
    g.pseudo_steps(2)
    with pytest.raises(IcsfError, match="not assembled"):
        g.solveForIncr()
    with pytest.raises(IcsfError, match="must run before"):
        g.createJacobian()
    g.residual()
    g.createJacobian()
    dW, _, _ = g.solveForIncr()
    assert np.all(np.isfinite(dW))


def test_case_after_timed_case_is_clean():
    """A case destroyed after event timing leaves no pending HIP error for
    the next case (its first step runs the autotuner, which checks)."""
    print("[lifecycle]", flush=True)
    c = box_small()
    g1 = dbnsCase.from_case(c)
    g1.pseudo_steps(2)
    g1.timing(0x3F)
    g1.pseudo_steps(1)
    g1.timing_read()
    g1.timing(0)
    del g1
    g2 = dbnsCase.from_case(_with(c, solver="FGMRES", nDirections=1, precision="mixed", innerDirections=20,
                                  orthogonalisation="CGS"))
    h = g2.pseudo_steps(3)
    assert np.all(np.isfinite(h)) and g2.tuning()["tuned"]
