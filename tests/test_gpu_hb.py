"""Harmonic balance on the GPU against the oracle (small annular sector,
3 and 7 instants).

* residual with the HB source: bitwise (same per-instance kernels as the
  single-instance path; the source sums in HBZone::addSource order);
* coupled SpMV and LU-SGS with the shared diagonal: bitwise (coupledMatrix
  order for the cross-instance terms);
* GMRES increment: relative 1e-10; pseudo-steps: first-step residuals 1e-12,
  histories and fields 1e-8 after 4 steps;
* one instant with D = 0 equals the single-instance GPU path (unfused
  Arnoldi step) bit for bit.
"""
import dataclasses

import numpy as np
import pytest

from icsfoam_amd import cases, dbnsCase
from icsfoam_amd.mesh import prepare_case
from icsfoam_amd.solver import hbCase
from oracle.oracle import OracleHB

from helpers import box_small, rel_err

pytestmark = pytest.mark.gpu


def _hb_cases(K, flux):
    insts, D, t = cases.case_hb(n=(8, 6, 5), n_harmonics=K, flux=flux, jitter=0.1)
    return [prepare_case(c, permute_seed=17) for c in insts], D


CASES = {
    "hb3_roe": lambda: _hb_cases(1, "ROE"),
    "hb7_ausm": lambda: _hb_cases(3, "AUSMPlusUp"),
}


@pytest.fixture(params=sorted(CASES))
def pair(request):
    print(f"[hb case {request.param}]", flush=True)
    insts, D = CASES[request.param]()
    return hbCase.from_cases(insts, D), OracleHB(insts, D)


def test_hb_residual_bitwise(pair):
    g, o = pair
    np.testing.assert_array_equal(g.residual(), o.residual())


def test_hb_spmv_lusgs_bitwise(pair):
    g, o = pair
    g.residual(); o.residual()
    g.createJacobian(); o.createJacobian()
    x = np.random.default_rng(4).standard_normal((g.nO, 5, g.nCells))
    np.testing.assert_array_equal(g.matrixMul(x), o.matrixMul(x))
    np.testing.assert_array_equal(g.precondition(x), o.precondition(x))


def test_hb_gmres_increment(pair):
    g, o = pair
    bg = g.residual(); bo = o.residual()
    g.createJacobian(); o.createJacobian()
    dWg, rg, _ = g.solveForIncr(bg)
    dWo, ro, _ = o.solveForIncr(bo)
    assert rel_err(dWg, dWo) < 1e-10
    assert rel_err(rg, ro) < 1e-10


def test_hb_pseudo_steps(pair):
    g, o = pair
    hg = g.pseudo_steps(4)
    ho = o.pseudo_steps(4)
    assert rel_err(hg[0], ho[0]) < 1e-12
    assert rel_err(hg, ho) < 1e-8
    for J in range(g.nO):
        fg, fo = g.fields(J), o.fields(J)
        for k in ("rho", "U", "T", "p", "W"):
            assert rel_err(fg[k], fo[k]) < 1e-8, (J, k)


def test_hb_cgs_matches_oracle_mgs():
    """CGS on the batched HB vectors spans the same Krylov space as the
    oracle's MGS: increments agree to rounding (1e-8)."""
    print("[hb cgs]", flush=True)
    insts, D, t = cases.case_hb(n=(8, 6, 5), n_harmonics=2, flux="ROE", orthogonalisation="CGS")
    insts = [prepare_case(c, permute_seed=17) for c in insts]
    insts_mgs = [dataclasses.replace(c, fvSolution=cases.solution_dict(co=20.0)) for c in insts]
    g, o = hbCase.from_cases(insts, D), OracleHB(insts_mgs, D)
    bg = g.residual(); bo = o.residual()
    g.createJacobian(); o.createJacobian()
    dWg, rg, _ = g.solveForIncr(bg)
    dWo, ro, _ = o.solveForIncr(bo)
    assert rel_err(dWg, dWo) < 1e-8
    assert rel_err(rg, ro) < 1e-8


def test_hb_fused_operator_matches_unfused():
    """The fused HB Arnoldi step (coupled SpMV + LU-SGS in three passes over
    all instances) equals the unfused one to 1e-12."""
    print("[hb fused]", flush=True)
    insts, D, t = cases.case_hb(n=(8, 6, 5), n_harmonics=2, flux="ROE")
    insts = [prepare_case(c, permute_seed=17) for c in insts]
    unf = [dataclasses.replace(c, fvSolution=cases.solution_dict(co=20.0, fusedSweeps=False)) for c in insts]
    gf, gu = hbCase.from_cases(insts, D), hbCase.from_cases(unf, D)
    assert gf.mesh.nFwdLevels == 2 and gf.mesh.nBwdLevels == 2
    bf = gf.residual(); bu = gu.residual()
    np.testing.assert_array_equal(bf, bu)
    gf.createJacobian(); gu.createJacobian()
    dWf, rf, _ = gf.solveForIncr(bf)
    dWu, ru, _ = gu.solveForIncr(bu)
    assert rel_err(dWf, dWu) < 1e-12
    assert rel_err(rf, ru) < 1e-12


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_hb_fused_cell_sums_give_the_same_bits(form):
    """The HB pseudo-step forms the instances' residual cell sums in the
    Jacobian kernel's face walk and adds the spectral source afterwards:
    same residual history and states bit for bit as the separate kernels."""
    print(f"[hb fused cell sums {form}]", flush=True)
    insts, D, t = cases.case_hb(n=(8, 6, 5), n_harmonics=2, flux="ROE")
    insts = [prepare_case(c, permute_seed=17) for c in insts]
    kw = dict(co=20.0)
    if form == "mixed":
        kw.update(solver="FGMRES", nDirections=1, precision="mixed", innerDirections=20)
    out = []
    for fused in (True, False):
        sol = cases.solution_dict(fusedCellSums=fused, **kw)
        g = hbCase.from_cases([dataclasses.replace(c, fvSolution=sol) for c in insts], D)
        out.append((g.pseudo_steps(4), [g.fields(J)["W"] for J in range(len(insts))]))
        del g
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        np.testing.assert_array_equal(a, b)


def test_single_instant_equals_single_case():
    """Same kernels, same order: bit-for-bit (the single-instance solver
    with the unfused Arnoldi step, which the HB solver uses)."""
    print("[hb single instant]", flush=True)
    c = box_small()
    c = dataclasses.replace(c, fvSolution=cases.solution_dict(fusedSweeps=False))
    s = dbnsCase.from_case(c)
    h = hbCase.from_cases([c], np.zeros((1, 1)), mesh=s.mesh)
    rs = s.pseudo_steps(3)
    rh = h.pseudo_steps(3)[:, 0, :]
    np.testing.assert_array_equal(rh, rs)
    np.testing.assert_array_equal(h.fields(0)["W"], s.fields()["W"])


def test_hb_graph_replay_equals_eager():
    """From the third step on the HB pseudo-step is replayed from a captured
    HIP graph: 6 steps (4 replays, each followed by a host read of the
    residuals) equal the eager path bit for bit."""
    print("[hb graph]", flush=True)
    insts, D, t = cases.case_hb(n=(8, 6, 5), n_harmonics=2, flux="ROE")
    insts = [prepare_case(c, permute_seed=17) for c in insts]
    eager = [dataclasses.replace(c, fvSolution=cases.solution_dict(co=20.0, graphCapture=False)) for c in insts]
    graph = [dataclasses.replace(c, fvSolution=cases.solution_dict(co=20.0, graphCapture=True)) for c in insts]
    ge, gg = hbCase.from_cases(eager, D), hbCase.from_cases(graph, D)
    np.testing.assert_array_equal(gg.pseudo_steps(6), ge.pseudo_steps(6))
    for J in range(len(insts)):
        np.testing.assert_array_equal(gg.fields(J)["W"], ge.fields(J)["W"])


def test_hb_mixed_precision_vs_fp64_oracle():
    """C3's solver form: outer FGMRES on the FP64 coupled vectors, FP32 HB
    Jacobian and an FP32 inner GMRES(30) + LU-SGS over all instances.  After
    4 pseudo-steps every instance stays within 1e-3 of the FP64 oracle and
    the nonlinear residuals fall at the FP64 pace."""
    from oracle.oracle import OracleHB
    print("[hb mixed]", flush=True)
    insts, D = _hb_cases(1, "ROE")
    mixed = [dataclasses.replace(c, fvSolution=cases.solution_dict(
        co=20.0, solver="FGMRES", nDirections=1, innerDirections=30, precision="mixed",
        orthogonalisation="CGS")) for c in insts]
    gm = hbCase.from_cases(mixed, D)
    o = OracleHB(insts, D)
    hm = gm.pseudo_steps(4)
    ho = o.pseudo_steps(4)
    for J in range(len(insts)):
        fm, fo = gm.fields(J), o.fields(J)
        errs = {k: rel_err(fm[k], fo[k]) for k in ("rho", "U", "T", "p")}
        print(f"instant {J}: mixed vs FP64 oracle after 4 steps {errs}")
        assert max(errs.values()) < 1e-3
    assert np.all(hm[-1][:, :2] < 1.5 * ho[-1][:, :2])


def test_fourier_reconstruction_of_the_instances():
    """reconstructHBFoam on the device: at an instant the fields of that
    instance, at any time sum_m w_m field_m in instance order (the same
    sums on the host, bit for bit)."""
    from icsfoam_amd.solver import hb_reconstruct_weights
    insts, D, t = cases.case_hb(n=(8, 6, 5), n_harmonics=3, flux="ROE")
    h = hbCase.from_cases(insts, D)
    h.pseudo_steps(2)
    om = [2 * np.pi * 100.0]
    F = [h.fields(J) for J in range(len(t))]
    for J in (0, 3):
        r = h.reconstruct(hb_reconstruct_weights(om, [3], t, t[J]))
        for k in ("rho", "U", "T", "p"):
            assert rel_err(r[k], F[J][k]) < 1e-13, (J, k)
    w = hb_reconstruct_weights(om, [3], t, 0.00377)
    r = h.reconstruct(w)
    for k in ("rho", "U", "T", "p"):
        ref = np.zeros_like(F[0][k])
        for J in range(len(t)):
            ref = ref + w[J] * F[J][k]
        np.testing.assert_array_equal(r[k], ref)
