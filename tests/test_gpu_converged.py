"""Converged-state parity: the HIP path run to a 1e-8 drop of ||R V||_2 on
the transonic bump (100x25, the tutorial's inflow patch, AUSM+up with
limited linearUpwind, GMRES(30) + LU-SGS, local CFL 50) ends within 1e-8
relative of the oracle's converged state (tests/golden/
bump_100x25_converged.npz, written by make_bump_converged.py from the same
inputs).  Steps from the third on are graph replays."""
import os

import numpy as np
import pytest

from icsfoam_amd import dbnsCase

from helpers import bump_tutorial_100x25, rel_err

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bump_100x25_converged.npz")


def test_converged_state_matches_oracle():
    print("[converged bump 100x25]", flush=True)
    gold = np.load(GOLD)
    steps = int(gold["steps"])
    g = dbnsCase.from_case(bump_tutorial_100x25())
    b0, _ = g.residual()
    r0 = np.sqrt(np.sum(b0 * b0))
    g.pseudo_steps(steps, residuals=False)
    b, _ = g.residual()
    drop = np.sqrt(np.sum(b * b)) / r0
    print(f"steps {steps}  GPU drop {drop:.3e}  oracle drop {gold['history'][-1][1]:.3e}", flush=True)
    assert drop < 1.2e-8
    W = g.fields()["W"]
    err = rel_err(W, gold["W"])
    print(f"converged W, GPU vs oracle: {err:.2e} relative", flush=True)
    assert err < 1e-8
    for q in range(5):
        assert rel_err(W[q], gold["W"][q]) < 1e-8, q


def _headline_box(n, form):
    """The 160^3 workload's configuration (jittered box, viscous Roe,
    linearUpwind + cellLimited, Mach 0.5-0.9 stream with a pressure jump,
    local CFL 50) on an n^3 mesh, in the FP64 (GMRES(30), CGS) or the
    mixed form (FGMRES(1) + FP32 GMRES(30))."""
    from icsfoam_amd import cases
    from icsfoam_amd.mesh import prepare_case

    if form == "fp64":
        kw = dict(solver="GMRES", nDirections=30, precision="double")
    else:
        kw = dict(solver="FGMRES", nDirections=1, precision="mixed", innerDirections=30)
    c = cases.case_box(n=(n, n, n), lengths=(1.0, 1.0, 1.0), jitter=0.1, seed=7, flux="ROE", mu=1.8e-5,
                       mach=(0.5, 0.9), p_jump=2.0, noise=0.02, co=50.0, orthogonalisation="CGS", **kw)
    return prepare_case(c, permute_seed=2024)


def test_mixed_and_fp64_forms_converge_to_the_same_fields():
    """The FP32 Jacobian and inner GMRES change the path, not the discrete
    solution: both forms of the headline configuration (24^3), run to a
    1e-11 drop of ||R V||_2, hold the same conserved fields to 1e-8
    relative (rho, rhoE against their own max, each momentum component
    against the max of |rho U|)."""
    n = 24
    print(f"[converged fp64 vs mixed {n}^3]", flush=True)
    Ws = {}
    for form in ("fp64", "mixed"):
        g = dbnsCase.from_case(_headline_box(n, form))
        b0, _ = g.residual()
        r0 = np.sqrt(np.sum(b0 * b0))
        steps, drop = 0, 1.0
        while steps < 4000 and drop >= 1e-11:
            g.pseudo_steps(100, residuals=False)
            steps += 100
            b, _ = g.residual()
            drop = np.sqrt(np.sum(b * b)) / r0
        print(f"{form}: {steps} steps, drop {drop:.3e}", flush=True)
        assert drop < 1e-11
        Ws[form] = g.fields()["W"]
    a, b = Ws["fp64"], Ws["mixed"]
    mom = np.sqrt(a[2] ** 2 + a[3] ** 2 + a[4] ** 2).max()
    errs = [np.abs(a[0] - b[0]).max() / np.abs(a[0]).max(), np.abs(a[1] - b[1]).max() / np.abs(a[1]).max()]
    errs += [np.abs(a[q] - b[q]).max() / mom for q in (2, 3, 4)]
    print("converged W, fp64 vs mixed (rho, rhoE, rhoUx, rhoUy, rhoUz): "
          + ", ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < 1e-8
