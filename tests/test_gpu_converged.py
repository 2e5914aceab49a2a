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
