"""The device-evaluated pseudotimeControl loop (icsf_pseudo_time_loop)
against the oracle's host loop: the same number of steps, the same
convergence verdict, and the end state within 1e-8 (the histories agree to
1e-8, so the test first checks that the deciding ratio is not that close to
the threshold).  Batched graph replays issue steps past the end, which must
leave the state unchanged: run with batch 8 and batch 1 and compare."""
import dataclasses

import numpy as np
import pytest

from icsfoam_amd import cases, dbnsCase
from icsfoam_amd.mesh import prepare_case
from oracle.oracle import OracleCase

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _bump(tol, tolRel, nCorrMin=1):
    c = prepare_case(cases.case_bump(40, 10), permute_seed=11)
    sol = c.fvSolution.replace("pseudoTol 1e-12;", f"pseudoTol {tol!r};").replace(
        "pseudoTolRel 1e-12;", f"pseudoTolRel {tolRel!r}; nPseudoCorrMin {nCorrMin};")
    return dataclasses.replace(c, fvSolution=sol)


@pytest.mark.parametrize("tol,tolRel", [(1e-12, 1e-3), (3e-4, 1e-12)])
def test_loop_matches_oracle(tol, tolRel):
    print(f"[loop tol={tol} tolRel={tolRel}]", flush=True)
    c = _bump(tol, tolRel)
    o = OracleCase(c)
    hist_case = OracleCase(c)
    run_o, conv_o = o.pseudo_time_loop(600)
    assert conv_o
    # margin of the deciding test against the GPU/oracle history difference
    h = np.array([hist_case.pseudo_steps(1)[0][:5] for _ in range(run_o)])
    r1 = h[0]
    last = h[run_o - 2]
    before = h[run_o - 3]
    m_last = min(abs(last.max() / tol - 1), abs((last / r1).max() / tolRel - 1))
    m_before = min(abs(before.max() / tol - 1), abs((before / r1).max() / tolRel - 1))
    assert m_last > 1e-6 and m_before > 1e-6
    for batch in (8, 1):
        g = dbnsCase.from_case(c)
        run_g, conv_g = g.pseudo_time_loop(600, batch=batch)
        assert (run_g, conv_g) == (run_o, conv_o), batch
        fg, fo = g.fields(), o.fields()
        for k in ("rho", "U", "T", "p", "W"):
            assert rel_err(fg[k], fo[k]) < 1e-8, (batch, k)
        assert rel_err(g.pseudoCo(), o.pseudoCo()) < 1e-8


def test_loop_unconverged_runs_max_steps_and_continues():
    c = _bump(1e-300, 1e-300)
    g = dbnsCase.from_case(c)
    o = OracleCase(c)
    assert g.pseudo_time_loop(7, batch=3) == (7, False)
    assert o.pseudo_time_loop(7) == (7, False)
    # later steps continue from there
    hg = g.pseudo_steps(2)
    ho = o.pseudo_steps(2)
    assert rel_err(hg, ho) < 1e-8
    assert rel_err(g.fields()["W"], o.fields()["W"]) < 1e-8


def test_loop_requires_tolerances():
    from icsfoam_amd import IcsfError
    c = prepare_case(cases.case_bump(20, 6), permute_seed=3)
    c = dataclasses.replace(c, fvSolution=c.fvSolution.replace("pseudoTol 1e-12;", ""))
    g = dbnsCase.from_case(c)
    with pytest.raises(IcsfError, match="pseudoTol"):
        g.pseudo_time_loop(5)
