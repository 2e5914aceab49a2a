"""Writes tests/golden/bump_100x25_converged.npz: the oracle (CPU FP64
restatement of the reference path, checker build) run of the transonic
bump channel, 100x25x1 cells, to a 1e-8 drop of ||R V||_2 (all five
components), checked every 25 steps.  The 200x50 bump's schemes and solver
(AUSM+up, linearUpwind + cellLimited<Venkatakrishnan>, GMRES(30) MGS with
LU-SGS, one cycle, local CFL 50), the seeded 1 % Fourier perturbation, the
circularArcBump tutorial's inflow patch (totalPressure /
pressureInletOutletVelocity / totalTemperature at the freestream's total
state) and fixed outlet pressure; cells permuted with seed 11, then RCM +
colour-major.  (With a fixed inflow velocity the channel chokes and the
run diverges after about 200 steps; on the 200x50 mesh the limited
second-order residual stalls near 1e-3, so the converged state is taken on
the coarser mesh.)  Stored: the
number of steps, the converged conserved state W [5][nC] and the residual
history.  Run from the repository root:

    python tests/golden/make_bump_converged.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from icsfoam_amd import cases  # noqa: E402
from icsfoam_amd.mesh import prepare_case  # noqa: E402
from oracle.oracle import OracleCase  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bump_100x25_converged.npz")
DROP, CHUNK, MAX_STEPS = 1e-8, 25, 20000


def case():
    return prepare_case(cases.case_bump(100, 25, bcs="tutorial"), permute_seed=11)


def main():
    o = OracleCase(case())
    b, _ = o.residual()
    r0 = float(np.sqrt(np.sum(b * b)))
    steps, hist = 0, [(0, 1.0)]
    t0 = time.time()
    while steps < MAX_STEPS:
        o.pseudo_steps(CHUNK)
        steps += CHUNK
        b, _ = o.residual()
        rel = float(np.sqrt(np.sum(b * b))) / r0
        hist.append((steps, rel))
        print(f"{steps} steps {time.time() - t0:.0f}s  ||R||/||R0|| = {rel:.3e}", flush=True)
        if rel < DROP:
            break
    np.savez_compressed(OUT, steps=steps, W=o.fields()["W"], history=np.array(hist), drop=DROP, chunk=CHUNK)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
