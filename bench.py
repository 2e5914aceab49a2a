"""Benchmark: implicit pseudo-steps/s of the coupled solver at 4.1 M cells.

Workload (default): 160^3 hex box (4 096 000 cells, 12 211 200 internal
faces), node jitter +-10%, random initial cell permutation then RCM then
colour-major numbering; viscous Roe flux (mu = 1.8e-5, Pr = 0.72), Mach
0.5 -> 0.9 ramp, planar tanh pressure jump (ratio 2), 2% smooth noise;
local pseudo-time CFL 50.  Two forms of the same pseudo-step:
  fp64  : GMRES(30) (left LU-SGS, one cycle) with FP64 Jacobian, factors,
          vectors and basis - the reference's precision throughout, with
          classical Gram-Schmidt (the reference orthogonalises with
          modified Gram-Schmidt; --orth MGS selects it);
  mixed : FP64 residual/update and FP64 outer FGMRES(1) around an FP32
          inner GMRES(30) with FP32 LU-SGS on the FP32 Jacobian.
`value` is the mixed form's pseudo-steps/s; both forms are reported.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 160]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_case(n: int, form: str, orth: str):
    from icsfoam_amd import cases
    from icsfoam_amd.mesh import prepare_case

    if form == "fp64":
        kw = dict(solver="GMRES", nDirections=30, precision="double")
    else:
        kw = dict(solver="FGMRES", nDirections=1, precision="mixed", innerDirections=30)
    c = cases.case_box(n=(n, n, n), lengths=(1.0, 1.0, 1.0), jitter=0.1, seed=7, flux="ROE", mu=1.8e-5,
                       mach=(0.5, 0.9), p_jump=2.0, noise=0.02, co=50.0, orthogonalisation=orth, **kw)
    return c


def prepared(n: int):
    from icsfoam_amd.mesh import prepare_case

    t0 = time.time()
    log(f"[bench] generating the {n}^3 case ...")
    base = make_case(n, "fp64", "CGS")
    log(f"[bench] mesh {n}^3 generated in {time.time() - t0:.1f}s")
    t0 = time.time()
    log("[bench] permute + RCM + colour-major ...")
    base = prepare_case(base, permute_seed=2024)
    log(f"[bench] permute + RCM + colour-major in {time.time() - t0:.1f}s")
    return base


def with_form(base, n, form, orth):
    import dataclasses

    c = make_case(4, form, orth)  # dictionaries only
    return dataclasses.replace(base, fvSolution=c.fvSolution)


def algorithmic_bytes(nC, nI, p, q, m=30):
    """Per-launch algorithmic bytes of the block SpMV: every off-diagonal
    block (2 per internal face) and diagonal block read once, column index
    per block slot, x read once per cell, y written once; the m Arnoldi
    launches of a GMRES(m) cycle also write the normalised basis vector
    (m of the m + 2 launches per pseudo-step), averaged in."""
    return (2 * nI + nC) * 25 * p + 2 * nI * 4 + nC * 5 * q * 2 + m / (m + 2) * nC * 5 * q


NGEN = 10  # generator values per off-diagonal block (jacobianStorage generators)


def algorithmic_bytes_fused(nC, nI, p, q, gen=True):
    """Per-group algorithmic bytes of the fused Arnoldi operator (one group
    per Arnoldi step: SpMV of the colour-0 rows; SpMV + forward sweep +
    first backward level of the colour-1 rows; backward sweep of the
    colour-0 rows).  Off-diagonal blocks (10 generators each with generator
    storage, else 25 entries): every block of a colour-0 row twice, of a
    colour-1 row once; diagonal blocks once; column indices per block read;
    vectors: x twice, y0 once more by the colour-1 rows, y once by the last
    pass, writes of y, the basis vector and the final colour-0 rows; rD
    twice for colour 0 (n0 = n1 = nC/2 on these meshes)."""
    per = NGEN if gen else 25
    return (3 * nI * per + nC * 25) * p + 3 * nI * 4 + 30 * q * nC + 12 * nC


def residual_kernel_bytes(nC, nI, nB, p=8, ng=7):
    """Algorithmic bytes per launch of the residual and Jacobian kernels of
    the viscous second-order workload (each array element the kernel needs
    counted once: a neighbour's row gathered by several threads is one read
    of that row; face records are read by both cells of their face):

    k_grad<ng>: per cell the field row 64, e, V 8 each, C 24, row length,
    lower count and offset 12 read, the limited gradients 24 ng and the
    viscous gradient row 96 written; per slot (2 nI) column 4 and slot
    geometry Sf, w, Cf 56; per boundary face 128 (face index, Sf, Cf, kind,
    boundary row, T).
    k_face_terms: per internal face owner, neighbour 8, Sf, magSf, Cf, w,
    nonOrthDeltaCoeffs, correction vector, deltaCoeffs 112 read and the
    192 B record written; per cell its rows gathered once: fields 64,
    gradients 192, viscous gradients 96, C 24, T 8, e 8; per boundary face
    356 (geometry, kind and condition types, boundary row, T, record,
    boundary mass flux).
    k_cell_jacobian<p, cell sums>: per slot the record 192, the slot's Sf
    and magSf 32, the face index 4 read and 10 generators of p bytes
    written; per cell offsets, lower count, V, pseudoCo 24 read and the
    diagonal block 25 p, rD, b (5), rDT 56 written; per boundary face 390
    (record, geometry, kinds, boundary row and T, the owner's row and T).
    k_cell_residual: per slot the record and the face index 196; per cell
    24 read, 48 written; per boundary face 196."""
    return {
        "k_grad": nC * (64 + 8 + 8 + 24 + 12 + 24 * ng + 96) + 2 * nI * (4 + 56) + nB * 128,
        "k_face_terms": nI * (112 + 192) + nC * (64 + 192 + 96 + 24 + 8 + 8) + nB * 356,
        "k_cell_jacobian": 2 * nI * (192 + 32 + 4 + NGEN * p) + nC * (24 + 25 * p + 56) + nB * 390,
        "k_cell_residual": 2 * nI * 196 + nC * (24 + 48) + nB * 196,
    }


def step_bytes(nC, nI, nB, form, m=30):
    """Algorithmic HBM bytes of one pseudo-step (one GMRES/FGMRES cycle of
    m Arnoldi steps), summed over the kernels' per-launch figures: the
    residual and Jacobian kernels as in residual_kernel_bytes (gradients,
    face terms, cell sums and Jacobian rows in one pass over the face
    records, generator storage); per Arnoldi step the fused operator group
    and the CGS projection (w and j + 1 basis vectors read, then w, j + 1
    basis vectors read and w written: 5 q nC (2 j + 5)); per cycle the
    normalisation pass (average, one SpMV, norm factors), the solution
    update (m + 1 vectors) and the final residual SpMV; the state update
    5 * 8 * 4 nC.  FP64 form: p = q = 8.  Mixed form: the inner GMRES(m) with
    p = q = 4 plus the FP64 outer FGMRES(1) (two SpMVs with FP32 blocks, a
    two-vector update)."""
    kb = residual_kernel_bytes(nC, nI, nB, 8 if form == "fp64" else 4)
    res = kb["k_grad"] + kb["k_face_terms"]
    jac = kb["k_cell_jacobian"]
    fused = lambda p, q: algorithmic_bytes_fused(nC, nI, p, q)
    spmv = lambda p, q: (2 * nI * NGEN + nC * 25) * p + 2 * nI * 4 + nC * 5 * q * 2
    orth = lambda q: sum(5 * q * nC * (2 * j + 5) for j in range(m))
    if form == "fp64":
        p = q = 8
        cycle = m * fused(p, q) + orth(q) + 3 * spmv(p, q) + 5 * q * nC * (m + 2)
    else:
        inner = m * fused(4, 4) + orth(4) + 5 * 4 * nC * (m + 2)
        outer = 3 * spmv(4, 8) + 5 * 8 * nC * 6
        cycle = inner + outer
    return res + jac + cycle + 5 * 8 * 4 * nC


def _class_frac(nC, nI, nB, form, r, cls):
    """HBM fraction of breakdown class 0 (gradients + face terms) or 1
    (cell sums + Jacobian rows) of one profiled pseudo-step."""
    kb = residual_kernel_bytes(nC, nI, nB, 8 if form == "fp64" else 4)
    b = kb["k_grad"] + kb["k_face_terms"] if cls == 0 else kb["k_cell_jacobian"]
    ms = r["breakdown_ms"][cls] / max(r["breakdown_count"][cls], 1)
    return b / (ms * 1e-3) / HBM_PEAK if ms > 0 else 0.0


def measured_traffic(n, form="fp64"):
    """HBM bytes per fused Arnoldi-operator group (FP64 or mixed form) from
    the committed FETCH_SIZE/WRITE_SIZE profile of the 160^3 workload (None
    otherwise)."""
    path = os.path.join(ROOT, "profiles", "arnoldi_operator_160cube_pmc.txt")
    if n != 160 or not os.path.exists(path):
        return None
    key = "traffic_bytes_per_group:" if form == "fp64" else "traffic_bytes_per_group_fp32:"
    for line in open(path):
        if line.startswith(key):
            return int(line.split(":")[1])
    return None


def time_form(case, steps, warmup, label):
    from icsfoam_amd import dbnsCase

    t0 = time.time()
    g = dbnsCase.from_case(case)
    log(f"[bench] {label}: case set up in {time.time() - t0:.1f}s")
    # the first two steps run eagerly; the third is captured as a HIP graph
    # (SER state on the device, one GMRES cycle) and later steps replay it
    g.pseudo_steps(warmup, residuals=False)
    g.synchronize()
    t0 = time.perf_counter()
    g.pseudo_steps(steps, residuals=False)
    g.synchronize()
    dt = time.perf_counter() - t0
    # the same number of steps eagerly with HIP events around every
    # Arnoldi-operator group (class 2) on the solver stream: the roofline
    g.timing(1 << 2)
    g.synchronize()
    t0 = time.perf_counter()
    g.pseudo_steps(steps, residuals=False)
    g.synchronize()
    dte = time.perf_counter() - t0
    ms, cnt = g.timing_read()
    g.timing(0)
    res = g.pseudo_steps(1, residuals=True)[0]
    log(f"[bench] {label}: {steps} steps in {dt:.3f}s -> {steps / dt:.3f} steps/s (graph replay), "
        f"{steps / dte:.3f} steps/s eager with events; operator {cnt[2]} groups avg {ms[2] / max(cnt[2], 1):.3f} ms; "
        f"final res {res[:5]}")
    # per-class breakdown from one extra profiled step
    g.timing(0x3F)
    g.pseudo_steps(1, residuals=False)
    bms, bcnt = g.timing_read()
    g.timing(0)
    fused = g.mesh.nFwdLevels == 2 and g.mesh.nBwdLevels == 2
    tuning = g.tuning()
    del g
    return dict(steps_per_s=steps / dt, ms_per_step=1e3 * dt / steps, eager_steps_per_s=steps / dte, tuning=tuning,
                spmv_ms=ms[2] / max(cnt[2], 1), fused=fused,
                spmv_count=int(cnt[2]), breakdown_ms=[round(float(x), 3) for x in bms],
                breakdown_count=[int(x) for x in bcnt])


def _oracle_rank(job):
    """One CPU rank of the baseline (runs in a child process that loads only
    the oracle library, never the HIP one): `warmup` untimed pseudo-steps,
    then `steps` timed ones of the prepared case; returns (cells, per-step
    seconds)."""
    case, steps, warmup, lib_path = job
    from oracle import oracle as O

    O.use_library(lib_path)
    o = O.OracleCase(case)
    o.pseudo_steps(warmup)
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        o.pseudo_steps(1)
        ts.append(time.perf_counter() - t0)
    return case.mesh.n_cells, ts


def cpu_baseline(n_sample: int, n_full: int, steps: int = 5, warmup: int = 2, ranks: int = 0):
    """The oracle (sequential C++ FP64 restatement of the reference path,
    timing build) on the host cores, in the reference's own parallel model:
    one single-threaded rank per core, each owning its share of the cells
    (dbnsFoam under mpirun, per-rank LU-SGS).  One rank alone gives the
    1-core figure; `ranks` ranks run concurrently (default: the host's
    cores, at most 16, the CPU share of a one-GPU host) for the all-core
    figure.
    Each rank runs the n_sample^3 box of the same workload (a bounded sample
    of the 160^3 case's per-rank share), `warmup` untimed steps, then the
    median of `steps` timed steps; steps/s of the full case = aggregate
    cells per second / full cell count (the work per step is linear in
    cells).  The oracle runs the reference's orthogonalisation (modified
    Gram-Schmidt)."""
    import concurrent.futures as cf
    import multiprocessing as mp

    from oracle.build import FAST_FLAGS, FAST_LIB
    from oracle.build import build as build_oracle
    from icsfoam_amd.mesh import prepare_case

    if not os.path.exists(FAST_LIB):
        build_oracle(fast=True)
    c = prepare_case(make_case(n_sample, "fp64", "MGS"), permute_seed=2024)
    ranks = ranks or min(16, os.cpu_count() or 1)
    ctx = mp.get_context("spawn")
    t_wall = time.perf_counter()
    with cf.ProcessPoolExecutor(max_workers=1, mp_context=ctx) as ex:
        cells1, ts1 = ex.submit(_oracle_rank, (c, steps, warmup, FAST_LIB)).result()
    one = float(np.median(ts1))
    log(f"[bench] cpu baseline: 1 rank, {n_sample}^3, median {one:.3f} s/step "
        f"({time.perf_counter() - t_wall:.0f} s)")
    t_wall = time.perf_counter()
    with cf.ProcessPoolExecutor(max_workers=ranks, mp_context=ctx) as ex:
        res = list(ex.map(_oracle_rank, [(c, steps, warmup, FAST_LIB)] * ranks))
    rate_all = sum(cells / float(np.median(ts)) for cells, ts in res)  # cells * steps / s
    med_all = float(np.median([np.median(ts) for _, ts in res]))
    log(f"[bench] cpu baseline: {ranks} concurrent ranks, median {med_all:.3f} s/step per rank "
        f"({time.perf_counter() - t_wall:.0f} s)")
    cpu = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    full = float(n_full) ** 3
    return dict(value=rate_all / full, unit="pseudo-steps/s", cores=ranks, kind="port",
                one_core_value=cells1 / one / full,
                sample=f"oracle (sequential C++ FP64 restatement of the reference path: MGS GMRES(30), "
                       f"sequential LU-SGS; g++ {' '.join(FAST_FLAGS)}) as {ranks} concurrent single-threaded "
                       f"ranks, one per host core, each on the {n_sample}^3 box of the same workload "
                       f"({cells1} cells; the reference's one-rank-per-core model); median of {steps} timed "
                       f"steps after {warmup} warm-up steps per rank: {med_all:.2f} s/step per rank "
                       f"concurrently, {one:.2f} s/step for one rank alone (one_core_value); steps/s of the "
                       f"{n_full}^3 case = aggregate cells/s / {int(full)}; host CPU: {cpu}; "
                       f"nproc {os.cpu_count()}")


def bench_bump(steps=50):
    """The reference's CPU-runnable case: 2-D transonic bump, 200x50x1
    cells, inviscid AUSM+up with linearUpwind + cellLimited<Venkatakrishnan>,
    GMRES(30) (MGS, the reference's order) with LU-SGS, local CFL 50, FP64,
    cells renumbered RCM + colour-major (renumberMesh's role; in blockMesh
    order the sequential LU-SGS sweep has 250 dependency levels, i.e. 500
    launches per application); `steps` pseudo-steps timed on the GPU (one
    residual read per step as the reference's loop does, steps from the
    third on replayed from the captured graph) and, beside them, the same
    steps of the oracle on one host core (the whole run, no sampling).  At
    10 000 cells the GPU step is launch-bound."""
    from icsfoam_amd import cases, dbnsCase
    from icsfoam_amd.mesh import prepare_case
    from oracle import oracle as O
    from oracle.build import FAST_LIB
    from oracle.build import build as build_oracle

    # the circularArcBump tutorial's inflow patch (totalPressure /
    # pressureInletOutletVelocity / totalTemperature at the freestream's
    # total state); with a fixed inflow velocity the channel chokes
    c = prepare_case(cases.case_bump(200, 50, seed=1, bcs="tutorial"))
    g = dbnsCase.from_case(c)
    g.pseudo_steps(3)
    g.synchronize()
    t0 = time.perf_counter()
    hg = g.pseudo_steps(steps)
    g.synchronize()
    dt = time.perf_counter() - t0
    log(f"[bench] bump GPU: {steps} steps in {dt:.3f}s; final residuals {hg[-1, :5]}")
    del g
    if not os.path.exists(FAST_LIB):
        build_oracle(fast=True)
    O.use_library(FAST_LIB)
    o = O.OracleCase(c)
    o.pseudo_steps(3)
    t0 = time.perf_counter()
    o.pseudo_steps(steps)
    dtc = time.perf_counter() - t0
    log(f"[bench] bump oracle: {steps} steps in {dtc:.3f}s")
    return {"metric": "implicit pseudo-steps/s, 2-D bump 200x50 (FP64)", "value": round(steps / dt, 3),
            "unit": "pseudo-steps/s", "n_gpus": 1, "steps": steps, "warmup": 3, "ms_per_step": round(1e3 * dt / steps, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (bump channel, M 0.675, 1% seeded 8-mode perturbation)",
            "config": {"workload": "200x50x1 bump (10000 cells), AUSM+up, GMRES(30) MGS + LU-SGS, CFL 50, "
                                   "total-pressure inflow patch, RCM + colour-major"},
            "cpu_baseline": {"value": round(steps / dtc, 3), "unit": "pseudo-steps/s", "cores": 1, "kind": "port",
                             "sample": f"oracle (fast build) on the same case, the same {steps} steps"}}


def bench_hb(steps, warmup, frame="none", mu=0.0):
    """Harmonic-balance pseudo-steps/s on a 64x64x32 annular sector with
    rotational cyclic theta patches and 7 instants (3 harmonics of 100 Hz),
    Roe inviscid, LU-SGS.  value: the mixed form (outer FGMRES(1) on the
    FP64 coupled vectors, FP32 HB Jacobian and FP32 inner GMRES(30) with CGS,
    hipGraph-captured steps); the FP64 form (GMRES(30), CGS) beside it.
    frame: "phase-lag" makes the cyclic pair phase-lagged (inter-blade
    phase angle 2 pi/9, cylCoords); "mrf" lets the whole sector rotate about
    its axis at 300 rad/s (every physical patch non-rotating).  mu > 0:
    viscous (the viscous terms across the cyclic faces included)."""
    from icsfoam_amd import cases
    from icsfoam_amd.mesh import prepare_case
    from icsfoam_amd.solver import hbCase

    t0 = time.time()
    insts, D, t = cases.case_hb(n=(64, 64, 32), n_harmonics=3, flux="ROE", orthogonalisation="CGS", mu=mu)
    m = insts[0].mesh
    if frame == "phase-lag":
        lo = next(i for i, p in enumerate(m.patches) if p.kind == cases.PATCH_CYCLIC)
        pair = cases.PhaseLagPair(owner=lo, ibpa=2 * np.pi / 9, frequencies=np.array([2 * np.pi * 100.0]),
                                  harmonics=np.array([3]), instants=t, cyl_coords=True)
        insts = [dataclasses.replace(c, phase_lag=pair) for c in insts]
    elif frame == "mrf":
        z = cases.MRFZone(cells=np.ones(m.n_cells, bool), origin=np.zeros(3), axis=np.array([0.0, 0.0, 1.0]),
                          omega=300.0, non_rotating=[i for i, p in enumerate(m.patches) if p.kind != cases.PATCH_CYCLIC])
        insts = [dataclasses.replace(c, mrf=z) for c in insts]
    insts = [prepare_case(c, permute_seed=2024) for c in insts]
    log(f"[bench] HB case (7 instants, {insts[0].mesh.n_cells} cells) prepared in {time.time() - t0:.1f}s")
    forms = {"fp64": cases.solution_dict(co=20.0, orthogonalisation="CGS"),
             "mixed": cases.solution_dict(co=20.0, solver="FGMRES", nDirections=1, innerDirections=30,
                                          precision="mixed", orthogonalisation="CGS")}
    rate = {}
    for form, sol in forms.items():
        g = hbCase.from_cases([dataclasses.replace(c, fvSolution=sol) for c in insts], D)
        g.pseudo_steps(max(warmup, 3), residuals=False)
        g.synchronize()
        t0 = time.perf_counter()
        g.pseudo_steps(steps, residuals=False)
        g.synchronize()
        dt = time.perf_counter() - t0
        res = g.pseudo_steps(1)[0]
        rate[form] = steps / dt
        log(f"[bench] HB {form}: {steps} steps in {dt:.3f}s; instant-0 residuals {res[0, :5]}")
        del g
    nC = insts[0].mesh.n_cells
    fdesc = {"none": "", "phase-lag": ", phase-lagged cyclic pair (IBPA 2 pi/9, cylCoords)",
             "mrf": ", whole sector rotating at 300 rad/s (MRF)"}[frame]
    return {"metric": "harmonic-balance pseudo-steps/s (7 instants)", "value": round(rate["mixed"], 4),
            "unit": "pseudo-steps/s", "n_gpus": 1, "steps": steps, "warmup": max(warmup, 3),
            "ms_per_step": round(1e3 / rate["mixed"], 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64 residual/outer + f32 Jacobian/inner",
            "data": "synthetic (seeded annular sector, periodic inflow)",
            "config": {"workload": f"64x64x32 annular sector with rotational cyclic patches ({nC} cells) x 7 "
                                   f"instants ({7 * nC} block rows), Roe "
                                   f"{'viscous (mu %g)' % mu if mu > 0 else 'inviscid'}, LU-SGS, hipGraph steps{fdesc}",
                       "mixed": "FGMRES(1) FP64 outer + FP32 GMRES(30, CGS) inner",
                       "fp64": "GMRES(30, CGS)", "fp64_steps_per_s": round(rate["fp64"], 4),
                       "mixed_over_fp64": round(rate["mixed"] / rate["fp64"], 3)}}


def bench_vjp(calls, flux="AUSMPlusUp"):
    """Residual VJP (inviscid AUSM+up, 2nd order) at 100^3: ms per call with
    device-resident W and v (torch tensors)."""
    import torch

    from icsfoam_amd import cases, dbnsCase
    from icsfoam_amd.adjoint import residual_device, residual_vjp_device
    from icsfoam_amd.mesh import prepare_case

    c = cases.case_box(n=(100, 100, 100), lengths=(1.0, 1.0, 1.0), jitter=0.1, seed=7, flux=flux, mu=0.0)
    c = prepare_case(c, permute_seed=2024)
    g = dbnsCase.from_case(c)
    W = torch.tensor(g.fields()["W"], device="cuda")
    v = torch.randn(W.shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    for _ in range(3):
        residual_vjp_device(g, W, v)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(calls):
        residual_vjp_device(g, W, v)
    dt = (time.perf_counter() - t0) / calls
    t0 = time.perf_counter()
    for _ in range(calls):
        residual_device(g, W)
    dr = (time.perf_counter() - t0) / calls
    log(f"[bench] VJP {1e3 * dt:.3f} ms/call, residual {1e3 * dr:.3f} ms/call")
    return {"metric": "residual VJP calls/s at 1M cells (AUSM+up 2nd order)", "value": round(1.0 / dt, 3),
            "unit": "calls/s", "n_gpus": 1, "steps": calls, "warmup": 3, "ms_per_step": round(1e3 * dt, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded jittered 100^3 box)",
            "config": {"workload": f"100^3 jittered box (1000000 cells), inviscid {flux}, linearUpwind + "
                                   "cellLimited<Venkatakrishnan>; gW, gSf, gCf per call",
                       "residual_ms": round(1e3 * dr, 3)}}


def bench_c5_vjp(n, max_steps, calls=10, n_dirs=32, chunk=25, drop=1e-8, time_limit=300.0):
    """FP64 residual VJP of the n^3 headline workload (viscous Roe) on its
    converged state: the mixed form run to a `drop` reduction of ||R V||_2
    (or the time limit), then ms per VJP call with device-resident W and v,
    and the finite-difference gate of icsfoam_amd.adjoint.fd_gate over
    n_dirs directions on that state with 1e-3 per-cell noise."""
    import torch

    from icsfoam_amd import cases, dbnsCase
    from icsfoam_amd.adjoint import fd_gate, residual_device, residual_vjp_device

    base = prepared(n)
    g = dbnsCase.from_case(with_form(base, n, "mixed", "CGS"))
    b, _ = g.residual()
    r0 = float(np.sqrt(np.sum(b * b)))
    steps, rel = 0, 1.0
    t0 = time.perf_counter()
    while steps < max_steps and rel >= drop and time.perf_counter() - t0 < time_limit:
        g.pseudo_steps(chunk, residuals=False)
        steps += chunk
        b, _ = g.residual()
        rel = float(np.sqrt(np.sum(b * b))) / r0
    t_conv = time.perf_counter() - t0
    log(f"[bench] {n}^3 mixed: {steps} steps in {t_conv:.1f}s, ||R||/||R0|| = {rel:.3e}")
    W = torch.tensor(g.fields()["W"], device="cuda")
    v = torch.randn(W.shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    for _ in range(2):
        residual_vjp_device(g, W, v, geometry=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(calls):
        residual_vjp_device(g, W, v, geometry=False)
    dt = (time.perf_counter() - t0) / calls
    t0 = time.perf_counter()
    for _ in range(calls):
        residual_device(g, W)
    dr = (time.perf_counter() - t0) / calls
    log(f"[bench] VJP {1e3 * dt:.3f} ms/call, residual {1e3 * dr:.3f} ms/call; FD gate ...")
    # the converged state is the uniform stream (limiter ties in every cell,
    # cross-stream velocity at rounding level): the residual has no
    # derivative there, so the finite-difference gate runs on it with 1e-3
    # per-cell noise, as the adjoint tests do
    rng = np.random.default_rng(11)
    Wn = W.cpu().numpy().copy()
    Wn[2:] += 1e-3 * np.abs(Wn[2:]).max() * rng.standard_normal(Wn[2:].shape)
    Wn[:2] *= 1 + 1e-3 * rng.standard_normal(Wn[:2].shape)
    X = torch.tensor(cases.approx_cell_centres(base.mesh), device="cuda")
    fdr = fd_gate(g, torch.tensor(Wn, device="cuda"), X, n_dirs=n_dirs)
    log(f"[bench] FD gate: {fdr}")
    return {"metric": f"residual VJP calls/s at {n}^3 on the converged state (viscous Roe, FP64)",
            "value": round(1.0 / dt, 3), "unit": "calls/s", "n_gpus": 1, "steps": calls, "warmup": 2,
            "ms_per_step": round(1e3 * dt, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic (seeded jittered box, converged by the solver)",
            "config": {"workload": f"{n}^3 jittered hex box, viscous Roe, linearUpwind + "
                                   f"cellLimited<Venkatakrishnan>; gW per call",
                       "converge_steps": steps, "converge_seconds": round(t_conv, 1), "converged_drop": rel,
                       "residual_ms": round(1e3 * dr, 3),
                       "fd_gate_on_noised_state": {k: (v if k != "unexplained" else len(v))
                                                   for k, v in fdr.items()}}}


def bench_converge(n, max_steps, chunk=25, drop=1e-8, time_limit=400.0, forms=("fp64", "mixed")):
    """Time to a `drop` reduction of ||R V||_2 (all five components, the
    norm dbnsFoam.C:136-137 prints for rho) for the FP64 and the mixed form
    of the n^3 workload."""
    base = prepared(n)
    out, Ws = {}, {}
    for form in forms:
        from icsfoam_amd import dbnsCase

        g = dbnsCase.from_case(with_form(base, n, form, "CGS"))
        b, _ = g.residual()
        r0 = float(np.sqrt(np.sum(b * b)))
        steps, hist = 0, [(0, 0.0, 1.0)]
        g.synchronize()
        t0 = time.perf_counter()
        while steps < max_steps:
            g.pseudo_steps(chunk, residuals=False)
            steps += chunk
            b, _ = g.residual()
            rel = float(np.sqrt(np.sum(b * b))) / r0
            el = time.perf_counter() - t0
            hist.append((steps, round(el, 3), rel))
            log(f"[bench] {form}: {steps} steps {el:.1f}s  ||R||/||R0|| = {rel:.3e}")
            if rel < drop or not np.isfinite(rel) or el > time_limit:
                break
        out[form] = dict(steps=steps, seconds=round(time.perf_counter() - t0, 3), drop=hist[-1][2],
                         history=hist[::max(1, len(hist) // 20)])
        Ws[form] = g.fields()["W"]
        del g
    if len(Ws) == 2:
        # converged states of the two forms: difference per conserved
        # component (reference order rho, rhoE, rhoUx, rhoUy, rhoUz),
        # max-norm and 2-norm, relative to the norm of rho, rhoE and of the
        # momentum vector (a cross-stream momentum component on its own sits
        # at rounding level in the converged stream)
        (fa, Wa), (fb, Wb) = Ws.items()
        d = Wa - Wb
        ref = [Wb[0], Wb[1]] + [np.sqrt(Wb[2] ** 2 + Wb[3] ** 2 + Wb[4] ** 2)] * 3
        out[f"W_rel_diff_{fa}_vs_{fb}"] = {
            name: {"max": float(np.abs(d[k]).max() / np.abs(ref[k]).max()),
                   "l2": float(np.linalg.norm(d[k]) / np.linalg.norm(ref[k]))}
            for k, name in enumerate(("rho", "rhoE", "rhoUx", "rhoUy", "rhoUz"))}
        log(f"[bench] converged W, {fa} vs {fb}: {out[f'W_rel_diff_{fa}_vs_{fb}']}")
    head = out[forms[-1]]
    return {"metric": f"time to a {drop:g} drop of ||R V||_2 at {n}^3 ({' vs '.join(forms)})",
            "value": head["seconds"], "unit": "s", "n_gpus": 1, "steps": head["steps"],
            "warmup": 0, "ms_per_step": round(1e3 * head["seconds"] / max(head["steps"], 1), 3),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64 residual/outer + f32 Jacobian/inner" if forms[-1] == "mixed" else "f64",
            "data": "synthetic",
            "config": dict({"workload": f"{n}^3 jittered hex box, viscous Roe, local CFL 50"}, **out)}


def bench_hb_converge(max_steps, chunk=25, drop=1e-8, time_limit=300.0):
    """Harmonic balance (64x64x32 cyclic sector, 7 instants): time to a
    `drop` reduction of the coupled residual ||R V||_2 (all instances, all
    five components, spectral source included) for the FP64 and the mixed
    form of the benchmark's HB case."""
    from icsfoam_amd import cases
    from icsfoam_amd.mesh import prepare_case
    from icsfoam_amd.solver import hbCase

    insts, D, _ = cases.case_hb(n=(64, 64, 32), n_harmonics=3, flux="ROE", orthogonalisation="CGS")
    insts = [prepare_case(c, permute_seed=2024) for c in insts]
    forms = {"fp64": cases.solution_dict(co=20.0, orthogonalisation="CGS"),
             "mixed": cases.solution_dict(co=20.0, solver="FGMRES", nDirections=1, innerDirections=30,
                                          precision="mixed", orthogonalisation="CGS")}
    out = {}
    for form, sol in forms.items():
        g = hbCase.from_cases([dataclasses.replace(c, fvSolution=sol) for c in insts], D)
        b = g.residual()
        r0 = float(np.sqrt(np.sum(b * b)))
        steps, hist = 0, [(0, 0.0, 1.0)]
        g.synchronize()
        t0 = time.perf_counter()
        while steps < max_steps:
            g.pseudo_steps(chunk, residuals=False)
            steps += chunk
            b = g.residual()
            rel = float(np.sqrt(np.sum(b * b))) / r0
            el = time.perf_counter() - t0
            hist.append((steps, round(el, 3), rel))
            log(f"[bench] HB {form}: {steps} steps {el:.1f}s  ||R||/||R0|| = {rel:.3e}")
            if rel < drop or not np.isfinite(rel) or el > time_limit:
                break
        out[form] = dict(steps=steps, seconds=round(time.perf_counter() - t0, 3), drop=hist[-1][2],
                         history=hist[::max(1, len(hist) // 20)])
        del g
    head = out["mixed"]
    return {"metric": f"time to a {drop:g} drop of the coupled HB residual (fp64 vs mixed)",
            "value": head["seconds"], "unit": "s", "n_gpus": 1, "steps": head["steps"], "warmup": 0,
            "ms_per_step": round(1e3 * head["seconds"] / max(head["steps"], 1), 3), "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64 residual/outer + f32 Jacobian/inner",
            "data": "synthetic (seeded annular sector, periodic inflow)",
            "config": dict({"workload": "64x64x32 annular sector with rotational cyclic patches x 7 instants, "
                                        "Roe inviscid, CFL 20"}, **out)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=160)
    ap.add_argument("--orth", default="CGS", choices=["CGS", "MGS"])
    ap.add_argument("--cpu-sample", type=int, default=48)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed", help="converge: forms to run, in order")
    ap.add_argument("--frame", default="none", choices=["none", "phase-lag", "mrf"],
                    help="hb: a phase-lagged cyclic pair or a rotating sector")
    ap.add_argument("--mu", type=float, default=0.0, help="hb: viscosity (0: inviscid)")
    ap.add_argument("--flux", default="AUSMPlusUp", choices=["AUSMPlusUp", "ROE", "HLLC"],
                    help="vjp: the flux scheme")
    ap.add_argument("--time-limit", type=float, default=400.0, help="converge: seconds per form")
    ap.add_argument("--workload", default="c5",
                    choices=["c5", "hb", "vjp", "c5vjp", "converge", "bump", "hbconverge"],
                    help="c5: the headline 160^3 pseudo-step (default); hb: 7-instant harmonic balance; "
                         "vjp: residual adjoint at 100^3; bump: the 200x50 bump, 50 steps, GPU and oracle; "
                         "hbconverge: HB time to a 1e-8 drop; c5vjp: FP64 VJP of the n^3 workload on its "
                         "converged state, with the finite-difference gate")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs the pseudo-step on exactly one GPU; --gpus {args.gpus} is not "
              "supported (multi-GPU sharding is designed, not built)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "hb":
        print(json.dumps(bench_hb(args.steps, args.warmup, args.frame, args.mu)), flush=True)
        return
    if args.workload == "hbconverge":
        print(json.dumps(bench_hb_converge(max(args.steps, 25), time_limit=args.time_limit)), flush=True)
        return
    if args.workload == "bump":
        print(json.dumps(bench_bump(max(args.steps, 50))), flush=True)
        return
    if args.workload == "vjp":
        print(json.dumps(bench_vjp(max(args.steps, 10), args.flux)), flush=True)
        return
    if args.workload == "c5vjp":
        print(json.dumps(bench_c5_vjp(args.n, 8000, time_limit=args.time_limit)), flush=True)
        return
    if args.workload == "converge":
        print(json.dumps(bench_converge(args.n, max(args.steps, 25), time_limit=args.time_limit,
                                        forms=tuple(args.forms.split(",")))), flush=True)
        return
    n = args.n
    base = prepared(n)
    nC, nI = base.mesh.n_cells, base.mesh.n_internal
    nB = base.mesh.n_faces - nI
    # at least 3 untimed steps so that the graph capture (third step) is
    # not inside the timed region
    warmup = max(args.warmup, 3)
    r64 = time_form(with_form(base, n, "fp64", args.orth), args.steps, warmup, "fp64")
    rmx = time_form(with_form(base, n, "mixed", args.orth), args.steps, warmup, "mixed")
    # dominant kernel group: the Arnoldi operator of the FP64 form (FP64
    # blocks and vectors, algorithmic bytes unambiguous): the fused SpMV +
    # LU-SGS group on two-colour numberings, else the block SpMV
    fused = r64["fused"]
    bytes64 = algorithmic_bytes_fused(nC, nI, 8, 8) if fused else algorithmic_bytes(nC, nI, 8, 8)
    ach = bytes64 / (r64["spmv_ms"] * 1e-3)
    bytes32 = algorithmic_bytes_fused(nC, nI, 4, 4) if fused else algorithmic_bytes(nC, nI, 4, 4)
    ach32 = bytes32 / (rmx["spmv_ms"] * 1e-3)
    cpu = None
    if not args.no_cpu:
        log(f"[bench] cpu baseline on {args.cpu_sample}^3 ...")
        cpu = cpu_baseline(args.cpu_sample, n)
    out = {
        "metric": "implicit pseudo-steps/s at 4.1M cells: FP32-precond vs FP64; HBM GB/s frac",
        "value": round(rmx["steps_per_s"], 4),
        "unit": "pseudo-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warmup,
        "ms_per_step": round(rmx["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 residual/outer + f32 Jacobian/inner",
        "data": "synthetic (seeded jittered 160^3 hex box, Roe viscous state with tanh pressure jump)",
        "config": {"workload": f"{n}^3 jittered hex box ({nC} cells, {nI} internal faces), viscous Roe, "
                               f"local CFL 50, LU-SGS, permuted+RCM+colour-major",
                   "fp64_steps_per_s": round(r64["steps_per_s"], 4),
                   "fp64_ms_per_step": round(r64["ms_per_step"], 3),
                   "mixed_over_fp64": round(rmx["steps_per_s"] / r64["steps_per_s"], 3),
                   "eager_steps_per_s": {"fp64": round(r64["eager_steps_per_s"], 4),
                                         "mixed": round(rmx["eager_steps_per_s"], 4)},
                   "orthogonalisation": args.orth,
                   "breakdown_classes": ["residual up to the face terms", "residual cell sums + Jacobian rows",
                                         "Arnoldi operator (SpMV + LU-SGS)", "preconditioner outside the operator",
                                         "orthogonalisation", "state update, solution combine, norms"],
                   "breakdown_ms_fp64": r64["breakdown_ms"], "breakdown_ms_mixed": rmx["breakdown_ms"],
                   "autotuned": {"fp64": r64["tuning"], "mixed": rmx["tuning"]},
                   "step_hbm_frac": {f: round(step_bytes(nC, nI, nB, f) * r["steps_per_s"] / HBM_PEAK, 4)
                                     for f, r in (("fp64", r64), ("mixed", rmx))},
                   "step_algorithmic_gb": {f: round(step_bytes(nC, nI, nB, f) / 1e9, 2)
                                           for f in ("fp64", "mixed")},
                   # residual (gradients + face terms) and Jacobian (cell sums +
                   # rows) classes of the profiled step: algorithmic bytes of
                   # their kernels / class time, as a fraction of the HBM peak
                   "residual_jacobian_hbm_frac": {
                       f: {"residual": round(_class_frac(nC, nI, nB, f, r, 0), 4),
                           "jacobian": round(_class_frac(nC, nI, nB, f, r, 1), 4)}
                       for f, r in (("fp64", r64), ("mixed", rmx))}},
        "roofline": {"bound": "hbm",
                     "kernel": ("fused Arnoldi operator, FP64 form, off-diagonal blocks as generators: "
                                "k_spmv_cells<double,double,false,true> + k_spmv_cells<double,double,true,true> "
                                "+ k_lusgs_bwd<double,double,true> (one group per Arnoldi step, timed as a "
                                "group)") if fused else
                               "k_spmv<double,double> (block SpMV, FP64 form)",
                     "achieved": round(ach / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK, 4), "traffic": measured_traffic(n) if fused else None,
                     "bytes_per_launch": int(bytes64), "avg_launch_ms": round(r64["spmv_ms"], 4)},
        # the same operator group in the form `value` is quoted on
        "roofline_mixed": {"bound": "hbm",
                           "kernel": "fused Arnoldi operator of the FP32 inner GMRES (FP32 generators and vectors)",
                           "achieved": round(ach32 / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                           "frac": round(ach32 / HBM_PEAK, 4),
                           "traffic": measured_traffic(n, "mixed") if fused else None,
                           "bytes_per_launch": int(bytes32), "avg_launch_ms": round(rmx["spmv_ms"], 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
