"""ORACLE — test infrastructure only.

ctypes wrapper of oracle/liboracle.so, the CPU FP64 restatement of the
reference path (see icsf_oracle.cpp for the reference files it follows).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
as the checker / baseline; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_vp, _i = C.c_void_p, C.c_int
_SIG = {
    "orc_mesh_new": (_vp, [_i, _vp, _i, _vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "orc_mesh_cyclic": (None, [_vp, _vp, _vp, _vp, _vp]),
    "orc_rcm_coupled": (None, [_i, _i, _vp, _vp, _i, _vp, _vp, _vp]),
    "orc_get_interfaces": (None, [_vp, _vp]),
    "orc_mesh_free": (None, [_vp]),
    "orc_mesh_ncells": (_i, [_vp]),
    "orc_mesh_geometry": (None, [_vp] + [_vp] * 8),
    "orc_rcm": (None, [_i, _i, _vp, _vp, _vp]),
    "orc_colour": (_i, [_i, _i, _vp, _vp, _vp]),
    "orc_renumber": (None, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "orc_case_new": (_vp, [_vp, _vp, _vp]),
    "orc_case_free": (None, [_vp]),
    "orc_set_bc": (None, [_vp, _i, _i, _i, _i, _vp, _vp, _vp]),
    "orc_set_bc_values": (None, [_vp, _i, _vp, _vp, _vp]),
    "orc_get_boundary": (None, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "orc_pseudo_time_loop": (_i, [_vp, _i, C.c_double, C.c_double, _i, _i, _vp, _vp]),
    "orc_set_fields": (None, [_vp, _vp, _vp, _vp]),
    "orc_get_fields": (None, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "orc_calc_flux": (None, [_vp, _vp, _vp, _vp]),
    "orc_residual": (None, [_vp, _vp]),
    "orc_local_dt": (None, [_vp, _vp]),
    "orc_assemble": (None, [_vp, _vp, _vp, _vp]),
    "orc_set_matrix": (None, [_vp, _vp, _vp, _vp]),
    "orc_amul": (None, [_vp, _vp, _vp]),
    "orc_precondition": (_i, [_vp, _vp]),
    "orc_solve": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "orc_pseudo_steps": (_i, [_vp, _i, _vp, _vp]),
    "orc_get_pseudo_co": (None, [_vp, _vp]),
    "orc_face_flux": (None, [_i, _i, C.c_double, C.c_double, C.c_double, _vp, _vp]),
    "orc_reconstruct": (None, [_vp, _i, _vp, _vp]),
    "orc_set_state_w": (None, [_vp, _vp]),
    "orc_residual_w": (None, [_vp, _vp, _vp, _vp, _vp]),
    "orc_residual_jvp": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "orc_residual_jac_dense": (None, [_vp, _vp, _vp, _vp, _vp]),
    "orc_set_mrf": (_i, [_vp, _i, _vp, _vp, _vp, C.c_double, _i, _vp]),
    "orc_set_translating": (_i, [_vp, _i, _vp, _vp, _i, _vp]),
    "orc_face_flux_mrf": (None, [_i, _i, C.c_double, C.c_double, C.c_double, _vp, C.c_double, _vp]),
    "orc_get_mrf": (None, [_vp, _vp, _vp, _vp, _vp]),
    "orc_hb_new": (_vp, [_i, _vp, _vp]),
    "orc_hb_free": (None, [_vp]),
    "orc_hb_residual": (None, [_vp, _vp]),
    "orc_hb_assemble": (_i, [_vp]),
    "orc_hb_amul": (None, [_vp, _vp, _vp]),
    "orc_hb_precondition": (None, [_vp, _vp]),
    "orc_hb_solve": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "orc_hb_pseudo_steps": (_i, [_vp, _i, _vp]),
    "orc_hb_operator": (_i, [_i, _vp, _i, _vp, _vp]),
    "orc_hb_set_phase_lag": (_i, [_vp, _i, _vp, _i, _vp, _vp]),
    "orc_hb_lag_operator": (_i, [_i, _vp, _vp, C.c_double, _vp]),
    "orc_hb_lag_far": (None, [_vp, _i, _i, _i, _vp]),
}
_lib = None


def use_library(path):
    """Load the given build of the oracle instead of the checker build (the
    CPU-baseline timing uses the -O3 build); only before the first use."""
    global LIB
    if _lib is not None and LIB != path:
        raise RuntimeError("the oracle library is already loaded")
    LIB = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from .build import build
            build()
        h = C.CDLL(LIB)
        for k, (r, a) in _SIG.items():
            f = getattr(h, k)
            f.restype = r
            f.argtypes = a
        _lib = h
    return _lib


def _p(a):
    # the ctypes pointer keeps the array (possibly a conversion temporary)
    # alive for the duration of the call it is passed to
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def parse_settings(thermo: str, schemes: str, solution: str):
    """Numeric settings from the case dictionaries (simple keyword scan)."""
    import re

    def num(text, key, default):
        m = re.search(r"\b" + re.escape(key) + r"\s+([-+0-9.eE]+)\s*;", text)
        return float(m.group(1)) if m else default

    def word(text, key, default):
        m = re.search(r"\b" + re.escape(key) + r"\s+([A-Za-z<>]+)", text)
        return m.group(1) if m else default

    flux = word(schemes, "fluxScheme", "AUSMPlusUp")
    lowMach = word(schemes, "lowMachAusm", "true") == "true"
    recon = 0 if re.search(r"reconstruct\(rho\)\s+upwind", schemes) else 1
    lim = 1 if "Venkatakrishnan" in schemes else (2 if "cellLimited<minmod>" in schemes else 0)
    mu = num(thermo, "mu", 0.0)
    precond = {"LUSGS": 0, "Jacobi": 1}[word(solution, "preconditioner", "LUSGS")]
    iset = np.array([{"AUSMPlusUp": 0, "ROE": 1, "HLLC": 2}[flux], int(lowMach), recon, lim, int(mu > 0), precond,
                     int(num(solution, "nDirections", 30)), int(num(solution, "maxIter", 1000)),
                     int(num(solution, "minIter", 0)), 1, 1], np.int32)
    co = num(solution, "pseudoCoNum", 1.0)
    rset = np.array([num(thermo, "R", 287.0), num(thermo, "gamma", 1.4), mu, num(thermo, "Pr", 1.0),
                     num(schemes, "entropyFixCoeff", 0.05), num(solution, "tolerance", 1e-12),
                     num(solution, "relTol", 1e-2), co, num(solution, "pseudoCoNumMin", 0.1),
                     num(solution, "pseudoCoNumMax", 25.0), 1.25, 0.1, 0.95, -1e15, 1e-15, 1e15], np.float64)
    return iset, rset


def pseudo_controls(solution: str):
    """(pseudoTol, pseudoTolRel, nPseudoCorr, nPseudoCorrMin) of pseudoTime."""
    import re

    def num(key, default):
        m = re.search(r"\b" + key + r"\s+([-+0-9.eE]+)\s*;", solution)
        return float(m.group(1)) if m else default
    return num("pseudoTol", float("nan")), num("pseudoTolRel", float("nan")), int(num("nPseudoCorr", 20)), \
        int(num("nPseudoCorrMin", 1))


class OracleMesh:
    def __init__(self, mesh):
        start, size, kind = mesh.patch_arrays()
        nbr, T = mesh.coupling_arrays()
        self._keep = [_f64(mesh.points), _i32(mesh.face_off), _i32(mesh.face_verts), _i32(mesh.owner),
                      _i32(mesh.neighbour), start, size, kind, _i32(nbr), _f64(T)]
        k = self._keep
        self.h = lib().orc_mesh_new(mesh.points.shape[0], _p(k[0]), mesh.n_faces, _p(k[1]), _p(k[2]),
                                    mesh.n_internal, _p(k[3]), _p(k[4]), len(mesh.patches), _p(start), _p(size),
                                    _p(kind), _p(k[8]), _p(k[9]))
        self.poly = mesh
        self.nCells = lib().orc_mesh_ncells(self.h)

    def geometry(self):
        nC, nF = self.nCells, self.poly.n_faces
        o = {k: np.empty(s) for k, s in [("C", (nC, 3)), ("V", nC), ("Sf", (nF, 3)), ("Cf", (nF, 3)),
                                        ("w", nF), ("dc", nF), ("nodc", nF), ("corr", (nF, 3))]}
        lib().orc_mesh_geometry(self.h, *[_p(o[k]) for k in ["C", "V", "Sf", "Cf", "w", "dc", "nodc", "corr"]])
        return o

    def cyclic(self):
        nF, nB = self.poly.n_faces, self.poly.n_faces - self.poly.n_internal
        w, dc, nodc, cell = np.empty(nF), np.empty(nF), np.empty(nF), np.empty(nB, np.int32)
        lib().orc_mesh_cyclic(self.h, _p(w), _p(dc), _p(nodc), _p(cell))
        return dict(w=w, dc=dc, nodc=nodc, cell=cell)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_free(self.h)
            self.h = None


def face_flux(scheme, S, left, right, lowMach=True, eFix=0.05, R=287.0, gamma=1.4, um=0.0):
    """Flux through one face; left/right = (rho, p, Ux, Uy, Uz, T); um: the
    face's MRF normal velocity (frame velocity . Sf/|Sf|)."""
    inp = np.ascontiguousarray(np.concatenate([S, left, right]), np.float64)
    out = np.empty(5)
    lib().orc_face_flux_mrf({"AUSMPlusUp": 0, "ROE": 1, "HLLC": 2}[scheme], int(lowMach), eFix, R, gamma, _p(inp),
                            float(um), _p(out))
    return out


def rcm(mesh):
    out = np.empty(mesh.n_cells, np.int32)
    if mesh.has_cyclic:
        co, cn = mesh.coupled_cells()
        lib().orc_rcm_coupled(mesh.n_cells, mesh.n_internal, _p(_i32(mesh.owner)), _p(_i32(mesh.neighbour)),
                              int(co.shape[0]), _p(_i32(co)), _p(_i32(cn)), _p(out))
    else:
        lib().orc_rcm(mesh.n_cells, mesh.n_internal, _p(_i32(mesh.owner)), _p(_i32(mesh.neighbour)), _p(out))
    return out


def colour(mesh):
    out = np.empty(mesh.n_cells, np.int32)
    n = lib().orc_colour(mesh.n_cells, mesh.n_internal, _p(_i32(mesh.owner)), _p(_i32(mesh.neighbour)), _p(out))
    return out, n


def renumber(mesh, new_to_old):
    nF, nI = mesh.n_faces, mesh.n_internal
    fv = np.empty(mesh.face_verts.shape[0], np.int32)
    own = np.empty(nF, np.int32)
    nei = np.empty(nI, np.int32)
    fm = np.empty(nF, np.int32)
    lib().orc_renumber(mesh.n_cells, nF, nI, _p(_i32(mesh.face_off)), _p(_i32(mesh.face_verts)),
                       _p(_i32(mesh.owner)), _p(_i32(mesh.neighbour)), _p(_i32(new_to_old)), _p(fv), _p(own),
                       _p(nei), _p(fm))
    return fv, own, nei, fm


class OracleCase:
    def __init__(self, case):
        self.mesh = OracleMesh(case.mesh)
        self.iset, self.rset = parse_settings(case.thermo, case.fvSchemes, case.fvSolution)
        self.ptol = pseudo_controls(case.fvSolution)
        self.h = lib().orc_case_new(self.mesh.h, _p(self.iset), _p(self.rset))
        self.nCells = self.mesh.nCells
        self.nInternal = case.mesh.n_internal
        self.nFaces = case.mesh.n_faces
        for i, bc in enumerate(case.bcs):
            if bc is None:  # constraint patch (cyclic)
                continue
            pv = _f64(bc.p_val) if bc.p_val is not None else None
            uv = _f64(bc.U_val) if bc.U_val is not None else None
            tv = _f64(bc.T_val) if bc.T_val is not None else None
            lib().orc_set_bc(self.h, i, bc.p_type, bc.U_type, bc.T_type, _p(pv), _p(uv), _p(tv))
            if bc.p_init is not None or bc.U_init is not None or bc.T_init is not None:
                init = [None if a is None else _f64(a) for a in (bc.p_init, bc.U_init, bc.T_init)]
                lib().orc_set_bc_values(self.h, i, *[_p(a) for a in init])
        for z in ([] if case.mrf is None else list(case.mrf) if isinstance(case.mrf, (list, tuple))
                  else [case.mrf]):
            self.set_mrf(z)
        if getattr(case, "translating", None) is not None:
            self.set_translating(case.translating)
        self.set_fields(case.rho, case.U, case.T)

    def set_mrf(self, z):
        """One rotating zone (cases.MRFZone); before set_fields, as the
        frame sets the rotating patches' U when the fields are created."""
        cells = _i32(np.flatnonzero(np.asarray(z.cells, bool)))
        nr = _i32(z.non_rotating)
        rc = lib().orc_set_mrf(self.h, int(cells.shape[0]), _p(cells), _p(_f64(z.origin)), _p(_f64(z.axis)),
                               float(z.omega), int(nr.shape[0]), _p(nr))
        if rc:
            raise ValueError("MRF: a patch that rotates with the zone needs a fixedValue U condition "
                             "(list it in non_rotating otherwise)")

    def set_translating(self, z):
        """One translating zone (cases.TranslatingZone); before set_fields."""
        cells = _i32(np.flatnonzero(np.asarray(z.cells, bool)))
        nm = _i32(z.non_moving)
        rc = lib().orc_set_translating(self.h, int(cells.shape[0]), _p(cells), _p(_f64(z.velocity)),
                                       int(nm.shape[0]), _p(nm))
        if rc:
            raise ValueError("translating frame: a moving patch needs a fixedValue U condition "
                             "(list it in non_moving otherwise)")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_case_free(self.h)
            self.h = None

    def set_fields(self, rho, U, T):
        lib().orc_set_fields(self.h, _p(_f64(rho)), _p(_f64(U)), _p(_f64(T)))

    def mrf(self):
        """The MRF set-up: face normal frame velocity [nF], Omega per cell
        [nC, 3], rotating boundary faces [nB] and their U [nB, 3]."""
        nF, nB = self.nFaces, self.nFaces - self.nInternal
        um, om, inc, ub = np.empty(nF), np.empty((self.nCells, 3)), np.empty(nB, np.int32), np.empty((nB, 3))
        lib().orc_get_mrf(self.h, _p(um), _p(om), _p(inc), _p(ub))
        return dict(umrf=um, omega=om, rotating=inc.astype(bool), U=ub)

    def fields(self):
        n = self.nCells
        rho, U, T, p, W = np.empty(n), np.empty((n, 3)), np.empty(n), np.empty(n), np.empty((5, n))
        lib().orc_get_fields(self.h, _p(rho), _p(U), _p(T), _p(p), _p(W))
        return dict(rho=rho, U=U, T=T, p=p, W=W)

    def boundary(self):
        """Boundary values [nB] (p, U [nB,3], T), the boundary mass flux phi
        of the last flux evaluation and the pressureInletOutletVelocity
        inflow switch of every boundary face."""
        nB = self.nFaces - self.nInternal
        p, U, T, phi, inflow = np.empty(nB), np.empty((nB, 3)), np.empty(nB), np.empty(nB), np.empty(nB, np.int32)
        lib().orc_get_boundary(self.h, _p(p), _p(U), _p(T), _p(phi), _p(inflow))
        return dict(p=p, U=U, T=T, phi=phi, inflow=inflow)

    def calcFlux(self):
        nF = self.nFaces
        phi, phiU, phiE = np.empty(nF), np.empty((nF, 3)), np.empty(nF)
        lib().orc_calc_flux(self.h, _p(phi), _p(phiU), _p(phiE))
        return phi, phiU, phiE

    def residual(self):
        b = np.empty((5, self.nCells))
        lib().orc_residual(self.h, _p(b))
        rdt = np.empty(self.nCells)
        lib().orc_local_dt(self.h, _p(rdt))
        return b, rdt

    def createJacobian(self):
        nC, nI = self.nCells, self.nInternal
        d, u, l = np.empty((nC, 5, 5)), np.empty((nI, 5, 5)), np.empty((nI, 5, 5))
        lib().orc_assemble(self.h, _p(d), _p(u), _p(l))
        return d, u, l

    def interfaces(self):
        """Interface (cyclic-face) blocks [nB, 5, 5] after createJacobian."""
        nB = self.nFaces - self.nInternal
        iu = np.empty((nB, 5, 5))
        lib().orc_get_interfaces(self.h, _p(iu))
        return iu

    def set_matrix(self, d, u, l):
        lib().orc_set_matrix(self.h, _p(_f64(d)), _p(_f64(u)), _p(_f64(l)))

    def matrixMul(self, x):
        y = np.empty((5, self.nCells))
        lib().orc_amul(self.h, _p(_f64(x)), _p(y))
        return y

    def precondition(self, r):
        r = np.array(r, dtype=np.float64, order="C", copy=True)
        rc = lib().orc_precondition(self.h, _p(r))
        if rc:
            raise RuntimeError(f"oracle precondition failed ({rc})")
        return r

    def solveForIncr(self, W, b):
        dW = np.empty((5, self.nCells))
        res = np.empty(10)
        it = np.zeros(1, np.int32)
        rc = lib().orc_solve(self.h, _p(_f64(W)), _p(_f64(b)), _p(dW), _p(res), _p(it))
        if rc:
            raise RuntimeError(f"oracle solve failed ({rc})")
        return dW, res, int(it[0])

    def pseudo_steps(self, n):
        res = np.empty((n, 10))
        it = np.zeros(n, np.int32)
        rc = lib().orc_pseudo_steps(self.h, n, _p(res), _p(it))
        if rc:
            raise RuntimeError(f"oracle pseudo-step failed ({rc})")
        return res

    def reconstruct(self, field):
        L, R = np.empty(self.nFaces), np.empty(self.nFaces)
        lib().orc_reconstruct(self.h, field, _p(L), _p(R))
        return L, R

    def pseudo_time_loop(self, max_steps):
        """dbnsFoam's steady loop under pseudotimeControl (criteria from the
        case's pseudoTime dictionary); returns (steps run, converged)."""
        run, conv = C.c_int(0), C.c_int(0)
        rc = lib().orc_pseudo_time_loop(self.h, int(max_steps), self.ptol[0], self.ptol[1], self.ptol[2],
                                        self.ptol[3], C.byref(run), C.byref(conv))
        if rc:
            raise RuntimeError(f"oracle pseudo-step failed ({rc})")
        return run.value, bool(conv.value)

    def pseudoCo(self):
        co = np.empty(self.nCells)
        lib().orc_get_pseudo_co(self.h, _p(co))
        return co

    # ---- tangent-linear restatement of the inviscid AUSM+up residual ----
    def set_state_w(self, W):
        """State from conserved W [5][nC] via updateFields (zero increment)."""
        lib().orc_set_state_w(self.h, _p(_f64(W)))

    def residual_w(self, W, Sf=None, Cf=None):
        """b(W, Sf, Cf) of the dual-number restatement evaluated in FP64."""
        b = np.empty((5, self.nCells))
        Sf = None if Sf is None else _f64(Sf)
        Cf = None if Cf is None else _f64(Cf)
        lib().orc_residual_w(self.h, _p(_f64(W)), _p(Sf), _p(Cf), _p(b))
        return b

    def residual_jvp(self, W, dW=None, dSf=None, dCf=None):
        """(b, J_W dW + J_S dSf + J_C dCf) by forward-mode dual numbers."""
        b, db = np.empty((5, self.nCells)), np.empty((5, self.nCells))
        a = [None if x is None else _f64(x) for x in (dW, dSf, dCf)]
        lib().orc_residual_jvp(self.h, _p(_f64(W)), _p(a[0]), _p(a[1]), _p(a[2]), _p(b), _p(db))
        return b, db

    def residual_jac_dense(self, W):
        """Dense J_W [5nC,5nC], J_S [5nC,3nF], J_C [5nC,3nF] (small meshes)."""
        n5, n3 = 5 * self.nCells, 3 * self.nFaces
        JW, JS, JC = np.empty((n5, n5)), np.empty((n5, n3)), np.empty((n5, n3))
        lib().orc_residual_jac_dense(self.h, _p(_f64(W)), _p(JW), _p(JS), _p(JC))
        return JW, JS, JC


def hb_operator(omegas, instants):
    """D [nO][nO] of the oracle's restatement of HBZone::updateHBOperators."""
    om, t = _f64(omegas), _f64(instants)
    D = np.empty((len(t), len(t)))
    rc = lib().orc_hb_operator(len(om), _p(om), len(t), _p(t), _p(D))
    if rc:
        raise RuntimeError(f"oracle HB operator failed ({rc})")
    return D


def hb_lag_operator(omegas, instants, ibpa):
    """Dl [nO][nO]: the oracle's restatement of a phase-lagged patch side's
    time shift Re(E^-1 M E) (omega list as hb_operator's)."""
    om, t = _f64(omegas), _f64(instants)
    Dl = np.empty((len(t), len(t)))
    rc = lib().orc_hb_lag_operator(len(t), _p(om), _p(t), float(ibpa), _p(Dl))
    if rc:
        raise RuntimeError(f"oracle lag operator failed ({rc})")
    return Dl


class OracleHB:
    """nO instances of one case coupled by the HB operator D."""

    def __init__(self, instances, D):
        self.inst = [OracleCase(c) for c in instances]
        self.nO = len(self.inst)
        self.nCells = self.inst[0].nCells
        self.D = _f64(D)
        arr = (C.c_void_p * self.nO)(*[o.h for o in self.inst])
        self._arr = arr
        self.h = lib().orc_hb_new(self.nO, C.cast(arr, C.c_void_p), _p(self.D))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_hb_free(self.h)
            self.h = None

    def set_phase_lag(self, patch, Dl, cyl=False, axis=(0.0, 0.0, 1.0), origin=(0.0, 0.0, 0.0)):
        """Make cyclic patch `patch` (one side of its pair) phase-lagged with
        the time shift Dl [nO][nO]; cyl: cylCoords about the patch's
        rotation axis and centre."""
        rc = lib().orc_hb_set_phase_lag(self.h, int(patch), _p(_f64(Dl)), int(bool(cyl)), _p(_f64(axis)),
                                        _p(_f64(origin)))
        if rc:
            raise RuntimeError(f"oracle phase lag set-up failed ({rc})")

    LAG_FIELDS = {"rho": 1, "p": 2, "U": 3, "E": 4, "H": 5, "c": 6, "gamma": 7}

    def lag_far(self, J, field, comp=0):
        """Far-side values of instance J's boundary faces on the
        phase-lagged faces (zero elsewhere): field in LAG_FIELDS; for "U"
        comp 0..2 the unrotated components, 3..5 the vector's (forwardT
        applied)."""
        nB = self.inst[0].nFaces - self.inst[0].nInternal
        out = np.empty(nB)
        lib().orc_hb_lag_far(self.h, int(J), self.LAG_FIELDS[field], int(comp), _p(out))
        return out

    def residual(self):
        b = np.empty((self.nO, 5, self.nCells))
        lib().orc_hb_residual(self.h, _p(b))
        return b

    def createJacobian(self):
        rc = lib().orc_hb_assemble(self.h)
        if rc:
            raise RuntimeError(f"oracle HB assemble failed ({rc})")

    def matrixMul(self, x):
        y = np.empty((self.nO, 5, self.nCells))
        lib().orc_hb_amul(self.h, _p(_f64(x)), _p(y))
        return y

    def precondition(self, r):
        r = np.array(r, dtype=np.float64, order="C", copy=True)
        lib().orc_hb_precondition(self.h, _p(r))
        return r

    def solveForIncr(self, b):
        dW = np.empty((self.nO, 5, self.nCells))
        res = np.empty((self.nO, 10))
        it = np.zeros(1, np.int32)
        rc = lib().orc_hb_solve(self.h, _p(_f64(b)), _p(dW), _p(res), _p(it))
        if rc:
            raise RuntimeError(f"oracle HB solve failed ({rc})")
        return dW, res, int(it[0])

    def pseudo_steps(self, n):
        res = np.empty((n, self.nO, 10))
        rc = lib().orc_hb_pseudo_steps(self.h, n, _p(res))
        if rc:
            raise RuntimeError(f"oracle HB pseudo-step failed ({rc})")
        return res

    def fields(self, J):
        return self.inst[J].fields()
