"""Host-side mirror of the reference's solver interfaces over libicsf.

  fvMesh(polyMesh)                         fvMesh (geometry, lduAddressing)
  dbnsCase(mesh, thermo, schemes, sol)     createFields.H + runtime selection of
                                           convectiveFluxScheme / viscousFluxScheme /
                                           coupledMatrix::solver / ::preconditioner
  .calcFlux()                              convectiveFluxScheme::calcFlux
  .residual()                              residualsUpdate.H (+ setCoAndDeltaT.H)
  .createJacobian()                        createConvectiveJacobian + createViscousJacobian
  .matrixMul(x) / .precondition(r)         coupledMatrix::matrixMul / preconditioner
  .solveForIncr()                          coupledMatrix::solveForIncr
  .pseudo_steps(n)                         dbnsFoam pseudo-time loop (outerLoop + updateFields)

Vectors use the reference's unknown order [rho, rhoE, rhoUx, rhoUy, rhoUz]
laid out [5][nCells].
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from ._lib import as_f64, as_i32, check, lib, ptr
from .cases import (BC_FIXED_VALUE, BC_PRESSURE_INLET_OUTLET_VELOCITY, BC_TOTAL_PRESSURE, BC_TOTAL_TEMPERATURE,
                    Case, PolyMesh)


def _bc_values(p_type, U_type, T_type, p_val, U_val, T_val):
    """The value arrays a condition kind takes: fixed values, p0
    (totalPressure), tangentialVelocity (pressureInletOutletVelocity, may be
    None), T0 (totalTemperature)."""
    def arr(v, wanted):
        return as_f64(v) if (wanted and v is not None) else None
    pv = arr(p_val, p_type in (BC_FIXED_VALUE, BC_TOTAL_PRESSURE))
    uv = arr(U_val, U_type in (BC_FIXED_VALUE, BC_PRESSURE_INLET_OUTLET_VELOCITY))
    tv = arr(T_val, T_type in (BC_FIXED_VALUE, BC_TOTAL_TEMPERATURE))
    return pv, uv, tv


def _cap(occ: int):
    """Register budget of an autotuned kernel class: None (the compiler's
    choice), 128 or 168 VGPRs."""
    return {0: None, 1: 128, 2: 168}.get(int(occ), None)


class fvMesh:
    def __init__(self, mesh: PolyMesh):
        self.poly = mesh
        start, size, kind = mesh.patch_arrays()
        nbr, T = mesh.coupling_arrays()
        h = C.c_void_p()
        check(lib().icsf_mesh_create(mesh.points.shape[0], ptr(as_f64(mesh.points)), mesh.n_faces,
                                     ptr(as_i32(mesh.face_off)), ptr(as_i32(mesh.face_verts)), mesh.n_internal,
                                     ptr(as_i32(mesh.owner)), ptr(as_i32(mesh.neighbour)), len(mesh.patches),
                                     ptr(start), ptr(size), ptr(kind), ptr(nbr), ptr(as_f64(T)), C.byref(h)))
        self._h = h
        info = np.zeros(8, np.int32)
        check(lib().icsf_mesh_info(h, ptr(info)))
        self.nCells, self.nFaces, self.nInternalFaces = int(info[0]), int(info[1]), int(info[2])
        self.nFwdLevels, self.nBwdLevels = int(info[3]), int(info[4])
        self.nSlices, self.sellSlots, self.emptyDir = int(info[5]), int(info[6]), int(info[7])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().icsf_mesh_destroy(h)
            self._h = None



def _zones(mrf):
    """A case's rotating zones as a list (cases.MRFZone, a list of them, or
    None)."""
    if mrf is None:
        return []
    return list(mrf) if isinstance(mrf, (list, tuple)) else [mrf]

class dbnsCase:
    """Fields + schemes + solution controls of one dbnsFoam case on the GPU."""

    def __init__(self, mesh: fvMesh, thermo: str, fvSchemes: str, fvSolution: str):
        self.mesh = mesh
        h = C.c_void_p()
        check(lib().icsf_case_create(mesh._h, thermo.encode(), fvSchemes.encode(), fvSolution.encode(),
                                     C.byref(h)))
        self._h = h
        self.nCells = mesh.nCells

    @classmethod
    def from_case(cls, case: Case, mesh: Optional[fvMesh] = None) -> "dbnsCase":
        mesh = mesh or fvMesh(case.mesh)
        self = cls(mesh, case.thermo, case.fvSchemes, case.fvSolution)
        for i, bc in enumerate(case.bcs):
            if bc is None:  # constraint patch (cyclic): no conditions to set
                continue
            self.set_boundary(i, bc.p_type, bc.U_type, bc.T_type, bc.p_val, bc.U_val, bc.T_val)
            if bc.p_init is not None or bc.U_init is not None or bc.T_init is not None:
                self.set_boundary_values(i, bc.p_init, bc.U_init, bc.T_init)
        for z in _zones(getattr(case, "mrf", None)):
            self.set_mrf(z)
        if getattr(case, "translating", None) is not None:
            self.set_translating(case.translating)
        self.set_fields(case.rho, case.U, case.T)
        return self

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().icsf_case_destroy(h)
            self._h = None

    # ---- state -------------------------------------------------------
    def set_boundary(self, patch, p_type, U_type, T_type, p_val=None, U_val=None, T_val=None):
        pv, uv, tv = _bc_values(p_type, U_type, T_type, p_val, U_val, T_val)
        check(lib().icsf_set_boundary(self._h, patch, p_type, U_type, T_type, ptr(pv), ptr(uv), ptr(tv)))

    def set_boundary_values(self, patch, p=None, U=None, T=None):
        """Initial boundary values of the derived conditions (`value` entries)."""
        pv, uv, tv = (None if a is None else as_f64(a) for a in (p, U, T))
        check(lib().icsf_set_boundary_values(self._h, patch, ptr(pv), ptr(uv), ptr(tv)))

    def set_fields(self, rho, U, T):
        check(lib().icsf_set_fields(self._h, ptr(as_f64(rho)), ptr(as_f64(U)), ptr(as_f64(T))))

    def set_mrf(self, zone):
        """Adds a rotating zone (cases.MRFZone: bool cell mask, origin,
        axis, omega, non-rotating patch indices; the zone list's next entry);
        after the boundary conditions and before the fields (icsf_set_mrf)."""
        cells = as_i32(np.flatnonzero(np.asarray(zone.cells, bool)))
        nr = as_i32(np.asarray(zone.non_rotating, np.int32).reshape(-1))
        check(lib().icsf_set_mrf(self._h, int(cells.shape[0]), ptr(cells), ptr(as_f64(zone.origin)),
                                 ptr(as_f64(zone.axis)), float(zone.omega), int(nr.shape[0]), ptr(nr)))

    def set_translating(self, zone):
        """One translating zone (cases.TranslatingZone: bool cell mask,
        frame velocity, non-moving patch indices); after the boundary
        conditions and before the fields (icsf_set_translating_frame)."""
        cells = as_i32(np.flatnonzero(np.asarray(zone.cells, bool)))
        nm = as_i32(np.asarray(zone.non_moving, np.int32).reshape(-1))
        check(lib().icsf_set_translating_frame(self._h, int(cells.shape[0]), ptr(cells), ptr(as_f64(zone.velocity)),
                                               int(nm.shape[0]), ptr(nm)))

    def fields(self):
        n = self.nCells
        rho, U, T, p, W = np.empty(n), np.empty((n, 3)), np.empty(n), np.empty(n), np.empty((5, n))
        check(lib().icsf_get_fields(self._h, ptr(rho), ptr(U), ptr(T), ptr(p), ptr(W)))
        return dict(rho=rho, U=U, T=T, p=p, W=W)

    def set_state(self, W):
        """Conserved state W [5][nC] through updateFields (zero increment)."""
        check(lib().icsf_set_state(self._h, ptr(as_f64(W))))

    def boundary(self):
        """Boundary values of every boundary face (p, U [nB,3], T) as of the
        last state update."""
        nB = self.mesh.nFaces - self.mesh.nInternalFaces
        p, U, T = np.empty(nB), np.empty((nB, 3)), np.empty(nB)
        check(lib().icsf_get_boundary_values(self._h, ptr(p), ptr(U), ptr(T)))
        return dict(p=p, U=U, T=T)

    # ---- flux / residual / Jacobian ----------------------------------
    def calcFlux(self):
        nF = self.mesh.nFaces
        phi, phiUp, phiEp = np.empty(nF), np.empty((nF, 3)), np.empty(nF)
        check(lib().icsf_calc_flux(self._h, ptr(phi), ptr(phiUp), ptr(phiEp)))
        return phi, phiUp, phiEp

    def residual(self):
        b, rdt = np.empty((5, self.nCells)), np.empty(self.nCells)
        check(lib().icsf_residual(self._h, ptr(b), ptr(rdt)))
        return b, rdt

    def residual_vjp(self, v, W=None, geometry=True):
        """(gW, gSf, gCf) = ((db/dW)^T v, (db/dSf)^T v, (db/dCf)^T v) of the
        residual (AUSM+up or Roe, inviscid or viscous) at W ([5][nCells];
        None = current state).  W becomes the case state.  gSf and gCf
        (geometry=True) cover inviscid residuals; with geometry=False they
        are None."""
        v = np.ascontiguousarray(v, np.float64)
        Wc = None if W is None else np.ascontiguousarray(W, np.float64)
        gW = np.empty((5, self.nCells))
        gS, gC = (np.empty((self.mesh.nFaces, 3)), np.empty((self.mesh.nFaces, 3))) if geometry else (None, None)
        check(lib().icsf_residual_vjp(self._h, ptr(Wc), ptr(v), ptr(gW), ptr(gS), ptr(gC)))
        return gW, gS, gC

    def createJacobian(self):
        check(lib().icsf_assemble_jacobian(self._h))

    def jacobian(self):
        nC, nI = self.nCells, self.mesh.nInternalFaces
        d, u, l = np.empty((nC, 5, 5)), np.empty((nI, 5, 5)), np.empty((nI, 5, 5))
        check(lib().icsf_get_jacobian(self._h, ptr(d), ptr(u), ptr(l)))
        return d, u, l

    def tuning(self) -> dict:
        """Launch parameters in use (chosen by the autotuner on the first
        pseudo-step unless `autotune false`)."""
        p = np.zeros(10, np.int32)
        check(lib().icsf_get_tuning(self._h, ptr(p)))
        return dict(op_workgroup=int(p[0]), face_workgroup=int(p[1]), cell_workgroup=int(p[2]),
                    jacobian_workgroup=int(p[3]), projection_chunk=int(p[4]), tuned=bool(p[5]),
                    op_vgpr_cap=_cap(p[6]), face_vgpr_cap=_cap(p[7]), cell_vgpr_cap=_cap(p[8]),
                    jacobian_vgpr_cap=_cap(p[9]))

    def set_tuning(self, params):
        """Pins the launch parameters (icsf_get_tuning's 10-int layout) and
        skips the autotuner."""
        check(lib().icsf_set_tuning(self._h, ptr(as_i32(params))))

    def tuning_params(self) -> np.ndarray:
        """The raw 10-int launch parameters (icsf_get_tuning)."""
        p = np.zeros(10, np.int32)
        check(lib().icsf_get_tuning(self._h, ptr(p)))
        return p

    def interfaces(self):
        """interfacesUpper blocks of the boundary faces [nB, 5, 5] (non-zero
        on cyclic faces only)."""
        nB = self.mesh.nFaces - self.mesh.nInternalFaces
        iu = np.empty((nB, 5, 5))
        check(lib().icsf_get_interface_blocks(self._h, ptr(iu)))
        return iu

    def set_jacobian(self, diag, upper, lower):
        check(lib().icsf_set_jacobian(self._h, ptr(as_f64(diag)), ptr(as_f64(upper)), ptr(as_f64(lower))))

    def matrixMul(self, x):
        y = np.empty((5, self.nCells))
        check(lib().icsf_matrix_mul(self._h, ptr(as_f64(x)), ptr(y)))
        return y

    def precondition(self, r):
        r = np.array(r, dtype=np.float64, order="C", copy=True)
        check(lib().icsf_precondition(self._h, ptr(r)))
        return r

    def solveForIncr(self, W=None, b=None):
        dW = np.empty((5, self.nCells))
        res = np.empty(10)
        it = np.zeros(1, np.int32)
        check(lib().icsf_solve_for_incr(self._h, ptr(None if W is None else as_f64(W)),
                                        ptr(None if b is None else as_f64(b)), ptr(dW), ptr(res), ptr(it)))
        return dW, res, int(it[0])

    # ---- pseudo-time loop ----------------------------------------------
    def pseudo_steps(self, n: int, residuals: bool = True):
        res = np.empty((n, 10)) if residuals else None
        check(lib().icsf_pseudo_steps(self._h, n, ptr(res), None))
        return res

    def pseudo_time_loop(self, max_steps: int, batch: int = 8):
        """dbnsFoam's steady loop under pseudotimeControl: steps until the
        pseudoTol / pseudoTolRel criteria end it (one final step after they
        hold) or max_steps; returns (steps run, converged)."""
        run, conv = C.c_int(0), C.c_int(0)
        check(lib().icsf_pseudo_time_loop(self._h, int(max_steps), int(batch), C.byref(run), C.byref(conv)))
        return run.value, bool(conv.value)

    def pseudoCo(self):
        co = np.empty(self.nCells)
        check(lib().icsf_get_pseudo_co(self._h, ptr(co)))
        return co

    def synchronize(self):
        check(lib().icsf_synchronize(self._h))

    def timing(self, mask: int = 0x3F):
        """Event timing of kernel classes (bit k = class k; 0 stops)."""
        check(lib().icsf_timing_enable(self._h, int(mask)))

    def timing_read(self):
        ms = np.zeros(6)
        cnt = np.zeros(6, np.int64)
        check(lib().icsf_timing_read(self._h, 6, ptr(ms), ptr(cnt)))
        return ms, cnt


def hb_reconstruct_weights(frequencies, harmonics, instants, time: float):
    """Weights of the Fourier reconstruction of the instants at `time`
    (HBZone::reconstruct); frequencies angular, as hb_operator."""
    f = as_f64(frequencies)
    h = as_i32(harmonics)
    t = as_f64(instants)
    w = np.empty(t.size)
    check(lib().icsf_hb_reconstruct_weights(int(f.size), ptr(f), ptr(h), int(t.size), ptr(t), float(time), ptr(w)))
    return w


def hb_operator(frequencies, harmonics, n_instants: int, selected_period: float = 0.0):
    """(D [nO, nO], instants [nO], condition number) of the harmonic-balance
    spectral operator (HBZone::updateHBOperators / HBZoneList::setInstants).
    Frequencies are angular (rad/s); selected_period <= 0 selects the
    instants by the condition-number scan."""
    f = as_f64(np.atleast_1d(frequencies))
    h = as_i32(np.atleast_1d(harmonics))
    D = np.empty((n_instants, n_instants))
    t = np.empty(n_instants)
    cond = np.zeros(1)
    check(lib().icsf_hb_operator(len(f), ptr(f), ptr(h), n_instants, float(selected_period), ptr(t), ptr(D),
                                 ptr(cond)))
    return D, t, float(cond[0])


def hb_phase_lag_operator(frequencies, harmonics, instants, ibpa: float):
    """Time shift Dl [nO, nO] of a phase-lagged cyclic patch side
    (phaseLagCyclicFvPatchField::phaseLaggedField): sum_m Dl[l, m] x(t_m)
    is x at t_l + ibpa / omega_1 for x in the span of the omega list."""
    f = as_f64(np.atleast_1d(frequencies))
    h = as_i32(np.atleast_1d(harmonics))
    t = as_f64(instants)
    Dl = np.empty((t.size, t.size))
    check(lib().icsf_hb_phase_lag_operator(int(f.size), ptr(f), ptr(h), int(t.size), ptr(t), float(ibpa), ptr(Dl)))
    return Dl


class hbCase:
    """dbnsFullyImplicitHBFoam on the GPU: nO instants of one case on one
    mesh coupled by the spectral operator D."""

    def __init__(self, mesh: fvMesh, thermo: str, fvSchemes: str, fvSolution: str, D):
        self.mesh = mesh
        self.D = as_f64(D)
        self.nO = int(self.D.shape[0])
        h = C.c_void_p()
        check(lib().icsf_hb_create(mesh._h, thermo.encode(), fvSchemes.encode(), fvSolution.encode(), self.nO,
                                   ptr(self.D), C.byref(h)))
        self._h = h
        self.nCells = mesh.nCells

    @classmethod
    def from_cases(cls, instances, D, mesh: Optional[fvMesh] = None) -> "hbCase":
        c0 = instances[0]
        mesh = mesh or fvMesh(c0.mesh)
        self = cls(mesh, c0.thermo, c0.fvSchemes, c0.fvSolution, D)
        for J, c in enumerate(instances):
            for i, bc in enumerate(c.bcs):
                if bc is None:  # constraint patch (cyclic)
                    continue
                self.set_boundary(J, i, bc.p_type, bc.U_type, bc.T_type, bc.p_val, bc.U_val, bc.T_val)
                if bc.p_init is not None or bc.U_init is not None or bc.T_init is not None:
                    self.set_boundary_values(J, i, bc.p_init, bc.U_init, bc.T_init)
        for z in _zones(getattr(c0, "mrf", None)):
            self.set_mrf(z)
        if getattr(c0, "translating", None) is not None:
            z = c0.translating
            cells = as_i32(np.flatnonzero(np.asarray(z.cells, bool)))
            nm = as_i32(np.asarray(z.non_moving, np.int32).reshape(-1))
            check(lib().icsf_hb_set_translating_frame(self._h, int(cells.shape[0]), ptr(cells),
                                                      ptr(as_f64(z.velocity)), int(nm.shape[0]), ptr(nm)))
        if getattr(c0, "phase_lag", None) is not None:
            self.set_phase_lag(c0.phase_lag, c0.mesh.patches[c0.phase_lag.owner].neighbour)
        for J, c in enumerate(instances):
            self.set_fields(J, c.rho, c.U, c.T)
        return self

    def set_phase_lag(self, pair, neighbour: int):
        """Both sides of a phase-lagged cyclic pair (cases.PhaseLagPair; the
        neighbour patch index of pair.owner): the owner side shifted by
        +IBPA, the neighbour side by -IBPA.  Returns their time shifts."""
        if neighbour <= pair.owner:
            raise ValueError("phase lag: the owner side is the pair's lower patch index")
        Do = hb_phase_lag_operator(pair.frequencies, pair.harmonics, pair.instants, pair.ibpa)
        Dn = hb_phase_lag_operator(pair.frequencies, pair.harmonics, pair.instants, -pair.ibpa)
        ax, ce = as_f64(pair.axis), as_f64(pair.centre)
        for patch, Dl in ((pair.owner, Do), (neighbour, Dn)):
            check(lib().icsf_hb_set_phase_lag(self._h, int(patch), ptr(as_f64(Dl)), int(bool(pair.cyl_coords)),
                                              ptr(ax), ptr(ce)))
        self.phase_lag_sides = [(pair.owner, Do), (neighbour, Dn)]
        return Do, Dn

    def set_mrf(self, zone):
        """Adds a rotating zone for all instances (cases.MRFZone)."""
        cells = as_i32(np.flatnonzero(np.asarray(zone.cells, bool)))
        nr = as_i32(np.asarray(zone.non_rotating, np.int32).reshape(-1))
        check(lib().icsf_hb_set_mrf(self._h, int(cells.shape[0]), ptr(cells), ptr(as_f64(zone.origin)),
                                    ptr(as_f64(zone.axis)), float(zone.omega), int(nr.shape[0]), ptr(nr)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().icsf_hb_destroy(h)
            self._h = None

    def set_boundary(self, J, patch, p_type, U_type, T_type, p_val=None, U_val=None, T_val=None):
        pv, uv, tv = _bc_values(p_type, U_type, T_type, p_val, U_val, T_val)
        check(lib().icsf_hb_set_boundary(self._h, J, patch, p_type, U_type, T_type, ptr(pv), ptr(uv), ptr(tv)))

    def set_boundary_values(self, J, patch, p=None, U=None, T=None):
        pv, uv, tv = (None if a is None else as_f64(a) for a in (p, U, T))
        check(lib().icsf_hb_set_boundary_values(self._h, J, patch, ptr(pv), ptr(uv), ptr(tv)))

    def set_fields(self, J, rho, U, T):
        check(lib().icsf_hb_set_fields(self._h, J, ptr(as_f64(rho)), ptr(as_f64(U)), ptr(as_f64(T))))

    def reconstruct(self, weights):
        """Fields at one time from the instances (reconstructHBFoam): rho,
        U, T, p = sum_m w[m] field_m, weights from hb_reconstruct_weights."""
        n = self.nCells
        rho, U, T, p = np.empty(n), np.empty((n, 3)), np.empty(n), np.empty(n)
        check(lib().icsf_hb_reconstruct(self._h, ptr(as_f64(weights)), ptr(rho), ptr(U), ptr(T), ptr(p)))
        return dict(rho=rho, U=U, T=T, p=p)

    def fields(self, J):
        n = self.nCells
        rho, U, T, p, W = np.empty(n), np.empty((n, 3)), np.empty(n), np.empty(n), np.empty((5, n))
        check(lib().icsf_hb_get_fields(self._h, J, ptr(rho), ptr(U), ptr(T), ptr(p), ptr(W)))
        return dict(rho=rho, U=U, T=T, p=p, W=W)

    def residual(self):
        b = np.empty((self.nO, 5, self.nCells))
        check(lib().icsf_hb_residual(self._h, ptr(b)))
        return b

    def createJacobian(self):
        check(lib().icsf_hb_assemble_jacobian(self._h))

    def matrixMul(self, x):
        y = np.empty((self.nO, 5, self.nCells))
        check(lib().icsf_hb_matrix_mul(self._h, ptr(as_f64(x)), ptr(y)))
        return y

    def precondition(self, r):
        r = np.array(r, dtype=np.float64, order="C", copy=True)
        check(lib().icsf_hb_precondition(self._h, ptr(r)))
        return r

    def solveForIncr(self, b):
        dW = np.empty((self.nO, 5, self.nCells))
        res = np.empty((self.nO, 10))
        it = np.zeros(1, np.int32)
        check(lib().icsf_hb_solve_for_incr(self._h, ptr(as_f64(b)), ptr(dW), ptr(res), ptr(it)))
        return dW, res, int(it[0])

    def pseudo_steps(self, n: int, residuals: bool = True):
        res = np.empty((n, self.nO, 10)) if residuals else None
        check(lib().icsf_hb_pseudo_steps(self._h, n, ptr(res)))
        return res

    def synchronize(self):
        check(lib().icsf_hb_synchronize(self._h))
