"""Seeded synthetic cases: polyMesh generators, initial fields, boundary
conditions and OpenFOAM-syntax dictionaries for the five workloads the
implicit coupled pseudo-step is measured on.

Meshes are written the way OpenFOAM's blockMesh writes them
(owner < neighbour, internal faces in upper-triangular order, boundary
faces grouped per patch, face normals pointing out of the owner), which is
the addressing ICSFoam's block matrices rely on
(src/blockFvMatrix/blockFvMatrix.C:220-234).

Everything here is input generation; no solver arithmetic lives here.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Tuple

import numpy as np

# patch kinds understood by the library (mirrors OpenFOAM polyPatch types)
PATCH_PATCH, PATCH_WALL, PATCH_EMPTY, PATCH_CYCLIC = 0, 1, 2, 3
# boundary-condition kinds (OpenFOAM fvPatchField type names)
BC_ZERO_GRADIENT, BC_FIXED_VALUE, BC_SLIP = 0, 1, 2
# derived inflow conditions: totalPressure (p), pressureInletOutletVelocity
# (U), totalTemperature (T)
BC_TOTAL_PRESSURE, BC_PRESSURE_INLET_OUTLET_VELOCITY, BC_TOTAL_TEMPERATURE = 3, 4, 5

R_GAS = 287.0
GAMMA = 1.4


@dataclasses.dataclass
class Patch:
    name: str
    kind: int
    start: int
    size: int
    # cyclic patches: index of the paired patch and forwardT (3x3, maps the
    # paired side's vectors into this side's frame); face i of the patch is
    # coupled to face i of the paired patch
    neighbour: int = -1
    transform: Optional[np.ndarray] = None
    # phaseLagCyclic patches as read from a boundary file: IBPA,
    # cyl_coords, axis, centre, zone (HBZoneName); None otherwise
    phase_lag: Optional[dict] = None


@dataclasses.dataclass
class PolyMesh:
    points: np.ndarray          # float64 [nPoints, 3]
    face_off: np.ndarray        # int32 [nFaces+1]
    face_verts: np.ndarray      # int32 [face_off[-1]]
    owner: np.ndarray           # int32 [nFaces]
    neighbour: np.ndarray       # int32 [nInternalFaces]
    patches: List[Patch]
    n_cells: int
    # vertex-average cell centres of a structured block in generation order
    centre_hint: Optional[np.ndarray] = None

    @property
    def n_faces(self) -> int:
        return int(self.owner.shape[0])

    @property
    def n_internal(self) -> int:
        return int(self.neighbour.shape[0])

    def patch_arrays(self):
        start = np.array([p.start for p in self.patches], np.int32)
        size = np.array([p.size for p in self.patches], np.int32)
        kind = np.array([p.kind for p in self.patches], np.int32)
        return start, size, kind

    def coupling_arrays(self):
        """(neighbour patch [nPatches], forwardT [nPatches, 9]) of the
        cyclic patches (-1 and zeros elsewhere)."""
        nbr = np.array([p.neighbour for p in self.patches], np.int32)
        T = np.zeros((len(self.patches), 9))
        for i, p in enumerate(self.patches):
            if p.transform is not None:
                T[i] = np.asarray(p.transform, np.float64).reshape(9)
        return nbr, T

    @property
    def has_cyclic(self) -> bool:
        return any(p.kind == PATCH_CYCLIC for p in self.patches)

    def coupled_cells(self):
        """Boundary faces of cyclic patches in face order: (owner cell,
        cell across the face), as decompositionMethod::calcCellCells adds
        them."""
        own, nbr = [], []
        for p in self.patches:
            if p.kind != PATCH_CYCLIC:
                continue
            q = self.patches[p.neighbour]
            own.append(self.owner[p.start:p.start + p.size])
            nbr.append(self.owner[q.start:q.start + q.size])
        if not own:
            return np.zeros(0, np.int32), np.zeros(0, np.int32)
        return np.concatenate(own).astype(np.int32), np.concatenate(nbr).astype(np.int32)


@dataclasses.dataclass
class BoundarySpec:
    """Per-patch p, U, T conditions; values per patch face: the fixed value,
    or p0 (totalPressure), tangentialVelocity (pressureInletOutletVelocity,
    None: zero), T0 (totalTemperature).  *_init: initial boundary values of
    the derived conditions (the field files' `value` entries; None: p0, the
    owner cell's U, T0)."""
    p_type: int
    U_type: int
    T_type: int
    p_val: Optional[np.ndarray] = None   # [size]
    U_val: Optional[np.ndarray] = None   # [size, 3]
    T_val: Optional[np.ndarray] = None   # [size]
    p_init: Optional[np.ndarray] = None  # [size]
    U_init: Optional[np.ndarray] = None  # [size, 3]
    T_init: Optional[np.ndarray] = None  # [size]


@dataclasses.dataclass
class MRFZone:
    """One rotating zone (the reference's MRFProperties entry, read by
    MRFCoupledZone): cells of the zone (bool [nC]), origin, axis (normalised
    on use), omega [rad/s], and the patches that do not rotate with it
    (nonRotatingPatches, patch indices).  The zone's other non-empty,
    non-coupled patch faces rotate with the frame: their U must be
    fixedValue, and the frame sets it to Omega ^ (Cf - origin)."""
    cells: np.ndarray
    origin: np.ndarray
    axis: np.ndarray
    omega: float
    non_rotating: List[int] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class TranslatingZone:
    """One translating zone (the reference's MRFTranslatingProperties entry,
    MRFTranslatingZone): cells (bool [nC]), frameVelocity, and the patches
    that do not move with it (nonMovingPatches, patch indices).  The zone's
    other non-empty, non-coupled patch faces move with the frame: their U
    must be fixedValue, and the frame sets it to the frame velocity."""
    cells: np.ndarray
    velocity: np.ndarray
    non_moving: List[int] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class PhaseLagPair:
    """A phase-lagged cyclic pair of a harmonic-balance case (the
    reference's phaseLagCyclic patch type: IBPA, cylCoords,
    phaseLagCyclicPolyPatch.C:79-90).  owner: the pair's lower patch index,
    whose IBPA holds (the neighbour side takes minus it); the HB zone's
    angular frequencies, harmonics and instants give the time shift; axis
    and centre: the pair's rotation axis and centre (cylCoords)."""
    owner: int
    ibpa: float
    frequencies: np.ndarray
    harmonics: np.ndarray
    instants: np.ndarray
    cyl_coords: bool = False
    axis: np.ndarray = dataclasses.field(default_factory=lambda: np.array([0.0, 0.0, 1.0]))
    centre: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(3))


@dataclasses.dataclass
class Case:
    name: str
    mesh: PolyMesh
    rho: np.ndarray     # [nC]
    U: np.ndarray       # [nC, 3]
    T: np.ndarray       # [nC]
    bcs: List[Optional[BoundarySpec]]   # None on cyclic (constraint) patches
    thermo: str
    fvSchemes: str
    fvSolution: str
    cell_order: Optional[np.ndarray] = None   # new -> original cell index
    mrf: Optional[MRFZone] = None           # or a list of zones (MRFCoupledZoneList order)
    translating: Optional[TranslatingZone] = None
    phase_lag: Optional[PhaseLagPair] = None


# --------------------------------------------------------------------------
# structured hex blocks
# --------------------------------------------------------------------------

_SIDES = ["-x", "+x", "-y", "+y", "-z", "+z"]


def _hex_topology(nx: int, ny: int, nz: int, patch_faces):
    """Faces of an nx*ny*nz block in blockMesh order (native generator of
    libicsf, icsf_block_mesh_hex)."""
    from ._lib import check, lib, ptr

    side_patch = np.full(6, -1, np.int32)
    for p, (_, _, sides) in enumerate(patch_faces):
        for sd in sides:
            side_patch[_SIDES.index(sd)] = p
    if (side_patch < 0).any():
        raise ValueError("every side of the block needs a patch")
    nI = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    nB = 2 * (ny * nz + nx * nz + nx * ny)
    fv = np.empty((nI + nB, 4), np.int32)
    own = np.empty(nI + nB, np.int32)
    nei = np.empty(nI, np.int32)
    psize = np.empty(len(patch_faces), np.int32)
    check(lib().icsf_block_mesh_hex(nx, ny, nz, ptr(side_patch), len(patch_faces), ptr(fv), ptr(own), ptr(nei),
                                    ptr(psize)))
    patches, start = [], nI
    for (name, kind, _), sz in zip(patch_faces, psize):
        patches.append(Patch(name, kind, start, int(sz)))
        start += int(sz)
    return fv, own, nei, patches


def _hex_topology_numpy(nx: int, ny: int, nz: int, patch_faces):
    """Faces of an nx*ny*nz block in blockMesh order (numpy statement of the
    same ordering, used to cross-check the native generator).

    patch_faces: list of (name, kind, sides) where sides is a list of
    '-x', '+x', '-y', '+y', '-z', '+z'.
    Returns face_verts [nF,4], owner, neighbour, patches.
    """
    NX, NY = nx + 1, ny + 1

    def pid(i, j, k):
        return i + NX * (j + NY * k)

    def cid(i, j, k):
        return i + nx * (j + ny * k)

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i = i.ravel(); j = j.ravel(); k = k.ravel()
    c = cid(i, j, k)

    def quad_x(ii, jj, kk):   # normal +x at x-index ii
        return np.stack([pid(ii, jj, kk), pid(ii, jj + 1, kk), pid(ii, jj + 1, kk + 1), pid(ii, jj, kk + 1)], 1)

    def quad_y(ii, jj, kk):   # normal +y at y-index jj
        return np.stack([pid(ii, jj, kk), pid(ii, jj, kk + 1), pid(ii + 1, jj, kk + 1), pid(ii + 1, jj, kk)], 1)

    def quad_z(ii, jj, kk):   # normal +z at z-index kk
        return np.stack([pid(ii, jj, kk), pid(ii + 1, jj, kk), pid(ii + 1, jj + 1, kk), pid(ii, jj + 1, kk)], 1)

    own_l, nei_l, fv_l = [], [], []
    m = i < nx - 1
    own_l.append(c[m]); nei_l.append(c[m] + 1); fv_l.append(quad_x(i[m] + 1, j[m], k[m]))
    m = j < ny - 1
    own_l.append(c[m]); nei_l.append(c[m] + nx); fv_l.append(quad_y(i[m], j[m] + 1, k[m]))
    m = k < nz - 1
    own_l.append(c[m]); nei_l.append(c[m] + nx * ny); fv_l.append(quad_z(i[m], j[m], k[m] + 1))
    own = np.concatenate(own_l); nei = np.concatenate(nei_l); fv = np.concatenate(fv_l)
    order = np.lexsort((nei, own))
    own, nei, fv = own[order], nei[order], fv[order]

    def side(s):
        if s == "-x":
            m = i == 0
            q = quad_x(i[m], j[m], k[m])[:, [0, 3, 2, 1]]
        elif s == "+x":
            m = i == nx - 1
            q = quad_x(i[m] + 1, j[m], k[m])
        elif s == "-y":
            m = j == 0
            q = quad_y(i[m], j[m], k[m])[:, [0, 3, 2, 1]]
        elif s == "+y":
            m = j == ny - 1
            q = quad_y(i[m], j[m] + 1, k[m])
        elif s == "-z":
            m = k == 0
            q = quad_z(i[m], j[m], k[m])[:, [0, 3, 2, 1]]
        elif s == "+z":
            m = k == nz - 1
            q = quad_z(i[m], j[m], k[m] + 1)
        else:
            raise ValueError(s)
        return c[m], q

    b_own, b_fv, patches = [], [], []
    start = own.shape[0]
    for name, kind, sides in patch_faces:
        size = 0
        for s in sides:
            o, q = side(s)
            b_own.append(o); b_fv.append(q); size += o.shape[0]
        patches.append(Patch(name, kind, start, size))
        start += size
    owner = np.concatenate([own] + b_own).astype(np.int32)
    face_verts = np.concatenate([fv] + b_fv).astype(np.int32)
    return face_verts, owner, nei.astype(np.int32), patches


def _mesh_from_quads(points, quads, owner, neighbour, patches, n_cells, grid=None) -> PolyMesh:
    nF = quads.shape[0]
    face_off = np.arange(0, 4 * (nF + 1), 4, dtype=np.int32)
    hint = None
    if grid is not None:
        # grid: points as [nz+1, ny+1, nx+1, 3]; average of the 8 vertices
        g = grid
        hint = (g[:-1, :-1, :-1] + g[:-1, :-1, 1:] + g[:-1, 1:, :-1] + g[:-1, 1:, 1:] +
                g[1:, :-1, :-1] + g[1:, :-1, 1:] + g[1:, 1:, :-1] + g[1:, 1:, 1:]).reshape(-1, 3) / 8.0
    return PolyMesh(points=np.ascontiguousarray(points, np.float64), face_off=face_off,
                    face_verts=np.ascontiguousarray(quads.ravel(), np.int32),
                    owner=owner, neighbour=neighbour, patches=patches, n_cells=n_cells, centre_hint=hint)


def hex_box(nx: int, ny: int, nz: int, lengths=(1.0, 1.0, 1.0), jitter: float = 0.0,
            seed: int = 0) -> PolyMesh:
    """Box [0,Lx]x[0,Ly]x[0,Lz]; interior nodes jittered by U(-jitter,jitter)*h
    along each axis in which they are interior (boundary nodes stay on their
    planes)."""
    Lx, Ly, Lz = lengths
    hx, hy, hz = Lx / nx, Ly / ny, Lz / nz
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    pts = np.stack([ii.ravel() * hx, jj.ravel() * hy, kk.ravel() * hz], 1).astype(np.float64)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        d = rng.uniform(-jitter, jitter, size=pts.shape)
        inner = np.stack([(ii.ravel() > 0) & (ii.ravel() < nx),
                          (jj.ravel() > 0) & (jj.ravel() < ny),
                          (kk.ravel() > 0) & (kk.ravel() < nz)], 1)
        pts += d * np.array([hx, hy, hz]) * inner
    fv, own, nei, patches = _hex_topology(
        nx, ny, nz,
        [("inlet", PATCH_PATCH, ["-x"]), ("outlet", PATCH_PATCH, ["+x"]),
         ("walls", PATCH_WALL, ["-y", "+y", "-z", "+z"])])
    return _mesh_from_quads(pts, fv, own, nei, patches, nx * ny * nz, pts.reshape(nz + 1, ny + 1, nx + 1, 3))


def bump_channel(nx: int = 200, ny: int = 50, height: float = 0.1) -> PolyMesh:
    """2-D channel x in [-1.5,1.5], y in [0,1], circular-arc bump of the given
    height on x in [-0.5,0.5] (geometry of the circularArcBump tutorial,
    uniform spacing), one cell thick in z in [-0.1, 0.1]."""
    chord = 1.0
    Rarc = (0.25 * chord * chord + height * height) / (2.0 * height)
    x = np.linspace(-1.5, 1.5, nx + 1)
    yb = np.where(np.abs(x) < 0.5 * chord,
                  np.sqrt(np.maximum(Rarc * Rarc - x * x, 0.0)) - (Rarc - height), 0.0)
    kk, jj, ii = np.meshgrid(np.arange(2), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    X = x[ii]
    Y = yb[ii] + (1.0 - yb[ii]) * jj / ny
    Z = np.where(kk == 0, -0.1, 0.1)
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float64)
    fv, own, nei, patches = _hex_topology(
        nx, ny, 1,
        [("inlet", PATCH_PATCH, ["-x"]), ("outlet", PATCH_PATCH, ["+x"]),
         ("lowerWall", PATCH_WALL, ["-y"]), ("upperWall", PATCH_WALL, ["+y"]),
         ("frontAndBack", PATCH_EMPTY, ["-z", "+z"])])
    return _mesh_from_quads(pts, fv, own, nei, patches, nx * ny, pts.reshape(2, ny + 1, nx + 1, 3))


def approx_cell_centres(mesh: PolyMesh) -> np.ndarray:
    """Approximate cell centres for seeding fields (input generation only):
    vertex average for generated blocks, else the average of face-vertex
    averages."""
    if mesh.centre_hint is not None:
        return mesh.centre_hint
    nF = mesh.n_faces
    fc = mesh.points[mesh.face_verts.reshape(nF, -1)].mean(axis=1) if \
        np.all(np.diff(mesh.face_off) == 4) else None
    if fc is None:
        raise ValueError("approx_cell_centres expects quad faces")
    n = mesh.n_cells
    idx = np.concatenate([mesh.owner, mesh.neighbour])
    cnt = np.bincount(idx, minlength=n).astype(np.float64)
    acc = np.stack([np.bincount(idx, weights=np.concatenate([fc[:, d], fc[:mesh.n_internal, d]]), minlength=n)
                    for d in range(3)], 1)
    return acc / cnt[:, None]


def face_centres_of(mesh: PolyMesh, start: int, size: int) -> np.ndarray:
    fv = mesh.face_verts.reshape(mesh.n_faces, -1)[start:start + size]
    return mesh.points[fv].mean(axis=1)


# --------------------------------------------------------------------------
# dictionaries (OpenFOAM syntax, the keyword subset the library reads)
# --------------------------------------------------------------------------

THERMO = """thermoType { type hePsiThermo; mixture pureMixture; transport const;
 thermo hConst; equationOfState perfectGas; specie specie; energy sensibleInternalEnergy; }
mixture { specie { molWeight 28.966; } thermodynamics { Cp %r; Hf 0; Tref 0; }
 transport { mu %r; Pr %r; } }
R %r; gamma %r;
"""


def thermo_dict(mu: float = 0.0, Pr: float = 0.72) -> str:
    return THERMO % (GAMMA * R_GAS / (GAMMA - 1.0), mu, Pr, R_GAS, GAMMA)


def schemes_dict(flux: str = "AUSMPlusUp", limiter: str = "cellLimited<Venkatakrishnan> Gauss linear 1",
                 recon: str = "linearUpwind limitedGrad", lowMach: bool = True,
                 entropyFix: float = 0.05) -> str:
    return f"""convectiveFluxScheme {{ fluxScheme {flux}; lowMachAusm {'true' if lowMach else 'false'}; entropyFixCoeff {entropyFix}; }}
viscousFluxScheme {{ LaxFriedrichJacobian true; }}
ddtSchemes {{ default dualTime rPseudoDeltaT steadyState; }}
gradSchemes {{ default Gauss linear; limitedGrad {limiter}; }}
divSchemes {{ default none; div(tauMC) Gauss linear; }}
laplacianSchemes {{ default Gauss linear corrected; }}
interpolationSchemes {{ default linear; reconstruct(rho) {recon}; reconstruct(U) {recon}; reconstruct(T) {recon}; }}
snGradSchemes {{ default corrected; }}
"""


def solution_dict(solver: str = "GMRES", precond: str = "LUSGS", nDirections: int = 30,
                  maxIter: int = 1, tol: float = 1e-12, relTol: float = 1e-2,
                  co: float = 50.0, coMin: Optional[float] = None, coMax: Optional[float] = None,
                  precision: str = "double", nPseudoCorr: int = 1, innerDirections: int = 30,
                  orthogonalisation: str = "MGS", fusedSweeps: bool = True, graphCapture: bool = True,
                  autotune: bool = True, jacobianStorage: str = "generators", fusedCellSums: bool = True) -> str:
    coMin = co if coMin is None else coMin
    coMax = co if coMax is None else coMax
    return f"""flowSolver {{ solver {solver}; {solver} {{ preconditioner {precond}; nDirections {nDirections};
 maxIter {maxIter}; minIter 0; tolerance {tol}; relTol {relTol}; precision {precision};
 innerDirections {innerDirections}; orthogonalisation {orthogonalisation};
 fusedSweeps {'true' if fusedSweeps else 'false'}; graphCapture {'true' if graphCapture else 'false'};
 autotune {'true' if autotune else 'false'}; jacobianStorage {jacobianStorage};
 fusedCellSums {'true' if fusedCellSums else 'false'}; }} }}
pseudoTime {{ nPseudoCorr {nPseudoCorr}; pseudoTol 1e-12; pseudoTolRel 1e-12; pseudoCoNum {co};
 pseudoCoNumMin {coMin}; pseudoCoNumMax {coMax}; localTimestepping true; localTimesteppingBounding true; }}
"""


# --------------------------------------------------------------------------
# seeded fields
# --------------------------------------------------------------------------

def fourier_noise(xy: np.ndarray, seed: int, n_modes: int = 8) -> np.ndarray:
    """sum_k sin(k.x + phi_k)/n_modes with seeded wave vectors and phases."""
    rng = np.random.default_rng(seed)
    kvec = rng.uniform(1.0, 3.0 * 2.0 * math.pi, size=(n_modes, xy.shape[1]))
    phi = rng.uniform(0.0, 2.0 * math.pi, size=n_modes)
    return np.sin(xy @ kvec.T + phi).sum(axis=1) / n_modes


def _state_from_mach(M, p, T):
    c = np.sqrt(GAMMA * R_GAS * T)
    rho = p / (R_GAS * T)
    return rho, M * c


def case_bump(nx: int = 200, ny: int = 50, seed: int = 1, flux: str = "AUSMPlusUp",
              co: float = 50.0, nDirections: int = 30, precision: str = "double",
              solver: str = "GMRES", precond: str = "LUSGS", bcs: str = "fixed") -> Case:
    """Inviscid transonic bump channel: M=0.675, p=1e5 Pa, T=300 K, plus a
    1% seeded 8-mode Fourier perturbation on rho, U and T.  bcs "fixed":
    fixed inflow velocity and temperature, fixed outlet pressure; "tutorial":
    the circularArcBump tutorial's set (tutorials/circularArcBump/transonic/
    0/{p,U,T}): totalPressure / pressureInletOutletVelocity / totalTemperature
    inlet at the total state of the freestream, fixed outlet pressure."""
    mesh = bump_channel(nx, ny)
    C = approx_cell_centres(mesh)
    s = 0.01 * fourier_noise(C[:, :2], seed)
    p0, T0, M0 = 1e5, 300.0, 0.675
    rho0, U0 = _state_from_mach(M0, p0, T0)
    rho = rho0 * (1 + s)
    T = T0 * (1 + s)
    U = np.zeros((mesh.n_cells, 3)); U[:, 0] = U0 * (1 + s)
    bcl = _bump_bcs(mesh, p0, U0, T0)
    if bcs == "tutorial":
        fac = 1.0 + 0.5 * (GAMMA - 1.0) * M0 * M0
        for i, pt in enumerate(mesh.patches):
            if pt.name == "inlet":
                Uv = np.zeros((pt.size, 3)); Uv[:, 0] = U0
                bcl[i] = total_inlet(pt.size, p0 * fac ** (GAMMA / (GAMMA - 1.0)), T0 * fac, U_init=Uv, T_init=T0)
    elif bcs != "fixed":
        raise ValueError(f"unknown boundary set {bcs!r}")
    bcs = bcl
    return Case("bump", mesh, rho, U, T, bcs, thermo_dict(0.0),
                schemes_dict(flux), solution_dict(solver=solver, precond=precond,
                                                  nDirections=nDirections, co=co,
                                                  precision=precision))


def _bump_bcs(mesh, p0, U0, T0):
    bcs = []
    for p in mesh.patches:
        if p.name == "inlet":
            Uv = np.zeros((p.size, 3)); Uv[:, 0] = U0
            bcs.append(BoundarySpec(BC_ZERO_GRADIENT, BC_FIXED_VALUE, BC_FIXED_VALUE,
                                    None, Uv, np.full(p.size, T0)))
        elif p.name == "outlet":
            bcs.append(BoundarySpec(BC_FIXED_VALUE, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT,
                                    np.full(p.size, p0), None, None))
        elif p.kind == PATCH_WALL:
            bcs.append(BoundarySpec(BC_ZERO_GRADIENT, BC_SLIP, BC_ZERO_GRADIENT))
        else:
            bcs.append(BoundarySpec(BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT))
    return bcs


def case_box(n=(32, 16, 16), lengths=(2.0, 1.0, 1.0), jitter: float = 0.1, seed: int = 7,
             flux: str = "ROE", mu: float = 1.8e-5, mach=(0.5, 0.9), p_jump: float = 2.0,
             noise: float = 0.02, co: float = 50.0, nDirections: int = 30,
             precision: str = "double", precond: str = "LUSGS", solver: str = "GMRES",
             innerDirections: int = 30, orthogonalisation: str = "MGS",
             name: str = "box") -> Case:
    """Jittered hex box with a linear Mach ramp along x, 2% smooth noise and
    an optional planar tanh pressure jump (ratio p_jump) at mid-length."""
    nx, ny, nz = n
    mesh = hex_box(nx, ny, nz, lengths, jitter, seed)
    C = approx_cell_centres(mesh)
    Lx = lengths[0]
    xi = C[:, 0] / Lx
    M = mach[0] + (mach[1] - mach[0]) * xi
    s = noise * fourier_noise(C / np.array(lengths), seed + 1)
    T0 = 300.0
    p = 1e5 * np.ones(mesh.n_cells)
    if p_jump != 1.0:
        delta = 0.05 * Lx
        p = 1e5 * (1.0 + 0.5 * (p_jump - 1.0) * (np.tanh((C[:, 0] - 0.5 * Lx) / delta) + 1.0))
    T = T0 * (1 + s)
    rho = p / (R_GAS * T)
    c = np.sqrt(GAMMA * R_GAS * T)
    U = np.zeros((mesh.n_cells, 3)); U[:, 0] = M * c * (1 + s)
    U[:, 1] = 0.05 * M * c * s
    rho0, U0 = _state_from_mach(mach[0], 1e5, T0)
    bcs = []
    for pt in mesh.patches:
        if pt.name == "inlet":
            Uv = np.zeros((pt.size, 3)); Uv[:, 0] = U0
            bcs.append(BoundarySpec(BC_ZERO_GRADIENT, BC_FIXED_VALUE, BC_FIXED_VALUE,
                                    None, Uv, np.full(pt.size, T0)))
        elif pt.name == "outlet":
            bcs.append(BoundarySpec(BC_FIXED_VALUE, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT,
                                    np.full(pt.size, 1e5 * p_jump), None, None))
        else:
            bcs.append(BoundarySpec(BC_ZERO_GRADIENT, BC_SLIP, BC_ZERO_GRADIENT))
    return Case(name, mesh, rho, U, T, bcs, thermo_dict(mu),
                schemes_dict(flux), solution_dict(solver=solver, precond=precond,
                                                  nDirections=nDirections, co=co,
                                                  precision=precision,
                                                  innerDirections=innerDirections,
                                                  orthogonalisation=orthogonalisation))


def total_inlet(size: int, p0, T0, U_init=None, T_init=None, p_init=None) -> BoundarySpec:
    """Inflow patch of the ICSFoam tutorials (tutorials/circularArcBump/
    transonic/0/{p,U,T}, INLE1): totalPressure p0, pressureInletOutletVelocity
    (zero tangential velocity), totalTemperature T0; p0/T0 scalars or per
    face, optional initial `value` entries."""
    def arr(v, shape):
        return None if v is None else np.broadcast_to(np.asarray(v, np.float64), shape).copy()
    return BoundarySpec(BC_TOTAL_PRESSURE, BC_PRESSURE_INLET_OUTLET_VELOCITY, BC_TOTAL_TEMPERATURE,
                        arr(p0, (size,)), None, arr(T0, (size,)),
                        p_init=arr(p_init, (size,)), U_init=arr(U_init, (size, 3)), T_init=arr(T_init, (size,)))


def permute_case(case: Case, new_to_old: np.ndarray, renumber) -> Case:
    """Apply a cell permutation (new -> old) through the supplied mesh
    renumbering routine and carry the cell fields along."""
    mesh = renumber(case.mesh, new_to_old)
    order = new_to_old if case.cell_order is None else case.cell_order[new_to_old]
    if case.mrf is None:
        mrf = None
    elif isinstance(case.mrf, (list, tuple)):
        mrf = [dataclasses.replace(z, cells=z.cells[new_to_old]) for z in case.mrf]
    else:
        mrf = dataclasses.replace(case.mrf, cells=case.mrf.cells[new_to_old])
    tr = None if case.translating is None else \
        dataclasses.replace(case.translating, cells=case.translating.cells[new_to_old])
    return dataclasses.replace(case, mesh=mesh, rho=case.rho[new_to_old],
                               U=case.U[new_to_old], T=case.T[new_to_old],
                               cell_order=order, mrf=mrf, translating=tr)


# --------------------------------------------------------------------------
# harmonic balance: annular sector with a periodic inflow
# --------------------------------------------------------------------------

def rotation_z(angle: float) -> np.ndarray:
    c, s = math.cos(angle), math.sin(angle)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def annulus_sector(nt: int, nr: int, nz: int, theta: float = np.pi / 8, r0: float = 0.5, r1: float = 1.0,
                   length: float = 1.0, jitter: float = 0.0, seed: int = 0, cyclic: bool = False,
                   jitter_period: Optional[int] = None, warp_planes: bool = False) -> PolyMesh:
    """Annular sector theta in [0, theta], r in [r0, r1], z in [0, length]
    (index directions r, theta, z: a right-handed map, cells ordered r
    fastest); interior nodes jittered in the index space.  Patches: inlet
    (z = 0), outlet (z = length), hub/shroud walls and the two theta planes:
    one slip-wall patch "sides", or (cyclic) the rotational cyclic pair
    cyclicLow (theta = 0, forwardT = rotation by -theta about z) and
    cyclicHigh (theta = theta, rotation by +theta).  With cyclic, the jitter
    repeats every jitter_period theta cells (default nt), so the theta = 0
    and theta = theta planes carry rotated copies of the same nodes; with
    warp_planes those planes are jittered in theta too (non-planar cyclic
    faces)."""
    kk, jj, ii = np.meshgrid(np.arange(nz + 1), np.arange(nt + 1), np.arange(nr + 1), indexing="ij")
    a, b, c = ii.astype(np.float64), jj.astype(np.float64), kk.astype(np.float64)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        tmask = (jj > 0) & (jj < nt)
        if cyclic:
            per = jitter_period or nt
            d = rng.uniform(-jitter, jitter, size=(3, nz + 1, per, nr + 1))[:, :, np.arange(nt + 1) % per, :]
            if warp_planes:
                tmask = np.ones_like(tmask)
            else:
                tmask &= (jj % per) != 0  # planes that are images of theta = 0 stay planar
        else:
            d = rng.uniform(-jitter, jitter, size=(3,) + a.shape)
        a = a + d[0] * ((ii > 0) & (ii < nr))
        b = b + d[1] * tmask
        c = c + d[2] * ((kk > 0) & (kk < nz))
    r = r0 + (r1 - r0) * a / nr
    th = theta * b / nt
    z = length * c / nz
    pts = np.stack([(r * np.cos(th)).ravel(), (r * np.sin(th)).ravel(), z.ravel()], 1).astype(np.float64)
    if cyclic:
        sides = [("cyclicLow", PATCH_CYCLIC, ["-y"]), ("cyclicHigh", PATCH_CYCLIC, ["+y"])]
    else:
        sides = [("sides", PATCH_WALL, ["-y", "+y"])]
    fv, own, nei, patches = _hex_topology(
        nr, nt, nz,
        [("inlet", PATCH_PATCH, ["-z"]), ("outlet", PATCH_PATCH, ["+z"]),
         ("hubShroud", PATCH_WALL, ["-x", "+x"])] + sides)
    if cyclic:
        lo, hi = len(patches) - 2, len(patches) - 1
        patches[lo] = dataclasses.replace(patches[lo], neighbour=hi, transform=rotation_z(-theta))
        patches[hi] = dataclasses.replace(patches[hi], neighbour=lo, transform=rotation_z(theta))
    return _mesh_from_quads(pts, fv, own, nei, patches, nt * nr * nz, pts.reshape(nz + 1, nt + 1, nr + 1, 3))


def case_sector(nt: int = 8, nr: int = 6, nz: int = 5, theta: float = np.pi / 8, copies: int = 1,
                jitter: float = 0.1, seed: int = 3, flux: str = "ROE", swirl: bool = True,
                recon: str = "linearUpwind limitedGrad", noise: float = 1.0, warp_planes: bool = False,
                mu: float = 0.0, **sol) -> Case:
    """Rotation-periodic state on an annular sector with rotational cyclic
    patches: `copies` rotated copies of a base sector of nt theta cells and
    angle theta (the mesh spans copies*theta).  Cell values of the base
    sector are seeded (density, temperature, radial/swirl/axial velocity in
    the local cylindrical frame); copy m carries the same values with the
    velocity rotated by m*theta, so every copy sees identical neighbours.
    Boundary conditions: fixed axial inflow velocity and temperature at z = 0,
    fixed pressure at z = length, slip hub/shroud walls; mu > 0: viscous."""
    mesh = annulus_sector(copies * nt, nr, nz, theta=copies * theta, jitter=jitter, seed=seed, cyclic=True,
                          jitter_period=nt, warp_planes=warp_planes)
    rng = np.random.default_rng(seed + 2)
    base = rng.standard_normal((nz, nt, nr, 5)) * noise
    k, j, i = np.meshgrid(np.arange(nz), np.arange(copies * nt), np.arange(nr), indexing="ij")
    v = base[k, j % nt, i]
    copy = j // nt
    rho = 1.2 * (1 + 0.05 * v[..., 0])
    T = 300.0 * (1 + 0.03 * v[..., 1])
    ur = 20.0 * v[..., 2] if swirl else 0.0 * v[..., 2]
    ut = 30.0 + 10.0 * v[..., 3] if swirl else 0.0 * v[..., 3]
    uz = 150.0 + 20.0 * v[..., 4]
    C = approx_cell_centres(mesh).reshape(nz, copies * nt, nr, 3)
    th0 = np.arctan2(C[..., 1], C[..., 0]) - theta * copy   # angle of the base-sector cell
    U = np.stack([ur * np.cos(th0) - ut * np.sin(th0), ur * np.sin(th0) + ut * np.cos(th0), uz], -1)
    for m_ in range(1, copies):
        sel = copy == m_
        U[sel] = U[sel] @ rotation_z(m_ * theta).T
    bcs: List[Optional[BoundarySpec]] = []
    for p in mesh.patches:
        if p.kind == PATCH_CYCLIC:
            bcs.append(None)
        elif p.name == "inlet":
            Uv = np.zeros((p.size, 3))
            Uv[:, 2] = 150.0
            bcs.append(BoundarySpec(BC_ZERO_GRADIENT, BC_FIXED_VALUE, BC_FIXED_VALUE, None, Uv,
                                    np.full(p.size, 300.0)))
        elif p.name == "outlet":
            bcs.append(BoundarySpec(BC_FIXED_VALUE, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT,
                                    np.full(p.size, 1.0e5), None, None))
        else:
            bcs.append(BoundarySpec(BC_ZERO_GRADIENT, BC_SLIP, BC_ZERO_GRADIENT))
    return Case(f"sector{copies}", mesh, rho.ravel(), U.reshape(-1, 3), T.ravel(), bcs, thermo_dict(mu),
                schemes_dict(flux, recon=recon), solution_dict(**sol))


def case_hb(n=(64, 64, 32), n_harmonics: int = 3, frequency_hz: float = 100.0, amplitude: float = 0.05,
            mach: float = 0.5, noise: float = 0.02, seed: int = 3, co: float = 20.0, nDirections: int = 30,
            jitter: float = 0.1, flux: str = "ROE", mu: float = 0.0, orthogonalisation: str = "MGS",
            cyclic: bool = True):
    """Harmonic-balance case on an annular sector with rotational cyclic
    theta patches (slip side walls with cyclic=False): 2 n_harmonics + 1
    instants over one period of frequency_hz (selectedPeriod = 1/f).  The
    inlet is the tutorials' inflow patch (totalPressure,
    pressureInletOutletVelocity, totalTemperature) with the total pressure of
    instant n perturbed by amplitude * sin(omega t_n + phi), phi seeded; the
    mean total state is that of the given axial Mach number at p = 1e5 Pa,
    T = 300 K, and the outlet holds the static pressure.  Initial state of
    every instant: axial flow at that Mach number with smooth noise (the
    density with its own phase of the perturbation).  Returns (instances,
    D, instants)."""
    from .solver import hb_operator

    omega = 2.0 * np.pi * frequency_hz
    nO = 2 * n_harmonics + 1
    D, t, _ = hb_operator([omega], [n_harmonics], nO, 1.0 / frequency_hz)
    mesh = annulus_sector(*n, jitter=jitter, seed=seed, cyclic=cyclic)
    C = approx_cell_centres(mesh)
    T0, p0 = 300.0, 1e5
    rho0, U0 = _state_from_mach(mach, p0, T0)
    fac = 1.0 + 0.5 * (GAMMA - 1.0) * mach * mach
    pt0, Tt0 = p0 * fac ** (GAMMA / (GAMMA - 1.0)), T0 * fac
    phase = np.random.default_rng(seed + 2).uniform(0.0, 2.0 * np.pi)
    s = noise * fourier_noise(C, seed + 1)
    instances = []
    for J in range(nO):
        ph = np.sin(omega * t[J])
        rho = rho0 * (1 + s * (1 + 0.5 * ph))
        T = T0 * (1 + 0.5 * s)
        U = np.zeros((mesh.n_cells, 3))
        U[:, 2] = U0 * (1 + s)
        U[:, 0] = 0.02 * U0 * s
        bcs: List[Optional[BoundarySpec]] = []
        for pt in mesh.patches:
            if pt.kind == PATCH_CYCLIC:
                bcs.append(None)
            elif pt.name == "inlet":
                Uv = np.zeros((pt.size, 3)); Uv[:, 2] = U0
                bcs.append(total_inlet(pt.size, pt0 * (1 + amplitude * np.sin(omega * t[J] + phase)), Tt0,
                                       U_init=Uv, T_init=T0))
            elif pt.name == "outlet":
                bcs.append(BoundarySpec(BC_FIXED_VALUE, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT,
                                        np.full(pt.size, p0), None, None))
            else:
                bcs.append(BoundarySpec(BC_ZERO_GRADIENT, BC_SLIP, BC_ZERO_GRADIENT))
        instances.append(Case(f"hb{J}", mesh, rho, U, T, bcs, thermo_dict(mu), schemes_dict(flux),
                              solution_dict(nDirections=nDirections, co=co,
                                            orthogonalisation=orthogonalisation)))
    return instances, D, t
