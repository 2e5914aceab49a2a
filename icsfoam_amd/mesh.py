"""Host-side mesh preprocessing through the C ABI: RCM renumbering, cell
colouring, colour-major ordering, polyMesh renumbering, sweep levels and
OpenFOAM geometry.  Integer work only (plus the static geometry); these
calls never touch the GPU."""
from __future__ import annotations

import dataclasses
from typing import Optional, Tuple

import numpy as np

from ._lib import as_f64, as_i32, check, lib, ptr
from .cases import Case, PolyMesh, permute_case


def rcm_order(mesh: PolyMesh) -> np.ndarray:
    """Reverse Cuthill-McKee order (new -> old) of the mesh cells; on meshes
    with cyclic patches the cells across cyclic faces count as neighbours
    (renumberMesh's cellCells)."""
    out = np.empty(mesh.n_cells, np.int32)
    if mesh.has_cyclic:
        co, cn = mesh.coupled_cells()
        check(lib().icsf_rcm_order_coupled(mesh.n_cells, mesh.n_internal, ptr(as_i32(mesh.owner)),
                                           ptr(as_i32(mesh.neighbour)), int(co.shape[0]), ptr(co), ptr(cn),
                                           ptr(out)))
    else:
        check(lib().icsf_rcm_order(mesh.n_cells, mesh.n_internal, ptr(as_i32(mesh.owner)),
                                   ptr(as_i32(mesh.neighbour)), ptr(out)))
    return out


def colour_cells(mesh: PolyMesh) -> Tuple[np.ndarray, int]:
    col = np.empty(mesh.n_cells, np.int32)
    n = np.zeros(1, np.int32)
    check(lib().icsf_colour_cells(mesh.n_cells, mesh.n_internal, ptr(as_i32(mesh.owner)),
                                  ptr(as_i32(mesh.neighbour)), ptr(col), ptr(n)))
    return col, int(n[0])


def colour_major_order(colour: np.ndarray, n_colours: int) -> np.ndarray:
    colour = as_i32(colour)
    out = np.empty(colour.shape[0], np.int32)
    check(lib().icsf_colour_major_order(colour.shape[0], ptr(colour), n_colours, ptr(out)))
    return out


def renumber_mesh(mesh: PolyMesh, new_to_old: np.ndarray) -> PolyMesh:
    """Apply a cell permutation (OpenFOAM renumberMesh semantics)."""
    new_to_old = as_i32(new_to_old)
    nF, nI = mesh.n_faces, mesh.n_internal
    fv = np.empty(mesh.face_verts.shape[0], np.int32)
    own = np.empty(nF, np.int32)
    nei = np.empty(nI, np.int32)
    fmap = np.empty(nF, np.int32)
    check(lib().icsf_renumber_mesh(mesh.n_cells, nF, nI, ptr(as_i32(mesh.face_off)), ptr(as_i32(mesh.face_verts)),
                                   ptr(as_i32(mesh.owner)), ptr(as_i32(mesh.neighbour)), ptr(new_to_old),
                                   ptr(fv), ptr(own), ptr(nei), ptr(fmap)))
    old = np.where(fmap < 0, -fmap - 1, fmap)
    sizes = np.diff(mesh.face_off)[old]
    off = np.zeros(nF + 1, np.int32)
    np.cumsum(sizes, out=off[1:])
    hint = None if mesh.centre_hint is None else mesh.centre_hint[new_to_old]
    return dataclasses.replace(mesh, face_off=off, face_verts=fv, owner=own, neighbour=nei, centre_hint=hint)


def sweep_levels(mesh: PolyMesh):
    f = np.empty(mesh.n_cells, np.int32)
    b = np.empty(mesh.n_cells, np.int32)
    nf = np.zeros(1, np.int32)
    nb = np.zeros(1, np.int32)
    check(lib().icsf_sweep_levels(mesh.n_cells, mesh.n_internal, ptr(as_i32(mesh.owner)),
                                  ptr(as_i32(mesh.neighbour)), ptr(f), ptr(b), ptr(nf), ptr(nb)))
    return f, b, int(nf[0]), int(nb[0])


def mesh_geometry(mesh: PolyMesh) -> dict:
    nC, nF = mesh.n_cells, mesh.n_faces
    out = {k: np.empty(s) for k, s in [("C", (nC, 3)), ("V", nC), ("Sf", (nF, 3)), ("Cf", (nF, 3)),
                                      ("w", nF), ("dc", nF), ("nodc", nF), ("corr", (nF, 3))]}
    check(lib().icsf_mesh_geometry(mesh.points.shape[0], ptr(as_f64(mesh.points)), nF, ptr(as_i32(mesh.face_off)),
                                   ptr(as_i32(mesh.face_verts)), mesh.n_internal, ptr(as_i32(mesh.owner)),
                                   ptr(as_i32(mesh.neighbour)), ptr(out["C"]), ptr(out["V"]), ptr(out["Sf"]),
                                   ptr(out["Cf"]), ptr(out["w"]), ptr(out["dc"]), ptr(out["nodc"]),
                                   ptr(out["corr"])))
    return out


def cyclic_geometry(mesh: PolyMesh) -> dict:
    """Interpolation weights, deltaCoeffs, nonOrthDeltaCoeffs, correction
    vectors of every face (cyclic faces with their coupled definitions) and
    the cell across each boundary face (-1 off cyclic patches)."""
    nF, nB = mesh.n_faces, mesh.n_faces - mesh.n_internal
    start, size, kind = mesh.patch_arrays()
    nbr, T = mesh.coupling_arrays()
    out = {k: np.empty(s) for k, s in [("w", nF), ("dc", nF), ("nodc", nF), ("corr", (nF, 3))]}
    out["cell"] = np.empty(nB, np.int32)
    check(lib().icsf_mesh_geometry_patches(mesh.points.shape[0], ptr(as_f64(mesh.points)), nF,
                                           ptr(as_i32(mesh.face_off)), ptr(as_i32(mesh.face_verts)),
                                           mesh.n_internal, ptr(as_i32(mesh.owner)), ptr(as_i32(mesh.neighbour)),
                                           len(mesh.patches), ptr(start), ptr(size), ptr(kind), ptr(nbr),
                                           ptr(as_f64(T)), ptr(out["w"]), ptr(out["dc"]), ptr(out["nodc"]),
                                           ptr(out["corr"]), ptr(out["cell"])))
    return out


def prepare_case(case: Case, permute_seed: Optional[int] = None, rcm: bool = True,
                 colour_major: bool = True) -> Case:
    """Random permutation (seeded), then RCM, then colour-major numbering.

    With colour-major numbering the sequential LU-SGS sweep of the reference
    has one dependency level per colour (two for hex meshes), so the level
    schedule on the GPU is exact and short."""
    if permute_seed is not None:
        perm = np.random.default_rng(permute_seed).permutation(case.mesh.n_cells).astype(np.int32)
        case = permute_case(case, perm, renumber_mesh)
    if rcm:
        case = permute_case(case, rcm_order(case.mesh), renumber_mesh)
    if colour_major:
        col, n = colour_cells(case.mesh)
        case = permute_case(case, colour_major_order(col, n), renumber_mesh)
    return case
