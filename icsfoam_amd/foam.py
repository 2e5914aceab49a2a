"""OpenFOAM case files: polyMesh and volFields (ASCII, optionally .gz).

Host mirror over the C ABI of libicsf.so (icsf_foam_*): what dbnsFoam reads
at start-up (createDynamicFvMesh.H: constant/polyMesh; createFields.H:51-145:
the p, U, T fields with their boundary conditions) and what it writes
(runTime.write(): the fields of a time directory).

  read_polymesh(dir)                  -> (PolyMesh, patch names)
  read_case(case_dir, ...)            -> Case (fields, boundary conditions,
                                         rotating zone)
  read_mrf(case_dir, n_cells, names)  -> MRFZone or None
                                         (constant/MRFProperties and the
                                         zone's polyMesh/cellZones entry)
  read_translating(case_dir, ...)     -> TranslatingZone or None
                                         (constant/MRFTranslatingProperties)
  write_polymesh(dir, mesh, names)
  write_fields(time_dir, mesh, names, fields, boundary)

Boundary-condition types mapped onto the library's kinds: zeroGradient,
fixedValue, slip, noSlip (U fixed at zero), totalPressure,
pressureInletOutletVelocity, totalTemperature, symmetryPlane/symmetry
(p, T zeroGradient; U slip); empty and cyclic patches take their constraint
types.  Other types raise unless the caller maps them (`overrides`).
"""
from __future__ import annotations

import ctypes as C
import os
import re
from typing import Dict, List, Optional, Tuple

import numpy as np

from ._lib import IcsfError, as_f64, as_i32, check, lib, ptr
from .cases import (BC_FIXED_VALUE, BC_PRESSURE_INLET_OUTLET_VELOCITY, BC_SLIP, BC_TOTAL_PRESSURE,
                    BC_TOTAL_TEMPERATURE, BC_ZERO_GRADIENT, PATCH_CYCLIC, PATCH_EMPTY, BoundarySpec, Case,
                    MRFZone, Patch, PhaseLagPair, PolyMesh, R_GAS, TranslatingZone)


def read_polymesh(poly_mesh_dir: str) -> Tuple[PolyMesh, List[str]]:
    """constant/polyMesh -> (PolyMesh, patch names)."""
    h = C.c_void_p()
    check(lib().icsf_foam_mesh_read(os.fsencode(poly_mesh_dir), C.byref(h)))
    try:
        sz = np.zeros(6, np.int32)
        check(lib().icsf_foam_mesh_sizes(h, ptr(sz)))
        nP, nF, nI, nC, nPa, nV = (int(x) for x in sz)
        pts = np.empty((nP, 3)); off = np.empty(nF + 1, np.int32); fv = np.empty(nV, np.int32)
        own = np.empty(nF, np.int32); nei = np.empty(nI, np.int32)
        ps, pz, pk, pn = (np.empty(nPa, np.int32) for _ in range(4))
        pT = np.empty((nPa, 9))
        check(lib().icsf_foam_mesh_get(h, ptr(pts), ptr(off), ptr(fv), ptr(own), ptr(nei), ptr(ps), ptr(pz),
                                       ptr(pk), ptr(pn), ptr(pT)))
        names = []
        lags = {}
        buf = C.create_string_buffer(256)
        for p in range(nPa):
            check(lib().icsf_foam_mesh_patch_name(h, p, buf, 256))
            names.append(buf.value.decode())
            on, cyl = np.zeros(1, np.int32), np.zeros(1, np.int32)
            ibpa, ax, ce = np.zeros(1), np.zeros(3), np.zeros(3)
            check(lib().icsf_foam_mesh_patch_phase_lag(h, p, ptr(on), ptr(ibpa), ptr(cyl), ptr(ax), ptr(ce), buf,
                                                       256))
            if on[0]:
                lags[p] = dict(ibpa=float(ibpa[0]), cyl_coords=bool(cyl[0]), axis=ax, centre=ce,
                               zone=buf.value.decode())
    finally:
        lib().icsf_foam_mesh_free(h)
    patches = [Patch(names[p], int(pk[p]), int(ps[p]), int(pz[p]), int(pn[p]),
                     pT[p].reshape(3, 3).copy() if pk[p] == PATCH_CYCLIC else None, lags.get(p))
               for p in range(nPa)]
    return PolyMesh(points=pts, face_off=off, face_verts=fv, owner=own, neighbour=nei, patches=patches,
                    n_cells=nC), names


def write_polymesh(poly_mesh_dir: str, mesh: PolyMesh, names: Optional[List[str]] = None) -> None:
    os.makedirs(poly_mesh_dir, exist_ok=True)
    names = names or [p.name for p in mesh.patches]
    start, size, kind = mesh.patch_arrays()
    nbr, T = mesh.coupling_arrays()
    cn = (C.c_char_p * len(names))(*[n.encode() for n in names])
    check(lib().icsf_foam_mesh_write(os.fsencode(poly_mesh_dir), mesh.points.shape[0], ptr(as_f64(mesh.points)),
                                     mesh.n_faces, ptr(as_i32(mesh.face_off)), ptr(as_i32(mesh.face_verts)),
                                     mesh.n_internal, ptr(as_i32(mesh.owner)), ptr(as_i32(mesh.neighbour)),
                                     len(names), C.cast(cn, C.c_void_p), ptr(start), ptr(size), ptr(kind),
                                     ptr(nbr), ptr(as_f64(T))))


class FoamField:
    """One volScalarField / volVectorField file."""

    def __init__(self, path: str, n_cells: int):
        h = C.c_void_p()
        check(lib().icsf_foam_field_read(os.fsencode(path), int(n_cells), C.byref(h)))
        self._h = h
        nc = C.c_int(0)
        check(lib().icsf_foam_field_internal(h, C.byref(nc), None))
        self.n_comp = nc.value
        v = np.empty(n_cells * self.n_comp)
        check(lib().icsf_foam_field_internal(h, None, ptr(v)))
        self.internal = v if self.n_comp == 1 else v.reshape(n_cells, 3)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().icsf_foam_field_free(h)
            self._h = None

    def patch_type(self, patch: str) -> str:
        buf = C.create_string_buffer(128)
        check(lib().icsf_foam_field_patch_type(self._h, patch.encode(), buf, 128))
        return buf.value.decode()

    def patch_entry(self, patch: str, key: str, n_faces: int) -> Optional[np.ndarray]:
        v = np.empty(n_faces * self.n_comp)
        try:
            check(lib().icsf_foam_field_patch_entry(self._h, patch.encode(), key.encode(), n_faces, ptr(v)))
        except IcsfError:
            return None
        return v if self.n_comp == 1 else v.reshape(n_faces, 3)


def _bc(pt: Patch, fp: FoamField, fU: FoamField, fT: FoamField, overrides: Dict[str, Dict[str, str]]):
    """BoundarySpec of one patch from its p, U, T entries."""
    n, name = pt.size, pt.name

    def typ(field, f):
        return overrides.get(field, {}).get(name) or f.patch_type(name)

    tp, tu, tt = typ("p", fp), typ("U", fU), typ("T", fT)
    if pt.kind == PATCH_CYCLIC:
        return None
    sym = ("symmetryPlane", "symmetry")
    spec = dict(p_type=None, U_type=None, T_type=None)
    vals = {}
    if tp in ("zeroGradient",) + sym:
        spec["p_type"] = BC_ZERO_GRADIENT
    elif tp == "fixedValue":
        spec["p_type"] = BC_FIXED_VALUE; vals["p_val"] = fp.patch_entry(name, "value", n)
    elif tp == "totalPressure":
        spec["p_type"] = BC_TOTAL_PRESSURE; vals["p_val"] = fp.patch_entry(name, "p0", n)
        vals["p_init"] = fp.patch_entry(name, "value", n)
    if tu in ("zeroGradient",):
        spec["U_type"] = BC_ZERO_GRADIENT
    elif tu in ("slip",) + sym:
        spec["U_type"] = BC_SLIP
    elif tu == "noSlip":
        spec["U_type"] = BC_FIXED_VALUE; vals["U_val"] = np.zeros((n, 3))
    elif tu == "fixedValue":
        spec["U_type"] = BC_FIXED_VALUE; vals["U_val"] = fU.patch_entry(name, "value", n)
    elif tu == "pressureInletOutletVelocity":
        spec["U_type"] = BC_PRESSURE_INLET_OUTLET_VELOCITY
        vals["U_val"] = fU.patch_entry(name, "tangentialVelocity", n)
        vals["U_init"] = fU.patch_entry(name, "value", n)
    if tt in ("zeroGradient",) + sym:
        spec["T_type"] = BC_ZERO_GRADIENT
    elif tt == "fixedValue":
        spec["T_type"] = BC_FIXED_VALUE; vals["T_val"] = fT.patch_entry(name, "value", n)
    elif tt == "totalTemperature":
        spec["T_type"] = BC_TOTAL_TEMPERATURE; vals["T_val"] = fT.patch_entry(name, "T0", n)
        vals["T_init"] = fT.patch_entry(name, "value", n)
    for field, t, k in (("p", tp, "p_type"), ("U", tu, "U_type"), ("T", tt, "T_type")):
        if spec[k] is None:
            raise IcsfError(-2, f"patch {name}: boundary condition {t} of {field} is not supported "
                                f"(map it with overrides)")
    for k, v in list(vals.items()):
        if v is None and k in ("p_val", "T_val") or (k == "U_val" and spec["U_type"] == BC_FIXED_VALUE and v is None):
            raise IcsfError(-1, f"patch {name}: missing value for {k}")
    return BoundarySpec(spec["p_type"], spec["U_type"], spec["T_type"], **vals)


def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _blocks(text: str) -> Dict[str, str]:
    """Top-level `name { ... }` entries of a dictionary file (FoamFile
    header excluded), brace-matched."""
    out, i = {}, 0
    for mt in re.finditer(r"([A-Za-z_][\w.:-]*)\s*\{", text):
        if mt.start() < i:
            continue
        depth, j = 1, mt.end()
        while j < len(text) and depth:
            depth += {"{": 1, "}": -1}.get(text[j], 0)
            j += 1
        if mt.group(1) != "FoamFile":
            out[mt.group(1)] = text[mt.end():j - 1]
        i = j
    return out


def _vector(body: str, key: str) -> np.ndarray:
    mt = re.search(r"\b" + key + r"\s*\(\s*([^)]*)\)\s*;", body)
    if not mt:
        raise IcsfError(-2, f"MRFProperties: missing {key}")
    return np.array([float(x) for x in mt.group(1).split()], np.float64)


def _cell_zone(poly_mesh_dir: str, name: str, n_cells: int) -> np.ndarray:
    """The cells of zone `name` of polyMesh/cellZones (cellLabels list)."""
    path = os.path.join(poly_mesh_dir, "cellZones")
    if not os.path.exists(path) and os.path.exists(path + ".gz"):
        import gzip
        with gzip.open(path + ".gz", "rt") as f:
            text = f.read()
    else:
        with open(path) as f:
            text = f.read()
    body = _blocks(_strip_comments(text)).get(name)
    if body is None:
        raise IcsfError(-2, f"cellZones: no zone {name}")
    mt = re.search(r"cellLabels\s+(?:List<label>\s*)?(\d+)\s*\(([^)]*)\)", body)
    if mt:
        labels = np.array(mt.group(2).split(), np.int64)
    else:  # uniform short form N{label}
        mt = re.search(r"cellLabels\s+(?:List<label>\s*)?(\d+)\s*\{\s*(\d+)\s*\}", body)
        if not mt:
            raise IcsfError(-2, f"cellZones: zone {name} has no readable cellLabels")
        labels = np.full(int(mt.group(1)), int(mt.group(2)), np.int64)
    if labels.shape[0] != int(mt.group(1)) or (labels.size and (labels.min() < 0 or labels.max() >= n_cells)):
        raise IcsfError(-1, f"cellZones: zone {name} has bad cell labels")
    cells = np.zeros(n_cells, bool)
    cells[labels] = True
    return cells


def _active_zones(path: str) -> List[Tuple[str, str]]:
    """The active top-level zone entries of a zone-list file, in file order:
    [(name, body)], empty without the file."""
    if not os.path.exists(path):
        return []
    with open(path) as f:
        text = _strip_comments(f.read())
    zones = []
    for zname, body in _blocks(text).items():
        mt = re.search(r"\bactive\s+(\w+)\s*;", body)
        if mt and mt.group(1).lower() in ("no", "false", "off", "n", "none"):
            continue
        zones.append((zname, body))
    return zones


def _active_zone(path: str, what: str):
    """The one active top-level zone entry of a zone-list file: (name,
    body), or None without the file or an active zone."""
    zones = _active_zones(path)
    if not zones:
        return None
    if len(zones) > 1:
        raise IcsfError(-1, f"{what}: one active zone is supported")
    return zones[0]


def _patches(body: str, key: str, names: List[str]) -> List[int]:
    """A wordRes entry (names or regular expressions) -> patch indices."""
    out = []
    mn = re.search(r"\b" + key + r"\s*\(([^)]*)\)\s*;", body)
    if mn:
        for pat in mn.group(1).split():
            pat = pat.strip('"')
            out += [i for i, n in enumerate(names) if re.fullmatch(pat, n) and i not in out]
    return sorted(out)


def read_translating(case_dir: str, n_cells: int, names: List[str]) -> Optional[TranslatingZone]:
    """constant/MRFTranslatingProperties (MRFTranslatingZone.C:240-296): the
    active zone's cellZone, frameVelocity (a constant: `frameVelocity (u v
    w);` or `frameVelocity constant (u v w);`) and nonMovingPatches; None
    without the file or an active zone."""
    z = _active_zone(os.path.join(case_dir, "constant", "MRFTranslatingProperties"), "MRFTranslatingProperties")
    if z is None:
        return None
    zname, body = z
    mt = re.search(r"\bcellZone\s+([\w.:-]+)\s*;", body)
    if not mt:
        raise IcsfError(-2, f"MRFTranslatingProperties {zname}: missing cellZone")
    cells = _cell_zone(os.path.join(case_dir, "constant", "polyMesh"), mt.group(1), n_cells)
    mv = re.search(r"\bframeVelocity\s+(?:constant\s+)?\(\s*([^)]*)\)\s*;", body)
    if not mv:
        raise IcsfError(-2, f"MRFTranslatingProperties {zname}: frameVelocity must be a constant vector")
    return TranslatingZone(cells=cells, velocity=np.array([float(x) for x in mv.group(1).split()], np.float64),
                           non_moving=_patches(body, "nonMovingPatches", names))


def read_mrf(case_dir: str, n_cells: int, names: List[str]):
    """constant/MRFProperties (IOMRFCoupledZoneList / MRFCoupledZone.C:
    240-296): each active zone's cellZone, origin, axis, omega (a constant:
    `omega 100;` or `omega constant 100;`) and nonRotatingPatches (names or
    regular expressions, matched against the patch names); None without the
    file or without an active zone, one MRFZone for one active zone, else
    the list of zones in file order (the zone list's order)."""
    zones = [_mrf_zone(case_dir, n_cells, names, zname, body)
             for zname, body in _active_zones(os.path.join(case_dir, "constant", "MRFProperties"))]
    if not zones:
        return None
    return zones[0] if len(zones) == 1 else zones


def _mrf_zone(case_dir: str, n_cells: int, names: List[str], zname: str, body: str) -> MRFZone:
    mt = re.search(r"\bcellZone\s+([\w.:-]+)\s*;", body)
    if not mt:
        raise IcsfError(-2, f"MRFProperties {zname}: missing cellZone")
    cells = _cell_zone(os.path.join(case_dir, "constant", "polyMesh"), mt.group(1), n_cells)
    mo = re.search(r"\bomega\s+(?:constant\s+)?([-+0-9.eE]+)\s*;", body)
    if not mo:
        raise IcsfError(-2, f"MRFProperties {zname}: omega must be a constant")
    return MRFZone(cells=cells, origin=_vector(body, "origin"), axis=_vector(body, "axis"),
                   omega=float(mo.group(1)), non_rotating=_patches(body, "nonRotatingPatches", names))


def read_hb_properties(case_dir: str) -> Dict[str, Dict]:
    """constant/HBProperties (IOHBZoneList; HBZone.C:46-91, 95-110): the
    active zones' frequenciesList (angular), harmonicsList and cylCoords,
    and system/fvSolution's harmonicBalance instantsNumber (default 3) and
    selectedPeriod (0 when absent: the condition-number scan).  {} without
    the file."""
    path = os.path.join(case_dir, "constant", "HBProperties")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        text = _strip_comments(f.read())
    sol_path = os.path.join(case_dir, "system", "fvSolution")
    hb = ""
    if os.path.exists(sol_path):
        with open(sol_path) as f:
            hb = _blocks(_strip_comments(f.read())).get("harmonicBalance", "")
    mi = re.search(r"\binstantsNumber\s+(\d+)\s*;", hb)
    mp = re.search(r"\bselectedPeriod\s+([-+0-9.eE]+)\s*;", hb)
    out = {}
    for zname, body in _blocks(text).items():
        mt = re.search(r"\bactive\s+(\w+)\s*;", body)
        if mt and mt.group(1).lower() in ("no", "false", "off", "n", "none"):
            continue
        mf = re.search(r"\bfrequenciesList\s*(?:\d+\s*)?\(([^)]*)\)\s*;", body)
        mh = re.search(r"\bharmonicsList\s*(?:\d+\s*)?\(([^)]*)\)\s*;", body)
        if not mf or not mh:
            raise IcsfError(-2, f"HBProperties {zname}: frequenciesList and harmonicsList are required")
        mc = re.search(r"\bcylCoords\s+(\w+)\s*;", body)
        out[zname] = dict(frequencies=np.array([float(x) for x in mf.group(1).split()], np.float64),
                          harmonics=np.array([int(x) for x in mh.group(1).split()], np.int32),
                          cyl_coords=bool(mc and mc.group(1).lower() in ("yes", "true", "on", "y")),
                          instants=int(mi.group(1)) if mi else 3,
                          selected_period=float(mp.group(1)) if mp else 0.0)
    return out


def read_phase_lag(case_dir: str, mesh: PolyMesh) -> Optional[PhaseLagPair]:
    """The phaseLagCyclic pair of the mesh (one pair supported) with its HB
    zone (HBZoneName of the pair's owner, the lower patch index): IBPA,
    cylCoords, rotation axis and centre of the owner; the zone's
    frequencies and harmonics and the instants of hb_operator.  None
    without phaseLagCyclic patches."""
    from .solver import hb_operator
    lag = [i for i, p in enumerate(mesh.patches) if p.phase_lag is not None]
    if not lag:
        return None
    owners = sorted({min(i, mesh.patches[i].neighbour) for i in lag})
    if len(owners) > 1:
        raise IcsfError(-1, "phaseLagCyclic: one phase-lagged pair is supported")
    own = mesh.patches[owners[0]]
    if own.phase_lag is None or mesh.patches[own.neighbour].phase_lag is None:
        raise IcsfError(-1, f"phaseLagCyclic: both sides of {own.name} must be phaseLagCyclic")
    zones = read_hb_properties(case_dir)
    z = zones.get(own.phase_lag["zone"])
    if z is None:
        raise IcsfError(-2, f"phaseLagCyclic {own.name}: HB zone {own.phase_lag['zone']!r} not in HBProperties")
    _, t, _ = hb_operator(z["frequencies"], z["harmonics"], z["instants"], z["selected_period"])
    pl = own.phase_lag
    return PhaseLagPair(owner=owners[0], ibpa=pl["ibpa"], frequencies=z["frequencies"], harmonics=z["harmonics"],
                        instants=t, cyl_coords=pl["cyl_coords"], axis=np.asarray(pl["axis"], np.float64),
                        centre=np.asarray(pl["centre"], np.float64))


def read_case(case_dir: str, time: str = "0", thermo: Optional[str] = None, fvSchemes: Optional[str] = None,
              fvSolution: Optional[str] = None, overrides: Optional[Dict[str, Dict[str, str]]] = None,
              R: float = R_GAS) -> Tuple[Case, List[str]]:
    """A case directory -> (Case, patch names): constant/polyMesh, the p, U,
    T files of `time` (rho = p/(R T), createFields.H), and the dictionaries
    (given text, else constant/thermophysicalProperties, system/fvSchemes,
    system/fvSolution).  overrides: {"T": {"outlet": "zeroGradient"}, ...}
    replaces boundary-condition types the library does not provide."""
    mesh, names = read_polymesh(os.path.join(case_dir, "constant", "polyMesh"))
    n = mesh.n_cells
    fp = FoamField(os.path.join(case_dir, time, "p"), n)
    fU = FoamField(os.path.join(case_dir, time, "U"), n)
    fT = FoamField(os.path.join(case_dir, time, "T"), n)
    overrides = overrides or {}
    bcs = [None if pt.kind == PATCH_EMPTY else _bc(pt, fp, fU, fT, overrides) for pt in mesh.patches]
    for i, pt in enumerate(mesh.patches):
        if pt.kind == PATCH_EMPTY:
            bcs[i] = BoundarySpec(BC_ZERO_GRADIENT, BC_ZERO_GRADIENT, BC_ZERO_GRADIENT)

    def text(given, *path):
        if given is not None:
            return given
        with open(os.path.join(case_dir, *path)) as f:
            return f.read()
    p = np.asarray(fp.internal, np.float64)
    T = np.asarray(fT.internal, np.float64)
    U = np.asarray(fU.internal, np.float64)
    case = Case(os.path.basename(os.path.abspath(case_dir)), mesh, p / (R * T), U.copy(), T.copy(), bcs,
                text(thermo, "constant", "thermophysicalProperties"), text(fvSchemes, "system", "fvSchemes"),
                text(fvSolution, "system", "fvSolution"), mrf=read_mrf(case_dir, n, names),
                translating=read_translating(case_dir, n, names), phase_lag=read_phase_lag(case_dir, mesh))
    return case, names


_TYPES_P = {BC_ZERO_GRADIENT: "zeroGradient", BC_FIXED_VALUE: "fixedValue", BC_TOTAL_PRESSURE: "totalPressure"}
_TYPES_U = {BC_ZERO_GRADIENT: "zeroGradient", BC_FIXED_VALUE: "fixedValue", BC_SLIP: "slip",
            BC_PRESSURE_INLET_OUTLET_VELOCITY: "pressureInletOutletVelocity"}
_TYPES_T = {BC_ZERO_GRADIENT: "zeroGradient", BC_FIXED_VALUE: "fixedValue", BC_TOTAL_TEMPERATURE: "totalTemperature"}
_DIMS = {"p": "[1 -1 -2 0 0 0 0]", "U": "[0 1 -1 0 0 0 0]", "T": "[0 0 0 1 0 0 0]", "rho": "[1 -3 0 0 0 0 0]"}


def _list(v, n, vec):
    """A nonuniform list entry body with 17 significant digits."""
    a = np.broadcast_to(np.asarray(v, np.float64), (n, 3) if vec else (n,))
    items = " ".join("(%s)" % " ".join(repr(float(x)) for x in r) for r in a) if vec else \
        " ".join(repr(float(x)) for x in a)
    return f"nonuniform List<{'vector' if vec else 'scalar'}> {n}({items})"


def _entries(key, typ, bc, n, gamma=1.4):
    """Keyword lines of the derived conditions (OpenFOAM names)."""
    if bc is None:
        return None
    if key == "p" and typ == "totalPressure":
        return (f"        p0              {_list(bc.p_val, n, False)};\n        rho             none;\n"
                f"        psi             thermo:psi;\n        gamma           {gamma!r};\n")
    if key == "U" and typ == "pressureInletOutletVelocity":
        tv = bc.U_val if bc.U_val is not None else np.zeros((n, 3))
        return f"        tangentialVelocity {_list(tv, n, True)};\n"
    if key == "T" and typ == "totalTemperature":
        return f"        T0              {_list(bc.T_val, n, False)};\n        gamma           {gamma!r};\n"
    return None


def write_fields(time_dir: str, mesh: PolyMesh, names: List[str], fields: Dict[str, np.ndarray],
                 bcs: Optional[List[Optional[BoundarySpec]]] = None,
                 boundary: Optional[Dict[str, np.ndarray]] = None) -> None:
    """Write p, U, T (and rho if given) of a state as volFields of a time
    directory (runTime.write()).  Patch types follow the case's boundary
    conditions (`calculated` where bcs is None); boundary values [nB] (as
    dbnsCase.boundary() returns) become the patches' value entries."""
    os.makedirs(time_dir, exist_ok=True)
    nI = mesh.n_internal
    for key, arr in fields.items():
        arr = as_f64(arr)
        ncomp = 3 if arr.ndim == 2 else 1
        types, values, sizes, entries = [], [], [], []
        for i, pt in enumerate(mesh.patches):
            sizes.append(pt.size)
            entries.append(None)
            bc = bcs[i] if bcs is not None else None
            if pt.kind == PATCH_EMPTY:
                types.append("empty"); values.append(None); continue
            if pt.kind == PATCH_CYCLIC:
                types.append("cyclic"); values.append(None); continue
            if bc is None:
                t = "calculated"
            elif key == "p":
                t = _TYPES_P[bc.p_type]
            elif key == "U":
                t = _TYPES_U[bc.U_type]
            elif key == "T":
                t = _TYPES_T[bc.T_type]
            else:
                t = "calculated"
            types.append(t)
            entries[-1] = _entries(key, t, bc, pt.size)
            bv = None
            if boundary is not None and key in boundary:
                b = np.asarray(boundary[key])
                bv = as_f64(b[pt.start - nI:pt.start - nI + pt.size])
            values.append(bv)
        cn = (C.c_char_p * len(names))(*[x.encode() for x in names])
        ct = (C.c_char_p * len(types))(*[x.encode() for x in types])
        keep = [v for v in values if v is not None]
        cv = (C.c_void_p * len(values))(*[None if v is None else v.ctypes.data for v in values])
        ce = (C.c_char_p * len(entries))(*[None if e is None else e.encode() for e in entries])
        check(lib().icsf_foam_field_write(os.fsencode(os.path.join(time_dir, key)), key.encode(),
                                          _DIMS.get(key, "[0 0 0 0 0 0 0]").encode(), ncomp, mesh.n_cells,
                                          ptr(arr), len(names), C.cast(cn, C.c_void_p), C.cast(ct, C.c_void_p),
                                          ptr(as_i32(sizes)), C.cast(cv, C.c_void_p), C.cast(ce, C.c_void_p)))
        del keep
