"""OpenFOAM case files (CPU): the polyMesh / volField reader and writer.

* The VKI-LS89 tutorial mesh (tests/golden/VKI-LS89, the tutorial's files
  gzip-compressed) reads with the sizes its own header note states, every
  cell closed with positive volume, and the translational cyclic pair
  matched face by face (equal areas, opposite normals, face centres one
  constant separation apart); where the reference tree is present the
  uncompressed original reads to the same arrays.
* Written meshes and fields read back identically (17 significant digits),
  including a rotational cyclic sector whose forwardT the reader recovers
  from the face normals.
* The LS89 0/ fields map onto the library's boundary conditions
  (totalPressure p0 160500, totalTemperature T0 420,
  pressureInletOutletVelocity, fixed outlet pressure, noSlip blade).
"""
import os

import numpy as np
import pytest

from icsfoam_amd import cases, foam
from icsfoam_amd.mesh import mesh_geometry

HERE = os.path.dirname(os.path.abspath(__file__))
LS89 = os.path.join(HERE, "golden", "VKI-LS89")
REF = "/root/reference/tutorials/VKI-LS89"


def test_ls89_mesh_reads_with_its_own_sizes():
    m, names = foam.read_polymesh(os.path.join(LS89, "constant", "polyMesh"))
    # owner header note: nPoints:57124 nCells:28059 nFaces:112739 nInternalFaces:55615
    assert (m.points.shape[0], m.n_cells, m.n_faces, m.n_internal) == (57124, 28059, 112739, 55615)
    assert names == ["blade", "inlet", "outlet", "Upper_periodicity", "Lower_periodicity", "front", "back"]
    kinds = [p.kind for p in m.patches]
    assert kinds == [cases.PATCH_WALL, cases.PATCH_PATCH, cases.PATCH_PATCH, cases.PATCH_CYCLIC, cases.PATCH_CYCLIC,
                     cases.PATCH_EMPTY, cases.PATCH_EMPTY]
    assert m.patches[3].neighbour == 4 and m.patches[4].neighbour == 3
    np.testing.assert_array_equal(m.patches[3].transform, np.eye(3))
    g = mesh_geometry(m)
    assert g["V"].min() > 0
    acc = np.zeros((m.n_cells, 3))
    np.add.at(acc, m.owner, g["Sf"])
    np.add.at(acc, m.neighbour, -g["Sf"][:m.n_internal])
    assert np.abs(acc).max() < 1e-12 * np.abs(g["Sf"]).max()
    up, lo = m.patches[3], m.patches[4]
    su, sl = g["Sf"][up.start:up.start + up.size], g["Sf"][lo.start:lo.start + lo.size]
    # coordinates carry 6 digits: the pair matches within the boundary file's matchTolerance (0.01)
    assert (np.linalg.norm(su + sl, axis=1) < 1e-3 * np.linalg.norm(su, axis=1)).all()
    d = g["Cf"][up.start:up.start + up.size] - g["Cf"][lo.start:lo.start + lo.size]
    np.testing.assert_allclose(d, np.tile(d.mean(axis=0), (up.size, 1)), rtol=0, atol=1e-6)
    np.testing.assert_allclose(d.mean(axis=0), [0, 0.0575, 0], atol=1e-6)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_ls89_gz_fixture_equals_original_files():
    a, na = foam.read_polymesh(os.path.join(LS89, "constant", "polyMesh"))
    b, nb = foam.read_polymesh(os.path.join(REF, "constant", "polyMesh"))
    assert na == nb
    for k in ("points", "face_off", "face_verts", "owner", "neighbour"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k))


def _same_mesh(a, b):
    for k in ("points", "face_off", "face_verts", "owner", "neighbour"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k))
    assert [(p.name, p.kind, p.start, p.size, p.neighbour) for p in a.patches] == \
        [(p.name, p.kind, p.start, p.size, p.neighbour) for p in b.patches]


@pytest.mark.parametrize("which", ["box", "sector"])
def test_mesh_write_read_round_trip(tmp_path, which):
    m = cases.hex_box(5, 4, 3, jitter=0.1, seed=2) if which == "box" else cases.annulus_sector(6, 4, 3)
    d = str(tmp_path / "polyMesh")
    foam.write_polymesh(d, m)
    r, names = foam.read_polymesh(d)
    assert names == [p.name for p in m.patches]
    _same_mesh(m, r)
    for p, q in zip(m.patches, r.patches):
        if p.kind == cases.PATCH_CYCLIC:
            np.testing.assert_allclose(q.transform, p.transform, rtol=0, atol=1e-13)


def test_fields_write_read_round_trip(tmp_path):
    c = cases.case_box(n=(4, 3, 3), jitter=0.1, noise=0.05)
    n0 = c.mesh.patches[0].size
    c.bcs[0] = cases.total_inlet(n0, 1.2e5, 310.0, U_init=np.tile([150.0, 0, 0], (n0, 1)), T_init=300.0)
    names = [p.name for p in c.mesh.patches]
    case_dir = tmp_path / "case"
    foam.write_polymesh(str(case_dir / "constant" / "polyMesh"), c.mesh)
    p = c.rho * cases.R_GAS * c.T
    nB = c.mesh.n_faces - c.mesh.n_internal
    rng = np.random.default_rng(0)
    bnd = {"p": rng.uniform(9e4, 1.1e5, nB), "U": rng.standard_normal((nB, 3)), "T": rng.uniform(290, 310, nB)}
    foam.write_fields(str(case_dir / "0"), c.mesh, names, {"p": p, "U": c.U, "T": c.T}, c.bcs, bnd)
    r, rn = foam.read_case(str(case_dir), thermo="", fvSchemes="", fvSolution="")
    assert rn == names
    np.testing.assert_array_equal(r.T, c.T)
    np.testing.assert_array_equal(r.U, c.U)
    np.testing.assert_allclose(r.rho, c.rho, rtol=1e-15)
    for a, b, pt in zip(r.bcs, c.bcs, c.mesh.patches):
        assert (a.p_type, a.U_type, a.T_type) == (b.p_type, b.U_type, b.T_type)
        sl = slice(pt.start - c.mesh.n_internal, pt.start - c.mesh.n_internal + pt.size)
        if b.p_type == cases.BC_FIXED_VALUE:
            np.testing.assert_array_equal(a.p_val, bnd["p"][sl])   # the written value entry
        if b.p_type == cases.BC_TOTAL_PRESSURE:
            np.testing.assert_array_equal(a.p_init, bnd["p"][sl])


def test_ls89_fields_map_to_boundary_conditions():
    c, names = foam.read_case(LS89, thermo=cases.thermo_dict(0.0), fvSchemes=cases.schemes_dict("ROE"),
                              fvSolution=cases.solution_dict(), overrides={"T": {"outlet": "zeroGradient"}})
    bc = dict(zip(names, c.bcs))
    assert (bc["inlet"].p_type, bc["inlet"].U_type, bc["inlet"].T_type) == \
        (cases.BC_TOTAL_PRESSURE, cases.BC_PRESSURE_INLET_OUTLET_VELOCITY, cases.BC_TOTAL_TEMPERATURE)
    np.testing.assert_array_equal(bc["inlet"].p_val, 160500.0)
    np.testing.assert_array_equal(bc["inlet"].T_val, 420.0)
    np.testing.assert_array_equal(bc["inlet"].U_init, np.tile([100.0, 0, 0], (105, 1)))
    np.testing.assert_array_equal(bc["outlet"].p_val, 82000.0)
    assert bc["blade"].U_type == cases.BC_FIXED_VALUE and not bc["blade"].U_val.any()
    np.testing.assert_array_equal(bc["blade"].T_val, 301.0)
    assert bc["Upper_periodicity"] is None
    np.testing.assert_array_equal(c.T, 400.0)
    np.testing.assert_allclose(c.rho, 1e5 / (cases.R_GAS * 400.0), rtol=1e-15)
    with pytest.raises(Exception, match="inletOutlet"):
        foam.read_case(LS89, thermo="", fvSchemes="", fvSolution="")


MRF_PROPS = """FoamFile { version 2.0; format ascii; class dictionary; object MRFProperties; }
// one rotating zone, one switched off
MRF1
{
    cellZone    rotor;
    active      yes;
    nonRotatingPatches (inlet "out.*" walls);
    origin    (0.1 0.2 -0.3);
    axis      (0 0 2);
    omega     constant 314.159;
}
MRF2
{
    cellZone    rotor;
    active      no;
    origin (0 0 0); axis (1 0 0); omega 1;
}
"""


def test_mrf_properties_and_cell_zones_read(tmp_path):
    """constant/MRFProperties and polyMesh/cellZones (the rotor zone's
    cellLabels over several lines) read into the rotating zone the library
    takes: the zone's cells, origin, axis, constant omega and the
    non-rotating patches (names and regular expressions); the case with
    the zone read from files gives the oracle the same residual as the
    zone given directly."""
    from oracle.oracle import OracleCase

    c = cases.case_box(n=(4, 3, 3), jitter=0.1, noise=0.05)
    names = [p.name for p in c.mesh.patches]
    case_dir = tmp_path / "case"
    pm = case_dir / "constant" / "polyMesh"
    foam.write_polymesh(str(pm), c.mesh)
    p = c.rho * cases.R_GAS * c.T
    nB = c.mesh.n_faces - c.mesh.n_internal
    bnd = {"p": np.full(nB, 1e5), "U": np.zeros((nB, 3)), "T": np.full(nB, 300.0)}
    foam.write_fields(str(case_dir / "0"), c.mesh, names, {"p": p, "U": c.U, "T": c.T}, c.bcs, bnd)
    zone = np.arange(c.mesh.n_cells) % 3 == 1
    labels = "\n".join(str(i) for i in np.flatnonzero(zone))
    (pm / "cellZones").write_text("FoamFile { version 2.0; format ascii; class regIOobject; object cellZones; }\n"
                                  f"1\n(\nrotor\n{{\n    type cellZone;\ncellLabels List<label>\n{zone.sum()}\n(\n"
                                  f"{labels}\n)\n;\n}}\n)\n")
    (case_dir / "constant" / "MRFProperties").write_text(MRF_PROPS)
    r, _ = foam.read_case(str(case_dir), thermo=c.thermo, fvSchemes=c.fvSchemes, fvSolution=c.fvSolution)
    z = r.mrf
    np.testing.assert_array_equal(z.cells, zone)
    np.testing.assert_array_equal(z.origin, [0.1, 0.2, -0.3])
    np.testing.assert_array_equal(z.axis, [0, 0, 2])
    assert z.omega == 314.159
    assert z.non_rotating == sorted(names.index(n) for n in ("inlet", "outlet", "walls"))
    # the same zone given directly: the same oracle residual
    direct = cases.MRFZone(cells=zone, origin=np.array([0.1, 0.2, -0.3]), axis=np.array([0, 0, 2.0]),
                           omega=314.159, non_rotating=z.non_rotating)
    b1, _ = OracleCase(r).residual()
    import dataclasses
    b2, _ = OracleCase(dataclasses.replace(r, mrf=direct)).residual()
    np.testing.assert_array_equal(b1, b2)
    b0, _ = OracleCase(dataclasses.replace(r, mrf=None)).residual()
    assert np.abs(b1 - b0).max() > 0
    assert r.translating is None
    # a translating zone beside it (MRFTranslatingProperties)
    (case_dir / "constant" / "MRFTranslatingProperties").write_text(
        "FoamFile { version 2.0; format ascii; class dictionary; object MRFTranslatingProperties; }\n"
        "slide { cellZone rotor; active true; nonMovingPatches (\"in.*\" outlet walls); "
        "frameVelocity constant (12.5 -3 0.25); }\n")
    r2, _ = foam.read_case(str(case_dir), thermo=c.thermo, fvSchemes=c.fvSchemes, fvSolution=c.fvSolution)
    t = r2.translating
    np.testing.assert_array_equal(t.cells, zone)
    np.testing.assert_array_equal(t.velocity, [12.5, -3.0, 0.25])
    assert t.non_moving == z.non_rotating
    bt, _ = OracleCase(r2).residual()
    assert np.abs(bt - b1).max() > 0
    # two active rotating zones: the zone list in file order
    (case_dir / "constant" / "MRFProperties").write_text(MRF_PROPS.replace("active      no;", "active yes;"))
    r3, _ = foam.read_case(str(case_dir), thermo=c.thermo, fvSchemes=c.fvSchemes, fvSolution=c.fvSolution)
    assert isinstance(r3.mrf, list) and len(r3.mrf) == 2
    assert r3.mrf[0].omega == 314.159 and r3.mrf[1].omega == 1.0
    np.testing.assert_array_equal(r3.mrf[1].axis, [1, 0, 0])


def test_phase_lag_cyclic_patches_and_hb_properties_read(tmp_path):
    """A boundary file whose rotational cyclic pair is phaseLagCyclic
    (IBPA, HBZoneName, cylCoords, rotationCentre), constant/HBProperties
    and fvSolution's harmonicBalance controls read into the phase-lagged
    pair the library takes: the owner side (the lower patch index) with its
    IBPA, cylCoords and rotation axis and centre, the zone's frequencies
    and harmonics, and the equispaced instants of selectedPeriod."""
    import re

    from icsfoam_amd.solver import hb_operator

    insts, D, t = cases.case_hb(n=(6, 5, 4), n_harmonics=2, frequency_hz=100.0)
    m = insts[0].mesh
    case_dir = tmp_path / "hb"
    pm = case_dir / "constant" / "polyMesh"
    foam.write_polymesh(str(pm), m)
    text = (pm / "boundary").read_text()
    text, n = re.subn(r"type\s+cyclic;", "type phaseLagCyclic; IBPA 0.6981; HBZoneName blade; cylCoords true;",
                      text)
    assert n == 2
    text, n = re.subn(r"rotationCentre\s*\(0 0 0\);", "rotationCentre (0 0 0.25);", text)
    assert n == 2
    (pm / "boundary").write_text(text)
    (case_dir / "system").mkdir()
    (case_dir / "system" / "fvSolution").write_text(
        "FoamFile { version 2.0; format ascii; class dictionary; object fvSolution; }\n"
        "harmonicBalance { instantsNumber 5; selectedPeriod 0.01; }\n")
    w = 2 * np.pi * 100.0
    (case_dir / "constant" / "HBProperties").write_text(
        "FoamFile { version 2.0; format ascii; class dictionary; object HBProperties; }\n"
        f"blade {{ active true; allMesh true; frequenciesList ({w!r}); harmonicsList (2); cylCoords true; }}\n"
        "other { active false; frequenciesList (1); harmonicsList (1); cylCoords false; }\n")
    r, names = foam.read_polymesh(str(pm))
    lo = next(i for i, p in enumerate(r.patches) if p.kind == cases.PATCH_CYCLIC)
    hi = r.patches[lo].neighbour
    assert hi > lo and r.patches[lo].phase_lag["zone"] == "blade"
    z = foam.read_hb_properties(str(case_dir))
    assert list(z) == ["blade"] and z["blade"]["instants"] == 5 and z["blade"]["selected_period"] == 0.01
    pl = foam.read_phase_lag(str(case_dir), r)
    assert pl.owner == lo and pl.ibpa == 0.6981 and pl.cyl_coords
    np.testing.assert_array_equal(pl.frequencies, [w])
    np.testing.assert_array_equal(pl.harmonics, [2])
    np.testing.assert_array_equal(pl.centre, [0, 0, 0.25])
    # the writer's rotationAxis (the pair's rotation about z)
    assert pl.axis[0] == 0 and pl.axis[1] == 0 and pl.axis[2] != 0
    np.testing.assert_allclose(pl.instants, hb_operator([w], [2], 5, 0.01)[1], rtol=0, atol=0)
    np.testing.assert_allclose(pl.instants, t, rtol=0, atol=1e-15)
    # a phaseLagCyclic patch needs cylCoords (phaseLagCyclicPolyPatch.C:89)
    (pm / "boundary").write_text(text.replace("cylCoords true;", "", 1))
    with pytest.raises(Exception, match="cylCoords"):
        foam.read_polymesh(str(pm))
