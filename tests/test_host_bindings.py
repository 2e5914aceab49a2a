"""Host bindings keep converted arrays alive through the C call.

The Python mirror converts inputs to the ABI's int32 / float64 C-contiguous
layout in the argument list of each call; the converted temporaries must
outlive the C call that receives their pointers.  Every case here passes
arrays that need a conversion (int64 addressing, float32 or Fortran-ordered
points) and checks the result against the call with native-layout inputs.
These are host-only entry points (no GPU use)."""
import dataclasses

import numpy as np

from icsfoam_amd import cases
from icsfoam_amd.mesh import colour_cells, mesh_geometry, rcm_order, renumber_mesh, sweep_levels


def _mesh():
    return cases.hex_box(14, 11, 9, jitter=0.1, seed=5)


def test_int64_addressing_matches_int32():
    m = _mesh()
    m64 = dataclasses.replace(m, owner=m.owner.astype(np.int64), neighbour=m.neighbour.astype(np.int64),
                              face_off=m.face_off.astype(np.int64), face_verts=m.face_verts.astype(np.int64))
    for _ in range(3):  # repeated conversions reuse freed memory if a temporary dies early
        assert np.array_equal(rcm_order(m64), rcm_order(m))
        c64, n64 = colour_cells(m64)
        c32, n32 = colour_cells(m)
        assert n64 == n32 and np.array_equal(c64, c32)
        f64, b64, nf64, nb64 = sweep_levels(m64)
        f32, b32, nf32, nb32 = sweep_levels(m)
        assert (nf64, nb64) == (nf32, nb32) and np.array_equal(f64, f32) and np.array_equal(b64, b32)
    perm = np.random.default_rng(2).permutation(m.n_cells)  # int64 permutation
    r64 = renumber_mesh(m64, perm)
    r32 = renumber_mesh(m, perm.astype(np.int32))
    assert np.array_equal(r64.owner, r32.owner) and np.array_equal(r64.neighbour, r32.neighbour)
    assert np.array_equal(r64.face_verts, r32.face_verts)


def test_converted_points_match_float64():
    m = _mesh()
    g64 = mesh_geometry(m)
    mf = dataclasses.replace(m, points=np.asfortranarray(m.points))  # needs a contiguity conversion
    gf = mesh_geometry(mf)
    for k in g64:
        assert np.array_equal(gf[k], g64[k]), k
    p32 = m.points.astype(np.float32)
    g32 = mesh_geometry(dataclasses.replace(m, points=p32))
    gref = mesh_geometry(dataclasses.replace(m, points=p32.astype(np.float64)))
    for k in gref:
        assert np.array_equal(g32[k], gref[k]), k
