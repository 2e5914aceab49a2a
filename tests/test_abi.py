"""The C-ABI library loads without a GPU and exports every entry point that
include/icsf.h declares (no compute calls)."""
import os
import re
import subprocess

from icsfoam_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "icsf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(icsf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "icsf_pseudo_steps" in syms and "icsf_mesh_create" in syms
    assert len(syms) >= 25


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (icsf_[a-z0-9_]+)", out))
    assert set(declared_symbols()) <= exported


def test_python_signatures_cover_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_no_device_is_reported_cleanly():
    """Without a GPU the device entry points fail with ICSF_ERR_NODEVICE
    (they never fall back to a CPU path)."""
    import ctypes as C

    import numpy as np

    from icsfoam_amd import cases

    try:
        import torch
        has_gpu = torch.cuda.device_count() > 0
    except Exception:  # pragma: no cover
        has_gpu = False
    if has_gpu:
        return
    m = cases.hex_box(2, 2, 2)
    start, size, kind = m.patch_arrays()
    h = C.c_void_p()
    lib = _lib.lib()
    rc = lib.icsf_mesh_create(m.points.shape[0], _lib.ptr(m.points), m.n_faces, _lib.ptr(m.face_off),
                              _lib.ptr(m.face_verts), m.n_internal, _lib.ptr(m.owner), _lib.ptr(m.neighbour),
                              len(m.patches), _lib.ptr(start), _lib.ptr(size), _lib.ptr(kind), None, None,
                              C.byref(h))
    assert rc == _lib.ERR_NODEVICE
    assert b"no HIP device" in lib.icsf_last_error()
    assert np.all(np.asarray([0]) == 0)
