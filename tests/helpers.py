"""Shared helpers of the tests: small seeded cases prepared (permuted, RCM,
colour-major) through the library's host-side renumbering."""
import numpy as np

from icsfoam_amd import cases
from icsfoam_amd.mesh import prepare_case


def bump_small(nx=40, ny=10, **kw):
    c = cases.case_bump(nx, ny, **kw)
    return prepare_case(c, permute_seed=11)


def box_small(n=(10, 6, 5), mu=1.8e-5, flux="ROE", **kw):
    c = cases.case_box(n=n, mu=mu, flux=flux, **kw)
    return prepare_case(c, permute_seed=13)


def rel_err(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
