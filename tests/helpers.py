"""Shared helpers of the tests: small seeded cases prepared (permuted, RCM,
colour-major) through the library's host-side renumbering."""
import numpy as np

from icsfoam_amd import cases
from oracle.oracle import OracleMesh
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


def ausm_inviscid(c, limiter=None, recon=None):
    """The same case with the inviscid AUSM+up scheme (optionally another
    limiter / reconstruction)."""
    import dataclasses
    kw = {}
    if limiter is not None:
        kw["limiter"] = limiter
    if recon is not None:
        kw["recon"] = recon
    return dataclasses.replace(c, thermo=cases.thermo_dict(0.0), fvSchemes=cases.schemes_dict("AUSMPlusUp", **kw))


def roe(c, mu=None, limiter=None, recon=None):
    """The same case with the Roe scheme: viscous with viscosity mu (None:
    the case's own), optionally another limiter / reconstruction."""
    import dataclasses
    kw = {}
    if limiter is not None:
        kw["limiter"] = limiter
    if recon is not None:
        kw["recon"] = recon
    c = dataclasses.replace(c, fvSchemes=cases.schemes_dict("ROE", **kw))
    if mu is not None:
        c = dataclasses.replace(c, thermo=cases.thermo_dict(mu))
    return c


def hllc(c, mu=None, limiter=None, recon=None):
    """The same case with the HLLC scheme (viscosity mu; None: the case's
    own), optionally another limiter / reconstruction."""
    import dataclasses
    kw = {}
    if limiter is not None:
        kw["limiter"] = limiter
    if recon is not None:
        kw["recon"] = recon
    c = dataclasses.replace(c, fvSchemes=cases.schemes_dict("HLLC", **kw))
    if mu is not None:
        c = dataclasses.replace(c, thermo=cases.thermo_dict(mu))
    return c


def with_mu(c, mu):
    """The same case with viscosity mu (0: inviscid)."""
    import dataclasses
    return dataclasses.replace(c, thermo=cases.thermo_dict(mu))


def generic_state(W, seed, amp=1e-3):
    """W with small uncorrelated per-cell noise on all five components, so
    that no two neighbouring values tie and no velocity component is exactly
    zero (ties and exact zeros are points where the residual has no
    derivative)."""
    rng = np.random.default_rng(seed)
    W = np.array(W, dtype=np.float64, copy=True)
    scale = np.abs(W[2:]).max()
    W[2:] += amp * scale * rng.standard_normal(W[2:].shape)
    W[:2] *= 1 + amp * rng.standard_normal(W[:2].shape)
    return W


def smooth_field(X, rng, ncomp, modes=4, k=3.0):
    """ncomp smooth random fields (sums of plane waves) at points X [n,3]."""
    out = np.zeros((ncomp, X.shape[0]))
    for q in range(ncomp):
        kv = rng.standard_normal((modes, 3)) * k
        ph = rng.uniform(0, 2 * np.pi, modes)
        out[q] = sum(np.sin(X @ kv[j] + ph[j]) for j in range(modes)) / modes
    return out


def linear_velocity_case(mu, R=287.0):
    """Uniform orthogonal box, constant rho and T, U = A x + u0 with a
    non-symmetric A, fixedValue U (the linear field at the face centres) and
    T on every patch; returns (case, oracle geometry, A)."""
    c = cases.case_box(n=(6, 5, 4), lengths=(1.0, 0.8, 0.6), jitter=0.0, flux="ROE", mu=mu, p_jump=1.0,
                       noise=0.0)
    geo = OracleMesh(c.mesh).geometry()
    A = np.array([[3.0, 7.0, -2.0], [-5.0, 1.0, 4.0], [6.0, -8.0, 2.0]])
    u0 = np.array([150.0, -40.0, 25.0])
    T0, p0 = 300.0, 1e5
    c.rho[:] = p0 / (R * T0)
    c.T[:] = T0
    c.U[:] = geo["C"] @ A.T + u0
    bcs = []
    for pt in c.mesh.patches:
        Cf = geo["Cf"][pt.start:pt.start + pt.size]
        bcs.append(cases.BoundarySpec(cases.BC_ZERO_GRADIENT, cases.BC_FIXED_VALUE, cases.BC_FIXED_VALUE,
                                      None, Cf @ A.T + u0, np.full(pt.size, T0)))
    c.bcs = bcs
    return c, geo, A


def total_inlet_box(n=(10, 6, 5), mu=1.8e-5, flux="ROE", mach=0.5, **kw):
    """Jittered box (permuted + RCM) with the tutorials' inflow patch on -x:
    totalPressure, pressureInletOutletVelocity, totalTemperature at the
    total state of the given Mach number; fixed static p at +x, slip walls."""
    c = cases.case_box(n=n, mu=mu, flux=flux, mach=(mach, mach), p_jump=1.0, **kw)
    fac = 1.0 + 0.5 * (cases.GAMMA - 1.0) * mach * mach
    for i, pt in enumerate(c.mesh.patches):
        if pt.name == "inlet":
            Uv = np.zeros((pt.size, 3)); Uv[:, 0] = mach * np.sqrt(cases.GAMMA * cases.R_GAS * 300.0)
            c.bcs[i] = cases.total_inlet(pt.size, 1e5 * fac ** (cases.GAMMA / (cases.GAMMA - 1.0)), 300.0 * fac,
                                         U_init=Uv, T_init=300.0)
    return prepare_case(c, permute_seed=17)


def bump_tutorial_100x25():
    """The converged-state parity case: 100x25 bump with the tutorial's
    inflow patch and the 200x50 bump's schemes and solver (the case of
    tests/golden/make_bump_converged.py)."""
    return prepare_case(cases.case_bump(100, 25, bcs="tutorial"), permute_seed=11)
