"""Pins of the oracle (CPU) against closed forms, invariants and brute force.

The reference ships no golden vectors and cannot be built here (no
OpenFOAM), so every oracle function is checked against what the mathematics
fixes: exact geometry of affine hexahedra, Euler-flux consistency and
supersonic upwinding, exact linear reconstruction, complex-step derivatives
of an independently written first-order Lax-Friedrichs residual model, and
dense linear algebra on tiny meshes.
"""
import numpy as np
import pytest

from icsfoam_amd import cases
from icsfoam_amd import mesh as M
from oracle import oracle as O

from helpers import linear_velocity_case, rel_err

GAMMA, R = 1.4, 287.0
CV = R / (GAMMA - 1)


# ------------------------------------------------------------- geometry ---
def test_geometry_unit_cubes():
    m = cases.hex_box(3, 2, 2, lengths=(3.0, 2.0, 2.0))
    g = O.OracleMesh(m).geometry()
    np.testing.assert_allclose(g["V"], 1.0, rtol=1e-14)
    np.testing.assert_allclose(np.linalg.norm(g["Sf"], axis=1), 1.0, rtol=1e-14)
    np.testing.assert_allclose(g["w"][:m.n_internal], 0.5, rtol=1e-14)
    np.testing.assert_allclose(g["dc"][:m.n_internal], 1.0, rtol=1e-14)
    np.testing.assert_allclose(g["nodc"][:m.n_internal], 1.0, rtol=1e-14)
    assert np.abs(g["corr"]).max() < 1e-14
    ctr = cases.approx_cell_centres(m)
    np.testing.assert_allclose(g["C"], ctr, rtol=0, atol=1e-14)


def test_geometry_affine_parallelepipeds():
    m = cases.hex_box(3, 3, 2)
    A = np.array([[1.2, 0.3, -0.1], [0.1, 0.9, 0.2], [-0.2, 0.15, 1.1]])
    b = np.array([0.3, -0.2, 1.0])
    m2 = cases.PolyMesh(points=m.points @ A.T + b, face_off=m.face_off, face_verts=m.face_verts,
                        owner=m.owner, neighbour=m.neighbour, patches=m.patches, n_cells=m.n_cells)
    g0, g = O.OracleMesh(m).geometry(), O.OracleMesh(m2).geometry()
    det = np.linalg.det(A)
    np.testing.assert_allclose(g["V"], det * g0["V"], rtol=1e-13)
    cof = det * np.linalg.inv(A).T
    np.testing.assert_allclose(g["Sf"], g0["Sf"] @ cof.T, rtol=0, atol=1e-14)
    np.testing.assert_allclose(g["C"], g0["C"] @ A.T + b, rtol=0, atol=1e-13)


def test_geometry_jittered_closed_cells_and_volume():
    m = cases.hex_box(6, 5, 4, lengths=(2.0, 1.0, 1.0), jitter=0.1, seed=3)
    g = O.OracleMesh(m).geometry()
    acc = np.zeros((m.n_cells, 3))
    np.add.at(acc, m.owner, g["Sf"])
    np.add.at(acc, m.neighbour, -g["Sf"][:m.n_internal])
    assert np.abs(acc).max() < 1e-14
    assert abs(g["V"].sum() - 2.0) < 1e-13
    assert (g["V"] > 0).all()


def test_product_geometry_matches_oracle_bitwise():
    c = M.prepare_case(cases.case_box(n=(6, 5, 4)), permute_seed=2)
    g1, g2 = M.mesh_geometry(c.mesh), O.OracleMesh(c.mesh).geometry()
    for k in g1:
        np.testing.assert_array_equal(g1[k], g2[k])


# ---------------------------------------------------------- renumbering ---
def _bandwidth(m, order=None):
    own, nei = m.owner[:m.n_internal], m.neighbour
    if order is not None:
        inv = np.empty_like(order); inv[order] = np.arange(order.size)
        own, nei = inv[own], inv[nei]
    return int(np.abs(own - nei).max())


def test_rcm_matches_oracle_and_scipy_bandwidth():
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    c = cases.case_box(n=(9, 7, 5))
    perm = np.random.default_rng(4).permutation(c.mesh.n_cells).astype(np.int32)
    m = M.renumber_mesh(c.mesh, perm)
    a, b = M.rcm_order(m), O.rcm(m)
    np.testing.assert_array_equal(a, b)
    assert sorted(a) == list(range(m.n_cells))
    n = m.n_cells
    A = coo_matrix((np.ones(m.n_internal), (m.owner[:m.n_internal], m.neighbour)), shape=(n, n))
    ref = reverse_cuthill_mckee((A + A.T).tocsr(), symmetric_mode=True)
    assert _bandwidth(m, a) <= 1.5 * _bandwidth(m, np.asarray(ref))
    assert _bandwidth(m, a) < _bandwidth(m) / 4


def test_renumber_invariants_and_oracle():
    c = cases.case_box(n=(6, 5, 4), jitter=0.1)
    perm = np.random.default_rng(9).permutation(c.mesh.n_cells).astype(np.int32)
    m = M.renumber_mesh(c.mesh, perm)
    fv, own, nei, fm = O.renumber(c.mesh, perm)
    np.testing.assert_array_equal(m.owner, own)
    np.testing.assert_array_equal(m.neighbour, nei)
    np.testing.assert_array_equal(m.face_verts, fv)
    nI = m.n_internal
    assert (m.owner[:nI] < m.neighbour).all()
    key = m.owner[:nI].astype(np.int64) * m.n_cells + m.neighbour
    assert (np.diff(key) > 0).all()
    # each new cell owns the same set of point-sets as its old cell
    inv = np.empty_like(perm); inv[perm] = np.arange(perm.size)

    def cell_faces(mm):
        out = [set() for _ in range(mm.n_cells)]
        fv4 = mm.face_verts.reshape(-1, 4)
        for f in range(mm.n_faces):
            out[mm.owner[f]].add(frozenset(fv4[f]))
            if f < mm.n_internal:
                out[mm.neighbour[f]].add(frozenset(fv4[f]))
        return out
    old, new = cell_faces(c.mesh), cell_faces(m)
    assert all(new[inv[o]] == old[o] for o in range(m.n_cells))
    # volumes stay positive after re-orientation
    assert (O.OracleMesh(m).geometry()["V"] > 0).all()


def test_colouring_two_colours_and_two_levels():
    c = M.prepare_case(cases.case_box(n=(7, 6, 5), jitter=0.1), permute_seed=3, colour_major=False)
    col, n = M.colour_cells(c.mesh)
    col2, n2 = O.colour(c.mesh)
    np.testing.assert_array_equal(col, col2)
    assert n == n2 == 2
    m = c.mesh
    assert (col[m.owner[:m.n_internal]] != col[m.neighbour]).all()
    cm = M.prepare_case(cases.case_box(n=(7, 6, 5), jitter=0.1), permute_seed=3)
    _, _, nf, nb = M.sweep_levels(cm.mesh)
    assert nf == 2 and nb == 2


# ---------------------------------------------------------------- fluxes --
def euler_flux(S, rho, p, U, T=None):
    """Analytic Euler flux through area vector S, [rho, rhoE, rhoU] order."""
    S = np.asarray(S)
    un = U @ S
    e = p / (rho * (GAMMA - 1))
    E = e + 0.5 * U @ U
    H = E + p / rho
    return np.concatenate([[rho * un, rho * H * un], rho * U * un + p * S])


def _state(rng, mach, n):
    T = rng.uniform(200, 400)
    p = rng.uniform(5e4, 2e5)
    rho = p / (R * T)
    c = np.sqrt(GAMMA * R * T)
    t = rng.standard_normal(3); t -= t @ n * n
    U = mach * c * n + 0.3 * c * t
    return rho, p, U, T


def _oflux(scheme, S, L, Rs):
    o = O.face_flux(scheme, S, [L[0], L[1], *L[2], L[3]], [Rs[0], Rs[1], *Rs[2], Rs[3]])
    return np.array([o[0], o[4], o[1], o[2], o[3]])


@pytest.mark.parametrize("scheme", ["AUSMPlusUp", "ROE", "HLLC"])
def test_flux_consistency(scheme):
    rng = np.random.default_rng(0)
    for _ in range(50):
        S = rng.standard_normal(3)
        n = S / np.linalg.norm(S)
        st = _state(rng, rng.uniform(-1.5, 1.5), n)
        f = _oflux(scheme, S, st, st)
        ref = euler_flux(S, st[0], st[1], st[2])
        assert rel_err(f, ref) < 2e-13


@pytest.mark.parametrize("scheme", ["AUSMPlusUp", "ROE", "HLLC"])
def test_flux_supersonic_upwinding(scheme):
    rng = np.random.default_rng(1)
    for sgn in (1, -1):
        for _ in range(20):
            S = rng.standard_normal(3)
            n = S / np.linalg.norm(S)
            L = _state(rng, sgn * rng.uniform(1.6, 3.0), n)
            Rs = _state(rng, sgn * rng.uniform(1.6, 3.0), n)
            up = L if sgn > 0 else Rs
            f = _oflux(scheme, S, L, Rs)
            assert rel_err(f, euler_flux(S, up[0], up[1], up[2])) < 1e-12


@pytest.mark.parametrize("scheme", ["AUSMPlusUp", "ROE", "HLLC"])
def test_flux_mirror_symmetry_and_rotation(scheme):
    rng = np.random.default_rng(2)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    for _ in range(20):
        S = rng.standard_normal(3)
        n = S / np.linalg.norm(S)
        L = _state(rng, rng.uniform(-0.9, 0.9), n)
        Rs = _state(rng, rng.uniform(-0.9, 0.9), n)
        f = _oflux(scheme, S, L, Rs)
        g = _oflux(scheme, -S, Rs, L)
        assert rel_err(f, -g) < 1e-12
        Lq = (L[0], L[1], Q @ L[2], L[3]); Rq = (Rs[0], Rs[1], Q @ Rs[2], Rs[3])
        h = _oflux(scheme, Q @ S, Lq, Rq)
        assert rel_err(h[:2], f[:2]) < 1e-11
        assert rel_err(h[2:], Q @ f[2:]) < 1e-11


# -------------------------------------------------------- reconstruction --
def _uniform_case(n=(6, 5, 4), jitter=0.0, mu=0.0, flux="AUSMPlusUp"):
    c = cases.case_box(n=n, jitter=jitter, mu=mu, flux=flux, p_jump=1.0, noise=0.0, mach=(0.5, 0.5))
    return c


def test_linear_reconstruction_exact_on_uniform_mesh():
    c = _uniform_case(n=(8, 6, 6))
    C = cases.approx_cell_centres(c.mesh)
    rho = 1.0 + 0.1 * C[:, 0] + 0.05 * C[:, 1] - 0.07 * C[:, 2]
    c.rho[:] = rho
    o = O.OracleCase(c)
    L, Rr = o.reconstruct(0)
    g = O.OracleMesh(c.mesh).geometry()
    exact = 1.0 + g["Cf"] @ np.array([0.1, 0.05, -0.07])
    m = c.mesh
    boundary_cells = np.zeros(m.n_cells, bool); boundary_cells[m.owner[m.n_internal:]] = True
    inner = ~(boundary_cells[m.owner[:m.n_internal]] | boundary_cells[m.neighbour])
    assert inner.sum() > 0
    np.testing.assert_allclose(L[:m.n_internal][inner], exact[:m.n_internal][inner], rtol=1e-14)
    np.testing.assert_allclose(Rr[:m.n_internal][inner], exact[:m.n_internal][inner], rtol=1e-14)


def test_limited_reconstruction_stays_within_neighbour_bounds():
    c = cases.case_box(n=(7, 6, 5), jitter=0.15, noise=0.2)
    o = O.OracleCase(c)
    m = c.mesh
    f = o.fields()
    L, Rr = o.reconstruct(1)  # pressure
    p = f["p"]
    nI = m.n_internal
    lo = p.copy(); hi = p.copy()
    np.minimum.at(lo, m.owner[:nI], p[m.neighbour]); np.maximum.at(hi, m.owner[:nI], p[m.neighbour])
    np.minimum.at(lo, m.neighbour, p[m.owner[:nI]]); np.maximum.at(hi, m.neighbour, p[m.owner[:nI]])
    # boundary values also enter the bounds
    pb_L, _ = L[nI:], None
    bown = m.owner[nI:]
    ok = np.ones(m.n_faces - nI, bool)
    np.minimum.at(lo, bown, np.where(ok, pb_L, np.inf)); np.maximum.at(hi, bown, np.where(ok, pb_L, -np.inf))
    tol = 1e-9 * p.max()
    assert (L[:nI] >= lo[m.owner[:nI]] - tol).all() and (L[:nI] <= hi[m.owner[:nI]] + tol).all()
    assert (Rr[:nI] >= lo[m.neighbour] - tol).all() and (Rr[:nI] <= hi[m.neighbour] + tol).all()


# ---------------------------------------------------------------- residual --
@pytest.mark.parametrize("flux", ["AUSMPlusUp", "ROE", "HLLC"])
def test_freestream_preservation_jittered(flux):
    c = cases.case_box(n=(7, 6, 5), jitter=0.15, flux=flux, mu=1.8e-5, p_jump=1.0, noise=0.0,
                       mach=(0.6, 0.6))
    # uniform state equal to the inlet/outlet values
    T0, p0 = 300.0, 1e5
    rho0 = p0 / (R * T0)
    U0 = 0.6 * np.sqrt(GAMMA * R * T0)
    c.rho[:] = rho0; c.T[:] = T0; c.U[:] = 0; c.U[:, 0] = U0
    c.bcs[1].p_val[:] = p0
    b, _ = O.OracleCase(c).residual()
    scale = rho0 * U0 * U0 * 0.1 * 0.1
    assert np.abs(b).max() < 1e-9 * scale


def test_viscous_terms_exact_for_linear_velocity():
    """Viscous residual terms (residualsUpdate.H: laplacian(muEff, U),
    div(muEff dev2(T(grad U))), div(sigmaDotU & Sf), laplacian(alphaEff, e))
    for a linear velocity field on a uniform orthogonal box, where the Gauss
    gradient, the linear face interpolation and snGrad are exact: with
    G = grad U (G_ij = dU_j/dx_i) constant, the momentum terms are surface
    integrals of constant tensors and vanish, and the energy term is
    V * M:G with M = mu G + mu (G^T - 2/3 tr(G) I) (divergence theorem on
    the face-midpoint rule, exact for the linear integrand); e is constant.
    A transposed stress, a wrong dev2 factor or a sigmaDotU built from the
    wrong tensor changes M:G for the non-symmetric G used here.  The
    difference of the residuals with and without viscosity isolates these
    terms (the convective flux does not depend on mu)."""
    mu = 2.0e-3
    cv, geo, A = linear_velocity_case(mu)
    c0, _, _ = linear_velocity_case(0.0)
    bv, _ = O.OracleCase(cv).residual()
    b0, _ = O.OracleCase(c0).residual()
    d = bv - b0
    G = A.T
    M = mu * G + mu * (G.T - 2.0 / 3.0 * np.trace(G) * np.eye(3))
    expect_e = geo["V"] * np.sum(M * G)
    assert np.array_equal(d[0], np.zeros_like(d[0]))
    scale = np.abs(expect_e).max()
    assert np.abs(d[2:]).max() < 1e-9 * scale
    np.testing.assert_allclose(d[1], expect_e, rtol=1e-8, atol=1e-9 * scale)


def test_discrete_conservation():
    c = cases.case_box(n=(6, 5, 4), jitter=0.1, flux="ROE", mu=0.0)
    o = O.OracleCase(c)
    b, _ = o.residual()
    phi, phiU, phiE = o.calcFlux()
    nI = c.mesh.n_internal
    tot = np.array([phi[nI:].sum(), phiE[nI:].sum(), *phiU[nI:].sum(axis=0)])
    np.testing.assert_allclose(b.sum(axis=1), -tot, rtol=1e-11, atol=1e-11 * np.abs(tot).max())


# ---------------------------------------------------------------- Jacobian -
def _flux_cplx(S, W):
    rho, rE, rU = W[0], W[1], W[2:]
    U = rU / rho
    p = (GAMMA - 1) * (rE - 0.5 * rho * (U @ U))
    un = U @ S
    return np.concatenate([[rho * un, (rE + p) * un], rU * un + p * S])


def _jac_cs(fun, W, h=1e-30):
    n = W.size
    J = np.zeros((fun(W).size, n))
    for j in range(n):
        Wc = W.astype(complex); Wc[j] += 1j * h
        J[:, j] = fun(Wc).imag / h
    return J


def _upwind_case(flux="AUSMPlusUp", mu=0.0, jitter=0.15):
    c = cases.case_box(n=(4, 3, 3), jitter=jitter, flux=flux, mu=mu, noise=0.1)
    c.fvSchemes = c.fvSchemes.replace("linearUpwind limitedGrad", "upwind")
    return c


def _dense(d, u, l, m):
    n = m.n_cells
    A = np.zeros((5 * n, 5 * n))
    for c in range(n):
        A[5 * c:5 * c + 5, 5 * c:5 * c + 5] = d[c]
    for f in range(m.n_internal):
        P, N = m.owner[f], m.neighbour[f]
        A[5 * P:5 * P + 5, 5 * N:5 * N + 5] = u[f]
        A[5 * N:5 * N + 5, 5 * P:5 * P + 5] = l[f]
    return A


@pytest.mark.parametrize("mu,bcset", [(0.0, "box"), (1.8e-5, "box"), (0.0, "total"), (1.8e-5, "total")])
def test_jacobian_is_derivative_of_first_order_lf_model(mu, bcset):
    """The assembled matrix equals the complex-step derivative of the
    first-order Lax-Friedrichs residual model with frozen wave speeds:
    internal faces  |S|/2 (F(W_P)+F(W_N)) - |S|/2 lam (W_N-W_P) - s (W_N-W_P) [rho,rhoE,rhoU rows]
    boundary faces  |S| F(W_b(W_P)) + |S|/2 lam_b W_P
    cells           V rDT W_c.
    bcset "total": the tutorials' inflow patch (totalPressure,
    pressureInletOutletVelocity, totalTemperature) after one pseudo-step, so
    that the inflow switch is set: p_b and T_b are constants of the state
    (corrected from lagged values), U_b = (n n) U_P on the axis-aligned
    inlet, where the reference's diagonal valueInternalCoeffs
    1 - sqrt(|1 - n_i^2|) equal n n exactly."""
    if bcset == "box":
        c = _upwind_case(mu=mu)
    else:
        c, _, _ = _total_inlet_case(n=(4, 3, 3), mu=mu)
        c.fvSchemes = c.fvSchemes.replace("linearUpwind limitedGrad", "upwind")
    o = O.OracleCase(c)
    if bcset == "total":
        o.pseudo_steps(1)
    bnd = o.boundary()
    b, rdt = o.residual()
    d, u, l = o.createJacobian()
    m = c.mesh
    g = O.OracleMesh(m).geometry()
    f = o.fields()
    W0 = f["W"].T.copy()  # [nC][5]
    T = f["T"]; Uc = f["U"]
    cs = np.sqrt(GAMMA * R * T)
    nC, nI = m.n_cells, m.n_internal
    lam = np.zeros(nI); s = np.zeros(nI)
    for fi in range(nI):
        P, N = m.owner[fi], m.neighbour[fi]
        w = g["w"][fi]
        cf = w * (cs[P] - cs[N]) + cs[N]
        Uf = w * (Uc[P] - Uc[N]) + Uc[N]
        nrm = g["Sf"][fi] / np.linalg.norm(g["Sf"][fi])
        lam[fi] = cf + abs(Uf @ nrm)
        if mu > 0:
            rf = w * (f["rho"][P] - f["rho"][N]) + f["rho"][N]
            lv = (mu + GAMMA * mu / 0.72) / rf
            s[fi] = 0.5 * lv * np.linalg.norm(g["Sf"][fi]) * g["dc"][fi]
    # boundary states as functions of the owner state (BCs of case_box)
    bstate = []
    for pi, p in enumerate(m.patches):
        bc = c.bcs[pi]
        for i in range(p.size):
            bstate.append((p.start + i, bc, i))

    def model(Wflat):
        W = Wflat.reshape(nC, 5)
        Rm = np.zeros((nC, 5), dtype=complex)
        for fi in range(nI):
            P, N = m.owner[fi], m.neighbour[fi]
            S = g["Sf"][fi]; ms = np.linalg.norm(S)
            F = 0.5 * (_flux_cplx(S, W[P]) + _flux_cplx(S, W[N])) - 0.5 * ms * lam[fi] * (W[N] - W[P]) \
                - s[fi] * (W[N] - W[P])
            Rm[P] += F; Rm[N] -= F
        for fi, bc, i in bstate:
            P = m.owner[fi]
            S = g["Sf"][fi]; ms = np.linalg.norm(S); nrm = S / ms
            rho, rE, rU = W[P][0], W[P][1], W[P][2:]
            U = rU / rho
            e = rE / rho - 0.5 * (U @ U)
            Tp = e / CV
            pp = rho * R * Tp
            pb = bc.p_val[i] if bc.p_type == cases.BC_FIXED_VALUE else pp
            Tb = bc.T_val[i] if bc.T_type == cases.BC_FIXED_VALUE else Tp
            if bc.U_type == cases.BC_FIXED_VALUE:
                Ub = bc.U_val[i].astype(complex)
            elif bc.U_type == cases.BC_SLIP:
                Ub = U - nrm * (nrm @ U)
            else:
                Ub = U
            bb = fi - m.n_internal
            if bc.p_type == cases.BC_TOTAL_PRESSURE:
                pb = bnd["p"][bb]
            if bc.T_type == cases.BC_TOTAL_TEMPERATURE:
                Tb = bnd["T"][bb]
            if bc.U_type == cases.BC_PRESSURE_INLET_OUTLET_VELOCITY:
                assert bnd["inflow"][bb] == 1
                Ub = nrm * (nrm @ U)
            rb = pb / (R * Tb)
            Wb = np.concatenate([[rb, rb * (CV * Tb + 0.5 * (Ub @ Ub))], rb * Ub])
            Ubr = np.real(Ub)
            lamb = np.sqrt(GAMMA * R * np.real(Tb)) + abs(Ubr @ nrm)
            Rm[P] += _flux_cplx(S, Wb) + 0.5 * ms * lamb * W[P]
        Rm += (g["V"] * rdt)[:, None] * W
        return Rm.ravel()

    J = _jac_cs(model, W0.ravel())
    A = _dense(d, u, l, m)
    assert rel_err(A, J) < 1e-9


# ------------------------------------------------------- linear algebra ---
def _assembled(flux="ROE", mu=1.8e-5, precond="LUSGS"):
    c = cases.case_box(n=(4, 3, 3), jitter=0.1, flux=flux, mu=mu, noise=0.1, precond=precond)
    o = O.OracleCase(c)
    o.residual()
    d, u, l = o.createJacobian()
    return c, o, d, u, l


def test_amul_is_dense_product():
    c, o, d, u, l = _assembled()
    A = _dense(d, u, l, c.mesh)
    x = np.random.default_rng(0).standard_normal((5, c.mesh.n_cells))
    y = o.matrixMul(x)
    np.testing.assert_allclose(y.T.ravel(), A @ x.T.ravel(), rtol=1e-13, atol=1e-13 * np.abs(A @ x.T.ravel()).max())


def test_lusgs_is_factored_inverse():
    c, o, d, u, l = _assembled()
    m = c.mesh
    n = m.n_cells
    A = _dense(d, u, l, m)
    rD = np.array([1.0 / max(abs(d[i][0, 0]), abs(d[i][1, 1]), abs(d[i][2, 2]), abs(d[i][3, 3]), abs(d[i][4, 4]))
                   for i in range(n)])
    D = np.kron(np.diag(1.0 / rD), np.eye(5))
    Lw = np.zeros_like(A); Up = np.zeros_like(A)
    for f in range(m.n_internal):
        P, N = m.owner[f], m.neighbour[f]
        Up[5 * P:5 * P + 5, 5 * N:5 * N + 5] = u[f]
        Lw[5 * N:5 * N + 5, 5 * P:5 * P + 5] = l[f]
    Mm = (D + Lw) @ np.linalg.inv(D) @ (D + Up)
    r = np.random.default_rng(1).standard_normal((5, n))
    z = o.precondition(r)
    zr = np.linalg.solve(Mm, r.T.ravel())
    assert rel_err(z.T.ravel(), zr) < 1e-11


def test_block_jacobi_is_block_inverse():
    c, o, d, u, l = _assembled(precond="Jacobi")
    n = c.mesh.n_cells
    r = np.random.default_rng(2).standard_normal((5, n))
    z = o.precondition(r)
    zr = np.stack([np.linalg.solve(d[i], r[:, i]) for i in range(n)], axis=1)
    assert rel_err(z, zr) < 1e-11


def test_gmres_full_space_is_direct_solve():
    c = cases.case_box(n=(3, 2, 2), jitter=0.1, flux="ROE", mu=1.8e-5, noise=0.1, nDirections=60)
    c.fvSolution = c.fvSolution.replace("tolerance 1e-12", "tolerance 1e-300").replace("relTol 0.01", "relTol 0")
    o = O.OracleCase(c)
    b, _ = o.residual()
    d, u, l = o.createJacobian()
    A = _dense(d, u, l, c.mesh)
    W = o.fields()["W"]
    dW, res, it = o.solveForIncr(W, b)
    x = np.linalg.solve(A, b.T.ravel())
    assert rel_err(dW.T.ravel(), x) < 1e-9


def test_native_block_mesh_matches_numpy_statement():
    pf = [("inlet", 0, ["-x"]), ("outlet", 0, ["+x"]), ("walls", 1, ["-y", "+y", "-z", "+z"])]
    a = cases._hex_topology(5, 4, 3, pf)
    b = cases._hex_topology_numpy(5, 4, 3, pf)
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)
    assert [(p.name, p.start, p.size) for p in a[3]] == [(p.name, p.start, p.size) for p in b[3]]


def test_inverted_cells_rejected():
    """checkMesh's negative-volume test: faces listed the wrong way round
    give negative volumes, which the library refuses (the sector generator
    maps (r, theta, z) right-handed; (theta, r, z) would invert every
    cell)."""
    import dataclasses

    from icsfoam_amd import IcsfError

    m = cases.hex_box(3, 2, 2)
    fo, fv = m.face_off, m.face_verts.copy()
    for f in range(m.n_faces):
        fv[fo[f]:fo[f + 1]] = fv[fo[f]:fo[f + 1]][::-1]
    bad = dataclasses.replace(m, face_verts=fv)
    with pytest.raises(IcsfError, match="non-positive volume"):
        M.mesh_geometry(bad)
    g = M.mesh_geometry(cases.annulus_sector(4, 3, 2))
    assert g["V"].min() > 0


# ------------------------------------------- closed-form pins, flux level ---
# Each pin below isolates one coefficient or branch of a flux scheme on a
# state where the published method has a closed-form answer, so that a wrong
# coefficient, sign or branch in the oracle fails it.
KSTAR = 2.0 * (GAMMA - 1.0) / (GAMMA + 1.0)


def _face(scheme, S, L, Rr, lowMach=True):
    """Oracle face flux: L/R = (rho, p, U, T); returns (phi, phiU, phiE)."""
    o = O.face_flux(scheme, S, [L[0], L[1], *L[2], L[3]], [Rr[0], Rr[1], *Rr[2], Rr[3]], lowMach=lowMach)
    return o[0], o[1:4], o[4]


def _cstar(rho, p, U, T):
    """Critical sound speed of Liou (2006): c*^2 = 2(gamma-1)/(gamma+1) H,
    H = e + |U|^2/2 + p/rho with e = Cv T."""
    return np.sqrt(KSTAR * (CV * T + 0.5 * (U @ U) + p / rho))


def _frame(seed):
    rng = np.random.default_rng(seed)
    S = rng.standard_normal(3) * rng.uniform(0.5, 2.0)
    n = S / np.linalg.norm(S)
    t = rng.standard_normal(3); t -= (t @ n) * n; t /= np.linalg.norm(t)
    return S, n, t


@pytest.mark.parametrize("scheme", ["AUSMPlusUp", "ROE", "HLLC"])
def test_flux_stationary_contact(scheme):
    """u = 0 on both sides, equal p, rho_L != rho_R: every scheme here keeps
    a stationary contact exactly (mass and energy flux 0, momentum p S).
    AUSM+up: M+(0) + M-(0) = 0.375 - 0.375 and P+(0) + P-(0) = 1 (Liou
    2006, beta = 1/8, alpha = 3/16); Roe: the jump is the entropy wave, whose
    eigenvalue u = 0 is replaced by the entropy fix, so the mass flux is
    not zero for Roe (pinned in test_roe_entropy_fix_on_contacts) but the
    momentum and energy fluxes are; HLLC: S_M = 0."""
    for seed in range(8):
        S, n, _ = _frame(seed)
        p = 1e5 * (1 + 0.5 * seed)
        rL, rR = 1.2, 0.4 + 0.2 * seed
        L = (rL, p, np.zeros(3), p / (rL * R)); Rr = (rR, p, np.zeros(3), p / (rR * R))
        phi, phiU, phiE = _face(scheme, S, L, Rr)
        np.testing.assert_allclose(phiU, p * S, rtol=1e-14, atol=1e-14 * p * np.abs(S).max())
        assert abs(phiE) <= 1e-9 * p * np.linalg.norm(S)
        if scheme != "ROE":
            assert phi == 0.0


@pytest.mark.parametrize("scheme", ["AUSMPlusUp", "HLLC"])
def test_flux_moving_contact_is_upwind(scheme):
    """Equal velocity U (normal part u, any tangential part) and equal p on
    both sides, rho_L != rho_R: AUSM+up (M+(M) + M-(M) = M, P+ + P- = 1, no
    pressure diffusion at dp = 0, no velocity diffusion at du = 0) and HLLC
    (S_M = u, star state = upwind state) return the exact upwind Euler flux
    of the side the contact comes from."""
    for seed in range(8):
        S, n, t = _frame(10 + seed)
        p = 8e4 + 1e4 * seed
        rL, rR = 1.1, 0.5 + 0.15 * seed
        c0 = np.sqrt(GAMMA * p / rL)
        for u in (0.35 * c0, -0.2 * c0):
            U = u * n + 0.3 * c0 * t
            L = (rL, p, U, p / (rL * R)); Rr = (rR, p, U, p / (rR * R))
            up = L if u > 0 else Rr
            rho_up, T_up = up[0], up[3]
            H_up = CV * T_up + 0.5 * U @ U + p / rho_up
            mdot = rho_up * (U @ S)
            phi, phiU, phiE = _face(scheme, S, L, Rr)
            assert abs(phi - mdot) <= 1e-13 * abs(mdot)
            np.testing.assert_allclose(phiU, mdot * U + p * S, rtol=0, atol=1e-13 * np.abs(mdot * U + p * S).max())
            assert abs(phiE - mdot * H_up) <= 1e-13 * abs(mdot * H_up)


def test_ausm_split_functions_known_values():
    """Left state at M_L = +-0.5 against a supersonic right state (M_R = 1.5,
    so M-(M_R) = P-(M_R) = 0) and mean Mach above 1 (no pressure diffusion):
    the mass flux is M+(M_L) c_1/2 rho_L |S| and the pressure flux P+(M_L) p_L.
    Liou (2006) polynomials: M+(4)(M) = M+(2)(1 - 16 beta M-(2)),
    P+(5)(M) = M+(2)((2 - M) - 16 alpha M M-(2)), M+-(2) = +-(M +- 1)^2/4,
    beta = 1/8, alpha = 3/16, evaluated by hand:
      M+(0.5)  = 0.5625 * 1.125   = 0.6328125
      P+(0.5)  = 0.5625 * 1.59375 = 0.896484375
      M+(-0.5) = 0.0625 * 2.125   = 0.1328125
      P+(-0.5) = 0.0625 * 1.65625 = 0.103515625
    c_1/2 = min(c*_L^2/max(c*_L, u_L), c*_R^2/max(c*_R, -u_R)) = c*_L here."""
    known = {0.5: (0.6328125, 0.896484375), -0.5: (0.1328125, 0.103515625)}
    for seed, (ML, (Mp, Pp)) in enumerate(sorted(known.items())):
        S, n, t = _frame(20 + seed)
        rL, pL, TL = 1.3, 1.1e5, 290.0
        # c*_L with |U_L| = |M_L| c*_L: c*^2 (1 - KSTAR M^2/2) = KSTAR (Cv T + p/rho)
        cs = np.sqrt(KSTAR * (CV * TL + pL / rL) / (1.0 - 0.5 * KSTAR * ML * ML))
        UL = ML * cs * n
        assert abs(_cstar(rL, pL, UL, TL) - cs) < 1e-12 * cs
        TR = 4.0 * TL
        rR, pR = 0.7, 0.9e5
        UR = 1.5 * cs * n
        assert _cstar(rR, pR, UR, TR) > cs
        phi, phiU, phiE = _face("AUSMPlusUp", S, (rL, pL, UL, TL), (rR, pR, UR, TR))
        ms = np.linalg.norm(S)
        np.testing.assert_allclose(phi, Mp * cs * rL * ms, rtol=1e-13)
        p12 = (phiU @ S - phi * (UL @ S)) / (S @ S)
        np.testing.assert_allclose(p12, Pp * pL, rtol=1e-13)


def test_ausm_pressure_diffusion_coefficients():
    """Equal velocity u n on both sides (M_L = M_R = M, so M+(M) + M-(M) = M),
    equal H (c*_L = c*_R = c), p_L != p_R: the mass flux is
    (M + M_p) c rho_up |S| with M_p = -K_p max(1 - sigma M^2, 0) (p_R - p_L) /
    (rho_1/2 c^2), rho_1/2 = (rho_L + rho_R)/2.  K_p and sigma are solved
    from two Mach numbers and must be the reference's 0.25 and 1
    (ausmPlusUpFluxScheme.C:231-232)."""
    S, n, t = _frame(30)
    ms = np.linalg.norm(S)
    rL, pL, T = 1.2, 1.0e5, 300.0
    pR = 1.002e5
    rR = rL * pR / pL          # p/rho equal: equal H, equal c*
    rows, rhs = [], []
    for M in (0.2, 0.6):
        H0 = CV * T + pL / rL
        c = np.sqrt(KSTAR * H0 / (1.0 - 0.5 * KSTAR * M * M))
        U = M * c * n
        assert abs(_cstar(rL, pL, U, T) - c) < 1e-12 * c and abs(_cstar(rR, pR, U, T) - c) < 1e-12 * c
        phi, _, _ = _face("AUSMPlusUp", S, (rL, pL, U, T), (rR, pR, U, T), lowMach=False)
        Mp = phi / (c * rL * ms) - M      # M12 > 0: upwind left
        # Mp = -(K_p - K_p sigma M^2) dp / (rho_1/2 c^2)
        k = -Mp * 0.5 * (rL + rR) * c * c / (pR - pL)
        rows.append([1.0, -M * M]); rhs.append(k)
    Kp, Kps = np.linalg.solve(np.array(rows), np.array(rhs))
    assert abs(Kp - 0.25) < 1e-9
    assert abs(Kps / Kp - 1.0) < 1e-8


def test_ausm_pressure_diffusion_damping_branch():
    """When the pressure diffusion would flip the sign of M_1/2 (M > 0,
    M + M_p <= 0) the reference adds 0.2 M_p instead of M_p
    (ausmPlusUpFluxScheme.C:236-256); otherwise the full M_p.  Same setup as
    the coefficient test, M = 0.005, a pressure rise across the face large
    enough to reverse the sign, and a small one that does not."""
    S, n, t = _frame(31)
    ms = np.linalg.norm(S)
    rL, pL, T, M = 1.2, 1.0e5, 300.0, 0.005
    c = np.sqrt(KSTAR * (CV * T + pL / rL) / (1.0 - 0.5 * KSTAR * M * M))
    U = M * c * n
    for dp_rel, damped in ((0.05, True), (0.002, False)):
        pR = pL * (1 + dp_rel)
        rR = rL * pR / pL
        Mp = -0.25 * (1 - M * M) * (pR - pL) / (0.5 * (rL + rR) * c * c)
        assert (M + Mp <= 0) == damped
        M12 = M + (0.2 * Mp if damped else Mp)
        rho_up = rL if M12 >= 0 else rR
        phi, _, _ = _face("AUSMPlusUp", S, (rL, pL, U, T), (rR, pR, U, T), lowMach=False)
        np.testing.assert_allclose(phi, M12 * c * rho_up * ms, rtol=1e-12)


def test_ausm_velocity_diffusion_coefficient():
    """The low-Mach pressure term p_u = -K_u P+(M_L) P-(M_R) (rho_L + rho_R)
    c_1/2 (u_R - u_L) is the difference of the pressure fluxes with
    lowMachAusm on and off (the mass flux does not depend on it).  At
    M_L = 0.5, M_R = -0.5 with equal H: P+(0.5) = P-(-0.5) = 0.896484375
    (Liou 2006, alpha = 3/16) and c_1/2 = c*, so K_u follows from the two
    fluxes and must be the reference's 0.25 (ausmPlusUpFluxScheme.C:261)."""
    S, n, t = _frame(32)
    rL, rR, p, T = 1.2, 0.8, 1.0e5, 300.0
    # equal H on both sides: equal p/rho and T
    pL, pR = p, p * rR / rL
    # |U|^2 = (0.25 + 0.01) c*^2 on both sides
    c = np.sqrt(KSTAR * (CV * T + pL / rL) / (1.0 - 0.5 * KSTAR * 0.26))
    UL, UR = 0.5 * c * n + 0.1 * c * t, -0.5 * c * n + 0.1 * c * t
    assert abs(_cstar(rL, pL, UL, T) - c) < 1e-12 * c and abs(_cstar(rR, pR, UR, T) - c) < 1e-12 * c
    on = _face("AUSMPlusUp", S, (rL, pL, UL, T), (rR, pR, UR, T), lowMach=True)
    off = _face("AUSMPlusUp", S, (rL, pL, UL, T), (rR, pR, UR, T), lowMach=False)
    assert on[0] == off[0]
    dp12 = (on[1] - off[1]) @ S / (S @ S)
    P = 0.896484375
    Ku = dp12 / (-P * P * (rL + rR) * c * (-c))
    assert abs(Ku - 0.25) < 1e-10


def test_roe_entropy_fix_on_contacts():
    """Roe flux of a contact (equal U and p, rho_L != rho_R): the jump is
    the entropy wave (1, U, |U|^2/2) d rho with eigenvalue u = U.n, so
    P|L|P^-1 dW = |lambda_1| dW and
      phi = |S| (u (rho_L + rho_R) - |lambda_1| d rho)/2.
    Harten's fix (roeFluxScheme.C:89-139, coefficient 0.05 of :255): with
    eps = 0.05 max(|u + c|, |u - c|) (c the Roe-averaged sound speed,
    c^2 = gamma p / sqrt(rho_L rho_R) for equal p), |lambda_1| = |u| when
    |u| >= eps, else (u^2 + eps^2)/(2 eps) -- eps/2 for a stationary contact."""
    for seed, (frac, fixed) in enumerate(((0.0, True), (0.02, True), (0.2, False), (-0.03, True))):
        S, n, t = _frame(40 + seed)
        ms = np.linalg.norm(S)
        p, rL, rR = 1.0e5, 1.2, 0.5
        cbar = np.sqrt(GAMMA * p / np.sqrt(rL * rR))
        u = frac * cbar
        U = u * n + 0.2 * cbar * t
        eps = 0.05 * (cbar + abs(u))
        assert (abs(u) < eps) == fixed
        lam = (u * u + eps * eps) / (2 * eps) if fixed else abs(u)
        dr = rR - rL
        phi, phiU, phiE = _face("ROE", S, (rL, p, U, p / (rL * R)), (rR, p, U, p / (rR * R)))
        expect = 0.5 * ms * (u * (rL + rR) - lam * dr)
        assert abs(phi - expect) <= 1e-11 * ms * (rL + rR) * cbar
        expU = expect * U + p * S
        np.testing.assert_allclose(phiU, expU, rtol=0, atol=1e-11 * np.abs(expU).max())
        HL, HR = GAMMA / (GAMMA - 1) * p / rL + 0.5 * U @ U, GAMMA / (GAMMA - 1) * p / rR + 0.5 * U @ U
        expE = 0.5 * ms * (u * (rL * HL + rR * HR) - lam * dr * 0.5 * (U @ U))
        assert abs(phiE - expE) <= 1e-11 * ms * (rL + rR) * cbar * abs(HL)


def test_hllc_star_states_textbook_form():
    """HLLC in the textbook star-state form (Toro, Riemann Solvers, 3rd ed.,
    eqs. 10.37-10.39): for S_L < 0 < S_M the flux is F*_L = F_L + S_L (U*_L -
    U_L) with U*_L = rho_L (S_L - u_L)/(S_L - S_M) [1, S_M n + U_L - u_L n,
    E_L + (S_M - u_L)(S_M + p_L/(rho_L (S_L - u_L)))]; for S_M < 0 < S_R the
    mirror F*_R.  The reference writes the same flux through the star
    pressure (hllcFluxScheme.C:170-240); its wave speeds are the Einfeldt
    estimates S_L = min(u_L - c_L, u~ - c~), S_R = max(u_R + c_R, u~ + c~)
    with Roe averages (:128-167), and S_M the contact speed."""
    rng = np.random.default_rng(50)
    for it in range(40):
        S, n, t = _frame(60 + it)
        ms = np.linalg.norm(S)

        def state():
            T = rng.uniform(250, 350); p = rng.uniform(8e4, 1.2e5)
            c = np.sqrt(GAMMA * R * T)
            return p / (R * T), p, rng.uniform(-0.6, 0.6) * c * n + rng.uniform(-0.3, 0.3) * c * t, T
        L, Rr = state(), state()
        rl, pl, Ul, Tl = L; rr, pr, Ur, Tr = Rr
        cl, cr = np.sqrt(GAMMA * R * Tl), np.sqrt(GAMMA * R * Tr)
        El, Er = CV * Tl + 0.5 * Ul @ Ul, CV * Tr + 0.5 * Ur @ Ur
        Hl, Hr = El + pl / rl, Er + pr / rr
        ul, ur = Ul @ n, Ur @ n
        k = np.sqrt(rr / rl)
        ua = ((k * Ur + Ul) / (k + 1)) @ n
        Ha = (k * Hr + Hl) / (k + 1)
        Uav = (k * Ur + Ul) / (k + 1)
        ca = np.sqrt((GAMMA - 1) * (Ha - 0.5 * Uav @ Uav))
        SL, SR = min(ul - cl, ua - ca), max(ur + cr, ua + ca)
        SM = (pr - pl + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))

        def cons(r, U, E):
            return np.concatenate([[r, r * E], r * U])

        def flux(r, p, U, H):
            un = U @ n
            return np.concatenate([[r * un, r * H * un], r * un * U + p * n])

        def star(r, p, U, E, u, Sk):
            f = r * (Sk - u) / (Sk - SM)
            Us = SM * n + (U - u * n)
            Es = E + (SM - u) * (SM + p / (r * (Sk - u)))
            return f * np.concatenate([[1.0, Es], Us])
        if SL >= 0 or SR <= 0:
            continue
        if SM >= 0:
            F = flux(rl, pl, Ul, Hl) + SL * (star(rl, pl, Ul, El, ul, SL) - cons(rl, Ul, El))
        else:
            F = flux(rr, pr, Ur, Hr) + SR * (star(rr, pr, Ur, Er, ur, SR) - cons(rr, Ur, Er))
        F *= ms
        phi, phiU, phiE = _face("HLLC", S, L, Rr)
        got = np.concatenate([[phi, phiE], phiU])
        assert rel_err(got, F) < 1e-12


# ------------------------------------------------- limiter function pins ---
def _line_case(values, limiter):
    """Five cells in a row (h = 1), empty patches on +-y and +-z, so each
    inner cell's gradient and limiter see exactly its two x-neighbours
    (empty faces are skipped by the Gauss gradient and the limiter)."""
    from icsfoam_amd.cases import PATCH_EMPTY, PATCH_PATCH
    nx = len(values)
    pts_k, pts_j, pts_i = np.meshgrid(np.arange(2), np.arange(2), np.arange(nx + 1), indexing="ij")
    pts = np.stack([pts_i.ravel() * 1.0, pts_j.ravel() * 1.0, pts_k.ravel() * 1.0], 1)
    fv, own, nei, patches = cases._hex_topology(
        nx, 1, 1, [("inlet", PATCH_PATCH, ["-x"]), ("outlet", PATCH_PATCH, ["+x"]),
                   ("sides", PATCH_EMPTY, ["-y", "+y", "-z", "+z"])])
    m = cases._mesh_from_quads(pts, fv, own, nei, patches, nx, pts.reshape(2, 2, nx + 1, 3))
    p0, T0 = 1e5, 300.0
    rho = np.asarray(values, np.float64)
    U = np.zeros((nx, 3)); U[:, 0] = 50.0
    T = p0 / (R * rho)
    bcs = [cases.BoundarySpec(cases.BC_ZERO_GRADIENT, cases.BC_ZERO_GRADIENT, cases.BC_ZERO_GRADIENT)
           for _ in m.patches]
    return cases.Case("line", m, rho, U, T, bcs, cases.thermo_dict(0.0),
                      cases.schemes_dict("AUSMPlusUp", limiter=limiter), cases.solution_dict())


@pytest.mark.parametrize("limiter", ["Venkatakrishnan", "minmod"])
def test_limiter_function_on_three_cell_stencil(limiter):
    """Middle cell of values (v_W, v_P, v_E) = (1, 1.1, 1.1 + 0.1 a): Gauss
    gradient g = (v_E - v_W)/2, extrapolations +-g/2 to the two faces,
    max/min neighbour differences a 0.1 and -0.1, so r_E = 0.4 a/(g/2) < 2
    and r_W = -0.1/(-g/2) > 2.  The limited east face value is
    v_P + psi g/2 with psi = min(1, l(r_E), l(r_W)), l the limiter function:
    Venkatakrishnan (r^2 + 2r)/(r^2 + r + 2), minmod min(r, 1)."""
    fn = {"Venkatakrishnan": lambda r: (r * r + 2 * r) / (r * r + r + 2), "minmod": lambda r: min(r, 1.0)}[limiter]
    for a in (0.1, 0.25, 0.5, 0.9):
        vals = [1.0, 1.0, 1.1, 1.1 + 0.1 * a, 1.1 + 0.1 * a]
        c = _line_case(vals, f"cellLimited<{limiter}> Gauss linear 1")
        o = O.OracleCase(c)
        L, Rr = o.reconstruct(0)
        m = c.mesh
        f = int(np.nonzero((m.owner[:m.n_internal] == 2) & (m.neighbour == 3))[0][0])
        g = (vals[3] - vals[1]) / 2.0
        rE = (vals[3] - vals[2]) / (0.5 * g)
        rW = (vals[1] - vals[2]) / (-0.5 * g)
        assert rE < 2.0 < rW
        psi = min(1.0, fn(rE), fn(rW))
        assert psi < 1.0 or (limiter == "minmod" and rE >= 1.0)
        np.testing.assert_allclose(L[f], vals[2] + psi * 0.5 * g, rtol=1e-14)
        # the west face of the middle cell (its neighbour side) uses the same psi
        fw = int(np.nonzero((m.owner[:m.n_internal] == 1) & (m.neighbour == 2))[0][0])
        np.testing.assert_allclose(Rr[fw], vals[2] - psi * 0.5 * g, rtol=1e-14)


# ------------------------------------------------------- heat-flux pins ---
def _still_gas_case(Tfun, mu):
    """Uniform orthogonal box, U = 0, uniform p, T(x) given; fixedValue T on
    inlet/outlet (exact face values), zeroGradient T on the walls (dT/dn = 0
    there for T = T(x)); returns (case, geometry)."""
    c = cases.case_box(n=(8, 4, 3), lengths=(1.0, 0.5, 0.375), jitter=0.0, flux="AUSMPlusUp", mu=mu,
                       p_jump=1.0, noise=0.0)
    geo = O.OracleMesh(c.mesh).geometry()
    p0 = 1e5
    c.T[:] = Tfun(geo["C"][:, 0])
    c.rho[:] = p0 / (R * c.T)
    c.U[:] = 0.0
    bcs = []
    for pt in c.mesh.patches:
        xf = geo["Cf"][pt.start:pt.start + pt.size, 0]
        if pt.name in ("inlet", "outlet"):
            bcs.append(cases.BoundarySpec(cases.BC_ZERO_GRADIENT, cases.BC_FIXED_VALUE, cases.BC_FIXED_VALUE,
                                          None, np.zeros((pt.size, 3)), Tfun(xf)))
        else:
            bcs.append(cases.BoundarySpec(cases.BC_ZERO_GRADIENT, cases.BC_SLIP, cases.BC_ZERO_GRADIENT))
    c.bcs = bcs
    return c, geo


@pytest.mark.parametrize("shape", ["linear", "quadratic"])
def test_heat_flux_term_closed_form(shape):
    """laplacian(alphaEff, e) of residualsUpdate.H:16-43 on a still gas
    (U = 0: the stress and sigma.U terms vanish) with e = Cv T(x) on a
    uniform orthogonal box, where corrected snGrad is exact: the viscous part
    of the energy residual (b with mu minus b without) is V alphaEff Cv T''
    with alphaEff = kappa/Cv = gamma mu / Pr (OpenFOAM's alphahe of
    sensible internal energy) -- zero for linear T in every cell, and
    V gamma mu Cv 2a / Pr for T = T0 + a x^2 in every cell away from the
    inlet/outlet (where the half-cell boundary snGrad is not exact for a
    quadratic).  A missing gamma, Cp instead of Cv, or a sign error fails."""
    mu, Pr = 2.0e-3, 0.72
    a = 60.0
    Tfun = {"linear": lambda x: 300.0 + 40.0 * x, "quadratic": lambda x: 300.0 + a * (x - 0.3) ** 2}[shape]
    cv, geo = _still_gas_case(Tfun, mu)
    c0, _ = _still_gas_case(Tfun, 0.0)
    d = O.OracleCase(cv).residual()[0] - O.OracleCase(c0).residual()[0]
    V = geo["V"]
    expect = np.zeros_like(V) if shape == "linear" else V * GAMMA * mu / Pr * CV * 2 * a
    scale = V.max() * GAMMA * mu / Pr * CV * 2 * a
    assert np.array_equal(d[0], np.zeros_like(d[0]))
    assert np.abs(d[2:]).max() < 1e-9 * scale
    x = geo["C"][:, 0]
    inner = (x > 1.0 / 8) & (x < 1.0 - 1.0 / 8)
    sel = np.ones_like(inner) if shape == "linear" else inner
    assert sel.sum() > 0
    np.testing.assert_allclose(d[1][sel], expect[sel], rtol=1e-8, atol=1e-9 * scale)


# ------------------------------------------------ pseudo-time step pins ---
def test_local_time_step_uniform_state_closed_form():
    """rPseudoDeltaT of setCoAndDeltaT.H:57-141 for a uniform state on an
    orthogonal box: each internal face gives nonOrthDeltaCoeffs (c + |U.n|)
    = (c + |U_d|)/h_d, each wall face 0.5 nonOrthDeltaCoeffs (c + |U_P.n|) =
    (c + |U_d|)/h_d (boundary delta h_d/2), inlet/outlet faces nothing; the
    cell value is the MAXIMUM over its faces divided by the CFL number.
    One cell thick in z (walls on +-z, h_z smallest), so the wall term sets
    every cell's value: (c + |U_z|)/h_z / CFL.  A factor-1 wall term, a sum
    over faces or deltaCoeffs of the wrong side fail."""
    co = 7.0
    c = cases.case_box(n=(4, 3, 1), lengths=(4.0, 3.0, 0.25), jitter=0.0, flux="ROE", mu=0.0, p_jump=1.0,
                       noise=0.0, co=co)
    T0, p0 = 300.0, 1e5
    cs = np.sqrt(GAMMA * R * T0)
    U0 = np.array([0.4 * cs, -0.25 * cs, 0.3 * cs])
    c.rho[:] = p0 / (R * T0); c.T[:] = T0; c.U[:] = U0
    _, rdt = O.OracleCase(c).residual()
    faces = [(cs + abs(U0[0])) / 1.0, (cs + abs(U0[1])) / 1.0, (cs + abs(U0[2])) / 0.25]
    np.testing.assert_allclose(rdt, max(faces) / co, rtol=1e-14)
    # three cells thick in z: the internal z faces give the same value
    c3 = cases.case_box(n=(4, 3, 3), lengths=(4.0, 3.0, 0.75), jitter=0.0, flux="ROE", mu=0.0, p_jump=1.0,
                        noise=0.0, co=co)
    c3.rho[:] = p0 / (R * T0); c3.T[:] = T0; c3.U[:] = U0
    _, rdt3 = O.OracleCase(c3).residual()
    np.testing.assert_allclose(rdt3, max(faces) / co, rtol=1e-14)


def test_switched_evolution_relaxation_recurrence():
    """SER of setCoAndDeltaT.H:3-37 read off the oracle's own history: the
    CFL field used in pseudo-step j >= 3 is the one of step j-1 times
    clamp(|r_{j-2}|/|r_{j-1}|, 0.1, 1.25), then clamped to
    [pseudoCoNumMin, pseudoCoNumMax]; |r| = sqrt(sInit_rho^2 + sInit_rhoE^2
    + |vInit_rhoU|^2) of the step's initial GMRES residuals; steps 1 and 2
    keep the initial CFL.  (No cell is bounded in this mild case, so the
    field stays uniform.)"""
    c = cases.case_box(n=(6, 5, 4), jitter=0.1, flux="ROE", mu=1.8e-5, noise=0.01, p_jump=1.2, co=5.0)
    c.fvSolution = cases.solution_dict(co=5.0, coMin=0.5, coMax=9.0)
    o = O.OracleCase(c)
    cos, norms = [], []
    hit_max = hit_incr = False
    for j in range(10):
        r = o.pseudo_steps(1)[0]
        co = o.pseudoCo()
        assert np.ptp(co) == 0.0
        cos.append(co[0])
        norms.append(np.sqrt(r[0] ** 2 + r[1] ** 2 + (r[2] ** 2 + r[3] ** 2 + r[4] ** 2)))
    assert cos[0] == cos[1] == 5.0
    for j in range(2, 10):
        ratio = min(max(norms[j - 2] / norms[j - 1], 0.1), 1.25)
        expect = min(max(cos[j - 1] * ratio, 0.5), 9.0)
        hit_incr |= norms[j - 2] / norms[j - 1] > 1.25
        hit_max |= cos[j - 1] * ratio > 9.0
        assert abs(cos[j] - expect) <= 1e-14 * expect
    assert hit_max or hit_incr


# ----------------------------------------------------- HB source sign pin --
def test_hb_source_is_minus_volume_times_time_derivative():
    """For instances that are each uniform in space (no spatial residual:
    zeroGradient conditions on every patch), the HB residual
    is the spectral source alone, b_J = -V (dW/dt)(t_J) (HBZone::addSource,
    HBZoneTemplates.C:39-93: the harmonic-balance time derivative moved to
    the right-hand side, where the dbnsFoam residual is b = R V with
    dW/dtau = R).  With rho = rho0 (1 + 0.1 sin wt), U = U0 (1 + 0.05 cos wt)
    and constant T the conserved variables hold harmonics up to 3, which
    D of 7 instants differentiates exactly; the test compares with the
    analytic time derivative."""
    from icsfoam_amd.solver import hb_operator as lib_hb_operator
    f = 100.0
    w = 2 * np.pi * f
    D, t, _ = lib_hb_operator([w], [3], 7, 1.0 / f)
    base = cases.case_box(n=(4, 3, 3), jitter=0.1, flux="ROE", mu=0.0, noise=0.0, p_jump=1.0)
    m = base.mesh
    bcs = [cases.BoundarySpec(cases.BC_ZERO_GRADIENT, cases.BC_ZERO_GRADIENT, cases.BC_ZERO_GRADIENT)
           for _ in m.patches]
    rho0, T0 = 1.1, 310.0
    U0 = np.array([120.0, -30.0, 15.0])
    insts = []
    import dataclasses
    for J in range(7):
        s, co = np.sin(w * t[J]), np.cos(w * t[J])
        rho = np.full(m.n_cells, rho0 * (1 + 0.1 * s))
        U = np.tile(U0 * (1 + 0.05 * co), (m.n_cells, 1))
        insts.append(dataclasses.replace(base, rho=rho, U=U, T=np.full(m.n_cells, T0), bcs=bcs))
    b = O.OracleHB(insts, D).residual()
    V = O.OracleMesh(m).geometry()["V"]
    for J in range(7):
        s, co = np.sin(w * t[J]), np.cos(w * t[J])
        r, dr = rho0 * (1 + 0.1 * s), rho0 * 0.1 * w * co
        u, du = U0 * (1 + 0.05 * co), -U0 * 0.05 * w * s
        E = CV * T0 + 0.5 * u @ u
        dW = np.concatenate([[dr, dr * E + r * (u @ du)], dr * u + r * du])
        expect = -V[None, :] * dW[:, None]
        scale = np.abs(expect).max()
        np.testing.assert_allclose(b[J], expect, rtol=0, atol=1e-9 * scale)


# ----------------------------------------------------- state update pins --
def test_state_update_primitive_recovery():
    """updateFields.H:7-105 on a conserved state W = [rho, rhoE, rhoU]:
    U = rhoU/rho, e = rhoE/rho - |U|^2/2 bounded below by Cv T_min, T = e/Cv,
    p = rho R T (perfectGas, psi = 1/(R T)); a cell whose internal energy is
    negative comes back at T_min (1e-15 K here, the bound the library's
    thermo uses)."""
    c = cases.case_box(n=(4, 3, 3), jitter=0.1, flux="ROE", mu=0.0, noise=0.05)
    o = O.OracleCase(c)
    rng = np.random.default_rng(3)
    n = c.mesh.n_cells
    rho = rng.uniform(0.5, 2.0, n)
    U = rng.uniform(-200, 200, (n, 3))
    e = rng.uniform(1e5, 3e5, n)
    e[5] = -1e3
    W = np.vstack([rho, rho * (e + 0.5 * np.sum(U * U, 1)), (rho[:, None] * U).T])
    o.set_state_w(W)
    f = o.fields()
    np.testing.assert_allclose(f["rho"], rho, rtol=1e-15)
    np.testing.assert_allclose(f["U"], U, rtol=1e-12)
    T = np.maximum(e, CV * 1e-15) / CV
    np.testing.assert_allclose(f["T"], T, rtol=1e-10)
    np.testing.assert_allclose(f["p"], rho * R * T, rtol=1e-10)
    assert f["T"][5] == pytest.approx(1e-15, rel=1e-12)


def test_time_step_bounding_factors():
    """boundLocalTimeStep.H: a cell whose internal energy is below SMALL (or
    whose rho / e fell under 0.95 of the previous iterate) has its CFL number
    multiplied by 0.5, and every cell sharing a face with it by 0.75; the
    others keep theirs.  The factor is evaluated before updateFields (the
    reference's order), so one step from a state with one cell at
    T = 1e-20 K (e < SMALL) flags exactly that cell."""
    co = 20.0
    c = cases.case_box(n=(5, 4, 3), jitter=0.1, flux="ROE", mu=0.0, noise=0.01, p_jump=1.0, co=co)
    bad = 27          # (i, j, k) = (2, 1, 1): an interior cell
    c.T[bad] = 1e-20
    o = O.OracleCase(c)
    o.pseudo_steps(1)
    got = o.pseudoCo()
    m = c.mesh
    nb = set(m.neighbour[m.owner[:m.n_internal] == bad]) | set(m.owner[:m.n_internal][m.neighbour == bad])
    expect = np.full(m.n_cells, co)
    expect[list(nb)] = 0.75 * co
    expect[bad] = 0.5 * co
    assert len(nb) == 6
    np.testing.assert_array_equal(got, expect)


# ------------------------------------------ derived inflow condition pins --
def _total_inlet_case(mach=0.5, flux="ROE", mu=0.0, n=(5, 4, 3), outflow=False):
    """Box with the tutorials' inflow patch on -x (totalPressure,
    pressureInletOutletVelocity, totalTemperature), fixed static p on +x,
    slip walls; uniform axial state of the given Mach number (reversed with
    outflow=True, so the inlet patch sees outflow)."""
    c = cases.case_box(n=n, jitter=0.1, flux=flux, mu=mu, noise=0.01, p_jump=1.0, mach=(mach, mach))
    T, p = 300.0, 1e5
    p0 = p * (1 + 0.5 * (GAMMA - 1) * mach ** 2) ** (GAMMA / (GAMMA - 1))
    T0 = T * (1 + 0.5 * (GAMMA - 1) * mach ** 2)
    if outflow:
        c.U[:, 0] = -np.abs(c.U[:, 0])
    for i, pt in enumerate(c.mesh.patches):
        if pt.name == "inlet":
            Uin = np.zeros((pt.size, 3)); Uin[:, 0] = (-1 if outflow else 1) * mach * np.sqrt(GAMMA * R * T)
            c.bcs[i] = cases.total_inlet(pt.size, p0, T0, U_init=Uin, T_init=T)
    return c, p0, T0


def test_total_conditions_isentropic_fixed_point():
    """totalPressure / totalTemperature / pressureInletOutletVelocity at an
    inflow patch (phi < 0), re-corrected with the cells frozen until the
    lagged values settle: the fixed point obeys the textbook relations of a
    steady isentropic inflow, T_b = T0 - |U_b|^2/(2 Cp) (total enthalpy)
    and p_b = p0 (T_b/T0)^(gamma/(gamma-1)) (isentropic), with U_b the
    normal part (n.U_P) n of the adjacent cell velocity (zero tangential
    velocity).  Initial state: p_b = p0 (no `value` entry), U_b and T_b from
    the `value` entries, phi = Sf.(rho_b U_b)."""
    c, p0, T0 = _total_inlet_case()
    o = O.OracleCase(c)
    m = c.mesh
    ip = [i for i, pt in enumerate(m.patches) if pt.name == "inlet"][0]
    pt = m.patches[ip]
    sl = slice(pt.start - m.n_internal, pt.start - m.n_internal + pt.size)
    b0 = o.boundary()
    np.testing.assert_array_equal(b0["p"][sl], p0)
    np.testing.assert_array_equal(b0["T"][sl], 300.0)
    assert (b0["phi"][sl] < 0).all() and (b0["inflow"][sl] == 0).all()
    W = o.fields()["W"]
    for _ in range(60):
        o.set_state_w(W)
    b = o.boundary()
    assert (b["inflow"][sl] == 1).all()
    geo = O.OracleMesh(m).geometry()
    n = geo["Sf"][pt.start:pt.start + pt.size] / np.linalg.norm(geo["Sf"][pt.start:pt.start + pt.size], axis=1)[:, None]
    Uc = o.fields()["U"][m.owner[pt.start:pt.start + pt.size]]
    Un = np.sum(n * Uc, axis=1)[:, None] * n
    np.testing.assert_allclose(b["U"][sl], Un, rtol=0, atol=1e-12 * np.abs(Uc).max())
    Cp = GAMMA * R / (GAMMA - 1)
    Tb = T0 - np.sum(b["U"][sl] ** 2, axis=1) / (2 * Cp)
    np.testing.assert_allclose(b["T"][sl], Tb, rtol=1e-13)
    np.testing.assert_allclose(b["p"][sl], p0 * (b["T"][sl] / T0) ** (GAMMA / (GAMMA - 1)), rtol=1e-13)


def test_total_conditions_at_outflow():
    """The same patch seeing outflow (phi >= 0): pos0(phi) = 1 switches the
    dynamic terms off (p_b = p0, T_b = T0) and pressureInletOutletVelocity is
    zeroGradient (U_b = U_P, valueFraction 0)."""
    c, p0, T0 = _total_inlet_case(outflow=True)
    o = O.OracleCase(c)
    m = c.mesh
    pt = [p for p in m.patches if p.name == "inlet"][0]
    sl = slice(pt.start - m.n_internal, pt.start - m.n_internal + pt.size)
    assert (o.boundary()["phi"][sl] > 0).all()
    o.set_state_w(o.fields()["W"])
    b = o.boundary()
    np.testing.assert_array_equal(b["p"][sl], p0)
    np.testing.assert_array_equal(b["T"][sl], T0)
    np.testing.assert_array_equal(b["U"][sl], o.fields()["U"][m.owner[pt.start:pt.start + pt.size]])
    assert (b["inflow"][sl] == 0).all()


# ------------------------------------------------ pseudotimeControl pins --
def _expected_stop(hist, tol, tolRel, nCorr=20, nCorrMin=1):
    """Step at which pseudotimeControl ends a steady run, restated from the
    reference's criteria on a recorded history of initial residuals:
    before step j (j >= 2, j != nCorr) the previous step's residuals r are
    tested, absolutely (max r < tol) and, from j = 3, relative to the
    residuals of step 1 (max r / r_1 < tolRel); the first step j at which a
    test holds (and j >= nCorrMin) is the last one run."""
    r1 = hist[0]
    for j in range(2, len(hist) + 1):
        if j == nCorr:
            continue
        r = hist[j - 2]
        ok = r.max() < tol or (j >= 3 and (r / (r1 + 1e-150)).max() < tolRel)
        if ok and j >= nCorrMin:
            return j
    return None


@pytest.mark.parametrize("tol,tolRel,nCorr,nCorrMin", [(1e-12, 1e-3, 20, 1), (3e-4, 1e-12, 20, 1),
                                                      (1e-12, 1e-2, 20, 40), (1e-12, 1e-2, 3, 1)])
def test_pseudotime_control_stops_where_criteria_hold(tol, tolRel, nCorr, nCorrMin):
    """The oracle's loop (pseudotimeControl::loop/criteriaSatisfied) runs
    exactly the steps the criteria restated on an independently recorded
    residual history allow, and ends in the same state; with neither
    criterion reachable it stops at maxSteps unconverged."""
    import dataclasses
    c = cases.case_bump(40, 10)
    sol = c.fvSolution.replace("pseudoTol 1e-12;", f"pseudoTol {tol!r};").replace(
        "pseudoTolRel 1e-12;", f"pseudoTolRel {tolRel!r}; nPseudoCorrMin {nCorrMin};").replace(
        "nPseudoCorr 1;", f"nPseudoCorr {nCorr};")
    c = dataclasses.replace(c, fvSolution=sol)
    a, b = O.OracleCase(c), O.OracleCase(c)
    run, conv = a.pseudo_time_loop(400)
    hist = np.array([b.pseudo_steps(1)[0][:5] for _ in range(max(run, 1))])
    stop = _expected_stop(hist, tol, tolRel, nCorr, nCorrMin)
    assert conv and stop == run
    for k in ("W", "p", "T"):
        np.testing.assert_array_equal(a.fields()[k], b.fields()[k])
    c2 = dataclasses.replace(c, fvSolution=sol.replace(f"pseudoTol {tol!r};", "pseudoTol 1e-300;").replace(
        f"pseudoTolRel {tolRel!r};", "pseudoTolRel 1e-300;"))
    assert O.OracleCase(c2).pseudo_time_loop(12) == (12, False)
