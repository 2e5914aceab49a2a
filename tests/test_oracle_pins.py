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


@pytest.mark.parametrize("mu", [0.0, 1.8e-5])
def test_jacobian_is_derivative_of_first_order_lf_model(mu):
    """The assembled matrix equals the complex-step derivative of the
    first-order Lax-Friedrichs residual model with frozen wave speeds:
    internal faces  |S|/2 (F(W_P)+F(W_N)) - |S|/2 lam (W_N-W_P) - s (W_N-W_P) [rho,rhoE,rhoU rows]
    boundary faces  |S| F(W_b(W_P)) + |S|/2 lam_b W_P
    cells           V rDT W_c."""
    c = _upwind_case(mu=mu)
    o = O.OracleCase(c)
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
