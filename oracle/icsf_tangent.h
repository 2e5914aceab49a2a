// ORACLE — test infrastructure only (included by icsf_oracle.cpp inside its
// anonymous namespace; nothing in the product uses it).
//
// Tangent-linear restatement of the inviscid convective residual
//     b(W, Sf, Cf) = -(sum_f +-F_f / V) * V
// as a function of the conserved state W = [rho, rhoE, rhoUx, rhoUy, rhoUz],
// the face area vectors Sf and the face centres Cf (cell centres, volumes
// and interpolation weights held fixed).  Every routine is a template over
// the scalar type: with T = double it evaluates the primal residual with the
// same operation order as residual() in icsf_oracle.cpp (a test checks the
// two bit for bit), with T = Dual it propagates one tangent direction
// (forward-mode automatic differentiation), i.e. the exact Jacobian-vector
// product J*d of the piecewise-smooth residual.  Branches (max/min, the
// |M| < 1 switch of AUSM+up, the limiter's min over faces) follow the primal
// value with these tie rules: max(a, b) takes b only if b > a, min(a, b)
// takes b only if b < a (the std::max / std::min convention).
//
// Chain followed (reference files):
//   W -> rho, U, e, T, p            updateFields.H:7-78 (zero increment)
//   boundary p, U, T, rho           updateFields.H:80-104 (fixedValue /
//                                   zeroGradient / slip)
//   E, H, c*                        ausmPlusUpFluxScheme.C:73-110
//   Gauss linear + cellLimited<Venkatakrishnan|minmod>, linearUpwind
//                                   ausmPlusUpFluxScheme.C:82-87 (reconstruct)
//   AUSM+up face flux               ausmPlusUpFluxScheme.C:112-299
//   Roe face flux (E, H, rho, p, U reconstructed; P|Lambda|P^-1 with the
//   entropy fix)                    roeFluxScheme.C:41-241,276-409
//   HLLC face flux (E, H, rho, p, U and max(sqrt(gamma/psi), VSMALL)
//   reconstructed)                  hllcFluxScheme.C:76-221
//   surfaceIntegrate                residualsUpdate.H:1-14
//   viscous terms (laplacian(muEff, U), div(tauMC), div(sigmaDotU),
//   laplacian(alphaEff, e) with Gauss linear grad(U), grad(e) and
//   corrected snGrad)               residualsUpdate.H:16-43
// With the viscous terms the derivatives in Sf and Cf are not defined
// (nonOrthDeltaCoeffs, deltaCoeffs and the non-orthogonal correction
// vectors, functions of the geometry, are held at the mesh's values).

struct Dual {
    double v, d;
    Dual(double x = 0.0, double t = 0.0) : v(x), d(t) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
inline Dual sqrtT(Dual a) {
    const double s = std::sqrt(a.v);
    return Dual(s, a.d / (2.0 * s));
}
inline double sqrtT(double a) { return std::sqrt(a); }
inline double val(double a) { return a; }
inline double val(Dual a) { return a.v; }
inline double tan_(double) { return 0.0; }
inline double tan_(Dual a) { return a.d; }
template <class T>
inline T tmax(T a, T b) { return val(a) < val(b) ? b : a; }
template <class T>
inline T tmin(T a, T b) { return val(b) < val(a) ? b : a; }
template <class T>
inline T sqrT(T a) { return a * a; }
// |a|: the sign of the primal value picks the branch (a == 0: +)
inline double fabsT(double a) { return std::fabs(a); }
inline Dual fabsT(Dual a) { return Dual(std::fabs(a.v), a.v < 0 ? -a.d : a.d); }

template <class T>
struct V3T {
    T x, y, z;
};
template <class T> inline V3T<T> operator+(V3T<T> a, V3T<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3T<T> operator-(V3T<T> a, V3T<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3T<T> operator*(T s, V3T<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3T<T> operator*(V3T<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3T<T> operator/(V3T<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dotT(V3T<T> a, V3T<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T magSqrT(V3T<T> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class T> inline V3T<T> ld(const T* p) { return {p[0], p[1], p[2]}; }
template <class T> inline V3T<T> lift(const double* p) { return {T(p[0]), T(p[1]), T(p[2])}; }
template <class T> inline V3T<T> crossT(V3T<T> a, V3T<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// 3x3 tensor, row-major, the oracle's T9 operations in the same order
template <class T>
struct T9T {
    T t[9];
};
template <class T> inline T9T<T> zero9() {
    T9T<T> r;
    for (int i = 0; i < 9; ++i) r.t[i] = T(0);
    return r;
}
template <class T> inline T9T<T> outerT(V3T<T> a, V3T<T> b) {
    return T9T<T>{{a.x * b.x, a.x * b.y, a.x * b.z, a.y * b.x, a.y * b.y, a.y * b.z, a.z * b.x, a.z * b.y, a.z * b.z}};
}
template <class T> inline T9T<T> operator+(const T9T<T>& a, const T9T<T>& b) {
    T9T<T> r;
    for (int i = 0; i < 9; ++i) r.t[i] = a.t[i] + b.t[i];
    return r;
}
template <class T> inline T9T<T> operator-(const T9T<T>& a, const T9T<T>& b) {
    T9T<T> r;
    for (int i = 0; i < 9; ++i) r.t[i] = a.t[i] - b.t[i];
    return r;
}
template <class T> inline T9T<T> operator*(T s, const T9T<T>& a) {
    T9T<T> r;
    for (int i = 0; i < 9; ++i) r.t[i] = s * a.t[i];
    return r;
}
template <class T> inline T9T<T> operator*(const T9T<T>& a, T s) {
    T9T<T> r;
    for (int i = 0; i < 9; ++i) r.t[i] = a.t[i] * s;
    return r;
}
template <class T> inline V3T<T> tdotvT(const T9T<T>& A, V3T<T> v) {
    return {A.t[0] * v.x + A.t[1] * v.y + A.t[2] * v.z, A.t[3] * v.x + A.t[4] * v.y + A.t[5] * v.z,
            A.t[6] * v.x + A.t[7] * v.y + A.t[8] * v.z};
}
template <class T> inline V3T<T> vdottT(V3T<T> v, const T9T<T>& A) {
    return {v.x * A.t[0] + v.y * A.t[3] + v.z * A.t[6], v.x * A.t[1] + v.y * A.t[4] + v.z * A.t[7],
            v.x * A.t[2] + v.y * A.t[5] + v.z * A.t[8]};
}
template <class T> inline T9T<T> tdottT(const T9T<T>& A, const T9T<T>& B) {
    T9T<T> r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.t[3 * i + j] = A.t[3 * i] * B.t[j] + A.t[3 * i + 1] * B.t[3 + j] + A.t[3 * i + 2] * B.t[6 + j];
    return r;
}
template <class T> inline T9T<T> transposeTT(const T9T<T>& A) {
    return T9T<T>{{A.t[0], A.t[3], A.t[6], A.t[1], A.t[4], A.t[7], A.t[2], A.t[5], A.t[8]}};
}
template <class T> inline T9T<T> dev2T(const T9T<T>& A) {
    T9T<T> r = A;
    const T tr = A.t[0] + A.t[4] + A.t[8];
    const T s = T(2.0 / 3.0) * tr;
    r.t[0] = A.t[0] - s;
    r.t[4] = A.t[4] - s;
    r.t[8] = A.t[8] - s;
    return r;
}

// Inviscid state of one evaluation: cell and boundary-face fields.
template <class T>
struct StateT {
    std::vector<T> rho, p, Tt, U, rhoE;  // cells (U: 3 per cell; rhoE rebuilt as updateFields does)
    std::vector<T> pB, TB, UB, rhoB, rhoEB;  // boundary faces (UB: 3 per face)
};

// updateFields.H:7-78 with a zero increment, then the boundary values of
// updateFields.H:80-104.
template <class T>
void primsFromW(const Case& k, const T* W, const T* Sf, StateT<T>& s) {
    const Mesh& m = *k.m;
    const int nC = m.nC, nB = m.nF - m.nI;
    s.rho.assign(nC, T(0)); s.p.assign(nC, T(0)); s.Tt.assign(nC, T(0)); s.U.assign(3 * (size_t)nC, T(0));
    s.rhoE.assign(nC, T(0));
    for (int c = 0; c < nC; ++c) {
        T rho = W[c] + T(0.0);
        V3T<T> rU{W[2 * (size_t)nC + c] + T(0.0), W[3 * (size_t)nC + c] + T(0.0), W[4 * (size_t)nC + c] + T(0.0)};
        T rE = W[(size_t)nC + c] + T(0.0);
        rho = tmax(rho, T(k.s.rhoMin));
        V3T<T> U{rU.x / rho, rU.y / rho, rU.z / rho};
        T e = rE / rho - T(0.5) * magSqrT(U);
        e = tmax(e, T(k.Cv * k.s.TMin));
        if (k.s.TMax < GREAT) e = tmin(e, T(k.Cv * k.s.TMax));
        T Tc = e / T(k.Cv);
        T psi = T(1.0) / (T(k.s.R) * Tc);
        s.rho[c] = rho;
        s.U[3 * c] = U.x; s.U[3 * c + 1] = U.y; s.U[3 * c + 2] = U.z;
        s.Tt[c] = Tc;
        s.p[c] = rho / psi;
        s.rhoE[c] = rho * (e + T(0.5) * magSqrT(U));
    }
    s.pB.assign(nB, T(0)); s.TB.assign(nB, T(0)); s.UB.assign(3 * (size_t)nB, T(0)); s.rhoB.assign(nB, T(0));
    s.rhoEB.assign(nB, T(0));
    for (size_t pi = 0; pi < m.pStart.size(); ++pi) {
        if (m.pKind[pi] == PK_EMPTY) continue;
        const BC& bc = k.bcs[pi];
        for (int i = 0; i < m.pSize[pi]; ++i) {
            const int f = m.pStart[pi] + i, b = f - m.nI, c = m.own[f];
            // the derived inflow conditions hold values set at the last state
            // update from lagged quantities: constants of this evaluation
            const T pB = bc.pT == BC_FIXED ? T(bc.pv[i]) : bc.pT == BC_TOTAL_P ? T(k.pB[b]) : s.p[c];
            const T TB = bc.tT == BC_FIXED ? T(bc.tv[i]) : bc.tT == BC_TOTAL_T ? T(k.TB[b]) : s.Tt[c];
            V3T<T> Uc{s.U[3 * c], s.U[3 * c + 1], s.U[3 * c + 2]};
            V3T<T> ub;
            if (bc.uT == BC_FIXED) ub = lift<T>(&bc.uv[3 * i]);
            else if (bc.uT == BC_PIOV) ub = lift<T>(&k.UB[3 * (size_t)b]);
            else if (bc.uT == BC_SLIP) {
                const V3T<T> S = ld(&Sf[3 * (size_t)f]);
                const V3T<T> n = S / sqrtT(magSqrT(S));
                // (U + (I - 2 n n) & U)/2, rows of (I - 2 n n) as in evalBoundary()
                const T r0x = T(1) - T(2.0) * n.x * n.x, r0y = T(-2.0) * n.x * n.y, r0z = T(-2.0) * n.x * n.z;
                const T r1x = T(-2.0) * n.y * n.x, r1y = T(1) - T(2.0) * n.y * n.y, r1z = T(-2.0) * n.y * n.z;
                const T r2x = T(-2.0) * n.z * n.x, r2y = T(-2.0) * n.z * n.y, r2z = T(1) - T(2.0) * n.z * n.z;
                const V3T<T> ref{r0x * Uc.x + r0y * Uc.y + r0z * Uc.z, r1x * Uc.x + r1y * Uc.y + r1z * Uc.z,
                                 r2x * Uc.x + r2y * Uc.y + r2z * Uc.z};
                ub = (Uc + ref) / T(2.0);
            } else ub = Uc;
            const T psi = T(1.0) / (T(k.s.R) * TB);
            s.pB[b] = pB;
            s.TB[b] = TB;
            s.UB[3 * b] = ub.x; s.UB[3 * b + 1] = ub.y; s.UB[3 * b + 2] = ub.z;
            s.rhoB[b] = psi * pB;
            s.rhoEB[b] = s.rhoB[b] * (T(k.Cv) * TB + T(0.5) * magSqrT(ub));
        }
    }
}

template <class T>
void gaussGradT(const Mesh& m, const T* vc, const T* vb, const T* Sf, std::vector<V3T<T>>& g) {
    g.assign(m.nC, V3T<T>{T(0), T(0), T(0)});
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        const T vf = T(m.w[f]) * (vc[o] - vc[n]) + vc[n];
        const V3T<T> S = ld(&Sf[3 * (size_t)f]) * vf;
        g[o] = g[o] + S;
        g[n] = g[n] - S;
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        g[m.own[f]] = g[m.own[f]] + ld(&Sf[3 * (size_t)f]) * vb[f - m.nI];
    }
    for (int c = 0; c < m.nC; ++c) g[c] = g[c] / T(m.V[c]);
}

template <class T>
inline void limitFaceT(T& lim, T maxD, T minD, T ext, int limiter) {
    T r = T(1);
    if (val(ext) > SMALL) r = maxD / ext;
    else if (val(ext) < -SMALL) r = minD / ext;
    T l;
    if (limiter == 1) l = (sqrT(r) + T(2) * r) / (sqrT(r) + r + T(2));
    else l = tmin(r, T(1.0));
    lim = tmin(lim, l);
}

template <class T>
void limitedGradT(const Case& k, const T* vc, const T* vb, const T* Sf, const T* Cf, std::vector<V3T<T>>& g) {
    const Mesh& m = *k.m;
    gaussGradT(m, vc, vb, Sf, g);
    if (k.s.limiter == 0) return;
    std::vector<T> mx(m.nC), mn(m.nC);
    for (int c = 0; c < m.nC; ++c) mx[c] = mn[c] = vc[c];
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        mx[o] = tmax(mx[o], vc[n]); mn[o] = tmin(mn[o], vc[n]);
        mx[n] = tmax(mx[n], vc[o]); mn[n] = tmin(mn[n], vc[o]);
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        const int o = m.own[f];
        const T b = vb[f - m.nI];
        mx[o] = tmax(mx[o], b); mn[o] = tmin(mn[o], b);
    }
    for (int c = 0; c < m.nC; ++c) { mx[c] = mx[c] - vc[c]; mn[c] = mn[c] - vc[c]; }
    std::vector<T> lim(m.nC, T(1.0));
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        const V3T<T> fc = ld(&Cf[3 * (size_t)f]);
        limitFaceT(lim[o], mx[o], mn[o], dotT(fc - lift<T>(&m.C[3 * o]), g[o]), k.s.limiter);
        limitFaceT(lim[n], mx[n], mn[n], dotT(fc - lift<T>(&m.C[3 * n]), g[n]), k.s.limiter);
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        const int o = m.own[f];
        limitFaceT(lim[o], mx[o], mn[o], dotT(ld(&Cf[3 * (size_t)f]) - lift<T>(&m.C[3 * o]), g[o]), k.s.limiter);
    }
    for (int c = 0; c < m.nC; ++c) g[c] = g[c] * lim[c];
}

template <class T>
void reconstructT(const Case& k, const T* vc, const T* vb, const T* Sf, const T* Cf, std::vector<T>& L,
                  std::vector<T>& R) {
    const Mesh& m = *k.m;
    std::vector<V3T<T>> g;
    if (k.s.recon == 1) limitedGradT(k, vc, vb, Sf, Cf, g);
    L.assign(m.nF, T(0));
    R.assign(m.nF, T(0));
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        T l = T(1.0) * (vc[o] - vc[n]) + vc[n];
        T r = T(0.0) * (vc[o] - vc[n]) + vc[n];
        if (k.s.recon == 1) {
            const V3T<T> fc = ld(&Cf[3 * (size_t)f]);
            l = l + dotT(fc - lift<T>(&m.C[3 * o]), g[o]);
            r = r + dotT(fc - lift<T>(&m.C[3 * n]), g[n]);
        }
        L[f] = l;
        R[f] = r;
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        L[f] = R[f] = vb[f - m.nI];
    }
}

// AUSM+up face flux (ausmPlusUpFluxScheme.C:112-299), same expression
// order as ausmFace() in icsf_oracle.cpp.
template <class T>
void ausmFaceT(int lowMach, V3T<T> S, T magS, V3T<T> UL, V3T<T> UR, T cL, T cR, T pL, T pR, T rL, T rR, T HL,
               T HR, T& phi, V3T<T>& phiU, T& phiE) {
    const T unL = dotT(UL, S) / magS;
    const T unR = dotT(UR, S) / magS;
    const T cl = sqrT(cL) / tmax(cL, unL);
    const T cr = sqrT(cR) / tmax(cR, -unR);
    const T cf = tmin(cl, cr);
    const T ML = unL / cf;
    T MpL, PpL;
    if (std::fabs(val(ML)) < 1.0) {
        const T p2 = T(0.25) * sqrT(ML + T(1)), m2 = T(-0.25) * sqrT(ML - T(1));
        MpL = p2 * (T(1) - T(2) * m2);
        PpL = p2 * (T(2) - ML - T(3) * ML * m2);
    } else {
        MpL = tmax(ML, T(0.0));
        PpL = T(val(ML) > 0 ? 1.0 : 0.0);
    }
    const T MR = unR / cf;
    T MmR, PmR;
    if (std::fabs(val(MR)) < 1.0) {
        const T m2 = T(-0.25) * sqrT(MR - T(1)), p2 = T(0.25) * sqrT(MR + T(1));
        MmR = m2 * (T(1) + T(2) * p2);
        PmR = m2 * (T(-2) - MR + T(3) * MR * p2);
    } else {
        MmR = tmin(MR, T(0.0));
        PmR = T(val(MR) < 0 ? 1.0 : 0.0);
    }
    T M12 = MpL + MmR;
    T p12 = PpL * pL + PmR * pR;
    const T Mmean = T(0.5) * (sqrT(unL) + sqrT(unR)) / sqrT(cf);
    const T MDiff = T(-0.25) * tmax(T(1.0) - Mmean, T(0.0)) * (pR - pL) / (T(0.5) * (rL + rR) * sqrT(cf));
    const double m12 = val(M12), m12d = val(M12 + MDiff);
    if ((m12 > 0.0 && m12d <= 0.0) || (m12 < 0.0 && m12d >= 0.0)) M12 = M12 + T(0.2) * MDiff;
    else M12 = M12 + MDiff;
    if (lowMach) p12 = p12 + T(-0.25) * PpL * PmR * (rL + rR) * cf * (unR - unL);
    const bool up = val(M12) >= 0;
    const V3T<T> Uf = up ? UL : UR;
    const T rhoa = M12 * cf * (up ? rL : rR);
    const V3T<T> rhoaU = rhoa * Uf;
    const T rhoah = rhoa * (up ? HL : HR);
    phi = rhoa * magS;
    phiU = rhoaU * magS + p12 * S;
    phiE = rhoah * magS;
}

// roeFluxScheme::getRoeDissipation (roeFluxScheme.C:41-241), same
// expression order as roeDissipation() in icsf_oracle.cpp.
template <class T>
struct RoeDissT {
    T CR, CE, ER, EE;
    V3T<T> CU, MR, ME, EU;
    T9T<T> MU;
};
template <class T>
RoeDissT<T> roeDissipationT(T rho, V3T<T> Vel, T c, double gI, T u, V3T<T> n, double eFix) {
    const double a1 = gI - 1;
    const T theta = T(0.5 * a1) * magSqrT(Vel);
    const T c2 = sqrT(c);
    const T a2 = T(1) / (rho * c * T(std::sqrt(2.0)));
    const T a3 = rho / (c * T(std::sqrt(2.0)));
    const T a4 = (theta + c2) / T(a1);
    const T a5 = T(1) - theta / c2;
    const T a6 = theta / T(a1);
    const V3T<T> iP11 = (n * a5) - crossT(Vel, n) / rho;
    T9T<T> iP12 = outerT(T(a1) / c2 * n, Vel);
    iP12.t[1] += n.z / rho; iP12.t[2] -= n.y / rho; iP12.t[3] -= n.z / rho;
    iP12.t[5] += n.x / rho; iP12.t[6] += n.y / rho; iP12.t[7] -= n.x / rho;
    const V3T<T> iP13 = T(-a1) / c2 * n;
    const T iP21 = a2 * (theta - c * u);
    const V3T<T> iP22 = (-a2) * (T(a1) * Vel - c * n);
    const T iP23 = T(a1) * a2;
    const T iP31 = a2 * (theta + c * u);
    const V3T<T> iP32 = (-a2) * (T(a1) * Vel + c * n);
    const T iP33 = T(a1) * a2;
    T L1 = fabsT(u), L2 = fabsT(u + c), L3 = fabsT(u - c);
    const T eps = T(eFix) * tmax(L2, L3);
    if (val(L1) < val(eps)) L1 = (sqrT(L1) + sqrT(eps)) / (T(2.0) * eps);
    if (val(L2) < val(eps)) L2 = (sqrT(L2) + sqrT(eps)) / (T(2.0) * eps);
    if (val(L3) < val(eps)) L3 = (sqrT(L3) + sqrT(eps)) / (T(2.0) * eps);
    const V3T<T> P11 = n;
    const T P12 = a3, P13 = a3;
    T9T<T> P21 = outerT(Vel, n);
    P21.t[1] -= n.z * rho; P21.t[2] += n.y * rho; P21.t[3] += n.z * rho;
    P21.t[5] -= n.x * rho; P21.t[6] -= n.y * rho; P21.t[7] += n.x * rho;
    const V3T<T> P22 = a3 * (Vel + c * n);
    const V3T<T> P23 = a3 * (Vel - c * n);
    const V3T<T> P31 = n * a6 + rho * crossT(Vel, n);
    const T P32 = a3 * (a4 + c * u);
    const T P33 = a3 * (a4 - c * u);
    RoeDissT<T> D;
    const V3T<T> P11L = P11 * L1;
    const T9T<T> P21L = P21 * L1;
    const V3T<T> P31L = P31 * L1;
    D.CR = dotT(P11L, iP11) + L2 * P12 * iP21 + L3 * P13 * iP31;
    D.CU = vdottT(P11L, iP12) + (L2 * P12) * iP22 + (L3 * P13) * iP32;
    D.CE = dotT(P11L, iP13) + L2 * P12 * iP23 + L3 * P13 * iP33;
    D.MR = tdotvT(P21L, iP11) + (L2 * P22) * iP21 + (L3 * P23) * iP31;
    D.MU = tdottT(P21L, iP12) + outerT(L2 * P22, iP22) + outerT(L3 * P23, iP32);
    D.ME = tdotvT(P21L, iP13) + (L2 * P22) * iP23 + (L3 * P23) * iP33;
    D.ER = dotT(P31L, iP11) + L2 * P32 * iP21 + L3 * P33 * iP31;
    D.EU = vdottT(P31L, iP12) + (L2 * P32) * iP22 + (L3 * P33) * iP32;
    D.EE = dotT(P31L, iP13) + L2 * P32 * iP23 + L3 * P33 * iP33;
    return D;
}

// roeFluxScheme::calcFlux (roeFluxScheme.C:276-409), one face, same order
// as roeFace() in icsf_oracle.cpp.
template <class T>
void roeFaceT(double gI, double eFix, V3T<T> S, T magS, T rl, T rr, T pl, T pr, V3T<T> Ul, V3T<T> Ur, T El, T Er,
              T Hl, T Hr, T& phi, V3T<T>& phiU, T& phiE) {
    const V3T<T> n = S / magS;
    const T coefR = sqrtT(tmax(T(VSMALL), rr) / tmax(T(VSMALL), rl));
    const T rhoRoe = coefR * rl;
    const V3T<T> URoe = (coefR * Ur + Ul) / (coefR + T(1.0));
    const T HRoe = (coefR * Hr + Hl) / (coefR + T(1.0));
    const T cRoe = sqrtT(fabsT(T(gI - 1.0) * (HRoe - T(0.5) * magSqrT(URoe))));
    const T uMl = dotT(Ul, n), uMr = dotT(Ur, n);
    const T uProj = dotT(URoe, n);
    const V3T<T> rUl = rl * Ul, rUr = rr * Ur;
    const T rEl = rl * El, rEr = rr * Er;
    const RoeDissT<T> D = roeDissipationT(rhoRoe, URoe, cRoe, gI, uProj, n, eFix);
    const T dR = rr - rl;
    const V3T<T> dRU = rUr - rUl;
    const T dRE = rEr - rEl;
    phi = T(-0.5) * magS * (D.CR * dR + dotT(D.CU, dRU) + D.CE * dRE);
    phiU = T(-0.5) * magS * (D.MR * dR + tdotvT(D.MU, dRU) + D.ME * dRE);
    phiE = T(-0.5) * magS * (D.ER * dR + dotT(D.EU, dRU) + D.EE * dRE);
    const T rnl = rl * uMl, rnr = rr * uMr;
    phi = phi + T(0.5) * magS * (rnl + rnr);
    phiU = phiU + T(0.5) * magS * (rnl * Ul + rnr * Ur + n * (pl + pr));
    phiE = phiE + T(0.5) * magS * (rnl * Hl + rnr * Hr);
}

// HLLC face flux (hllcFluxScheme.C:76-221), same expression order as
// hllcFace() in icsf_oracle.cpp (no frame velocity).
template <class T>
void hllcFaceT(double gIv, V3T<T> S, T magS, T rl, T rr, T pl, T pr, V3T<T> Ul, V3T<T> Ur, T cl, T cr, T El,
               T Er, T Hl, T Hr, T& phi, V3T<T>& phiU, T& phiE) {
    const T gI(gIv), one(1.0);
    const V3T<T> n = S / magS;
    const T coefR = sqrtT(tmax(T(VSMALL), rr) / tmax(T(VSMALL), rl));
    const V3T<T> uAvg = (coefR * Ur + Ul) / (coefR + one);
    const T HAvg = (coefR * Hr + Hl) / (coefR + one);
    const T cAvg = sqrtT(fabsT((gI - one) * (HAvg - T(0.5) * magSqrT(uAvg))));
    const T ul = dotT(Ul, n), ur = dotT(Ur, n), ua = dotT(uAvg, n);
    const T Sl = tmin(ul - cl, ua - cAvg);
    const T Sr = tmax(ur + cr, ua + cAvg);
    const T Sm = (rr * ur * (Sr - ur) - rl * ul * (Sl - ul) + pl - pr) / (rr * (Sr - ur) - rl * (Sl - ul));
    const T cSl(val(Sl) > 0 ? 1.0 : 0.0), cSr(val(Sr) < 0 ? 1.0 : 0.0), cSm(val(Sm) > 0 ? 1.0 : 0.0);
    const T cSlm = (one - cSl) * cSm, cSmr = (one - cSm) * (one - cSr);
    const T frl = Sm / (Sl - Sm) * ((Sl - ul) * rl);
    const T frr = Sm / (Sr - Sm) * ((Sr - ur) * rr);
    phi = (cSl * rl * ul + cSlm * frl + cSmr * frr + cSr * rr * ur) * magS;
    const T psl = rl * (ul - Sl) * (ul - Sm) + pl;
    const T psr = rr * (ur - Sr) * (ur - Sm) + pr;
    const V3T<T> rUsl = one / (Sl - Sm) * ((Sl - ul) * rl * Ul + (psl - pl) * n);
    const V3T<T> rUsr = one / (Sr - Sm) * ((Sr - ur) * rr * Ur + (psr - pr) * n);
    const V3T<T> fUl = Sm * rUsl + psl * n;
    const V3T<T> fUr = Sm * rUsr + psr * n;
    phiU = (cSl * (rl * ul * Ul + pl * n) + cSlm * fUl + cSmr * fUr + cSr * (rr * ur * Ur + pr * n)) * magS;
    const T rEsl = one / (Sl - Sm) * ((Sl - ul) * (rl * El) - pl * ul + psl * Sm);
    const T rEsr = one / (Sr - Sm) * ((Sr - ur) * (rr * Er) - pr * ur + psr * Sm);
    const T fEl = Sm * (rEsl + psl);
    const T fEr = Sm * (rEsr + psr);
    phiE = (cSl * rl * Hl * ul + cSlm * fEl + cSmr * fEr + cSr * rr * Hr * ur) * magS;
}

// Viscous terms of residualsUpdate.H:16-43 as the oracle's
// viscousResidual() forms them: per face |S| mu snGrad(U), Sf & tauMC,
// (mu grad(U) + tauMC) & U & Sf and |S| alpha snGrad(e), each summed per
// cell (surfaceIntegrate, divided by V) into RU [3 nC] and RE [nC].
template <class T>
void viscousT(const Case& k, const StateT<T>& s, const T* Sf, std::vector<T>& RU, std::vector<T>& RE) {
    const Mesh& m = *k.m;
    const int nC = m.nC, nB = m.nF - m.nI;
    const double mu = k.s.mu;
    const double alpha = k.s.gamma * mu / k.s.Pr;
    auto Uc = [&](int c) { return V3T<T>{s.U[3 * c], s.U[3 * c + 1], s.U[3 * c + 2]}; };
    auto Ub = [&](int b) { return V3T<T>{s.UB[3 * b], s.UB[3 * b + 1], s.UB[3 * b + 2]}; };
    auto magSf = [&](int f) { return sqrtT(magSqrT(ld(&Sf[3 * (size_t)f]))); };
    std::vector<T9T<T>> gU(nC, zero9<T>());
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        const V3T<T> Uf = T(m.w[f]) * (Uc(o) - Uc(n)) + Uc(n);
        const T9T<T> t = outerT(ld(&Sf[3 * (size_t)f]), Uf);
        gU[o] = gU[o] + t;
        gU[n] = gU[n] - t;
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        const T9T<T> t = outerT(ld(&Sf[3 * (size_t)f]), Ub(f - m.nI));
        gU[m.own[f]] = gU[m.own[f]] + t;
    }
    for (int c = 0; c < nC; ++c)
        for (int i = 0; i < 9; ++i) gU[c].t[i] = gU[c].t[i] / T(m.V[c]);
    std::vector<V3T<T>> snU(nB, V3T<T>{T(0), T(0), T(0)});
    std::vector<T9T<T>> gUb(nB, zero9<T>());
    for (size_t pi = 0; pi < m.pStart.size(); ++pi) {
        if (m.pKind[pi] == PK_EMPTY) continue;
        for (int i = 0; i < m.pSize[pi]; ++i) {
            const int f = m.pStart[pi] + i, b = f - m.nI, c = m.own[f];
            const V3T<T> n = ld(&Sf[3 * (size_t)f]) / magSf(f);
            const V3T<T> U = Uc(c);
            const int ut = k.bcs[pi].uT;
            if (ut == BC_ZG) snU[b] = V3T<T>{T(0), T(0), T(0)};
            else if (ut == BC_SLIP) {
                const T9T<T> Tr{{T(1) - T(2.0) * n.x * n.x, T(-2.0) * n.x * n.y, T(-2.0) * n.x * n.z,
                                 T(-2.0) * n.y * n.x, T(1) - T(2.0) * n.y * n.y, T(-2.0) * n.y * n.z,
                                 T(-2.0) * n.z * n.x, T(-2.0) * n.z * n.y, T(1) - T(2.0) * n.z * n.z}};
                snU[b] = (tdotvT(Tr, U) - U) * (T(m.dc[f]) / T(2.0));
            } else snU[b] = T(m.dc[f]) * (Ub(b) - U);
            gUb[b] = gU[c] + outerT(n, snU[b] - vdottT(n, gU[c]));
        }
    }
    std::vector<T9T<T>> tau(nC), taub(nB, zero9<T>());
    for (int c = 0; c < nC; ++c) tau[c] = T(mu) * dev2T(transposeTT(gU[c]));
    for (int b = 0; b < nB; ++b) taub[b] = T(mu) * dev2T(transposeTT(gUb[b]));
    std::vector<T> ec(nC), eb(nB, T(0));
    for (int c = 0; c < nC; ++c) ec[c] = s.rhoE[c] / s.rho[c] - T(0.5) * magSqrT(Uc(c));
    for (int b = 0; b < nB; ++b)
        if (val(s.rhoB[b]) != 0) eb[b] = s.rhoEB[b] / s.rhoB[b] - T(0.5) * magSqrT(Ub(b));
    std::vector<V3T<T>> ge;
    gaussGradT(m, ec.data(), eb.data(), Sf, ge);
    std::vector<T> lapU(3 * (size_t)m.nF, T(0)), divTau(3 * (size_t)m.nF, T(0)), sDotU(m.nF, T(0)),
        lapE(m.nF, T(0));
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        const double w = m.w[f];
        const V3T<T> S = ld(&Sf[3 * (size_t)f]);
        const V3T<T> k3 = lift<T>(&m.corr[3 * (size_t)f]);
        const T9T<T> gUf = T(w) * (gU[o] - gU[n]) + gU[n];
        const T9T<T> tauf = T(w) * (tau[o] - tau[n]) + tau[n];
        const V3T<T> Uf = T(w) * (Uc(o) - Uc(n)) + Uc(n);
        const T muf = T(w * (mu - mu) + mu);
        const T alf = T(w * (alpha - alpha) + alpha);
        const V3T<T> sn = T(m.nodc[f]) * (Uc(n) - Uc(o)) + vdottT(k3, gUf);
        const V3T<T> l = (muf * sn) * magSf(f);
        lapU[3 * f] = l.x; lapU[3 * f + 1] = l.y; lapU[3 * f + 2] = l.z;
        const V3T<T> dt = vdottT(S, tauf);
        divTau[3 * f] = dt.x; divTau[3 * f + 1] = dt.y; divTau[3 * f + 2] = dt.z;
        const V3T<T> sig = tdotvT(muf * gUf + tauf, Uf);
        sDotU[f] = dotT(sig, S);
        const V3T<T> gef = T(w) * (ge[o] - ge[n]) + ge[n];
        const T sne = T(m.nodc[f]) * (ec[n] - ec[o]) + dotT(k3, gef);
        lapE[f] = (alf * sne) * magSf(f);
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        const int b = f - m.nI, c = m.own[f];
        const V3T<T> S = ld(&Sf[3 * (size_t)f]);
        const V3T<T> l = (T(mu) * snU[b]) * magSf(f);
        lapU[3 * f] = l.x; lapU[3 * f + 1] = l.y; lapU[3 * f + 2] = l.z;
        const V3T<T> dt = vdottT(S, taub[b]);
        divTau[3 * f] = dt.x; divTau[3 * f + 1] = dt.y; divTau[3 * f + 2] = dt.z;
        const V3T<T> sig = tdotvT(T(mu) * gUb[b] + taub[b], Ub(b));
        sDotU[f] = dotT(sig, S);
        const T sne = T(m.dc[f]) * (eb[b] - ec[c]);
        lapE[f] = (T(alpha) * sne) * magSf(f);
    }
    // surfaceIntegrate of one component (stride, comp), divided by V
    auto integ = [&](const std::vector<T>& sf, int stride, int comp, std::vector<T>& out) {
        out.assign(nC, T(0));
        for (int f = 0; f < m.nI; ++f) {
            out[m.own[f]] = out[m.own[f]] + sf[(size_t)f * stride + comp];
            out[m.nei[f]] = out[m.nei[f]] - sf[(size_t)f * stride + comp];
        }
        for (int f = m.nI; f < m.nF; ++f) {
            if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
            out[m.own[f]] = out[m.own[f]] + sf[(size_t)f * stride + comp];
        }
        for (int c = 0; c < nC; ++c) out[c] = out[c] / T(m.V[c]);
    };
    std::vector<T> t;
    for (int d = 0; d < 3; ++d) {
        integ(lapU, 3, d, t);
        for (int c = 0; c < nC; ++c) RU[3 * (size_t)c + d] = RU[3 * (size_t)c + d] + t[c];
    }
    for (int d = 0; d < 3; ++d) {
        integ(divTau, 3, d, t);
        for (int c = 0; c < nC; ++c) RU[3 * (size_t)c + d] = RU[3 * (size_t)c + d] + t[c];
    }
    integ(sDotU, 1, 0, t);
    for (int c = 0; c < nC; ++c) RE[c] = RE[c] + t[c];
    integ(lapE, 1, 0, t);
    for (int c = 0; c < nC; ++c) RE[c] = RE[c] + t[c];
}

// b(W, Sf, Cf) of the residual in [rho, rhoE, rhoU] order: AUSM+up or Roe,
// inviscid or viscous.
template <class T>
void residualT(const Case& k, const T* W, const T* Sf, const T* Cf, T* b) {
    const Mesh& m = *k.m;
    const int nC = m.nC, nB = m.nF - m.nI;
    const double g = k.s.gamma;
    StateT<T> s;
    primsFromW(k, W, Sf, s);
    // E, H, c* on cells and boundary faces (derive() in icsf_oracle.cpp)
    std::vector<T> H(nC), cs(nC), Hb(nB, T(0)), csb(nB, T(0));
    for (int c = 0; c < nC; ++c) {
        const V3T<T> U{s.U[3 * c], s.U[3 * c + 1], s.U[3 * c + 2]};
        const T E = T(k.Cv) * s.Tt[c] + T(0.5) * magSqrT(U);
        H[c] = tmax(E, T(SMALL)) + tmax(s.p[c] / s.rho[c], T(SMALL));
        cs[c] = sqrtT(T(2.0 * (g - 1.0) / (g + 1.0)) * H[c]);
    }
    for (int bb = 0; bb < nB; ++bb) {
        if (val(s.TB[bb]) == 0) continue;  // empty
        const V3T<T> U{s.UB[3 * bb], s.UB[3 * bb + 1], s.UB[3 * bb + 2]};
        const T E = T(k.Cv) * s.TB[bb] + T(0.5) * magSqrT(U);
        Hb[bb] = tmax(E, T(SMALL)) + tmax(s.pB[bb] / s.rhoB[bb], T(SMALL));
        csb[bb] = sqrtT(T(2.0 * (g - 1.0) / (g + 1.0)) * Hb[bb]);
    }
    std::vector<T> Ux(nC), Uy(nC), Uz(nC), Ubx(nB), Uby(nB), Ubz(nB);
    for (int c = 0; c < nC; ++c) { Ux[c] = s.U[3 * c]; Uy[c] = s.U[3 * c + 1]; Uz[c] = s.U[3 * c + 2]; }
    for (int bb = 0; bb < nB; ++bb) { Ubx[bb] = s.UB[3 * bb]; Uby[bb] = s.UB[3 * bb + 1]; Ubz[bb] = s.UB[3 * bb + 2]; }
    std::vector<T> UxL, UxR, UyL, UyR, UzL, UzR, pL, pR, rL, rR, HL, HR, cL, cR;
    reconstructT(k, Ux.data(), Ubx.data(), Sf, Cf, UxL, UxR);
    reconstructT(k, Uy.data(), Uby.data(), Sf, Cf, UyL, UyR);
    reconstructT(k, Uz.data(), Ubz.data(), Sf, Cf, UzL, UzR);
    reconstructT(k, s.p.data(), s.pB.data(), Sf, Cf, pL, pR);
    reconstructT(k, s.rho.data(), s.rhoB.data(), Sf, Cf, rL, rR);
    reconstructT(k, H.data(), Hb.data(), Sf, Cf, HL, HR);
    const bool roe = k.s.flux == FLUX_ROE, hllc = k.s.flux == FLUX_HLLC;
    std::vector<T> E(nC), Eb(nB, T(0)), EL, ER;
    if (roe || hllc) {
        for (int c = 0; c < nC; ++c)
            E[c] = T(k.Cv) * s.Tt[c] + T(0.5) * magSqrT(V3T<T>{s.U[3 * c], s.U[3 * c + 1], s.U[3 * c + 2]});
        for (int bb = 0; bb < nB; ++bb)
            if (val(s.TB[bb]) != 0)
                Eb[bb] = T(k.Cv) * s.TB[bb] + T(0.5) * magSqrT(V3T<T>{s.UB[3 * bb], s.UB[3 * bb + 1], s.UB[3 * bb + 2]});
        reconstructT(k, E.data(), Eb.data(), Sf, Cf, EL, ER);
    }
    if (hllc) {
        // HLLC's speed: max(sqrt(gamma/psi), VSMALL), psi = 1/(R T)
        std::vector<T> ch(nC), chb(nB, T(0));
        for (int c = 0; c < nC; ++c) {
            const T psi = T(1.0) / (T(k.s.R) * s.Tt[c]);
            ch[c] = tmax(sqrtT(T(g) / psi), T(VSMALL));
        }
        for (int bb = 0; bb < nB; ++bb)
            if (val(s.TB[bb]) != 0) {
                const T psi = T(1.0) / (T(k.s.R) * s.TB[bb]);
                chb[bb] = tmax(sqrtT(T(g) / psi), T(VSMALL));
            }
        reconstructT(k, ch.data(), chb.data(), Sf, Cf, cL, cR);
    } else if (!roe)
        reconstructT(k, cs.data(), csb.data(), Sf, Cf, cL, cR);
    std::vector<T> acc(5 * (size_t)nC, T(0));
    auto add = [&](int c, T sgn, T ph, V3T<T> pu, T pe) {
        acc[c] = acc[c] + sgn * ph;
        acc[(size_t)nC + c] = acc[(size_t)nC + c] + sgn * pe;
        acc[2 * (size_t)nC + c] = acc[2 * (size_t)nC + c] + sgn * pu.x;
        acc[3 * (size_t)nC + c] = acc[3 * (size_t)nC + c] + sgn * pu.y;
        acc[4 * (size_t)nC + c] = acc[4 * (size_t)nC + c] + sgn * pu.z;
    };
    std::vector<T> PH(m.nF, T(0)), PE(m.nF, T(0));
    std::vector<V3T<T>> PU(m.nF, V3T<T>{T(0), T(0), T(0)});
    for (int f = 0; f < m.nF; ++f) {
        if (f >= m.nI && m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        const V3T<T> S = ld(&Sf[3 * (size_t)f]);
        const T magS = sqrtT(magSqrT(S));
        const V3T<T> UL{UxL[f], UyL[f], UzL[f]}, UR{UxR[f], UyR[f], UzR[f]};
        if (roe)
            roeFaceT(k.s.gamma, k.s.entropyFix, S, magS, rL[f], rR[f], pL[f], pR[f], UL, UR, EL[f], ER[f], HL[f],
                     HR[f], PH[f], PU[f], PE[f]);
        else if (hllc)
            hllcFaceT(k.s.gamma, S, magS, rL[f], rR[f], pL[f], pR[f], UL, UR, cL[f], cR[f], EL[f], ER[f], HL[f],
                      HR[f], PH[f], PU[f], PE[f]);
        else
            ausmFaceT(k.s.lowMach, S, magS, UL, UR, cL[f], cR[f], pL[f], pR[f], rL[f], rR[f], HL[f], HR[f], PH[f],
                      PU[f], PE[f]);
    }
    // surfaceIntegrate: the internal-face loop, then the boundary faces
    for (int f = 0; f < m.nI; ++f) {
        add(m.own[f], T(1.0), PH[f], PU[f], PE[f]);
        add(m.nei[f], T(-1.0), PH[f], PU[f], PE[f]);
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        add(m.own[f], T(1.0), PH[f], PU[f], PE[f]);
    }
    std::vector<T> RU, RE;
    if (k.s.viscous) {
        // the convective part of R, then the viscous sums in the oracle's order
        RU.assign(3 * (size_t)nC, T(0));
        RE.assign(nC, T(0));
        for (int c = 0; c < nC; ++c) {
            RE[c] = -(acc[(size_t)nC + c] / T(m.V[c]));
            for (int d = 0; d < 3; ++d) RU[3 * (size_t)c + d] = -(acc[(size_t)(2 + d) * nC + c] / T(m.V[c]));
        }
        viscousT(k, s, Sf, RU, RE);
    }
    for (int c = 0; c < nC; ++c)
        for (int q = 0; q < 5; ++q) {
            T R = -(acc[(size_t)q * nC + c] / T(m.V[c]));
            if (k.s.viscous && q == 1) R = RE[c];
            if (k.s.viscous && q >= 2) R = RU[3 * (size_t)c + q - 2];
            b[(size_t)q * nC + c] = R * T(m.V[c]);
        }
}
