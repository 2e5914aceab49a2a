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
//   surfaceIntegrate                residualsUpdate.H:1-14

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

// Inviscid state of one evaluation: cell and boundary-face fields.
template <class T>
struct StateT {
    std::vector<T> rho, p, Tt, U;        // cells (U: 3 per cell)
    std::vector<T> pB, TB, UB, rhoB;     // boundary faces (UB: 3 per face)
};

// updateFields.H:7-78 with a zero increment, then the boundary values of
// updateFields.H:80-104.
template <class T>
void primsFromW(const Case& k, const T* W, const T* Sf, StateT<T>& s) {
    const Mesh& m = *k.m;
    const int nC = m.nC, nB = m.nF - m.nI;
    s.rho.assign(nC, T(0)); s.p.assign(nC, T(0)); s.Tt.assign(nC, T(0)); s.U.assign(3 * (size_t)nC, T(0));
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
    }
    s.pB.assign(nB, T(0)); s.TB.assign(nB, T(0)); s.UB.assign(3 * (size_t)nB, T(0)); s.rhoB.assign(nB, T(0));
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

// b(W, Sf, Cf) of the inviscid AUSM+up residual in [rho, rhoE, rhoU] order.
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
        ausmFaceT(k.s.lowMach, S, magS, V3T<T>{UxL[f], UyL[f], UzL[f]}, V3T<T>{UxR[f], UyR[f], UzR[f]}, cL[f], cR[f],
                  pL[f], pR[f], rL[f], rR[f], HL[f], HR[f], PH[f], PU[f], PE[f]);
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
    for (int c = 0; c < nC; ++c)
        for (int q = 0; q < 5; ++q) {
            const T R = -(acc[(size_t)q * nC + c] / T(m.V[c]));
            b[(size_t)q * nC + c] = R * T(m.V[c]);
        }
}
