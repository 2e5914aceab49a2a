// ORACLE — test infrastructure only.
//
// A CPU, FP64, single-threaded restatement of ICSFoam's implicit coupled
// pseudo-time step, written to follow the reference's own operation order
// (OpenFOAM-style face loops, blockFvMatrix LDU sub-blocks, left-
// preconditioned MGS GMRES, sequential scalar-diagonal LU-SGS).  It is the
// checker for the HIP path and the timed CPU baseline; nothing in the
// product links it.  Build: oracle/build.py (g++ -O2 -ffp-contract=off).
//
// Reference files followed (paths relative to the ICSFoam tree):
//   applications/solvers/dbnsFoam/{outerLoop.H,residualsUpdate.H,
//       setCoAndDeltaT.H,boundLocalTimeStep.H,updateFields.H,dbnsFoam.C}
//   src/convectiveFluxSchemes/{convectiveFluxScheme/convectiveFluxScheme.C,
//       ausmPlusUp/ausmPlusUpFluxScheme.C,roe/roeFluxScheme.C}
//   src/viscousFluxScheme/viscousFluxScheme.C
//   src/blockFvMatrix/{blockFvMatrix.C,blockFvOperatorsTemplates.C}
//   src/coupledMatrix/{coupledMatrix.C,coupledMatrixSolver.C,
//       solvers/gmres/gmres.C,preconditioners/lusgs/lusgs.C,
//       smoothers/Jacobi/JacobiSmoother.C,preconditioners/Jacobi/Jacobi.C}
// OpenFOAM-v2112 operators the reference calls (restated from their
// published definitions; not vendored in the reference): primitiveMesh
// geometry, surfaceInterpolation weights/deltaCoeffs, gaussGrad,
// cellLimitedGrad<Venkatakrishnan>, linearUpwind, correctedSnGrad,
// gaussLaplacianScheme, surfaceIntegrate, meshTools::bandCompression.
//
// Unknown ordering everywhere at this interface: [rho, rhoE, rhoUx, rhoUy, rhoUz]
// (coupledMatrix with nScal=2, nVect=1; outerLoop.H:53).
//
// Pins: every function below is checked against closed forms, invariants
// or dense brute force by tests/test_oracle_pins.py, test_hb_oracle.py,
// test_cyclic_oracle.py and test_adjoint_oracle.py (geometry, renumbering,
// limiter function, each flux scheme's coefficients and branches, viscous
// and heat-flux terms, local time step, SER, bounding, state update, LF
// Jacobian, SpMV, LU-SGS, Jacobi, GMRES, HB operator and source).
// Parity unpinned (only an OpenFOAM-v2112 build could fix them): the
// gaussGrad boundary correction on non-orthogonal boundary faces, the
// limited-scheme values on coupled faces beyond the sector/double-sector
// identity, and slip's valueInternalCoeffs on walls that are not
// axis-aligned.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <deque>
#include <functional>
#include <numeric>
#include <vector>

namespace {

const double SMALL = 1.0e-15;
const double VSMALL = 1.0e-300;
const double ROOTVSMALL = 1.0e-150;
const double GREAT = 1.0e15;

typedef std::vector<double> vec;

inline double sqr(double x) { return x * x; }

struct V3 {
    double x, y, z;
};
inline V3 v3(const double* p) { return V3{p[0], p[1], p[2]}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return V3{a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double magSqr(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double mag(V3 a) { return std::sqrt(magSqr(a)); }
inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// 3x3 tensor, row-major xx xy xz yx yy yz zx zy zz
struct T9 {
    double t[9];
};
inline T9 outer(V3 a, V3 b) {
    return T9{{a.x * b.x, a.x * b.y, a.x * b.z, a.y * b.x, a.y * b.y, a.y * b.z, a.z * b.x, a.z * b.y, a.z * b.z}};
}
inline T9 operator+(const T9& a, const T9& b) {
    T9 r;
    for (int i = 0; i < 9; ++i) r.t[i] = a.t[i] + b.t[i];
    return r;
}
inline T9 operator-(const T9& a, const T9& b) {
    T9 r;
    for (int i = 0; i < 9; ++i) r.t[i] = a.t[i] - b.t[i];
    return r;
}
inline T9 operator*(double s, const T9& a) {
    T9 r;
    for (int i = 0; i < 9; ++i) r.t[i] = s * a.t[i];
    return r;
}
inline T9 operator*(const T9& a, double s) {
    T9 r;
    for (int i = 0; i < 9; ++i) r.t[i] = a.t[i] * s;
    return r;
}
// tensor & vector
inline V3 tdotv(const T9& T, V3 v) {
    return V3{T.t[0] * v.x + T.t[1] * v.y + T.t[2] * v.z, T.t[3] * v.x + T.t[4] * v.y + T.t[5] * v.z,
              T.t[6] * v.x + T.t[7] * v.y + T.t[8] * v.z};
}
// vector & tensor
inline V3 vdott(V3 v, const T9& T) {
    return V3{v.x * T.t[0] + v.y * T.t[3] + v.z * T.t[6], v.x * T.t[1] + v.y * T.t[4] + v.z * T.t[7],
              v.x * T.t[2] + v.y * T.t[5] + v.z * T.t[8]};
}
// tensor & tensor
inline T9 tdott(const T9& A, const T9& B) {
    T9 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.t[3 * i + j] = A.t[3 * i] * B.t[j] + A.t[3 * i + 1] * B.t[3 + j] + A.t[3 * i + 2] * B.t[6 + j];
    return r;
}
inline T9 transposeT(const T9& A) {
    return T9{{A.t[0], A.t[3], A.t[6], A.t[1], A.t[4], A.t[7], A.t[2], A.t[5], A.t[8]}};
}
inline T9 dev2(const T9& A) {
    T9 r = A;
    const double tr = A.t[0] + A.t[4] + A.t[8];
    const double s = (2.0 / 3.0) * tr;
    r.t[0] = A.t[0] - s;
    r.t[4] = A.t[4] - s;
    r.t[8] = A.t[8] - s;
    return r;
}

enum { PK_PATCH = 0, PK_WALL = 1, PK_EMPTY = 2, PK_CYCLIC = 3 };
// p: zeroGradient, fixedValue, totalPressure; U: zeroGradient, fixedValue,
// slip, pressureInletOutletVelocity; T: zeroGradient, fixedValue,
// totalTemperature
enum { BC_ZG = 0, BC_FIXED = 1, BC_SLIP = 2, BC_TOTAL_P = 3, BC_PIOV = 4, BC_TOTAL_T = 5 };
enum { FLUX_AUSM = 0, FLUX_ROE = 1, FLUX_HLLC = 2 };

// ---------------------------------------------------------------- mesh ----
struct Mesh {
    int nC = 0, nF = 0, nI = 0, nP = 0;
    vec pts;
    std::vector<int> fOff, fV, own, nei;
    std::vector<int> pStart, pSize, pKind;
    vec Sf, magSf, Cf, C, V, w, dc, nodc, corr;
    std::vector<std::vector<int>> cells;  // OpenFOAM cells(): owned faces, then neighboured faces
    int emptyDir = -1;
    bool isEmptyFace(int f) const {
        for (size_t p = 0; p < pStart.size(); ++p)
            if (f >= pStart[p] && f < pStart[p] + pSize[p]) return pKind[p] == PK_EMPTY;
        return false;
    }
    std::vector<int> facePatch;  // boundary face -> patch
    // cyclic patches: neighbour patch and forwardT (maps vectors of the
    // neighbour side into this side's frame), per patch
    std::vector<int> pNbr;
    vec pT;
    // cyclic faces: per boundary face b, cycCell[b] = cell on the far side
    // (owner of the paired face), -1 if not cyclic; cycR[9b], cycDelta[3b]
    std::vector<int> cycCell;
    vec cycR, cycDelta;
    bool cyclicFace(int f) const { return f >= nI && cycCell[f - nI] >= 0; }
    bool anyCyclic = false;
};

inline V3 transformV(const double* R, V3 v) {
    return V3{R[0] * v.x + R[1] * v.y + R[2] * v.z, R[3] * v.x + R[4] * v.y + R[5] * v.z,
              R[6] * v.x + R[7] * v.y + R[8] * v.z};
}

// primitiveMesh::makeFaceCentresAndAreas / makeCellCentresAndVols and
// surfaceInterpolation::makeWeights / makeDeltaCoeffs /
// makeNonOrthDeltaCoeffs / makeNonOrthCorrectionVectors.
void makeGeometry(Mesh& m) {
    m.Sf.assign(3 * m.nF, 0);
    m.Cf.assign(3 * m.nF, 0);
    m.magSf.assign(m.nF, 0);
    for (int f = 0; f < m.nF; ++f) {
        const int n = m.fOff[f + 1] - m.fOff[f];
        const int* fv = &m.fV[m.fOff[f]];
        V3 S, Cc;
        if (n == 3) {
            V3 a = v3(&m.pts[3 * fv[0]]), b = v3(&m.pts[3 * fv[1]]), c = v3(&m.pts[3 * fv[2]]);
            Cc = (1.0 / 3.0) * (a + b + c);
            S = 0.5 * cross(b - a, c - a);
        } else {
            V3 fc = v3(&m.pts[3 * fv[0]]);
            for (int k = 1; k < n; ++k) fc = fc + v3(&m.pts[3 * fv[k]]);
            fc = fc / (double)n;
            V3 sumN{0, 0, 0}, sumAc{0, 0, 0};
            double sumA = 0;
            for (int k = 0; k < n; ++k) {
                V3 tp = v3(&m.pts[3 * fv[k]]);
                V3 np_ = v3(&m.pts[3 * fv[(k + 1) % n]]);
                V3 c = tp + np_ + fc;
                V3 nn = cross(np_ - tp, fc - tp);
                double a = mag(nn);
                sumN = sumN + nn;
                sumA += a;
                sumAc = sumAc + a * c;
            }
            if (sumA < ROOTVSMALL) {
                Cc = fc;
                S = V3{0, 0, 0};
            } else {
                Cc = (1.0 / 3.0) * sumAc / sumA;
                S = 0.5 * sumN;
            }
        }
        m.Sf[3 * f] = S.x; m.Sf[3 * f + 1] = S.y; m.Sf[3 * f + 2] = S.z;
        m.Cf[3 * f] = Cc.x; m.Cf[3 * f + 1] = Cc.y; m.Cf[3 * f + 2] = Cc.z;
        m.magSf[f] = mag(S);
    }
    std::vector<V3> cEst(m.nC, V3{0, 0, 0});
    std::vector<int> nCF(m.nC, 0);
    for (int f = 0; f < m.nF; ++f) { cEst[m.own[f]] = cEst[m.own[f]] + v3(&m.Cf[3 * f]); nCF[m.own[f]]++; }
    for (int f = 0; f < m.nI; ++f) { cEst[m.nei[f]] = cEst[m.nei[f]] + v3(&m.Cf[3 * f]); nCF[m.nei[f]]++; }
    for (int c = 0; c < m.nC; ++c) cEst[c] = cEst[c] / (double)nCF[c];
    std::vector<V3> cc(m.nC, V3{0, 0, 0});
    vec vol(m.nC, 0);
    for (int f = 0; f < m.nF; ++f) {
        int o = m.own[f];
        V3 fc = v3(&m.Cf[3 * f]);
        double pyr3 = dot(v3(&m.Sf[3 * f]), fc - cEst[o]);
        V3 pc = (3.0 / 4.0) * fc + (1.0 / 4.0) * cEst[o];
        cc[o] = cc[o] + pyr3 * pc;
        vol[o] += pyr3;
    }
    for (int f = 0; f < m.nI; ++f) {
        int o = m.nei[f];
        V3 fc = v3(&m.Cf[3 * f]);
        double pyr3 = dot(v3(&m.Sf[3 * f]), cEst[o] - fc);
        V3 pc = (3.0 / 4.0) * fc + (1.0 / 4.0) * cEst[o];
        cc[o] = cc[o] + pyr3 * pc;
        vol[o] += pyr3;
    }
    m.C.assign(3 * m.nC, 0);
    m.V.assign(m.nC, 0);
    for (int c = 0; c < m.nC; ++c) {
        V3 r = std::fabs(vol[c]) > VSMALL ? cc[c] / vol[c] : cEst[c];
        m.C[3 * c] = r.x; m.C[3 * c + 1] = r.y; m.C[3 * c + 2] = r.z;
        m.V[c] = vol[c] * (1.0 / 3.0);
    }
    m.w.assign(m.nF, 1.0);
    m.dc.assign(m.nF, 0);
    m.nodc.assign(m.nF, 0);
    m.corr.assign(3 * m.nF, 0);
    for (int f = 0; f < m.nI; ++f) {
        V3 S = v3(&m.Sf[3 * f]), fc = v3(&m.Cf[3 * f]);
        V3 Co = v3(&m.C[3 * m.own[f]]), Cn = v3(&m.C[3 * m.nei[f]]);
        double so = std::fabs(dot(S, fc - Co)), sn = std::fabs(dot(S, Cn - fc));
        m.w[f] = sn / (so + sn);
        V3 d = Cn - Co;
        m.dc[f] = 1.0 / mag(d);
        V3 ua = S / m.magSf[f];
        m.nodc[f] = 1.0 / std::max(dot(ua, d), 0.05 * mag(d));
        V3 k = ua - d * m.nodc[f];
        m.corr[3 * f] = k.x; m.corr[3 * f + 1] = k.y; m.corr[3 * f + 2] = k.z;
    }
    for (int f = m.nI; f < m.nF; ++f) {
        V3 S = v3(&m.Sf[3 * f]), fc = v3(&m.Cf[3 * f]);
        V3 Co = v3(&m.C[3 * m.own[f]]);
        V3 n = S / m.magSf[f];
        V3 d = n * dot(n, fc - Co);
        m.dc[f] = 1.0 / mag(d);
        m.nodc[f] = 1.0 / std::max(dot(n, d), 0.05 * mag(d));
    }
    // cells(): OpenFOAM primitiveMesh::calcCells order
    m.cells.assign(m.nC, {});
    for (int f = 0; f < m.nF; ++f) m.cells[m.own[f]].push_back(f);
    for (int f = 0; f < m.nI; ++f) m.cells[m.nei[f]].push_back(f);
    m.facePatch.assign(m.nF - m.nI, -1);
    for (size_t p = 0; p < m.pStart.size(); ++p)
        for (int f = m.pStart[p]; f < m.pStart[p] + m.pSize[p]; ++f) m.facePatch[f - m.nI] = (int)p;
    m.emptyDir = -1;
    double acc[3] = {0, 0, 0};
    bool any = false;
    for (int f = m.nI; f < m.nF; ++f)
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) {
            any = true;
            for (int d = 0; d < 3; ++d) acc[d] += std::fabs(m.Sf[3 * f + d]);
        }
    if (any) m.emptyDir = (int)(std::max_element(acc, acc + 3) - acc);
    // cyclic faces: face i of a cyclic patch couples to face i of its
    // neighbour patch.  cyclicFvPatch::makeWeights: w = dn/(d + dn) with
    // d = nf & (Cf - C) on each side; cyclicFvPatch::delta:
    // (Cf - C) - forwardT & (Cf' - C'); deltaCoeffs, nonOrthDeltaCoeffs and
    // the correction vectors from that delta as for internal faces.
    const int nB = m.nF - m.nI;
    m.cycCell.assign(nB, -1);
    m.cycR.assign(9 * (size_t)nB, 0.0);
    m.cycDelta.assign(3 * (size_t)nB, 0.0);
    m.anyCyclic = false;
    for (size_t p = 0; p < m.pStart.size(); ++p) {
        if (m.pKind[p] != PK_CYCLIC) continue;
        m.anyCyclic = true;
        const int q = m.pNbr[p];
        const double* R = &m.pT[9 * p];
        for (int i = 0; i < m.pSize[p]; ++i) {
            const int f = m.pStart[p] + i, g = m.pStart[q] + i;
            const int o = m.own[f], n = m.own[g];
            V3 nf = v3(&m.Sf[3 * f]) / m.magSf[f], ng = v3(&m.Sf[3 * g]) / m.magSf[g];
            V3 d = v3(&m.Cf[3 * f]) - v3(&m.C[3 * o]);
            V3 dn = v3(&m.Cf[3 * g]) - v3(&m.C[3 * n]);
            const double di = dot(nf, d), dni = dot(ng, dn);
            m.w[f] = dni / (di + dni);
            V3 delta = d - transformV(R, dn);
            m.dc[f] = 1.0 / mag(delta);
            m.nodc[f] = 1.0 / std::max(dot(nf, delta), 0.05 * mag(delta));
            V3 k = nf - delta * m.nodc[f];
            m.corr[3 * f] = k.x; m.corr[3 * f + 1] = k.y; m.corr[3 * f + 2] = k.z;
            const int b = f - m.nI;
            m.cycCell[b] = n;
            for (int e = 0; e < 9; ++e) m.cycR[9 * b + e] = R[e];
            m.cycDelta[3 * b] = delta.x; m.cycDelta[3 * b + 1] = delta.y; m.cycDelta[3 * b + 2] = delta.z;
        }
    }
}

// ---------------------------------------------------------------- case ----
struct Settings {
    int flux = FLUX_AUSM, lowMach = 1, recon = 1, limiter = 1, viscous = 0, precond = 0;
    int nDirs = 30, maxIter = 1, minIter = 0, localTS = 1, bounding = 1;
    double R = 287.0, gamma = 1.4, mu = 0.0, Pr = 0.72, entropyFix = 0.05;
    double tol = 1e-12, relTol = 1e-2, co = 50, coMin = 50, coMax = 50, coIncr = 1.25, coDecr = 0.1;
    double lowerBound = 0.95, rhoMin = -GREAT, TMin = SMALL, TMax = GREAT;
};

// pv: fixed p or totalPressure p0; uv: fixed U or the tangentialVelocity of
// pressureInletOutletVelocity; tv: fixed T or totalTemperature T0.
// pInit/uInit/tInit: initial boundary values of the derived conditions
// (the `value` entries of the field files; empty: p0, the owner cell's U,
// T0).
struct BC {
    int pT = BC_ZG, uT = BC_ZG, tT = BC_ZG;
    vec pv, uv, tv;
    vec pInit, uInit, tInit;
};

// Sub-block of a blockFvMatrix: diag[nC], upper/lower[nI], each ncomp
// wide; iu[nB] the interfacesUpper blocks of the boundary faces (used on
// cyclic faces only; empty when no cyclic face contributed).
struct LduBlock {
    int ncomp = 1;
    vec diag, upper, lower, iu;
    void init(int nC, int nI, int nc) {
        ncomp = nc;
        diag.assign((size_t)nC * nc, 0);
        upper.assign((size_t)nI * nc, 0);
        lower.assign((size_t)nI * nc, 0);
        iu.clear();
    }
};

// One frame zone's contribution (set-up of the rotating or the translating
// zone): faceU & Sf/|Sf| per face, the moving boundary faces and their U.
struct FramePart {
    bool on = false;
    vec un;                // faceU & Sf/|Sf| per face
    std::vector<char> incl;  // moving boundary faces
    vec Ub;                // their U
    std::vector<char> set;   // faces the zone assigns (types 1 and 2)
};

struct Case {
    Mesh* m;
    Settings s;
    std::vector<BC> bcs;
    double Cv = 0;
    // state (cells)
    vec rho, U, T, p, e, rhoU, rhoE;
    // boundary values (per boundary face index f - nI)
    vec pB, UB, TB, rhoB, eB, rhoUB, rhoEB;
    // mass flux of the last flux evaluation on the boundary faces (the
    // registered surfaceScalarField phi the derived conditions look up) and
    // the valueFraction switch of pressureInletOutletVelocity (1: inflow at
    // its last update)
    vec phiB;
    std::vector<int> inflowB;
    // pseudo-time state
    vec pseudoCo, rDT;
    bool haveInit = false, havePrev = false;
    double initS[2], initV[3], prevS[2], prevV[3];
    int corr = 0;
    // pseudotimeControl state: firstIteration_, initResiduals_
    bool ptFirst = true;
    double ptInit[5] = {0, 0, 0, 0, 0};
    // coupled matrix: sub-blocks SS[i][j], SV[i] (i scalar row, vector col),
    // VS[j] (vector row, scalar col j), VV
    LduBlock SS[2][2], SV[2], VS[2], VV;
    vec source;  // [5][nC]
    vec rD;      // LU-SGS reciprocal scalar diagonal
    // MRF, one rotating zone (MRFCoupledZone, MRFCoupledZone.C:42-675):
    // umrf[f] = faceU & Sf/|Sf| (zero outside the zone), omegaC[3c] = Omega
    // in zone cells (zero elsewhere), mrfIncl[b] = 1 on the boundary faces
    // whose U the frame sets (correctBoundaryVelocity), to mrfUb[3b]
    bool mrf = false;
    vec umrf, omegaC, mrfUb;
    std::vector<char> mrfIncl;
    // the rotating zones in list order, their combination and the
    // translating zone's part (set-up)
    std::vector<FramePart> rotZones;
    FramePart frot, ftr;
    // harmonic balance with phase-lagged cyclic patches: the instances'
    // shared set-up and this case's instance index
    const struct PhaseLag* pl = nullptr;
    int inst = 0;
};

inline double gammaOf(const Case& c) { return c.s.gamma; }

// Phase-lagged cyclic patches of harmonic balance (phaseLagCyclicPolyPatch.C,
// phaseLagCyclicFvPatchField.C:303-396, phaseLagCyclicFvPatchFields.C:
// 41-702).  On such a face the far-side value of a lagged field for
// instance I is sum_J Dl[I][J] x_J(far cell), accumulated from zero in J
// order, with Dl = Re(E^-1 M E) the patch side's time shift (M = diag(1,
// e^{i n IBPA}, ..., e^{-i n IBPA}); the owner side takes the patch's IBPA,
// the neighbour side minus the owner's).  Per side: Dl, cylCoords, the
// patch's unit rotation axis and rotation centre.
enum { PL_NONE = 0, PL_RHO, PL_P, PL_U, PL_E, PL_H, PL_C, PL_GAMMA };
struct PhaseLag {
    struct Side {
        vec D;  // [nO][nO]
        bool cyl = false;
        V3 axis{0, 0, 1}, origin{0, 0, 0};
    };
    int nO = 0;
    std::vector<Case*> inst;
    std::vector<int> side;  // per boundary face: its side, -1 off the lagged patches
    std::vector<Side> sides;
};

// Boundary values of p, U, T and the derived rho = psi*p, e = Cv*T fields
// (updateFields.H:90-105: p, U and T corrected in that order).
// fixedValue / zeroGradient / slip from the current cells; the derived
// inflow conditions (OpenFOAM-v2112 finiteVolume, restated):
//   totalPressure (psi, gamma > 1): p0 / (1 + 0.5 psi (gamma-1)/gamma
//       (1 - pos0(phi)) |U_b|^2)^(gamma/(gamma-1)), with psi = 1/(R T_b)
//       and U_b the boundary values before this correction (thermo.correct()
//       runs first and U is corrected after p);
//   pressureInletOutletVelocity: directionMixed with valueFraction
//       neg(phi) (I - n n) and refValue the tangential part of
//       tangentialVelocity: U_b = vf & ref + (I - vf) & U_P;
//   totalTemperature: T0 / (1 + 0.5 psi (gamma-1)/gamma (1 - pos0(phi))
//       |U_b|^2) with the corrected U_b;
// phi the boundary mass flux of the last flux evaluation.  init: the
// createFields state (createFields.H:51-145): the derived conditions take
// their `value` entries (defaults p0, the owner cell's U, T0),
// valueFraction starts at zero and phi = Sf & (rho_b U_b).
void evalBoundary(Case& k, bool init) {
    Mesh& m = *k.m;
    const int nB = m.nF - m.nI;
    const vec UOld = k.UB, TOld = k.TB;
    k.pB.assign(nB, 0); k.UB.assign(3 * nB, 0); k.TB.assign(nB, 0);
    k.rhoB.assign(nB, 0); k.eB.assign(nB, 0); k.rhoUB.assign(3 * nB, 0); k.rhoEB.assign(nB, 0);
    if (init) { k.phiB.assign(nB, 0.0); k.inflowB.assign(nB, 0); }
    const double gM1ByG = (k.s.gamma - 1.0) / k.s.gamma;
    for (size_t pi = 0; pi < m.pStart.size(); ++pi) {
        if (m.pKind[pi] == PK_EMPTY || m.pKind[pi] == PK_CYCLIC) continue;
        const BC& bc = k.bcs[pi];
        for (int i = 0; i < m.pSize[pi]; ++i) {
            const int f = m.pStart[pi] + i, b = f - m.nI, c = m.own[f];
            const double npos = 1.0 - (k.phiB[b] >= 0 ? 1.0 : 0.0);
            // p
            if (bc.pT == BC_FIXED) k.pB[b] = bc.pv[i];
            else if (bc.pT == BC_TOTAL_P) {
                if (init) k.pB[b] = bc.pInit.empty() ? bc.pv[i] : bc.pInit[i];
                else {
                    const double psip = 1.0 / (k.s.R * TOld[b]);
                    const V3 Up = v3(&UOld[3 * b]);
                    k.pB[b] = bc.pv[i] / std::pow(1.0 + 0.5 * psip * gM1ByG * npos * magSqr(Up), 1.0 / gM1ByG);
                }
            } else k.pB[b] = k.p[c];
            // U
            V3 Uc = v3(&k.U[3 * c]);
            V3 ub;
            if (bc.uT == BC_FIXED) ub = v3(&bc.uv[3 * i]);
            else if (bc.uT == BC_SLIP) {
                V3 n = v3(&m.Sf[3 * f]) / m.magSf[f];
                // (U + (I - 2 n n) & U)/2
                T9 Tr{{1 - 2.0 * n.x * n.x, -2.0 * n.x * n.y, -2.0 * n.x * n.z,
                       -2.0 * n.y * n.x, 1 - 2.0 * n.y * n.y, -2.0 * n.y * n.z,
                       -2.0 * n.z * n.x, -2.0 * n.z * n.y, 1 - 2.0 * n.z * n.z}};
                ub = (Uc + tdotv(Tr, Uc)) / 2.0;
            } else if (bc.uT == BC_PIOV) {
                if (init) ub = bc.uInit.empty() ? Uc : v3(&bc.uInit[3 * i]);
                else if (k.phiB[b] < 0) {
                    k.inflowB[b] = 1;
                    const V3 n = v3(&m.Sf[3 * f]) / m.magSf[f];
                    const V3 tv = v3(&bc.uv[3 * i]);
                    const V3 ref = tv - n * dot(n, tv);
                    const T9 vf{{1 - n.x * n.x, -n.x * n.y, -n.x * n.z,
                                 -n.y * n.x, 1 - n.y * n.y, -n.y * n.z,
                                 -n.z * n.x, -n.z * n.y, 1 - n.z * n.z}};
                    const T9 I{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
                    ub = tdotv(vf, ref) + tdotv(I - vf, Uc);
                } else {
                    k.inflowB[b] = 0;
                    ub = Uc;
                }
            } else ub = Uc;
            // MRFCoupledZone::correctBoundaryVelocity: the frame's velocity
            // Omega ^ (Cf - origin) on the rotating patches (all fixedValue),
            // assigned at the start of every outer iteration
            if (k.mrf && k.mrfIncl[b]) ub = v3(&k.mrfUb[3 * b]);
            k.UB[3 * b] = ub.x; k.UB[3 * b + 1] = ub.y; k.UB[3 * b + 2] = ub.z;
            // T
            if (bc.tT == BC_FIXED) k.TB[b] = bc.tv[i];
            else if (bc.tT == BC_TOTAL_T) {
                if (init) k.TB[b] = bc.tInit.empty() ? bc.tv[i] : bc.tInit[i];
                else {
                    const double psip = 1.0 / (k.s.R * TOld[b]);
                    k.TB[b] = bc.tv[i] / (1.0 + 0.5 * psip * gM1ByG * npos * magSqr(ub));
                }
            } else k.TB[b] = k.T[c];
            const double psi = 1.0 / (k.s.R * k.TB[b]);
            k.eB[b] = k.Cv * k.TB[b];
            k.rhoB[b] = psi * k.pB[b];
            for (int d = 0; d < 3; ++d) k.rhoUB[3 * b + d] = k.rhoB[b] * k.UB[3 * b + d];
            k.rhoEB[b] = k.rhoB[b] * (k.eB[b] + 0.5 * magSqr(ub));
            if (init) k.phiB[b] = dot(v3(&m.Sf[3 * f]), v3(&k.rhoUB[3 * b]));
        }
    }
}

// ------------------------------------------------------ reconstruction ----
// Gauss linear gradient (gaussGrad::gradf) of a cell field with boundary
// face values vb (indexed by boundary face).
// On cyclic faces the face value is the coupled linear interpolation
// w*v_P + (1 - w)*vn (surfaceInterpolationScheme::interpolate), vn being
// the far-side value as cyclicFvPatchField::patchNeighbourField gives it
// (vn[b], per boundary face).
void gaussGrad(const Mesh& m, const double* vc, const double* vb, int stride, std::vector<V3>& g,
               const double* vn = nullptr) {
    g.assign(m.nC, V3{0, 0, 0});
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        const double vf = m.w[f] * (vc[o * stride] - vc[n * stride]) + vc[n * stride];
        V3 S = v3(&m.Sf[3 * f]) * vf;
        g[o] = g[o] + S;
        g[n] = g[n] - S;
    }
    for (int f = m.nI; f < m.nF; ++f) {
        const int pi = m.facePatch[f - m.nI];
        if (m.pKind[pi] == PK_EMPTY) continue;
        const int o = m.own[f], b = f - m.nI;
        double vf;
        if (m.cyclicFace(f)) vf = m.w[f] * vc[o * stride] + (1.0 - m.w[f]) * vn[b];
        else vf = vb[b * stride];
        g[o] = g[o] + v3(&m.Sf[3 * f]) * vf;
    }
    for (int c = 0; c < m.nC; ++c) g[c] = g[c] / m.V[c];
}

inline void limitFace(double& lim, double maxD, double minD, double ext, int limiter) {
    double r = 1;
    if (ext > SMALL) r = maxD / ext;
    else if (ext < -SMALL) r = minD / ext;
    double l;
    if (limiter == 1) l = (sqr(r) + 2 * r) / (sqr(r) + r + 2);   // Venkatakrishnan
    else l = std::min(r, 1.0);                                   // minmod
    lim = std::min(lim, l);
}

// cellLimitedGrad<limiter> Gauss linear 1 (k = 1).
void limitedGrad(const Case& k, const double* vc, const double* vb, int stride, std::vector<V3>& g,
                 const double* vn = nullptr) {
    const Mesh& m = *k.m;
    gaussGrad(m, vc, vb, stride, g, vn);
    if (k.s.limiter == 0) return;
    vec mx(m.nC), mn(m.nC);
    for (int c = 0; c < m.nC; ++c) mx[c] = mn[c] = vc[c * stride];
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        mx[o] = std::max(mx[o], vc[n * stride]); mn[o] = std::min(mn[o], vc[n * stride]);
        mx[n] = std::max(mx[n], vc[o * stride]); mn[n] = std::min(mn[n], vc[o * stride]);
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        const int o = m.own[f];
        const double b = m.cyclicFace(f) ? vn[f - m.nI] : vb[(f - m.nI) * stride];
        mx[o] = std::max(mx[o], b); mn[o] = std::min(mn[o], b);
    }
    for (int c = 0; c < m.nC; ++c) { mx[c] -= vc[c * stride]; mn[c] -= vc[c * stride]; }
    vec lim(m.nC, 1.0);
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        V3 fc = v3(&m.Cf[3 * f]);
        limitFace(lim[o], mx[o], mn[o], dot(fc - v3(&m.C[3 * o]), g[o]), k.s.limiter);
        limitFace(lim[n], mx[n], mn[n], dot(fc - v3(&m.C[3 * n]), g[n]), k.s.limiter);
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        const int o = m.own[f];
        limitFace(lim[o], mx[o], mn[o], dot(v3(&m.Cf[3 * f]) - v3(&m.C[3 * o]), g[o]), k.s.limiter);
    }
    for (int c = 0; c < m.nC; ++c) g[c] = g[c] * lim[c];
}

// fvc::interpolate(vf, pos_/neg_, "reconstruct(X)") with linearUpwind (or
// upwind) weights: returns L (owner side) and R (neighbour side) values on
// every face; boundary faces carry the patch value on both sides.  Cyclic
// faces (linearUpwind::correction, coupled branch): L = v_P + (Cf - C_P) &
// grad_P; R = vn + ((Cf - delta) - C_P) & (forwardT & grad_N), the
// far-side gradient transformed as a vector (gradient fields are not
// U-named, so cyclicFvPatchField transforms them).
void reconstruct(const Case& k, const double* vc, const double* vb, int stride, vec& L, vec& R,
                 const double* vn = nullptr) {
    const Mesh& m = *k.m;
    std::vector<V3> g;
    if (k.s.recon == 1) limitedGrad(k, vc, vb, stride, g, vn);
    L.assign(m.nF, 0);
    R.assign(m.nF, 0);
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        double l = 1.0 * (vc[o * stride] - vc[n * stride]) + vc[n * stride];
        double r = 0.0 * (vc[o * stride] - vc[n * stride]) + vc[n * stride];
        if (k.s.recon == 1) {
            V3 fc = v3(&m.Cf[3 * f]);
            l += dot(fc - v3(&m.C[3 * o]), g[o]);
            r += dot(fc - v3(&m.C[3 * n]), g[n]);
        }
        L[f] = l;
        R[f] = r;
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        const int b = f - m.nI;
        if (m.cyclicFace(f)) {
            const int o = m.own[f], n = m.cycCell[b];
            double l = vc[o * stride], r = vn[b];
            if (k.s.recon == 1) {
                V3 fc = v3(&m.Cf[3 * f]);
                l += dot(fc - v3(&m.C[3 * o]), g[o]);
                V3 gR = transformV(&m.cycR[9 * b], g[n]);
                r += dot((fc - v3(&m.cycDelta[3 * b])) - v3(&m.C[3 * o]), gR);
            }
            L[f] = l;
            R[f] = r;
        } else {
            L[f] = R[f] = vb[b * stride];
        }
    }
}

// Far-side values on cyclic faces (cyclicFvPatchField::patchNeighbourField
// as patched in originalOFFiles/constraintFvPatchFields/cyclic/
// cyclicFvPatchField.C:127-188): a scalar field's value at the far cell;
// for the components of U (fields "U.component(i)") component i of
// forwardT & U at the far cell.  comp = -1: scalar field vc.  lag: the
// field's phase-lag rule (PL_*) on phase-lagged faces.
double lagScalar(const Case& k, int b, int field, int comp);
vec cyclicNbr(const Case& k, const double* vc, int stride, int comp, int lag = PL_NONE) {
    const Mesh& m = *k.m;
    const int nB = m.nF - m.nI;
    vec vn(nB, 0.0);
    if (!m.anyCyclic) return vn;
    for (int b = 0; b < nB; ++b) {
        const int n = m.cycCell[b];
        if (n < 0) continue;
        if (lag != PL_NONE && k.pl && k.pl->side[b] >= 0) vn[b] = lagScalar(k, b, lag, comp < 0 ? 0 : comp);
        else if (comp < 0) vn[b] = vc[n * stride];
        else {
            V3 u = transformV(&m.cycR[9 * b], v3(&k.U[3 * n]));
            vn[b] = (&u.x)[comp];
        }
    }
    return vn;
}

// Coupled linear interpolation of a cell scalar on cyclic face f.
inline double interpCyclic(const Mesh& m, int f, double vP, double vN) {
    return m.w[f] * vP + (1.0 - m.w[f]) * vN;
}
// ... and of U: w*U_P + (1 - w)*(forwardT & U_N)
inline V3 lagU(const Case& k, int b);
inline bool lagFace(const Case& k, int b);
inline V3 interpCyclicU(const Case& k, int f) {
    const Mesh& m = *k.m;
    const int b = f - m.nI;
    V3 UN = lagFace(k, b) ? lagU(k, b) : transformV(&m.cycR[9 * b], v3(&k.U[3 * m.cycCell[b]]));
    return m.w[f] * v3(&k.U[3 * m.own[f]]) + (1.0 - m.w[f]) * UN;
}

// ---- phase-lagged far-side values ----
// The fields patchNeighbourField lags (phaseLagCyclicFvPatchField.C:
// 351-358): names starting with U, p, rho, gamma, c, sqrt((gamma|thermo:psi)),
// (he+(0.5*magSqr(U))) and H.  Unregistered ones are formed per instance
// from its registered U, p, rho, e and psi (phaseLagCyclicFvPatchFields.C:
// 128-243): E = e + 0.5|U|^2, H = E + p/rho (without the flux schemes'
// SMALL bounds), c = sqrt(gamma/psi) (also for AUSM+up's renamed c*), the
// velocity components unrotated.  Every other field (gradients, HLLC's
// max(sqrt(gamma/psi), c0), e) keeps the plain cyclic rule.
inline bool lagFace(const Case& k, int b) { return k.pl && k.pl->side[b] >= 0; }

// The lagged velocity vector of boundary face b (the vector "U" field,
// forwardT applied).  cylCoords (phaseLagCyclicFvPatchFields.C:582-640):
// each instance's far velocity in the (r, theta, axis) components at the
// paired face's centre, combined, back to Cartesian there, then forwardT.
V3 lagU(const Case& k, int b) {
    const Mesh& m = *k.m;
    const PhaseLag& L = *k.pl;
    const PhaseLag::Side& sd = L.sides[L.side[b]];
    const double* D = &sd.D[(size_t)k.inst * L.nO];
    const int n = m.cycCell[b];
    const double* R = &m.cycR[9 * (size_t)b];
    if (!sd.cyl) {
        V3 a{0, 0, 0};
        for (int J = 0; J < L.nO; ++J) a = a + D[J] * v3(&L.inst[J]->U[3 * n]);
        return transformV(R, a);
    }
    const int p = m.facePatch[b];
    const int g = m.pStart[m.pNbr[p]] + (m.nI + b - m.pStart[p]);
    const V3 ah = sd.axis;
    V3 r = v3(&m.Cf[3 * (size_t)g]) - sd.origin;
    const double ar = dot(ah, r);
    r = r - ar * ah;
    const double mr = std::sqrt(magSqr(r));
    const V3 rh{r.x / mr, r.y / mr, r.z / mr};
    const V3 th = cross(ah, rh);
    V3 sc{0, 0, 0};
    for (int J = 0; J < L.nO; ++J) {
        const V3 u = v3(&L.inst[J]->U[3 * n]);
        sc = sc + D[J] * V3{dot(u, rh), dot(u, th), dot(u, ah)};
    }
    const V3 c = (sc.x * rh + sc.y * th) + sc.z * ah;
    return transformV(R, c);
}

// The lagged value of a scalar field (comp: the velocity component for
// PL_U) on boundary face b.
double lagScalar(const Case& k, int b, int field, int comp) {
    const Mesh& m = *k.m;
    const PhaseLag& L = *k.pl;
    const PhaseLag::Side& sd = L.sides[L.side[b]];
    if (field == PL_U && sd.cyl) {
        const V3 u = lagU(k, b);
        return (&u.x)[comp];
    }
    const double* D = &sd.D[(size_t)k.inst * L.nO];
    const int n = m.cycCell[b];
    double a = 0.0;
    for (int J = 0; J < L.nO; ++J) {
        const Case& q = *L.inst[J];
        const V3 u = v3(&q.U[3 * n]);
        double x;
        switch (field) {
            case PL_RHO: x = q.rho[n]; break;
            case PL_P: x = q.p[n]; break;
            case PL_U: x = (&u.x)[comp]; break;
            case PL_E: x = q.e[n] + 0.5 * magSqr(u); break;
            case PL_H: x = (q.e[n] + 0.5 * magSqr(u)) + q.p[n] / q.rho[n]; break;
            case PL_C: x = std::sqrt(q.s.gamma / (1.0 / (q.s.R * q.T[n]))); break;
            default: x = q.s.gamma; break;
        }
        a = a + D[J] * x;
    }
    return a;
}

// fvc::interpolate of c = sqrt(gamma/psi) and of gamma on cyclic face f
// (the far side lagged on phase-lagged faces)
inline double interpCyclicC(const Case& k, int f, const double* c) {
    const Mesh& m = *k.m;
    const int b = f - m.nI;
    return interpCyclic(m, f, c[m.own[f]], lagFace(k, b) ? lagScalar(k, b, PL_C, 0) : c[m.cycCell[b]]);
}
inline double interpCyclicGamma(const Case& k, int f) {
    const int b = f - k.m->nI;
    return interpCyclic(*k.m, f, k.s.gamma, lagFace(k, b) ? lagScalar(k, b, PL_GAMMA, 0) : k.s.gamma);
}

inline double interpLinear(const Mesh& m, int f, const double* vc, int stride) {
    return m.w[f] * (vc[m.own[f] * stride] - vc[m.nei[f] * stride]) + vc[m.nei[f] * stride];
}

// Derived cell/boundary fields used by the flux schemes.
struct Derived {
    vec E, H, cstar, Eb, Hb, cstarb, c, cb;  // c = sqrt(gamma/psi)
    vec chl, chlb;                            // HLLC: max(sqrt(gamma/psi), VSMALL)
};
void derive(const Case& k, Derived& d) {
    const Mesh& m = *k.m;
    const double g = k.s.gamma;
    const int nB = m.nF - m.nI;
    d.E.resize(m.nC); d.H.resize(m.nC); d.cstar.resize(m.nC); d.c.resize(m.nC);
    d.Eb.assign(nB, 0); d.Hb.assign(nB, 0); d.cstarb.assign(nB, 0); d.cb.assign(nB, 0);
    d.chl.resize(m.nC); d.chlb.assign(nB, 0);
    for (int c = 0; c < m.nC; ++c) {
        d.E[c] = k.Cv * k.T[c] + 0.5 * magSqr(v3(&k.U[3 * c]));
        d.H[c] = std::max(d.E[c], SMALL) + std::max(k.p[c] / k.rho[c], SMALL);
        d.cstar[c] = std::sqrt(2.0 * (g - 1.0) / (g + 1.0) * d.H[c]);
        const double psi = 1.0 / (k.s.R * k.T[c]);
        d.c[c] = std::sqrt(g / psi);
        d.chl[c] = std::max(d.c[c], VSMALL);
    }
    for (int b = 0; b < nB; ++b) {
        if (k.TB[b] == 0) continue;  // empty
        d.Eb[b] = k.Cv * k.TB[b] + 0.5 * magSqr(v3(&k.UB[3 * b]));
        d.Hb[b] = std::max(d.Eb[b], SMALL) + std::max(k.pB[b] / k.rhoB[b], SMALL);
        d.cstarb[b] = std::sqrt(2.0 * (g - 1.0) / (g + 1.0) * d.Hb[b]);
        const double psi = 1.0 / (k.s.R * k.TB[b]);
        d.cb[b] = std::sqrt(g / psi);
        d.chlb[b] = std::max(d.cb[b], VSMALL);
    }
}

// ---------------------------------------------------------------- fluxes --
// AUSM+up (ausmPlusUpFluxScheme.C:73-299) on one face.
// um: the face's MRF normal velocity (ausmPlusUpFluxScheme.C:101-102, 293):
// the normal velocities relative to the frame, and p_1/2 um |S| in the
// energy flux.
void ausmFace(const Case& k, V3 S, double magS, V3 UL, V3 UR, double cL, double cR, double pL,
              double pR, double rL, double rR, double HL, double HR, double um, double& phi, V3& phiU,
              double& phiE) {
    const double unL = dot(UL, S) / magS - um;
    const double unR = dot(UR, S) / magS - um;
    const double cl = sqr(cL) / std::max(cL, unL);
    const double cr = sqr(cR) / std::max(cR, -unR);
    const double cf = std::min(cl, cr);
    const double ML = unL / cf;
    double MpL, PpL;
    if (std::fabs(ML) < 1.0) {
        const double p2 = 0.25 * sqr(ML + 1), m2 = -0.25 * sqr(ML - 1);
        MpL = p2 * (1 - 2 * m2);
        PpL = p2 * (2 - ML - 3 * ML * m2);
    } else {
        MpL = std::max(ML, 0.0);
        PpL = ML > 0 ? 1.0 : 0.0;
    }
    const double MR = unR / cf;
    double MmR, PmR;
    if (std::fabs(MR) < 1.0) {
        const double m2 = -0.25 * sqr(MR - 1), p2 = 0.25 * sqr(MR + 1);
        MmR = m2 * (1 + 2 * p2);
        PmR = m2 * (-2 - MR + 3 * MR * p2);
    } else {
        MmR = std::min(MR, 0.0);
        PmR = MR < 0 ? 1.0 : 0.0;
    }
    double M12 = MpL + MmR;
    double p12 = PpL * pL + PmR * pR;
    const double Mmean = 0.5 * (sqr(unL) + sqr(unR)) / sqr(cf);
    const double MDiff = -0.25 * std::max(1.0 - Mmean, 0.0) * (pR - pL) / (0.5 * (rL + rR) * sqr(cf));
    if ((M12 > 0.0 && M12 + MDiff <= 0.0) || (M12 < 0.0 && M12 + MDiff >= 0.0)) M12 += 0.2 * MDiff;
    else M12 += MDiff;
    if (k.s.lowMach) p12 += -0.25 * PpL * PmR * (rL + rR) * cf * (unR - unL);
    const bool up = M12 >= 0;
    const V3 Uf = up ? UL : UR;
    const double rhoa = M12 * cf * (up ? rL : rR);
    const V3 rhoaU = rhoa * Uf;
    const double rhoah = rhoa * (up ? HL : HR);
    phi = rhoa * magS;
    phiU = rhoaU * magS + p12 * S;
    phiE = rhoah * magS + p12 * um * magS;
}

struct RoeDiss {
    double CR, CE, ER, EE;
    V3 CU, MR, ME, EU;
    T9 MU;
};

// roeFluxScheme::getRoeDissipation (roeFluxScheme.C:41-241), one face.
RoeDiss roeDissipation(double rho, V3 Vel, double c, double gI, double u, V3 n, double eFix) {
    const double a1 = gI - 1;
    const double theta = 0.5 * a1 * magSqr(Vel);
    const double c2 = sqr(c);
    const double a2 = 1 / (rho * c * std::sqrt(2.0));
    const double a3 = rho / (c * std::sqrt(2.0));
    const double a4 = (theta + c2) / a1;
    const double a5 = 1 - theta / c2;
    const double a6 = theta / a1;
    V3 iP11 = (n * a5) - cross(Vel, n) / rho;
    T9 iP12 = outer(a1 / c2 * n, Vel);
    iP12.t[1] += n.z / rho; iP12.t[2] -= n.y / rho; iP12.t[3] -= n.z / rho;
    iP12.t[5] += n.x / rho; iP12.t[6] += n.y / rho; iP12.t[7] -= n.x / rho;
    V3 iP13 = (-a1) / c2 * n;
    double iP21 = a2 * (theta - c * u);
    V3 iP22 = (-a2) * (a1 * Vel - c * n);
    double iP23 = a1 * a2;
    double iP31 = a2 * (theta + c * u);
    V3 iP32 = (-a2) * (a1 * Vel + c * n);
    double iP33 = a1 * a2;
    double L1 = std::fabs(u), L2 = std::fabs(u + c), L3 = std::fabs(u - c);
    const double eps = eFix * std::max(L2, L3);
    if (L1 < eps) L1 = (sqr(L1) + sqr(eps)) / (2.0 * eps);
    if (L2 < eps) L2 = (sqr(L2) + sqr(eps)) / (2.0 * eps);
    if (L3 < eps) L3 = (sqr(L3) + sqr(eps)) / (2.0 * eps);
    V3 P11 = n;
    double P12 = a3, P13 = a3;
    T9 P21 = outer(Vel, n);
    P21.t[1] -= n.z * rho; P21.t[2] += n.y * rho; P21.t[3] += n.z * rho;
    P21.t[5] -= n.x * rho; P21.t[6] -= n.y * rho; P21.t[7] += n.x * rho;
    V3 P22 = a3 * (Vel + c * n);
    V3 P23 = a3 * (Vel - c * n);
    V3 P31 = n * a6 + rho * cross(Vel, n);
    double P32 = a3 * (a4 + c * u);
    double P33 = a3 * (a4 - c * u);
    RoeDiss D;
    V3 P11L = P11 * L1;
    T9 P21L = P21 * L1;
    V3 P31L = P31 * L1;
    D.CR = dot(P11L, iP11) + L2 * P12 * iP21 + L3 * P13 * iP31;
    D.CU = vdott(P11L, iP12) + (L2 * P12) * iP22 + (L3 * P13) * iP32;
    D.CE = dot(P11L, iP13) + L2 * P12 * iP23 + L3 * P13 * iP33;
    D.MR = tdotv(P21L, iP11) + (L2 * P22) * iP21 + (L3 * P23) * iP31;
    D.MU = tdott(P21L, iP12) + outer(L2 * P22, iP22) + outer(L3 * P23, iP32);
    D.ME = tdotv(P21L, iP13) + (L2 * P22) * iP23 + (L3 * P23) * iP33;
    D.ER = dot(P31L, iP11) + L2 * P32 * iP21 + L3 * P33 * iP31;
    D.EU = vdott(P31L, iP12) + (L2 * P32) * iP22 + (L3 * P33) * iP32;
    D.EE = dot(P31L, iP13) + L2 * P32 * iP23 + L3 * P33 * iP33;
    return D;
}

// roeFluxScheme::calcFlux (roeFluxScheme.C:276-409), one face.
// um: the face's MRF normal velocity (roeFluxScheme.C:361-363, 404-408): the
// Roe-averaged normal velocity of the dissipation relative to the frame,
// and -1/2 |S| um (Q_l + Q_r) after the central flux.
void roeFace(const Case& k, double gI, V3 S, double magS, double rl, double rr, double pl, double pr, V3 Ul, V3 Ur,
             double El, double Er, double Hl, double Hr, double um, double& phi, V3& phiU, double& phiE) {
    const V3 n = S / magS;
    const double coefR = std::sqrt(std::max(VSMALL, rr) / std::max(VSMALL, rl));
    const double rhoRoe = coefR * rl;
    const V3 URoe = (coefR * Ur + Ul) / (coefR + 1.0);
    const double HRoe = (coefR * Hr + Hl) / (coefR + 1.0);
    const double cRoe = std::sqrt(std::fabs((gI - 1.0) * (HRoe - 0.5 * magSqr(URoe))));
    const double uMl = dot(Ul, n), uMr = dot(Ur, n);
    const double uProj = dot(URoe, n) - um;
    const V3 rUl = rl * Ul, rUr = rr * Ur;
    const double rEl = rl * El, rEr = rr * Er;
    RoeDiss D = roeDissipation(rhoRoe, URoe, cRoe, gI, uProj, n, k.s.entropyFix);
    const double dR = rr - rl;
    const V3 dRU = rUr - rUl;
    const double dRE = rEr - rEl;
    phi = -0.5 * magS * (D.CR * dR + dot(D.CU, dRU) + D.CE * dRE);
    phiU = -0.5 * magS * (D.MR * dR + tdotv(D.MU, dRU) + D.ME * dRE);
    phiE = -0.5 * magS * (D.ER * dR + dot(D.EU, dRU) + D.EE * dRE);
    const double rnl = rl * uMl, rnr = rr * uMr;
    phi += 0.5 * magS * (rnl + rnr);
    phiU = phiU + 0.5 * magS * (rnl * Ul + rnr * Ur + n * (pl + pr));
    phiE += 0.5 * magS * (rnl * Hl + rnr * Hr);
    const double hm = 0.5 * magS * um;
    phi -= hm * (rl + rr);
    phiU = phiU - hm * (rUl + rUr);
    phiE -= hm * (rEl + rEr);
}

// hllcFluxScheme::calcFlux (hllcFluxScheme.C:70-240), one face; wave
// speeds from Roe averages, switches pos(x) = x > 0, neg(x) = x < 0.
// um: the face's MRF normal velocity (hllcFluxScheme.C:157-161, 217-218):
// normal velocities relative to the frame, p* um in the star energy fluxes.
void hllcFace(const Case& k, double gI, V3 S, double magS, double rl, double rr, double pl, double pr, V3 Ul,
              V3 Ur, double cl, double cr, double El, double Er, double Hl, double Hr, double um, double& phi,
              V3& phiU, double& phiE) {
    const V3 n = S / magS;
    const double coefR = std::sqrt(std::max(VSMALL, rr) / std::max(VSMALL, rl));
    const V3 uAvg = (coefR * Ur + Ul) / (coefR + 1.0);
    const double HAvg = (coefR * Hr + Hl) / (coefR + 1.0);
    (void)k;
    const double cAvg = std::sqrt(std::fabs((gI - 1.0) * (HAvg - 0.5 * magSqr(uAvg))));
    const double ul = dot(Ul, n) - um, ur = dot(Ur, n) - um, ua = dot(uAvg, n) - um;
    const double Sl = std::min(ul - cl, ua - cAvg);
    const double Sr = std::max(ur + cr, ua + cAvg);
    const double Sm = (rr * ur * (Sr - ur) - rl * ul * (Sl - ul) + pl - pr) / (rr * (Sr - ur) - rl * (Sl - ul));
    const double cSl = Sl > 0 ? 1.0 : 0.0, cSr = Sr < 0 ? 1.0 : 0.0, cSm = Sm > 0 ? 1.0 : 0.0;
    const double cSlm = (1.0 - cSl) * cSm, cSmr = (1.0 - cSm) * (1.0 - cSr);
    const double frl = Sm / (Sl - Sm) * ((Sl - ul) * rl);
    const double frr = Sm / (Sr - Sm) * ((Sr - ur) * rr);
    phi = (cSl * rl * ul + cSlm * frl + cSmr * frr + cSr * rr * ur) * magS;
    const double psl = rl * (ul - Sl) * (ul - Sm) + pl;
    const double psr = rr * (ur - Sr) * (ur - Sm) + pr;
    const V3 rUsl = 1.0 / (Sl - Sm) * ((Sl - ul) * rl * Ul + (psl - pl) * n);
    const V3 rUsr = 1.0 / (Sr - Sm) * ((Sr - ur) * rr * Ur + (psr - pr) * n);
    const V3 fUl = Sm * rUsl + psl * n;
    const V3 fUr = Sm * rUsr + psr * n;
    phiU = (cSl * (rl * ul * Ul + pl * n) + cSlm * fUl + cSmr * fUr + cSr * (rr * ur * Ur + pr * n)) * magS;
    const double rEsl = 1.0 / (Sl - Sm) * ((Sl - ul) * (rl * El) - pl * ul + psl * Sm);
    const double rEsr = 1.0 / (Sr - Sm) * ((Sr - ur) * (rr * Er) - pr * ur + psr * Sm);
    const double fEl = Sm * (rEsl + psl) + psl * um;
    const double fEr = Sm * (rEsr + psr) + psr * um;
    phiE = (cSl * rl * Hl * ul + cSlm * fEl + cSmr * fEr + cSr * rr * Hr * ur) * magS;
}

// convectiveFluxScheme::calcFlux for the selected scheme on all faces.
void calcFlux(Case& k, vec& phi, vec& phiU, vec& phiE) {
    Mesh& m = *k.m;
    Derived d;
    derive(k, d);
    const int nB = m.nF - m.nI;
    vec Ux(m.nC), Uy(m.nC), Uz(m.nC), Ubx(nB), Uby(nB), Ubz(nB);
    for (int c = 0; c < m.nC; ++c) { Ux[c] = k.U[3 * c]; Uy[c] = k.U[3 * c + 1]; Uz[c] = k.U[3 * c + 2]; }
    for (int b = 0; b < nB; ++b) { Ubx[b] = k.UB[3 * b]; Uby[b] = k.UB[3 * b + 1]; Ubz[b] = k.UB[3 * b + 2]; }
    vec UxL, UxR, UyL, UyR, UzL, UzR, pL, pR, rL, rR, HL, HR, cL, cR, EL, ER;
    auto rec = [&](const double* vc, const double* vb, int comp, vec& L, vec& R, int lag) {
        const vec vn = cyclicNbr(k, vc, 1, comp, lag);
        reconstruct(k, vc, vb, 1, L, R, vn.data());
    };
    rec(Ux.data(), Ubx.data(), 0, UxL, UxR, PL_U);
    rec(Uy.data(), Uby.data(), 1, UyL, UyR, PL_U);
    rec(Uz.data(), Ubz.data(), 2, UzL, UzR, PL_U);
    rec(k.p.data(), k.pB.data(), -1, pL, pR, PL_P);
    rec(k.rho.data(), k.rhoB.data(), -1, rL, rR, PL_RHO);
    rec(d.H.data(), d.Hb.data(), -1, HL, HR, PL_H);
    // AUSM+up's c* is named "c": lagged as sqrt(gamma/psi); HLLC's
    // max(sqrt(gamma/psi), c0) is not lagged
    if (k.s.flux == FLUX_AUSM) rec(d.cstar.data(), d.cstarb.data(), -1, cL, cR, PL_C);
    else rec(d.E.data(), d.Eb.data(), -1, EL, ER, PL_E);
    if (k.s.flux == FLUX_HLLC) rec(d.chl.data(), d.chlb.data(), -1, cL, cR, PL_NONE);
    phi.assign(m.nF, 0);
    phiU.assign(3 * m.nF, 0);
    phiE.assign(m.nF, 0);
    for (int f = 0; f < m.nF; ++f) {
        if (f >= m.nI && m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        V3 S = v3(&m.Sf[3 * f]);
        V3 UL{UxL[f], UyL[f], UzL[f]}, UR{UxR[f], UyR[f], UzR[f]};
        double ph, pe;
        V3 pu;
        // fvc::interpolate(gamma): exact on internal faces (w (g - g) + g),
        // w g + (1 - w) g on cyclic faces (the far gamma lagged on
        // phase-lagged faces)
        const double gI = m.cyclicFace(f) ? interpCyclicGamma(k, f) : k.s.gamma;
        const double um = k.mrf ? k.umrf[f] : 0.0;
        if (k.s.flux == FLUX_AUSM)
            ausmFace(k, S, m.magSf[f], UL, UR, cL[f], cR[f], pL[f], pR[f], rL[f], rR[f], HL[f], HR[f], um, ph, pu,
                     pe);
        else if (k.s.flux == FLUX_ROE)
            roeFace(k, gI, S, m.magSf[f], rL[f], rR[f], pL[f], pR[f], UL, UR, EL[f], ER[f], HL[f], HR[f], um, ph,
                    pu, pe);
        else
            hllcFace(k, gI, S, m.magSf[f], rL[f], rR[f], pL[f], pR[f], UL, UR, cL[f], cR[f], EL[f], ER[f], HL[f],
                     HR[f], um, ph, pu, pe);
        phi[f] = ph;
        if (f >= m.nI && !m.cyclicFace(f)) k.phiB[f - m.nI] = ph;
        phiU[3 * f] = pu.x; phiU[3 * f + 1] = pu.y; phiU[3 * f + 2] = pu.z;
        phiE[f] = pe;
    }
}

// fvc::surfaceIntegrate of one scalar surface field (internal+boundary),
// divided by V.
void surfaceIntegrate(const Mesh& m, const double* sf, int stride, int comp, vec& out) {
    out.assign(m.nC, 0);
    for (int f = 0; f < m.nI; ++f) {
        out[m.own[f]] += sf[f * stride + comp];
        out[m.nei[f]] -= sf[f * stride + comp];
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        out[m.own[f]] += sf[f * stride + comp];
    }
    for (int c = 0; c < m.nC; ++c) out[c] /= m.V[c];
}

// Viscous residual terms (residualsUpdate.H:16-43), added to R (per-volume).
void viscousResidual(Case& k, vec& RU /*3nC*/, vec& RE) {
    Mesh& m = *k.m;
    const int nB = m.nF - m.nI;
    const double mu = k.s.mu;
    const double alpha = k.s.gamma * mu / k.s.Pr;  // alphahe = Cp/Cv * kappa/Cp
    // grad(U) Gauss linear: tensor gU_ij = d U_j / d x_i
    std::vector<T9> gU(m.nC, T9{{0}});
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        V3 Uf = m.w[f] * (v3(&k.U[3 * o]) - v3(&k.U[3 * n])) + v3(&k.U[3 * n]);
        T9 t = outer(v3(&m.Sf[3 * f]), Uf);
        gU[o] = gU[o] + t;
        gU[n] = gU[n] - t;
    }
    // (cyclic faces: the coupled interpolation w U_P + (1 - w) U_far)
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        const V3 Ub = m.cyclicFace(f) ? interpCyclicU(k, f) : v3(&k.UB[3 * (f - m.nI)]);
        T9 t = outer(v3(&m.Sf[3 * f]), Ub);
        gU[m.own[f]] = gU[m.own[f]] + t;
    }
    for (int c = 0; c < m.nC; ++c)
        for (int i = 0; i < 9; ++i) gU[c].t[i] /= m.V[c];
    // boundary snGrad of U and gradient boundary values (physical patches)
    std::vector<V3> snU(nB, V3{0, 0, 0});
    std::vector<T9> gUb(nB, T9{{0}});
    for (size_t pi = 0; pi < m.pStart.size(); ++pi) {
        if (m.pKind[pi] == PK_EMPTY || m.pKind[pi] == PK_CYCLIC) continue;
        for (int i = 0; i < m.pSize[pi]; ++i) {
            const int f = m.pStart[pi] + i, b = f - m.nI, c = m.own[f];
            V3 n = v3(&m.Sf[3 * f]) / m.magSf[f];
            V3 Uc = v3(&k.U[3 * c]);
            const int ut = k.bcs[pi].uT;
            if (ut == BC_ZG) snU[b] = V3{0, 0, 0};
            else if (ut == BC_SLIP) {
                T9 Tr{{1 - 2.0 * n.x * n.x, -2.0 * n.x * n.y, -2.0 * n.x * n.z,
                       -2.0 * n.y * n.x, 1 - 2.0 * n.y * n.y, -2.0 * n.y * n.z,
                       -2.0 * n.z * n.x, -2.0 * n.z * n.y, 1 - 2.0 * n.z * n.z}};
                snU[b] = (tdotv(Tr, Uc) - Uc) * (m.dc[f] / 2.0);
            } else snU[b] = m.dc[f] * (v3(&k.UB[3 * b]) - Uc);
            gUb[b] = gU[c] + outer(n, snU[b] - vdott(n, gU[c]));
        }
    }
    std::vector<T9> tau(m.nC), taub(nB, T9{{0}});
    for (int c = 0; c < m.nC; ++c) tau[c] = mu * dev2(transposeT(gU[c]));
    for (int b = 0; b < nB; ++b) taub[b] = mu * dev2(transposeT(gUb[b]));
    // eCalc and its Gauss gradient
    vec ec(m.nC), eb(nB, 0);
    for (int c = 0; c < m.nC; ++c) ec[c] = k.rhoE[c] / k.rho[c] - 0.5 * magSqr(v3(&k.U[3 * c]));
    for (int b = 0; b < nB; ++b)
        if (k.rhoB[b] != 0) eb[b] = k.rhoEB[b] / k.rhoB[b] - 0.5 * magSqr(v3(&k.UB[3 * b]));
    std::vector<V3> ge;
    const vec ecFar = cyclicNbr(k, ec.data(), 1, -1, PL_NONE);
    gaussGrad(m, ec.data(), eb.data(), 1, ge, ecFar.data());
    // surface fields
    vec lapU(3 * m.nF, 0), divTau(3 * m.nF, 0), sDotU(m.nF, 0), lapE(m.nF, 0);
    for (int f = 0; f < m.nI; ++f) {
        const int o = m.own[f], n = m.nei[f];
        const double w = m.w[f];
        V3 S = v3(&m.Sf[3 * f]);
        V3 k3 = v3(&m.corr[3 * f]);
        T9 gUf = w * (gU[o] - gU[n]) + gU[n];
        T9 tauf = w * (tau[o] - tau[n]) + tau[n];
        V3 Uf = w * (v3(&k.U[3 * o]) - v3(&k.U[3 * n])) + v3(&k.U[3 * n]);
        const double muf = w * (mu - mu) + mu;
        const double alf = w * (alpha - alpha) + alpha;
        V3 sn = m.nodc[f] * (v3(&k.U[3 * n]) - v3(&k.U[3 * o])) + vdott(k3, gUf);
        V3 l = (muf * sn) * m.magSf[f];
        lapU[3 * f] = l.x; lapU[3 * f + 1] = l.y; lapU[3 * f + 2] = l.z;
        V3 dt = vdott(S, tauf);
        divTau[3 * f] = dt.x; divTau[3 * f + 1] = dt.y; divTau[3 * f + 2] = dt.z;
        V3 sig = tdotv(muf * gUf + tauf, Uf);
        sDotU[f] = dot(sig, S);
        V3 gef = w * (ge[o] - ge[n]) + ge[n];
        const double sne = m.nodc[f] * (ec[n] - ec[o]) + dot(k3, gef);
        lapE[f] = (alf * sne) * m.magSf[f];
    }
    for (int f = m.nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - m.nI]] == PK_EMPTY) continue;
        const int b = f - m.nI, c = m.own[f];
        V3 S = v3(&m.Sf[3 * f]);
        if (m.cyclicFace(f)) {
            // coupled faces: every interpolation w v_P + (1 - w) v_far with
            // the far value through patchNeighbourField: U (lagged on a
            // phase-lagged face), grad(U) and tauMC as tensors (forwardT &
            // T & forwardT^T), grad(eCalc) as a vector; the snGrad
            // correction interpolates the component gradients
            // grad(U.component(i)) (correctedSnGrad<vector>, each rotated as
            // a vector: forwardT & grad(U)); nonOrthDeltaCoeffs and the
            // correction vectors of the coupled face
            const int n = m.cycCell[b];
            const double w = m.w[f];
            const double* R9 = &m.cycR[9 * (size_t)b];
            const T9 Rt{{R9[0], R9[1], R9[2], R9[3], R9[4], R9[5], R9[6], R9[7], R9[8]}};
            const V3 k3 = v3(&m.corr[3 * f]);
            const V3 UP = v3(&k.U[3 * c]);
            const V3 UN = lagFace(k, b) ? lagU(k, b) : transformV(R9, v3(&k.U[3 * n]));
            const T9 gUc = w * gU[c] + (1.0 - w) * tdott(Rt, gU[n]);
            const T9 gUf = w * gU[c] + (1.0 - w) * tdott(tdott(Rt, gU[n]), transposeT(Rt));
            const T9 tauf = w * tau[c] + (1.0 - w) * tdott(tdott(Rt, tau[n]), transposeT(Rt));
            const V3 Uf = w * UP + (1.0 - w) * UN;
            const double muf = w * mu + (1.0 - w) * mu;
            const double alf = w * alpha + (1.0 - w) * alpha;
            const V3 sn = m.nodc[f] * (UN - UP) + vdott(k3, gUc);
            V3 l = (muf * sn) * m.magSf[f];
            lapU[3 * f] = l.x; lapU[3 * f + 1] = l.y; lapU[3 * f + 2] = l.z;
            V3 dt = vdott(S, tauf);
            divTau[3 * f] = dt.x; divTau[3 * f + 1] = dt.y; divTau[3 * f + 2] = dt.z;
            V3 sig = tdotv(muf * gUf + tauf, Uf);
            sDotU[f] = dot(sig, S);
            const V3 gef = w * ge[c] + (1.0 - w) * transformV(R9, ge[n]);
            const double sne = m.nodc[f] * (ec[n] - ec[c]) + dot(k3, gef);
            lapE[f] = (alf * sne) * m.magSf[f];
            continue;
        }
        V3 l = (mu * snU[b]) * m.magSf[f];
        lapU[3 * f] = l.x; lapU[3 * f + 1] = l.y; lapU[3 * f + 2] = l.z;
        V3 dt = vdott(S, taub[b]);
        divTau[3 * f] = dt.x; divTau[3 * f + 1] = dt.y; divTau[3 * f + 2] = dt.z;
        V3 sig = tdotv(mu * gUb[b] + taub[b], v3(&k.UB[3 * b]));
        sDotU[f] = dot(sig, S);
        const double sne = m.dc[f] * (eb[b] - ec[c]);
        lapE[f] = (alpha * sne) * m.magSf[f];
    }
    vec t;
    for (int d = 0; d < 3; ++d) {
        surfaceIntegrate(m, lapU.data(), 3, d, t);
        for (int c = 0; c < m.nC; ++c) RU[3 * c + d] += t[c];
    }
    for (int d = 0; d < 3; ++d) {
        surfaceIntegrate(m, divTau.data(), 3, d, t);
        for (int c = 0; c < m.nC; ++c) RU[3 * c + d] += t[c];
    }
    surfaceIntegrate(m, sDotU.data(), 1, 0, t);
    for (int c = 0; c < m.nC; ++c) RE[c] += t[c];
    surfaceIntegrate(m, lapE.data(), 1, 0, t);
    for (int c = 0; c < m.nC; ++c) RE[c] += t[c];
}

// residualsUpdate.H: b = R*V in [rho, rhoE, rhoU] order.
void residual(Case& k, vec& b) {
    Mesh& m = *k.m;
    vec phi, phiU, phiE;
    calcFlux(k, phi, phiU, phiE);
    vec Rr, RU(3 * m.nC), RE, t;
    surfaceIntegrate(m, phi.data(), 1, 0, Rr);
    for (int c = 0; c < m.nC; ++c) Rr[c] = -Rr[c];
    for (int d = 0; d < 3; ++d) {
        surfaceIntegrate(m, phiU.data(), 3, d, t);
        for (int c = 0; c < m.nC; ++c) RU[3 * c + d] = -t[c];
    }
    surfaceIntegrate(m, phiE.data(), 1, 0, RE);
    for (int c = 0; c < m.nC; ++c) RE[c] = -RE[c];
    if (k.s.viscous) viscousResidual(k, RU, RE);
    const int nC = m.nC;
    b.assign(5 * (size_t)nC, 0);
    for (int c = 0; c < nC; ++c) {
        b[c] = Rr[c] * m.V[c];
        b[nC + c] = RE[c] * m.V[c];
        for (int d = 0; d < 3; ++d) b[(2 + d) * (size_t)nC + c] = RU[3 * c + d] * m.V[c];
    }
    // MRF: the momentum source -(Omega ^ rho U) V that addMRFSource adds to
    // the coupled system's source before every solve
    // (convectiveFluxScheme.C:123-129); a rotating zone only
    if (k.mrf && k.frot.on)
        for (int c = 0; c < nC; ++c) {
            const V3 rU{k.rho[c] * k.U[3 * c], k.rho[c] * k.U[3 * c + 1], k.rho[c] * k.U[3 * c + 2]};
            const V3 s = cross(v3(&k.omegaC[3 * c]), rU) * m.V[c];
            b[2 * (size_t)nC + c] -= s.x;
            b[3 * (size_t)nC + c] -= s.y;
            b[4 * (size_t)nC + c] -= s.z;
        }
}

// setCoAndDeltaT.H:39-142 (local time stepping): rPseudoDeltaT.
void localDeltaT(Case& k) {
    Mesh& m = *k.m;
    Derived d;
    derive(k, d);
    k.rDT.assign(m.nC, 0.0);
    for (int f = 0; f < m.nI; ++f) {
        const double cf = interpLinear(m, f, d.c.data(), 1);
        V3 Uf = m.w[f] * (v3(&k.U[3 * m.own[f]]) - v3(&k.U[3 * m.nei[f]])) + v3(&k.U[3 * m.nei[f]]);
        const double lam = cf + std::fabs(dot(Uf, v3(&m.Sf[3 * f]) / m.magSf[f]) - (k.mrf ? k.umrf[f] : 0.0));
        const double frdt = m.nodc[f] * lam;
        k.rDT[m.own[f]] = std::max(k.rDT[m.own[f]], frdt);
        k.rDT[m.nei[f]] = std::max(k.rDT[m.nei[f]], frdt);
    }
    for (size_t pi = 0; pi < m.pStart.size(); ++pi) {
        if (m.pKind[pi] == PK_CYCLIC) {
            // coupled patches: nonOrthDeltaCoeffs * lambda of the coupled
            // interpolation (setCoAndDeltaT.H:78-89)
            for (int f = m.pStart[pi]; f < m.pStart[pi] + m.pSize[pi]; ++f) {
                const int b = f - m.nI, c = m.own[f];
                const double cf = interpCyclicC(k, f, d.c.data());
                const V3 Uf = interpCyclicU(k, f);
                const double lam = cf + std::fabs(dot(Uf, v3(&m.Sf[3 * f]) / m.magSf[f]) - (k.mrf ? k.umrf[f] : 0.0));
                k.rDT[c] = std::max(k.rDT[c], m.nodc[f] * lam);
            }
            continue;
        }
        if (m.pKind[pi] != PK_WALL) continue;
        for (int f = m.pStart[pi]; f < m.pStart[pi] + m.pSize[pi]; ++f) {
            const int c = m.own[f];
            const double pl = 0.5 * m.nodc[f] *
                              (d.c[c] + std::fabs(dot(v3(&k.U[3 * c]), v3(&m.Sf[3 * f]) / m.magSf[f]) -
                                                  (k.mrf ? k.umrf[f] : 0.0)));
            k.rDT[c] = std::max(k.rDT[c], pl);
        }
    }
    for (int c = 0; c < m.nC; ++c) k.rDT[c] /= k.pseudoCo[c];
}

// ------------------------------------------------------------- Jacobian ---
// addFluxTerms / addDissipationJacobian / addBoundaryTerms /
// addTemporalTerms (convectiveFluxScheme.C:47-546), viscous LF term
// (viscousFluxScheme.C:220-247).  Block layout follows blockFvMatrix:
// insertBlock (blockFvMatrix.C:212-268) and insertDissipationBlock (:272-326)
// build a temporary matrix whose diagonal comes from negSumDiag, then add /
// subtract it.
void addBlockTo(LduBlock& A, const LduBlock& B, double sign) {
    for (size_t i = 0; i < A.diag.size(); ++i) A.diag[i] += sign * B.diag[i];
    for (size_t i = 0; i < A.upper.size(); ++i) A.upper[i] += sign * B.upper[i];
    for (size_t i = 0; i < A.lower.size(); ++i) A.lower[i] += sign * B.lower[i];
    if (!B.iu.empty()) {
        if (A.iu.empty()) A.iu.assign(B.iu.size(), 0.0);
        for (size_t i = 0; i < A.iu.size(); ++i) A.iu[i] += sign * B.iu[i];
    }
}
void negSumDiag(const Mesh& m, LduBlock& B) {
    const int nc = B.ncomp;
    for (int f = 0; f < m.nI; ++f)
        for (int q = 0; q < nc; ++q) {
            B.diag[m.own[f] * nc + q] -= B.lower[f * nc + q];
            B.diag[m.nei[f] * nc + q] -= B.upper[f * nc + q];
        }
}

void assemble(Case& k) {
    Mesh& m = *k.m;
    const int nC = m.nC, nI = m.nI, nB = m.nF - m.nI;
    const double g = k.s.gamma;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) k.SS[i][j].init(nC, nI, 1);
    for (int i = 0; i < 2; ++i) { k.SV[i].init(nC, nI, 3); k.VS[i].init(nC, nI, 3); }
    k.VV.init(nC, nI, 9);
    Derived d;
    derive(k, d);
    // left/right states for the Jacobian: U (reconstruct(U)) and E (reconstruct(T))
    vec Ux(nC), Uy(nC), Uz(nC), Ubx(nB), Uby(nB), Ubz(nB);
    for (int c = 0; c < nC; ++c) { Ux[c] = k.U[3 * c]; Uy[c] = k.U[3 * c + 1]; Uz[c] = k.U[3 * c + 2]; }
    for (int b = 0; b < nB; ++b) { Ubx[b] = k.UB[3 * b]; Uby[b] = k.UB[3 * b + 1]; Ubz[b] = k.UB[3 * b + 2]; }
    vec UxL, UxR, UyL, UyR, UzL, UzR, EL, ER;
    {
        vec n0 = cyclicNbr(k, Ux.data(), 1, 0, PL_U), n1 = cyclicNbr(k, Uy.data(), 1, 1, PL_U),
            n2 = cyclicNbr(k, Uz.data(), 1, 2, PL_U);
        vec nE = cyclicNbr(k, d.E.data(), 1, -1, PL_E);
        reconstruct(k, Ux.data(), Ubx.data(), 1, UxL, UxR, n0.data());
        reconstruct(k, Uy.data(), Uby.data(), 1, UyL, UyR, n1.data());
        reconstruct(k, Uz.data(), Ubz.data(), 1, UzL, UzR, n2.data());
        reconstruct(k, d.E.data(), d.Eb.data(), 1, EL, ER, nE.data());
    }
    // insertBlock for each sub-block: temporaries with negSumDiag then +=
    LduBlock tS, tV, tT;
    auto insert = [&](LduBlock& target, int nc, auto leftFn, auto rightFn) {
        LduBlock mx;
        mx.init(nC, nI, nc);
        double lv[9], rv[9];
        for (int f = 0; f < nI; ++f) {
            leftFn(f, lv);
            rightFn(f, rv);
            const double ms = m.magSf[f];
            for (int q = 0; q < nc; ++q) {
                mx.upper[f * nc + q] = 0.5 * ms * rv[q];
                mx.lower[f * nc + q] = -0.5 * ms * lv[q];
            }
        }
        negSumDiag(m, mx);
        // coupled patches: interfacesUpper = 1/2|S| right, interfacesLower =
        // -1/2|S| left, diag[own] -= interfacesLower (blockFvMatrix.C:241-265)
        if (m.anyCyclic) {
            mx.iu.assign((size_t)nB * nc, 0.0);
            for (int f = nI; f < m.nF; ++f) {
                if (!m.cyclicFace(f)) continue;
                leftFn(f, lv);
                rightFn(f, rv);
                const double ms = m.magSf[f];
                const int b = f - nI, o = m.own[f];
                for (int q = 0; q < nc; ++q) {
                    mx.iu[b * nc + q] = 0.5 * ms * rv[q];
                    mx.diag[o * nc + q] -= -0.5 * ms * lv[q];
                }
            }
        }
        addBlockTo(target, mx, 1.0);
    };
    auto state = [&](int f, bool left, V3& U, double& E) {
        if (left) { U = V3{UxL[f], UyL[f], UzL[f]}; E = EL[f]; }
        else { U = V3{UxR[f], UyR[f], UzR[f]}; E = ER[f]; }
    };
    auto nrm = [&](int f) { return v3(&m.Sf[3 * f]) / m.magSf[f]; };
    // linear interpolation of a uniform gamma: exact on internal faces,
    // w g + (1 - w) g on cyclic faces (the far gamma lagged on
    // phase-lagged faces)
    auto gam = [&](int f) { return m.cyclicFace(f) ? interpCyclicGamma(k, f) : g; };
    auto theta = [&](int f, V3 U) { return 0.5 * (gam(f) - 1) * magSqr(U); };
    // dContByRhoU: n, n
    insert(k.SV[0], 3, [&](int f, double* o) { V3 n = nrm(f); o[0] = n.x; o[1] = n.y; o[2] = n.z; },
           [&](int f, double* o) { V3 n = nrm(f); o[0] = n.x; o[1] = n.y; o[2] = n.z; });
    auto momByRho = [&](bool left) {
        return [&, left](int f, double* o) {
            V3 U; double E; state(f, left, U, E);
            V3 n = nrm(f);
            V3 r = (n * theta(f, U)) - (U * dot(U, n));
            o[0] = r.x; o[1] = r.y; o[2] = r.z;
        };
    };
    insert(k.VS[0], 3, momByRho(true), momByRho(false));
    auto momByRhoU = [&](bool left) {
        return [&, left](int f, double* o) {
            V3 U; double E; state(f, left, U, E);
            V3 n = nrm(f);
            const double a2 = gam(f) - 1;
            T9 r = outer(U, n) - outer(a2 * n, U);
            const double pu = dot(U, n);
            r.t[0] += pu; r.t[4] += pu; r.t[8] += pu;
            for (int q = 0; q < 9; ++q) o[q] = r.t[q];
        };
    };
    insert(k.VV, 9, momByRhoU(true), momByRhoU(false));
    auto momByRhoE = [&](bool left) {
        return [&, left](int f, double* o) {
            V3 n = nrm(f);
            V3 r = n * (gam(f) - 1);
            (void)left;
            o[0] = r.x; o[1] = r.y; o[2] = r.z;
        };
    };
    insert(k.VS[1], 3, momByRhoE(true), momByRhoE(false));
    auto enByRho = [&](bool left) {
        return [&, left](int f, double* o) {
            V3 U; double E; state(f, left, U, E);
            V3 n = nrm(f);
            const double th = theta(f, U);
            const double a1 = gam(f) * E - th;
            o[0] = dot(U, n) * (th - a1);
        };
    };
    insert(k.SS[1][0], 1, enByRho(true), enByRho(false));
    auto enByRhoU = [&](bool left) {
        return [&, left](int f, double* o) {
            V3 U; double E; state(f, left, U, E);
            V3 n = nrm(f);
            const double th = theta(f, U);
            const double a1 = gam(f) * E - th;
            const double a2 = gam(f) - 1;
            V3 r = (n * a1) - ((a2 * U) * dot(U, n));
            o[0] = r.x; o[1] = r.y; o[2] = r.z;
        };
    };
    insert(k.SV[1], 3, enByRhoU(true), enByRhoU(false));
    auto enByRhoE = [&](bool left) {
        return [&, left](int f, double* o) {
            V3 U; double E; state(f, left, U, E);
            o[0] = gam(f) * dot(U, nrm(f));
        };
    };
    insert(k.SS[1][1], 1, enByRhoE(true), enByRhoE(false));
    // MRF: fvj::div(w, um |S|) (blockFvOperatorsTemplates.C:148-200: upper
    // sf (1 - w), lower -sf w, negSumDiag, coupled patches diag -=
    // interfacesLower) subtracted from the rho, rhoU (times I) and rhoE
    // diagonal sub-blocks (convectiveFluxScheme.C:477-481)
    if (k.mrf) {
        LduBlock mx;
        mx.init(nC, nI, 1);
        for (int f = 0; f < nI; ++f) {
            const double sf = k.umrf[f] * m.magSf[f];
            mx.upper[f] = sf * (1 - m.w[f]);
            mx.lower[f] = -sf * m.w[f];
        }
        negSumDiag(m, mx);
        if (m.anyCyclic) {
            mx.iu.assign((size_t)nB, 0.0);
            for (int f = nI; f < m.nF; ++f) {
                if (!m.cyclicFace(f)) continue;
                const double sf = k.umrf[f] * m.magSf[f];
                mx.iu[f - nI] = sf * (1 - m.w[f]);
                mx.diag[m.own[f]] -= -sf * m.w[f];
            }
        }
        LduBlock mt;
        mt.init(nC, nI, 9);
        auto toI = [](const vec& s, vec& t) {
            for (size_t i = 0; i < s.size(); ++i) t[9 * i] = t[9 * i + 4] = t[9 * i + 8] = s[i];
        };
        toI(mx.diag, mt.diag); toI(mx.upper, mt.upper); toI(mx.lower, mt.lower);
        if (!mx.iu.empty()) { mt.iu.assign(9 * mx.iu.size(), 0.0); toI(mx.iu, mt.iu); }
        addBlockTo(k.SS[0][0], mx, -1.0);
        addBlockTo(k.VV, mt, -1.0);
        addBlockTo(k.SS[1][1], mx, -1.0);
    }
    // Dissipation: lambdaConv = interp(c) + |interp(U).n - um|
    vec lamI(nI), lamB(nB, 0);
    for (int f = 0; f < nI; ++f) {
        const double cf = interpLinear(m, f, d.c.data(), 1);
        V3 Uf = m.w[f] * (v3(&k.U[3 * m.own[f]]) - v3(&k.U[3 * m.nei[f]])) + v3(&k.U[3 * m.nei[f]]);
        lamI[f] = cf + std::fabs(dot(Uf, nrm(f)) - (k.mrf ? k.umrf[f] : 0.0));
    }
    for (int f = nI; f < m.nF; ++f) {
        if (m.pKind[m.facePatch[f - nI]] == PK_EMPTY) continue;
        if (m.cyclicFace(f)) {
            const double cf = interpCyclicC(k, f, d.c.data());
            lamB[f - nI] = cf + std::fabs(dot(interpCyclicU(k, f), nrm(f)) - (k.mrf ? k.umrf[f] : 0.0));
        } else {
            lamB[f - nI] = d.cb[f - nI] + std::fabs(dot(v3(&k.UB[3 * (f - nI)]), nrm(f)) - (k.mrf ? k.umrf[f] : 0.0));
        }
    }
    auto dissip = [&](LduBlock& target, int nc) {
        LduBlock mx;
        mx.init(nC, nI, nc);
        for (int f = 0; f < nI; ++f) {
            const double v = 0.5 * m.magSf[f] * lamI[f];
            if (nc == 1) { mx.upper[f] = v; mx.lower[f] = v; }
            else {
                for (int q = 0; q < 9; ++q) { mx.upper[9 * f + q] = 0; mx.lower[9 * f + q] = 0; }
                mx.upper[9 * f] = mx.upper[9 * f + 4] = mx.upper[9 * f + 8] = v;
                mx.lower[9 * f] = mx.lower[9 * f + 4] = mx.lower[9 * f + 8] = v;
            }
        }
        negSumDiag(m, mx);
        if (m.anyCyclic) mx.iu.assign((size_t)nB * nc, 0.0);
        for (int f = nI; f < m.nF; ++f) {
            if (m.pKind[m.facePatch[f - nI]] == PK_EMPTY) continue;
            const double v = 0.5 * m.magSf[f] * lamB[f - nI];
            const int o = m.own[f];
            if (nc == 1) mx.diag[o] -= v;
            else { mx.diag[9 * o] -= v; mx.diag[9 * o + 4] -= v; mx.diag[9 * o + 8] -= v; }
            if (m.cyclicFace(f)) {
                const int b = f - nI;
                if (nc == 1) mx.iu[b] = v;
                else { mx.iu[9 * b] = v; mx.iu[9 * b + 4] = v; mx.iu[9 * b + 8] = v; }
            }
        }
        addBlockTo(target, mx, -1.0);
    };
    dissip(k.SS[0][0], 1);
    dissip(k.VV, 9);
    dissip(k.SS[1][1], 1);
    // addBoundaryTerms (convectiveFluxScheme.C:219-355), non-coupled patches
    for (size_t pi = 0; pi < m.pStart.size(); ++pi) {
        if (m.pKind[pi] == PK_EMPTY || m.pKind[pi] == PK_CYCLIC) continue;
        const BC& bc = k.bcs[pi];
        for (int i = 0; i < m.pSize[pi]; ++i) {
            const int f = m.pStart[pi] + i, b = f - nI, c = m.own[f];
            V3 Sf = v3(&m.Sf[3 * f]);
            V3 n = Sf / m.magSf[f];
            // valueInternalCoeffs: 0 for fixedValue and the fixedValue-derived
            // totalPressure / totalTemperature, 1 for zeroGradient; for the
            // transform conditions 1 - snGradTransformDiag: slip 1 - |n|,
            // pressureInletOutletVelocity 1 - sqrt(|diag(valueFraction)|)
            // with valueFraction = (I - n n) at inflow, 0 otherwise
            const double pVIC = (bc.pT == BC_FIXED || bc.pT == BC_TOTAL_P) ? 0.0 : 1.0;
            const double tVIC = (bc.tT == BC_FIXED || bc.tT == BC_TOTAL_T) ? 0.0 : 1.0;
            V3 uVIC;
            if (bc.uT == BC_FIXED) uVIC = V3{0, 0, 0};
            else if (bc.uT == BC_SLIP) uVIC = V3{1 - std::fabs(n.x), 1 - std::fabs(n.y), 1 - std::fabs(n.z)};
            else if (bc.uT == BC_PIOV && k.inflowB[b])
                uVIC = V3{1 - std::sqrt(std::fabs(1 - n.x * n.x)), 1 - std::sqrt(std::fabs(1 - n.y * n.y)),
                          1 - std::sqrt(std::fabs(1 - n.z * n.z))};
            else uVIC = V3{1, 1, 1};
            const double rhoB = k.rhoB[b], pB = k.pB[b], TB = k.TB[b];
            V3 UB = v3(&k.UB[3 * b]);
            double UrelSf = dot(UB, Sf);
            if (k.mrf) UrelSf -= k.umrf[f] * m.magSf[f];
            const double rhoEB = rhoB * (k.Cv * TB + 0.5 * magSqr(UB));
            const double cvB = k.Cv;
            V3 SfV{Sf.x * uVIC.x, Sf.y * uVIC.y, Sf.z * uVIC.z};
            const double dCdp = rhoB / pB * UrelSf * pVIC;
            V3 dCdU = rhoB * SfV;
            const double dCdT = -rhoB / TB * UrelSf * tVIC;
            V3 dMdp = (rhoB / pB * UB * UrelSf + Sf) * pVIC;
            T9 dMdU = outer(rhoB * UB, SfV);
            V3 dMdUDiag = rhoB * UrelSf * uVIC;
            V3 dMdT = -rhoB / TB * UB * UrelSf * tVIC;
            const double dEdp = (rhoEB / pB * UrelSf + dot(UB, Sf)) * pVIC;
            V3 UBV{UB.x * uVIC.x, UB.y * uVIC.y, UB.z * uVIC.z};
            V3 dEdU = SfV * (rhoEB + pB) + rhoB * UrelSf * UBV;
            const double dEdT = UrelSf * (rhoB * cvB - rhoEB / TB) * tVIC;
            dMdU.t[0] += dMdUDiag.x; dMdU.t[4] += dMdUDiag.y; dMdU.t[8] += dMdUDiag.z;
            const double rhoI = k.rho[c];
            V3 UI = v3(&k.U[3 * c]);
            const double rhoEI = rhoI * (k.Cv * k.T[c] + 0.5 * magSqr(UI));
            const double gammaI = g, cvI = k.Cv;
            const double dPdRho = 0.5 * (gammaI - 1) * magSqr(UI);
            V3 dUdRho = V3{-UI.x / rhoI, -UI.y / rhoI, -UI.z / rhoI};
            const double dTdRho = -1.0 / (cvI * rhoI) * (rhoEI / rhoI - magSqr(UI));
            V3 dPdRhoU = -(gammaI - 1) * UI;
            const double dUdRhoU = 1.0 / rhoI;
            V3 dTdRhoU = V3{-UI.x / (cvI * rhoI), -UI.y / (cvI * rhoI), -UI.z / (cvI * rhoI)};
            const double dPdRhoE = gammaI - 1;
            const double dTdRhoE = 1.0 / (cvI * rhoI);
            // continuity rows
            double& cr = k.SS[0][0].diag[c];
            cr += dCdp * dPdRho; cr += dot(dCdU, dUdRho); cr += dCdT * dTdRho;
            double* cu = &k.SV[0].diag[3 * c];
            V3 t1 = dCdp * dPdRhoU, t2 = dCdU * dUdRhoU, t3 = dCdT * dTdRhoU;
            cu[0] += t1.x; cu[1] += t1.y; cu[2] += t1.z;
            cu[0] += t2.x; cu[1] += t2.y; cu[2] += t2.z;
            cu[0] += t3.x; cu[1] += t3.y; cu[2] += t3.z;
            double& ce = k.SS[0][1].diag[c];
            ce += dCdp * dPdRhoE; ce += 0.0; ce += dCdT * dTdRhoE;
            // momentum rows
            double* mr = &k.VS[0].diag[3 * c];
            V3 a = dMdp * dPdRho, bb = tdotv(dMdU, dUdRho), cc = dMdT * dTdRho;
            mr[0] += a.x; mr[1] += a.y; mr[2] += a.z;
            mr[0] += bb.x; mr[1] += bb.y; mr[2] += bb.z;
            mr[0] += cc.x; mr[1] += cc.y; mr[2] += cc.z;
            double* mu_ = &k.VV.diag[9 * c];
            T9 A = outer(dMdp, dPdRhoU), B = dMdU * dUdRhoU, Cc = outer(dMdT, dTdRhoU);
            for (int q = 0; q < 9; ++q) mu_[q] += A.t[q];
            for (int q = 0; q < 9; ++q) mu_[q] += B.t[q];
            for (int q = 0; q < 9; ++q) mu_[q] += Cc.t[q];
            double* me = &k.VS[1].diag[3 * c];
            V3 e1 = dMdp * dPdRhoE, e3 = dMdT * dTdRhoE;
            me[0] += e1.x; me[1] += e1.y; me[2] += e1.z;
            me[0] += 0.0; me[1] += 0.0; me[2] += 0.0;
            me[0] += e3.x; me[1] += e3.y; me[2] += e3.z;
            // energy rows
            double& er = k.SS[1][0].diag[c];
            er += dEdp * dPdRho; er += dot(dEdU, dUdRho); er += dEdT * dTdRho;
            double* eu = &k.SV[1].diag[3 * c];
            V3 f1 = dEdp * dPdRhoU, f2 = dEdU * dUdRhoU, f3 = dEdT * dTdRhoU;
            eu[0] += f1.x; eu[1] += f1.y; eu[2] += f1.z;
            eu[0] += f2.x; eu[1] += f2.y; eu[2] += f2.z;
            eu[0] += f3.x; eu[1] += f3.y; eu[2] += f3.z;
            double& ee = k.SS[1][1].diag[c];
            ee += dEdp * dPdRhoE; ee += 0.0; ee += dEdT * dTdRhoE;
        }
    }
    // addTemporalTerms: ddtCoeff = max(diag(fvm::ddt))/V with dualTime diag = rDT*V
    for (int c = 0; c < nC; ++c) {
        const double ddtCoeff = (k.rDT[c] * m.V[c]) / m.V[c];
        const double dc = ddtCoeff * m.V[c];
        k.SS[0][0].diag[c] += dc;
        k.VV.diag[9 * c] += dc; k.VV.diag[9 * c + 4] += dc; k.VV.diag[9 * c + 8] += dc;
        k.SS[1][1].diag[c] += dc;
    }
    // addMRFSource (convectiveFluxScheme.C:123-138): the derivative of
    // -(Omega ^ rhoU) V on the rhoU diagonal block
    if (k.mrf && k.frot.on)
        for (int c = 0; c < nC; ++c) {
            const double* om = &k.omegaC[3 * c];
            double* t = &k.VV.diag[9 * c];
            t[1] -= om[2] * m.V[c];
            t[2] += om[1] * m.V[c];
            t[3] += om[2] * m.V[c];
            t[5] -= om[0] * m.V[c];
            t[6] -= om[1] * m.V[c];
            t[7] += om[0] * m.V[c];
        }
    // viscous LF Jacobian: laplacian(0.5*lambdaVisc, 1) subtracted
    if (k.s.viscous) {
        const double mu = k.s.mu, alpha = k.s.gamma * mu / k.s.Pr;
        auto stab = [&](LduBlock& target, int nc) {
            LduBlock mx;
            mx.init(nC, nI, nc);
            for (int f = 0; f < nI; ++f) {
                const double rf = interpLinear(m, f, k.rho.data(), 1);
                const double muf = m.w[f] * (mu - mu) + mu;
                const double alf = m.w[f] * (alpha - alpha) + alpha;
                const double lv = (muf + alf) / rf;
                const double s = 0.5 * lv * m.magSf[f] * m.dc[f];
                if (nc == 1) { mx.upper[f] = s; mx.lower[f] = s; }
                else {
                    mx.upper[9 * f] = mx.upper[9 * f + 4] = mx.upper[9 * f + 8] = s;
                    mx.lower[9 * f] = mx.lower[9 * f + 4] = mx.lower[9 * f + 8] = s;
                }
            }
            negSumDiag(m, mx);
            // coupled faces (blockFvOperatorsTemplates.C:501-556):
            // interfacesUpper = interfacesLower = the face's coefficient,
            // diag[own] -= interfacesLower; rho_f = w rho_P + (1 - w)
            // rho_far (lagged on a phase-lagged face)
            if (m.anyCyclic) {
                mx.iu.assign((size_t)(m.nF - nI) * nc, 0.0);
                for (int f = nI; f < m.nF; ++f) {
                    if (!m.cyclicFace(f)) continue;
                    const int b = f - nI, o = m.own[f];
                    const double rN = lagFace(k, b) ? lagScalar(k, b, PL_RHO, 0) : k.rho[m.cycCell[b]];
                    const double rf = interpCyclic(m, f, k.rho[o], rN);
                    const double muf = m.w[f] * mu + (1.0 - m.w[f]) * mu;
                    const double alf = m.w[f] * alpha + (1.0 - m.w[f]) * alpha;
                    const double lv = (muf + alf) / rf;
                    const double s = 0.5 * lv * m.magSf[f] * m.dc[f];
                    if (nc == 1) { mx.iu[b] = s; mx.diag[o] -= s; }
                    else
                        for (int d = 0; d < 3; ++d) { mx.iu[9 * b + 4 * d] = s; mx.diag[9 * o + 4 * d] -= s; }
                }
            }
            addBlockTo(target, mx, -1.0);
        };
        stab(k.SS[0][0], 1);
        stab(k.VV, 9);
        stab(k.SS[1][1], 1);
    }
}

// ---------------------------------------------------------- linear alg ---
// Vectors at this level: s[0] (rho), s[1] (rhoE): nC each; v: 3*nC interleaved.
struct CVec {
    vec s0, s1, v;
    void init(int nC) { s0.assign(nC, 0); s1.assign(nC, 0); v.assign(3 * (size_t)nC, 0); }
};
void toC(const double* x, int nC, CVec& c) {
    c.init(nC);
    for (int i = 0; i < nC; ++i) {
        c.s0[i] = x[i];
        c.s1[i] = x[nC + i];
        for (int d = 0; d < 3; ++d) c.v[3 * i + d] = x[(2 + d) * (size_t)nC + i];
    }
}
void fromC(const CVec& c, int nC, double* x) {
    for (int i = 0; i < nC; ++i) {
        x[i] = c.s0[i];
        x[nC + i] = c.s1[i];
        for (int d = 0; d < 3; ++d) x[(2 + d) * (size_t)nC + i] = c.v[3 * i + d];
    }
}

// blockFvMatrix::Amul for (sourceType, blockType) pairs.
// Cyclic faces add interfacesUpper & patchNeighbourField(x) to the owner
// row after the internal faces (blockFvMatrix.C:398-420); vector columns
// are transformed by forwardT, scalar columns are not.
void amulSS(const Mesh& m, const LduBlock& A, const vec& x, vec& y) {
    y.assign(m.nC, 0);
    for (int c = 0; c < m.nC; ++c) y[c] = A.diag[c] * x[c];
    for (int f = 0; f < m.nI; ++f) {
        y[m.nei[f]] += A.lower[f] * x[m.own[f]];
        y[m.own[f]] += A.upper[f] * x[m.nei[f]];
    }
    if (!A.iu.empty())
        for (int f = m.nI; f < m.nF; ++f)
            if (m.cyclicFace(f)) y[m.own[f]] += A.iu[f - m.nI] * x[m.cycCell[f - m.nI]];
}
void amulSV(const Mesh& m, const LduBlock& A, const vec& x, vec& y) {  // scalar row, vector col
    y.assign(m.nC, 0);
    for (int c = 0; c < m.nC; ++c) y[c] = dot(v3(&A.diag[3 * c]), v3(&x[3 * c]));
    for (int f = 0; f < m.nI; ++f) {
        y[m.nei[f]] += dot(v3(&A.lower[3 * f]), v3(&x[3 * m.own[f]]));
        y[m.own[f]] += dot(v3(&A.upper[3 * f]), v3(&x[3 * m.nei[f]]));
    }
    if (!A.iu.empty())
        for (int f = m.nI; f < m.nF; ++f) {
            if (!m.cyclicFace(f)) continue;
            const int b = f - m.nI;
            y[m.own[f]] += dot(v3(&A.iu[3 * b]), transformV(&m.cycR[9 * b], v3(&x[3 * m.cycCell[b]])));
        }
}
void amulVS(const Mesh& m, const LduBlock& A, const vec& x, vec& y) {  // vector row, scalar col
    y.assign(3 * (size_t)m.nC, 0);
    for (int c = 0; c < m.nC; ++c)
        for (int d = 0; d < 3; ++d) y[3 * c + d] = A.diag[3 * c + d] * x[c];
    for (int f = 0; f < m.nI; ++f)
        for (int d = 0; d < 3; ++d) {
            y[3 * m.nei[f] + d] += A.lower[3 * f + d] * x[m.own[f]];
            y[3 * m.own[f] + d] += A.upper[3 * f + d] * x[m.nei[f]];
        }
    if (!A.iu.empty())
        for (int f = m.nI; f < m.nF; ++f) {
            if (!m.cyclicFace(f)) continue;
            const int b = f - m.nI;
            for (int d = 0; d < 3; ++d) y[3 * m.own[f] + d] += A.iu[3 * b + d] * x[m.cycCell[b]];
        }
}
void amulVV(const Mesh& m, const LduBlock& A, const vec& x, vec& y) {
    y.assign(3 * (size_t)m.nC, 0);
    auto T = [&](const vec& a, int i) { T9 t; for (int q = 0; q < 9; ++q) t.t[q] = a[9 * i + q]; return t; };
    for (int c = 0; c < m.nC; ++c) {
        V3 r = tdotv(T(A.diag, c), v3(&x[3 * c]));
        y[3 * c] = r.x; y[3 * c + 1] = r.y; y[3 * c + 2] = r.z;
    }
    for (int f = 0; f < m.nI; ++f) {
        V3 a = tdotv(T(A.lower, f), v3(&x[3 * m.own[f]]));
        V3 b = tdotv(T(A.upper, f), v3(&x[3 * m.nei[f]]));
        for (int d = 0; d < 3; ++d) y[3 * m.nei[f] + d] += (&a.x)[d];
        for (int d = 0; d < 3; ++d) y[3 * m.own[f] + d] += (&b.x)[d];
    }
    if (!A.iu.empty())
        for (int f = m.nI; f < m.nF; ++f) {
            if (!m.cyclicFace(f)) continue;
            const int bf = f - m.nI;
            V3 r = tdotv(T(A.iu, bf), transformV(&m.cycR[9 * bf], v3(&x[3 * m.cycCell[bf]])));
            for (int d = 0; d < 3; ++d) y[3 * m.own[f] + d] += (&r.x)[d];
        }
}

// coupledMatrix::matrixMul (coupledMatrix.C:66-123)
void matrixMul(const Case& k, const CVec& x, CVec& y) {
    const Mesh& m = *k.m;
    y.init(m.nC);
    vec t;
    vec* S[2] = {&y.s0, &y.s1};
    const vec* X[2] = {&x.s0, &x.s1};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            amulSS(m, k.SS[i][j], *X[j], t);
            for (int c = 0; c < m.nC; ++c) (*S[i])[c] += t[c];
        }
    for (int i = 0; i < 2; ++i) {
        amulSV(m, k.SV[i], x.v, t);
        for (int c = 0; c < m.nC; ++c) (*S[i])[c] += t[c];
    }
    for (int j = 0; j < 2; ++j) {
        amulVS(m, k.VS[j], *X[j], t);
        for (size_t c = 0; c < t.size(); ++c) y.v[c] += t[c];
    }
    amulVV(m, k.VV, x.v, t);
    for (size_t c = 0; c < t.size(); ++c) y.v[c] += t[c];
}

// lusgs::lusgs (lusgs.C:50-125): scalar reciprocal of the max diagonal.
int lusgsSetup(Case& k) {
    const Mesh& m = *k.m;
    k.rD.assign(m.nC, GREAT);
    for (int c = 0; c < m.nC; ++c) {
        double r = k.rD[c];
        r = 1.0 / std::max(1.0 / r, std::fabs(k.SS[0][0].diag[c]));
        r = 1.0 / std::max(1.0 / r, std::fabs(k.SS[1][1].diag[c]));
        r = 1.0 / std::max(1.0 / r, std::fabs(k.VV.diag[9 * c]));
        r = 1.0 / std::max(1.0 / r, std::fabs(k.VV.diag[9 * c + 4]));
        r = 1.0 / std::max(1.0 / r, std::fabs(k.VV.diag[9 * c + 8]));
        k.rD[c] = r;
        if (r < VSMALL) return -3;
    }
    return 0;
}

// lusgs::precondition (lusgs.C:220-382)
void lusgs(const Case& k, CVec& r) {
    const Mesh& m = *k.m;
    const int nC = m.nC;
    // forward
    for (int c = 0; c < nC; ++c) {
        const double rd = k.rD[c];
        const double dS0 = rd * r.s0[c], dS1 = rd * r.s1[c];
        const V3 dV = rd * v3(&r.v[3 * c]);
        const double dS[2] = {dS0, dS1};
        vec* R[2] = {&r.s0, &r.s1};
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int f : m.cells[c])
                    if (f < m.nI && m.nei[f] > c) (*R[i])[m.nei[f]] -= k.SS[i][j].lower[f] * dS[j];
        for (int i = 0; i < 2; ++i)
            for (int f : m.cells[c])
                if (f < m.nI && m.nei[f] > c) (*R[i])[m.nei[f]] -= dot(v3(&k.SV[i].lower[3 * f]), dV);
        for (int j = 0; j < 2; ++j)
            for (int f : m.cells[c])
                if (f < m.nI && m.nei[f] > c)
                    for (int d = 0; d < 3; ++d) r.v[3 * m.nei[f] + d] -= k.VS[j].lower[3 * f + d] * dS[j];
        for (int f : m.cells[c])
            if (f < m.nI && m.nei[f] > c) {
                T9 t; for (int q = 0; q < 9; ++q) t.t[q] = k.VV.lower[9 * f + q];
                V3 a = tdotv(t, dV);
                for (int d = 0; d < 3; ++d) r.v[3 * m.nei[f] + d] -= (&a.x)[d];
            }
    }
    // backward
    for (int c = nC - 1; c >= 0; --c) {
        const double rd = k.rD[c];
        const double dS0 = rd * r.s0[c], dS1 = rd * r.s1[c];
        const V3 dV = rd * v3(&r.v[3 * c]);
        r.s0[c] = dS0; r.s1[c] = dS1;
        r.v[3 * c] = dV.x; r.v[3 * c + 1] = dV.y; r.v[3 * c + 2] = dV.z;
        const double dS[2] = {dS0, dS1};
        vec* R[2] = {&r.s0, &r.s1};
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int f : m.cells[c])
                    if (f < m.nI && m.own[f] < c) (*R[i])[m.own[f]] -= k.SS[i][j].upper[f] * dS[j];
        for (int i = 0; i < 2; ++i)
            for (int f : m.cells[c])
                if (f < m.nI && m.own[f] < c) (*R[i])[m.own[f]] -= dot(v3(&k.SV[i].upper[3 * f]), dV);
        for (int j = 0; j < 2; ++j)
            for (int f : m.cells[c])
                if (f < m.nI && m.own[f] < c)
                    for (int d = 0; d < 3; ++d) r.v[3 * m.own[f] + d] -= k.VS[j].upper[3 * f + d] * dS[j];
        for (int f : m.cells[c])
            if (f < m.nI && m.own[f] < c) {
                T9 t; for (int q = 0; q < 9; ++q) t.t[q] = k.VV.upper[9 * f + q];
                V3 a = tdotv(t, dV);
                for (int d = 0; d < 3; ++d) r.v[3 * m.own[f] + d] -= (&a.x)[d];
            }
    }
}

// Block-Jacobi (Jacobi.C:55-132, JacobiSmoother.C:42-203): z = D^-1 r with
// the 5x5 diagonal block inverted by LU with partial pivoting.
void diagBlock(const Case& k, int c, double J[5][5]) {
    J[0][0] = k.SS[0][0].diag[c]; J[0][1] = k.SS[0][1].diag[c];
    J[1][0] = k.SS[1][0].diag[c]; J[1][1] = k.SS[1][1].diag[c];
    for (int d = 0; d < 3; ++d) {
        J[0][2 + d] = k.SV[0].diag[3 * c + d];
        J[1][2 + d] = k.SV[1].diag[3 * c + d];
        J[2 + d][0] = k.VS[0].diag[3 * c + d];
        J[2 + d][1] = k.VS[1].diag[3 * c + d];
        for (int e = 0; e < 3; ++e) J[2 + d][2 + e] = k.VV.diag[9 * c + 3 * d + e];
    }
}
void invert5(double A[5][5], double inv[5][5]) {
    int piv[5];
    double a[5][5];
    std::memcpy(a, A, sizeof(a));
    for (int i = 0; i < 5; ++i) piv[i] = i;
    for (int col = 0; col < 5; ++col) {
        int p = col;
        for (int r = col + 1; r < 5; ++r)
            if (std::fabs(a[r][col]) > std::fabs(a[p][col])) p = r;
        if (p != col) {
            for (int q = 0; q < 5; ++q) std::swap(a[p][q], a[col][q]);
            std::swap(piv[p], piv[col]);
        }
        for (int r = col + 1; r < 5; ++r) {
            a[r][col] /= a[col][col];
            for (int q = col + 1; q < 5; ++q) a[r][q] -= a[r][col] * a[col][q];
        }
    }
    for (int e = 0; e < 5; ++e) {
        double y[5];
        for (int i = 0; i < 5; ++i) {
            double s = piv[i] == e ? 1.0 : 0.0;
            for (int q = 0; q < i; ++q) s -= a[i][q] * y[q];
            y[i] = s;
        }
        for (int i = 4; i >= 0; --i) {
            double s = y[i];
            for (int q = i + 1; q < 5; ++q) s -= a[i][q] * y[q];
            y[i] = s / a[i][i];
        }
        for (int i = 0; i < 5; ++i) inv[i][e] = y[i];
    }
}
void jacobi(const Case& k, CVec& r) {
    const Mesh& m = *k.m;
    for (int c = 0; c < m.nC; ++c) {
        double J[5][5], Ji[5][5];
        diagBlock(k, c, J);
        invert5(J, Ji);
        double v[5] = {r.s0[c], r.s1[c], r.v[3 * c], r.v[3 * c + 1], r.v[3 * c + 2]}, o[5] = {0, 0, 0, 0, 0};
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) o[i] += Ji[i][j] * v[j];
        r.s0[c] = o[0]; r.s1[c] = o[1];
        for (int d = 0; d < 3; ++d) r.v[3 * c + d] = o[2 + d];
    }
}

void precondition(const Case& k, CVec& r) {
    if (k.s.precond == 0) lusgs(k, r);
    else jacobi(k, r);
}

// gmres::givensRotation (gmres.C:44-69)
void givens(double h, double beta, double& c, double& s) {
    if (beta == 0) { c = 1; s = 0; }
    else if (std::fabs(beta) > std::fabs(h)) {
        double tau = -h / beta;
        s = 1.0 / std::sqrt(1.0 + sqr(tau));
        c = s * tau;
    } else {
        double tau = -beta / h;
        c = 1.0 / std::sqrt(1.0 + sqr(tau));
        s = c * tau;
    }
}

struct Res {
    double sInit[2], sFinal[2], vInit[3], vFinal[3];
    int nIter;
};

double vsumMag(const vec& a) { double s = 0; for (double x : a) s += std::fabs(x); return s; }

// gmres::solveDelta (6-arg) (gmres.C:772-1108) with coupledMatrixSolver::stop
Res gmresSolve(Case& k, const CVec& W, const CVec& b, CVec& dW) {
    const Mesh& m = *k.m;
    const int nC = m.nC, N = k.s.nDirs;
    Res res{};
    dW = W;
    double avg0 = 0, avg1 = 0; V3 avgv{0, 0, 0};
    for (int c = 0; c < nC; ++c) { avg0 += W.s0[c]; avg1 += W.s1[c]; avgv = avgv + v3(&W.v[3 * c]); }
    avg0 /= nC; avg1 /= nC; avgv = avgv / (double)nC;
    for (int c = 0; c < nC; ++c) {
        dW.s0[c] -= avg0; dW.s1[c] -= avg1;
        dW.v[3 * c] -= avgv.x; dW.v[3 * c + 1] -= avgv.y; dW.v[3 * c + 2] -= avgv.z;
    }
    CVec t;
    matrixMul(k, dW, t);
    dW.init(nC);
    double sNF[2], vNF;
    {
        double a = 0, bb = 0;
        for (int c = 0; c < nC; ++c) a += std::fabs(t.s0[c]) + std::fabs(b.s0[c]);
        for (int c = 0; c < nC; ++c) bb += std::fabs(t.s1[c]) + std::fabs(b.s1[c]);
        sNF[0] = a + VSMALL; sNF[1] = bb + VSMALL;
        res.sInit[0] = vsumMag(b.s0) / sNF[0];
        res.sInit[1] = vsumMag(b.s1) / sNF[1];
        double v = 0;
        for (int c = 0; c < nC; ++c) v += mag(v3(&t.v[3 * c])) + mag(v3(&b.v[3 * c]));
        vNF = v + VSMALL;
        double cm[3] = {0, 0, 0};
        for (int c = 0; c < nC; ++c) for (int d = 0; d < 3; ++d) cm[d] += std::fabs(b.v[3 * c + d]);
        for (int d = 0; d < 3; ++d) res.vInit[d] = cm[d] / vNF;
        res.sFinal[0] = res.sInit[0]; res.sFinal[1] = res.sInit[1];
        for (int d = 0; d < 3; ++d) res.vFinal[d] = res.vInit[d];
    }
    CVec r = b;
    std::vector<std::vector<double>> H(N, std::vector<double>(N, 0.0));
    vec yh(N, 0), bh(N + 1, 0), cs(N, 0), sn(N, 0);
    std::vector<CVec> Vk(N);
    auto norm2 = [&](const CVec& x) {
        double s = 0, a = 0;
        for (double q : x.s0) a += q * q;
        s += a; a = 0;
        for (double q : x.s1) a += q * q;
        s += a; a = 0;
        for (int c = 0; c < nC; ++c) a += magSqr(v3(&x.v[3 * c]));
        s += a;
        return std::sqrt(s);
    };
    res.nIter = 0;
    while (true) {
        precondition(k, r);
        double beta = norm2(r);
        std::fill(bh.begin(), bh.end(), 0.0);
        bh[0] = beta;
        for (int i = 0; i < N; ++i) {
            CVec& v = Vk[i];
            v.init(nC);
            for (int c = 0; c < nC; ++c) { v.s0[c] = r.s0[c] / beta; v.s1[c] = r.s1[c] / beta; }
            for (size_t q = 0; q < v.v.size(); ++q) v.v[q] = r.v[q] / beta;
            matrixMul(k, v, r);
            precondition(k, r);
            for (int j = 0; j <= i; ++j) {
                double bb = 0, a = 0;
                for (int c = 0; c < nC; ++c) a += r.s0[c] * Vk[j].s0[c];
                bb += a; a = 0;
                for (int c = 0; c < nC; ++c) a += r.s1[c] * Vk[j].s1[c];
                bb += a; a = 0;
                for (int c = 0; c < nC; ++c) a += dot(v3(&r.v[3 * c]), v3(&Vk[j].v[3 * c]));
                bb += a;
                H[j][i] = bb;
                for (int c = 0; c < nC; ++c) { r.s0[c] -= H[j][i] * Vk[j].s0[c]; r.s1[c] -= H[j][i] * Vk[j].s1[c]; }
                for (size_t q = 0; q < r.v.size(); ++q) r.v[q] -= H[j][i] * Vk[j].v[q];
            }
            beta = norm2(r);
            for (int j = 0; j < i; ++j) {
                const double Hji = H[j][i];
                H[j][i] = cs[j] * Hji - sn[j] * H[j + 1][i];
                H[j + 1][i] = sn[j] * Hji + cs[j] * H[j + 1][i];
            }
            givens(H[i][i], beta, cs[i], sn[i]);
            const double bhi = bh[i];
            bh[i] = cs[i] * bhi - sn[i] * bh[i + 1];
            bh[i + 1] = sn[i] * bhi + cs[i] * bh[i + 1];
            H[i][i] = cs[i] * H[i][i] - sn[i] * beta;
        }
        for (int i = N - 1; i >= 0; --i) {
            double s = bh[i];
            for (int j = i + 1; j < N; ++j) s -= H[i][j] * yh[j];
            double hd = H[i][i];
            // stabilise(H, VSMALL)
            if (hd >= 0) hd = hd + VSMALL; else hd = hd - VSMALL;
            yh[i] = s / hd;
        }
        for (int i = 0; i < N; ++i) {
            for (int c = 0; c < nC; ++c) { dW.s0[c] += yh[i] * Vk[i].s0[c]; dW.s1[c] += yh[i] * Vk[i].s1[c]; }
            for (size_t q = 0; q < dW.v.size(); ++q) dW.v[q] += yh[i] * Vk[i].v[q];
        }
        matrixMul(k, dW, r);
        for (int c = 0; c < nC; ++c) { r.s0[c] = b.s0[c] - r.s0[c]; r.s1[c] = b.s1[c] - r.s1[c]; }
        for (size_t q = 0; q < r.v.size(); ++q) r.v[q] = b.v[q] - r.v[q];
        res.sFinal[0] = vsumMag(r.s0) / sNF[0];
        res.sFinal[1] = vsumMag(r.s1) / sNF[1];
        double cm[3] = {0, 0, 0};
        for (int c = 0; c < nC; ++c) for (int d = 0; d < 3; ++d) cm[d] += std::fabs(r.v[3 * c + d]);
        for (int d = 0; d < 3; ++d) res.vFinal[d] = (d == m.emptyDir) ? 0.0 : cm[d] / vNF;
        res.nIter++;
        // stop
        if (res.nIter < k.s.minIter) continue;
        double mx = std::max(res.sFinal[0], res.sFinal[1]);
        for (int d = 0; d < 3; ++d) mx = std::max(mx, res.vFinal[d]);
        const double RVS = 1.0e-150;
        double rel = -1e300;
        rel = std::max(rel, res.sFinal[0] / (res.sInit[0] + RVS));
        rel = std::max(rel, res.sFinal[1] / (res.sInit[1] + RVS));
        for (int d = 0; d < 3; ++d)
            if (d != m.emptyDir) rel = std::max(rel, res.vFinal[d] / (res.vInit[d] + RVS));
        if (res.nIter >= k.s.maxIter || mx < k.s.tol || rel < k.s.relTol) break;
    }
    if (m.emptyDir >= 0)
        for (int c = 0; c < nC; ++c) dW.v[3 * c + m.emptyDir] = 0.0;
    return res;
}

// conserved variables from primitives
void makeConserved(Case& k) {
    const int nC = k.m->nC;
    k.rhoU.resize(3 * nC);
    k.rhoE.resize(nC);
    for (int c = 0; c < nC; ++c) {
        for (int d = 0; d < 3; ++d) k.rhoU[3 * c + d] = k.rho[c] * k.U[3 * c + d];
        k.rhoE[c] = k.rho[c] * (k.e[c] + 0.5 * magSqr(v3(&k.U[3 * c])));
    }
}

// updateFields.H:7-104 (rho bounding with the default rhoMin is inactive
// for positive densities; the fvc::average branch of bound() is not
// restated).
void updateFields(Case& k, const CVec& dW) {
    const int nC = k.m->nC;
    for (int c = 0; c < nC; ++c) {
        k.rho[c] += dW.s0[c];
        for (int d = 0; d < 3; ++d) k.rhoU[3 * c + d] += dW.v[3 * c + d];
        k.rhoE[c] += dW.s1[c];
    }
    for (int c = 0; c < nC; ++c) {
        k.rho[c] = std::max(k.rho[c], k.s.rhoMin);
        for (int d = 0; d < 3; ++d) k.U[3 * c + d] = k.rhoU[3 * c + d] / k.rho[c];
        k.e[c] = k.rhoE[c] / k.rho[c] - 0.5 * magSqr(v3(&k.U[3 * c]));
        const double eMin = k.Cv * k.s.TMin;
        k.e[c] = std::max(k.e[c], eMin);
        if (k.s.TMax < GREAT) k.e[c] = std::min(k.e[c], k.Cv * k.s.TMax);
        k.T[c] = k.e[c] / k.Cv;
        const double psi = 1.0 / (k.s.R * k.T[c]);
        k.p[c] = k.rho[c] / psi;
        for (int d = 0; d < 3; ++d) k.rhoU[3 * c + d] = k.rho[c] * k.U[3 * c + d];
        k.rhoE[c] = k.rho[c] * (k.e[c] + 0.5 * magSqr(v3(&k.U[3 * c])));
    }
    evalBoundary(k, false);
}

void setFields(Case& k, const double* rho, const double* U, const double* T) {
    const int nC = k.m->nC;
    k.rho.assign(rho, rho + nC);
    k.U.assign(U, U + 3 * nC);
    k.T.assign(T, T + nC);
    k.e.resize(nC);
    k.p.resize(nC);
    for (int c = 0; c < nC; ++c) {
        k.e[c] = k.Cv * k.T[c];
        const double psi = 1.0 / (k.s.R * k.T[c]);
        k.p[c] = k.rho[c] / psi;
    }
    makeConserved(k);
    evalBoundary(k, true);
}

// One dbnsFoam steady pseudo-step: outerLoop.H + updateFields.H, with the
// SER update of setCoAndDeltaT.H:3-37 and boundLocalTimeStep.H.
int pseudoStep(Case& k, Res& out) {
    Mesh& m = *k.m;
    const int nC = m.nC;
    k.corr++;
    // SER
    if (k.haveInit) {
        if (k.corr != 1 && k.havePrev) {
            const double nI = std::sqrt(sqr(k.initS[0]) + sqr(k.initS[1]) + (sqr(k.initV[0]) + sqr(k.initV[1]) + sqr(k.initV[2])));
            const double nP = std::sqrt(sqr(k.prevS[0]) + sqr(k.prevS[1]) + (sqr(k.prevV[0]) + sqr(k.prevV[1]) + sqr(k.prevV[2])));
            double ratio = nP / nI;
            ratio = std::max(std::min(ratio, k.s.coIncr), k.s.coDecr);
            for (int c = 0; c < nC; ++c) {
                k.pseudoCo[c] *= ratio;
                k.pseudoCo[c] = std::max(std::min(k.pseudoCo[c], k.s.coMax), k.s.coMin);
            }
        }
        std::copy(k.initS, k.initS + 2, k.prevS);
        std::copy(k.initV, k.initV + 3, k.prevV);
        k.havePrev = true;
    }
    vec b;
    residual(k, b);
    localDeltaT(k);
    assemble(k);
    int rc = k.s.precond == 0 ? lusgsSetup(k) : 0;
    if (rc) return rc;
    CVec W, B, dW;
    {
        vec w(5 * (size_t)nC);
        for (int c = 0; c < nC; ++c) {
            w[c] = k.rho[c];
            w[nC + c] = k.rhoE[c];
            for (int d = 0; d < 3; ++d) w[(2 + d) * (size_t)nC + c] = k.rhoU[3 * c + d];
        }
        toC(w.data(), nC, W);
        toC(b.data(), nC, B);
    }
    Res r = gmresSolve(k, W, B, dW);
    out = r;
    k.haveInit = true;
    std::copy(r.sInit, r.sInit + 2, k.initS);
    std::copy(r.vInit, r.vInit + 3, k.initV);
    // boundLocalTimeStep (evaluated before the update, as the reference does)
    if (k.s.localTS && k.s.bounding) {
        vec fac(nC, 1.0);
        auto bad = [&](int c) {
            const double rmin = k.s.lowerBound * W.s0[c];
            const double Uo[3] = {W.v[3 * c] / W.s0[c], W.v[3 * c + 1] / W.s0[c], W.v[3 * c + 2] / W.s0[c]};
            const double emin = k.s.lowerBound * (W.s1[c] / W.s0[c] - 0.5 * (Uo[0] * Uo[0] + Uo[1] * Uo[1] + Uo[2] * Uo[2]));
            const double Un[3] = {k.rhoU[3 * c] / k.rho[c], k.rhoU[3 * c + 1] / k.rho[c], k.rhoU[3 * c + 2] / k.rho[c]};
            const double et = k.rhoE[c] / k.rho[c] - 0.5 * (Un[0] * Un[0] + Un[1] * Un[1] + Un[2] * Un[2]);
            return k.rho[c] < rmin || et < emin || et < SMALL;
        };
        for (int f = 0; f < m.nI; ++f) {
            const int o = m.own[f], n = m.nei[f];
            if (bad(o)) { fac[o] = std::min(0.5, fac[o]); fac[n] = std::min(0.75, fac[n]); }
            if (bad(n)) { fac[n] = std::min(0.5, fac[n]); fac[o] = std::min(0.75, fac[o]); }
        }
        // coupled patches: the far cell through patchNeighbourField
        for (int f = m.nI; f < m.nF; ++f) {
            if (!m.cyclicFace(f)) continue;
            const int o = m.own[f], n = m.cycCell[f - m.nI];
            if (bad(o)) fac[o] = std::min(0.5, fac[o]);
            if (bad(n)) fac[o] = std::min(0.75, fac[o]);
        }
        for (int c = 0; c < nC; ++c) k.pseudoCo[c] *= fac[c];
    }
    updateFields(k, dW);
    return 0;
}

#include "icsf_tangent.h"

// ----------------------------------------------------- harmonic balance ---
// dbnsFullyImplicitHBFoam (outerLoop.H:1-230, residualsUpdate.H:72-74) with
// the spectral operator of HBZone (HBZone.C:270-518, HBZoneTemplates.C:
// 39-93).  nO instances of the single-instance case on one mesh, coupled
// by the source -V sum_K D_JK W_K and by diagonal-only blocks V D_JK I.
// Coupled-vector variable order as coupledMatrix(2 nO scalars, nO vectors):
// rho_0, rhoE_0, rho_1, rhoE_1, ..., then rhoU_0, rhoU_1, ...
struct HB {
    int nO = 0;
    std::vector<Case*> k;
    vec D;   // [nO][nO]
    vec rD;  // LU-SGS diagonal shared by all instances
    int corr = 0;
    bool haveInit = false, havePrev = false;
    vec initS, initV, prevS, prevV;  // [2 nO], [3 nO]
    PhaseLag pl;  // phase-lagged cyclic patch sides (none: pl.sides empty)
};

// HBZone::addSource (HBZoneTemplates.C:39-93): source_J = -sum_K V D_JK W_K
void hbSource(const HB& h, int J, vec& src) {
    const Mesh& m = *h.k[0]->m;
    const int nC = m.nC;
    src.assign(5 * (size_t)nC, 0.0);
    for (int c = 0; c < nC; ++c) {
        double s0 = 0, s1 = 0, v[3] = {0, 0, 0};
        for (int K = 0; K < h.nO; ++K) {
            const Case& q = *h.k[K];
            const double VD = m.V[c] * h.D[(size_t)J * h.nO + K];
            s0 -= VD * q.rho[c];
            s1 -= VD * q.rhoE[c];
            for (int d = 0; d < 3; ++d) v[d] -= VD * q.rhoU[3 * c + d];
        }
        src[c] = s0;
        src[nC + c] = s1;
        for (int d = 0; d < 3; ++d) src[(2 + d) * (size_t)nC + c] = v[d];
    }
}

// b_J = R_J V + source_J  (residualsUpdate.H:72-74)
void hbResidual(HB& h, std::vector<vec>& b) {
    b.resize(h.nO);
    std::vector<vec> src(h.nO);
    for (int J = 0; J < h.nO; ++J) hbSource(h, J, src[J]);
    for (int J = 0; J < h.nO; ++J) {
        residual(*h.k[J], b[J]);
        for (size_t i = 0; i < b[J].size(); ++i) b[J][i] = b[J][i] + src[J][i];
    }
}

// Per-instance Jacobians (after localDeltaT) with V D_JJ added to the
// diagonal (HB.addBlock, outerLoop.H:159-177), then the shared LU-SGS
// diagonal of lusgs::lusgs over all 2 nO scalar and nO vector diagonals.
int hbAssemble(HB& h) {
    const Mesh& m = *h.k[0]->m;
    for (int J = 0; J < h.nO; ++J) {
        Case& q = *h.k[J];
        localDeltaT(q);
        assemble(q);
        const double dJJ = h.D[(size_t)J * h.nO + J];
        for (int c = 0; c < m.nC; ++c) {
            const double VD = m.V[c] * dJJ;
            q.SS[0][0].diag[c] += VD;
            q.SS[1][1].diag[c] += VD;
            q.VV.diag[9 * c] += VD;
            q.VV.diag[9 * c + 4] += VD;
            q.VV.diag[9 * c + 8] += VD;
        }
    }
    h.rD.assign(m.nC, GREAT);
    for (int c = 0; c < m.nC; ++c) {
        double r = h.rD[c];
        for (int J = 0; J < h.nO; ++J) {
            r = 1.0 / std::max(1.0 / r, std::fabs(h.k[J]->SS[0][0].diag[c]));
            r = 1.0 / std::max(1.0 / r, std::fabs(h.k[J]->SS[1][1].diag[c]));
        }
        for (int J = 0; J < h.nO; ++J) {
            r = 1.0 / std::max(1.0 / r, std::fabs(h.k[J]->VV.diag[9 * c]));
            r = 1.0 / std::max(1.0 / r, std::fabs(h.k[J]->VV.diag[9 * c + 4]));
            r = 1.0 / std::max(1.0 / r, std::fabs(h.k[J]->VV.diag[9 * c + 8]));
        }
        h.rD[c] = r;
        if (r < VSMALL) return -3;
    }
    for (int J = 0; J < h.nO; ++J) h.k[J]->rD = h.rD;
    return 0;
}

// coupledMatrix::matrixMul (coupledMatrix.C:66-123) over the HB blocks, in
// its order: scalar rows sum the scalar columns (rho_K, rhoE_K for K =
// 0..nO-1) then the vector columns; vector rows the scalar columns then the
// vector columns.  Cross-instance blocks are diagonal V D_JK (I); the
// instance-diagonal blocks are the instance's Jacobian; all-zero
// cross-instance blocks add nothing.
void hbAmul(const HB& h, const std::vector<CVec>& x, std::vector<CVec>& y) {
    const Mesh& m = *h.k[0]->m;
    const int nC = m.nC;
    y.resize(h.nO);
    vec t;
    for (int J = 0; J < h.nO; ++J) {
        const Case& q = *h.k[J];
        CVec& yJ = y[J];
        yJ.init(nC);
        vec* S[2] = {&yJ.s0, &yJ.s1};
        const vec* XJ[2] = {&x[J].s0, &x[J].s1};
        for (int i = 0; i < 2; ++i) {
            for (int K = 0; K < h.nO; ++K) {
                if (K == J) {
                    for (int j = 0; j < 2; ++j) {
                        amulSS(m, q.SS[i][j], *XJ[j], t);
                        for (int c = 0; c < nC; ++c) (*S[i])[c] += t[c];
                    }
                } else {
                    const vec& xk = i == 0 ? x[K].s0 : x[K].s1;
                    const double dJK = h.D[(size_t)J * h.nO + K];
                    for (int c = 0; c < nC; ++c) (*S[i])[c] += (m.V[c] * dJK) * xk[c];
                }
            }
            amulSV(m, q.SV[i], x[J].v, t);
            for (int c = 0; c < nC; ++c) (*S[i])[c] += t[c];
        }
        for (int j = 0; j < 2; ++j) {
            amulVS(m, q.VS[j], *XJ[j], t);
            for (size_t c = 0; c < t.size(); ++c) yJ.v[c] += t[c];
        }
        for (int K = 0; K < h.nO; ++K) {
            if (K == J) {
                amulVV(m, q.VV, x[J].v, t);
                for (size_t c = 0; c < t.size(); ++c) yJ.v[c] += t[c];
            } else {
                const double dJK = h.D[(size_t)J * h.nO + K];
                for (int c = 0; c < nC; ++c)
                    for (int d = 0; d < 3; ++d) yJ.v[3 * c + d] += (m.V[c] * dJK) * x[K].v[3 * c + d];
            }
        }
    }
}

// lusgs::precondition: the cross-instance blocks have no faces, so the
// sweeps act per instance, with the shared diagonal.
void hbPrecondition(const HB& h, std::vector<CVec>& r) {
    for (int J = 0; J < h.nO; ++J) precondition(*h.k[J], r[J]);
}

struct HBRes {
    vec sInit, sFinal, vInit, vFinal;  // [2 nO], [3 nO]
    int nIter = 0;
};

// gmres::solveDelta over the HB coupled vectors (same algorithm as
// gmresSolve(): left preconditioning, MGS, per-variable norm factors).
HBRes hbGmres(HB& h, const std::vector<CVec>& W, const std::vector<CVec>& b, std::vector<CVec>& dW) {
    const Mesh& m = *h.k[0]->m;
    const int nC = m.nC, nO = h.nO, N = h.k[0]->s.nDirs;
    const Settings& st = h.k[0]->s;
    HBRes res;
    res.sInit.assign(2 * nO, 0); res.sFinal.assign(2 * nO, 0); res.vInit.assign(3 * nO, 0); res.vFinal.assign(3 * nO, 0);
    dW = W;
    for (int J = 0; J < nO; ++J) {
        double a0 = 0, a1 = 0; V3 av{0, 0, 0};
        for (int c = 0; c < nC; ++c) { a0 += W[J].s0[c]; a1 += W[J].s1[c]; av = av + v3(&W[J].v[3 * c]); }
        a0 /= nC; a1 /= nC; av = av / (double)nC;
        for (int c = 0; c < nC; ++c) {
            dW[J].s0[c] -= a0; dW[J].s1[c] -= a1;
            dW[J].v[3 * c] -= av.x; dW[J].v[3 * c + 1] -= av.y; dW[J].v[3 * c + 2] -= av.z;
        }
    }
    std::vector<CVec> t;
    hbAmul(h, dW, t);
    vec sNF(2 * nO), vNF(nO);
    for (int J = 0; J < nO; ++J) {
        double a = 0, bb = 0, v = 0;
        for (int c = 0; c < nC; ++c) a += std::fabs(t[J].s0[c]) + std::fabs(b[J].s0[c]);
        for (int c = 0; c < nC; ++c) bb += std::fabs(t[J].s1[c]) + std::fabs(b[J].s1[c]);
        for (int c = 0; c < nC; ++c) v += mag(v3(&t[J].v[3 * c])) + mag(v3(&b[J].v[3 * c]));
        sNF[2 * J] = a + VSMALL; sNF[2 * J + 1] = bb + VSMALL; vNF[J] = v + VSMALL;
        res.sInit[2 * J] = vsumMag(b[J].s0) / sNF[2 * J];
        res.sInit[2 * J + 1] = vsumMag(b[J].s1) / sNF[2 * J + 1];
        double cm[3] = {0, 0, 0};
        for (int c = 0; c < nC; ++c) for (int d = 0; d < 3; ++d) cm[d] += std::fabs(b[J].v[3 * c + d]);
        for (int d = 0; d < 3; ++d) res.vInit[3 * J + d] = cm[d] / vNF[J];
    }
    res.sFinal = res.sInit; res.vFinal = res.vInit;
    for (int J = 0; J < nO; ++J) dW[J].init(nC);
    std::vector<CVec> r = b;
    auto dotv = [&](const std::vector<CVec>& x, const std::vector<CVec>& y) {
        double s = 0, a;
        for (int J = 0; J < nO; ++J) {
            a = 0; for (int c = 0; c < nC; ++c) a += x[J].s0[c] * y[J].s0[c];
            s += a;
            a = 0; for (int c = 0; c < nC; ++c) a += x[J].s1[c] * y[J].s1[c];
            s += a;
        }
        for (int J = 0; J < nO; ++J) {
            a = 0; for (int c = 0; c < nC; ++c) a += dot(v3(&x[J].v[3 * c]), v3(&y[J].v[3 * c]));
            s += a;
        }
        return s;
    };
    std::vector<std::vector<double>> H(N, std::vector<double>(N, 0.0));
    vec yh(N, 0), bh(N + 1, 0), cs(N, 0), sn(N, 0);
    std::vector<std::vector<CVec>> Vk(N);
    while (true) {
        hbPrecondition(h, r);
        double beta = std::sqrt(dotv(r, r));
        std::fill(bh.begin(), bh.end(), 0.0);
        bh[0] = beta;
        for (int i = 0; i < N; ++i) {
            Vk[i].resize(nO);
            for (int J = 0; J < nO; ++J) {
                CVec& v = Vk[i][J];
                v.init(nC);
                for (int c = 0; c < nC; ++c) { v.s0[c] = r[J].s0[c] / beta; v.s1[c] = r[J].s1[c] / beta; }
                for (size_t q = 0; q < v.v.size(); ++q) v.v[q] = r[J].v[q] / beta;
            }
            hbAmul(h, Vk[i], r);
            hbPrecondition(h, r);
            for (int j = 0; j <= i; ++j) {
                H[j][i] = dotv(r, Vk[j]);
                for (int J = 0; J < nO; ++J) {
                    for (int c = 0; c < nC; ++c) { r[J].s0[c] -= H[j][i] * Vk[j][J].s0[c]; r[J].s1[c] -= H[j][i] * Vk[j][J].s1[c]; }
                    for (size_t q = 0; q < r[J].v.size(); ++q) r[J].v[q] -= H[j][i] * Vk[j][J].v[q];
                }
            }
            beta = std::sqrt(dotv(r, r));
            for (int j = 0; j < i; ++j) {
                const double Hji = H[j][i];
                H[j][i] = cs[j] * Hji - sn[j] * H[j + 1][i];
                H[j + 1][i] = sn[j] * Hji + cs[j] * H[j + 1][i];
            }
            givens(H[i][i], beta, cs[i], sn[i]);
            const double bhi = bh[i];
            bh[i] = cs[i] * bhi - sn[i] * bh[i + 1];
            bh[i + 1] = sn[i] * bhi + cs[i] * bh[i + 1];
            H[i][i] = cs[i] * H[i][i] - sn[i] * beta;
        }
        for (int i = N - 1; i >= 0; --i) {
            double s = bh[i];
            for (int j = i + 1; j < N; ++j) s -= H[i][j] * yh[j];
            double hd = H[i][i];
            if (hd >= 0) hd = hd + VSMALL; else hd = hd - VSMALL;
            yh[i] = s / hd;
        }
        for (int J = 0; J < nO; ++J)
            for (int i = 0; i < N; ++i) {
                for (int c = 0; c < nC; ++c) { dW[J].s0[c] += yh[i] * Vk[i][J].s0[c]; dW[J].s1[c] += yh[i] * Vk[i][J].s1[c]; }
                for (size_t q = 0; q < dW[J].v.size(); ++q) dW[J].v[q] += yh[i] * Vk[i][J].v[q];
            }
        hbAmul(h, dW, r);
        for (int J = 0; J < nO; ++J) {
            for (int c = 0; c < nC; ++c) { r[J].s0[c] = b[J].s0[c] - r[J].s0[c]; r[J].s1[c] = b[J].s1[c] - r[J].s1[c]; }
            for (size_t q = 0; q < r[J].v.size(); ++q) r[J].v[q] = b[J].v[q] - r[J].v[q];
            res.sFinal[2 * J] = vsumMag(r[J].s0) / sNF[2 * J];
            res.sFinal[2 * J + 1] = vsumMag(r[J].s1) / sNF[2 * J + 1];
            double cm[3] = {0, 0, 0};
            for (int c = 0; c < nC; ++c) for (int d = 0; d < 3; ++d) cm[d] += std::fabs(r[J].v[3 * c + d]);
            for (int d = 0; d < 3; ++d) res.vFinal[3 * J + d] = (d == m.emptyDir) ? 0.0 : cm[d] / vNF[J];
        }
        res.nIter++;
        if (res.nIter < st.minIter) continue;
        double mx = -1e300, rel = -1e300;
        const double RVS = 1.0e-150;
        for (int i = 0; i < 2 * nO; ++i) { mx = std::max(mx, res.sFinal[i]); rel = std::max(rel, res.sFinal[i] / (res.sInit[i] + RVS)); }
        for (int i = 0; i < 3 * nO; ++i) {
            mx = std::max(mx, res.vFinal[i]);
            if (i % 3 != m.emptyDir) rel = std::max(rel, res.vFinal[i] / (res.vInit[i] + RVS));
        }
        if (res.nIter >= st.maxIter || mx < st.tol || rel < st.relTol) break;
    }
    if (m.emptyDir >= 0)
        for (int J = 0; J < nO; ++J)
            for (int c = 0; c < nC; ++c) dW[J].v[3 * c + m.emptyDir] = 0.0;
    return res;
}

void hbConserved(const Case& q, CVec& W) {
    const int nC = q.m->nC;
    W.init(nC);
    W.s0 = q.rho;
    W.s1 = q.rhoE;
    W.v = q.rhoU;
}

// One HB pseudo-step (outerLoop.H:1-230 + updateFields per instance).
int hbPseudoStep(HB& h, HBRes& out) {
    const Mesh& m = *h.k[0]->m;
    const int nC = m.nC, nO = h.nO;
    const Settings& st = h.k[0]->s;
    h.corr++;
    if (h.haveInit) {
        if (h.corr != 1 && h.havePrev) {
            double nI = 0, nP = 0;
            for (int J = 0; J < nO; ++J) {
                nI += sqr(h.initS[2 * J]) + sqr(h.initS[2 * J + 1]) +
                      (sqr(h.initV[3 * J]) + sqr(h.initV[3 * J + 1]) + sqr(h.initV[3 * J + 2]));
                nP += sqr(h.prevS[2 * J]) + sqr(h.prevS[2 * J + 1]) +
                      (sqr(h.prevV[3 * J]) + sqr(h.prevV[3 * J + 1]) + sqr(h.prevV[3 * J + 2]));
            }
            double ratio = std::sqrt(nP) / std::sqrt(nI);
            ratio = std::max(std::min(ratio, st.coIncr), st.coDecr);
            for (int J = 0; J < nO; ++J)
                for (int c = 0; c < nC; ++c) {
                    h.k[J]->pseudoCo[c] *= ratio;
                    h.k[J]->pseudoCo[c] = std::max(std::min(h.k[J]->pseudoCo[c], st.coMax), st.coMin);
                }
        }
        h.prevS = h.initS;
        h.prevV = h.initV;
        h.havePrev = true;
    }
    std::vector<vec> b;
    hbResidual(h, b);
    int rc = hbAssemble(h);
    if (rc) return rc;
    std::vector<CVec> W(nO), B(nO), dW;
    for (int J = 0; J < nO; ++J) {
        hbConserved(*h.k[J], W[J]);
        toC(b[J].data(), nC, B[J]);
    }
    out = hbGmres(h, W, B, dW);
    h.haveInit = true;
    h.initS = out.sInit;
    h.initV = out.vInit;
    for (int J = 0; J < nO; ++J) {
        Case& k = *h.k[J];
        if (k.s.localTS && k.s.bounding) {
            vec fac(nC, 1.0);
            auto bad = [&](int c) {
                const double rmin = k.s.lowerBound * W[J].s0[c];
                const double Uo[3] = {W[J].v[3 * c] / W[J].s0[c], W[J].v[3 * c + 1] / W[J].s0[c], W[J].v[3 * c + 2] / W[J].s0[c]};
                const double emin = k.s.lowerBound * (W[J].s1[c] / W[J].s0[c] - 0.5 * (Uo[0] * Uo[0] + Uo[1] * Uo[1] + Uo[2] * Uo[2]));
                const double Un[3] = {k.rhoU[3 * c] / k.rho[c], k.rhoU[3 * c + 1] / k.rho[c], k.rhoU[3 * c + 2] / k.rho[c]};
                const double et = k.rhoE[c] / k.rho[c] - 0.5 * (Un[0] * Un[0] + Un[1] * Un[1] + Un[2] * Un[2]);
                return k.rho[c] < rmin || et < emin || et < SMALL;
            };
            for (int f = 0; f < m.nI; ++f) {
                const int o = m.own[f], n = m.nei[f];
                if (bad(o)) { fac[o] = std::min(0.5, fac[o]); fac[n] = std::min(0.75, fac[n]); }
                if (bad(n)) { fac[n] = std::min(0.5, fac[n]); fac[o] = std::min(0.75, fac[o]); }
            }
            // phase-lagged faces: rho_.patchNeighbourField() is lagged over
            // the instances as they stand (those before J already updated,
            // outerLoop.H:221-229); rhoMin, eTemp and eMin are not
            auto badFar = [&](int f) {
                const int b = f - m.nI, n = m.cycCell[b];
                if (!lagFace(k, b)) return bad(n);
                const double rmin = k.s.lowerBound * W[J].s0[n];
                const double Uo[3] = {W[J].v[3 * n] / W[J].s0[n], W[J].v[3 * n + 1] / W[J].s0[n], W[J].v[3 * n + 2] / W[J].s0[n]};
                const double emin = k.s.lowerBound * (W[J].s1[n] / W[J].s0[n] - 0.5 * (Uo[0] * Uo[0] + Uo[1] * Uo[1] + Uo[2] * Uo[2]));
                const double Un[3] = {k.rhoU[3 * n] / k.rho[n], k.rhoU[3 * n + 1] / k.rho[n], k.rhoU[3 * n + 2] / k.rho[n]};
                const double et = k.rhoE[n] / k.rho[n] - 0.5 * (Un[0] * Un[0] + Un[1] * Un[1] + Un[2] * Un[2]);
                return lagScalar(k, b, PL_RHO, 0) < rmin || et < emin || et < SMALL;
            };
            for (int f = m.nI; f < m.nF; ++f) {
                if (!m.cyclicFace(f)) continue;
                const int o = m.own[f];
                if (bad(o)) fac[o] = std::min(0.5, fac[o]);
                if (badFar(f)) fac[o] = std::min(0.75, fac[o]);
            }
            for (int c = 0; c < nC; ++c) k.pseudoCo[c] *= fac[c];
        }
        updateFields(k, dW[J]);
    }
    return 0;
}

// HBZone::setOmegaList / updateHBOperators restated: E^-1 = [exp(i w_k t_n)],
// E = its inverse (square case), D = -(Im(E^-1) A Re(E) + Re(E^-1) A Im(E)).
typedef std::complex<double> cx;
// E^-1 = [exp(i w_k t_n)] and E = its inverse (Gauss-Jordan, partial
// pivoting); false if singular.
bool hbFourier(int nT, const double* omega, const double* t, std::vector<cx>& Einv, std::vector<cx>& I) {
    std::vector<cx> A((size_t)nT * nT);
    I.assign((size_t)nT * nT, cx(0, 0));
    for (int n = 0; n < nT; ++n)
        for (int k = 0; k < nT; ++k) A[(size_t)n * nT + k] = cx(std::cos(omega[k] * t[n]), std::sin(omega[k] * t[n]));
    Einv = A;
    for (int i = 0; i < nT; ++i) I[(size_t)i * nT + i] = 1.0;
    for (int col = 0; col < nT; ++col) {
        int piv = col;
        for (int r = col + 1; r < nT; ++r)
            if (std::abs(A[(size_t)r * nT + col]) > std::abs(A[(size_t)piv * nT + col])) piv = r;
        if (std::abs(A[(size_t)piv * nT + col]) < 1e-13) return false;
        for (int q = 0; q < nT; ++q) { std::swap(A[(size_t)piv * nT + q], A[(size_t)col * nT + q]); std::swap(I[(size_t)piv * nT + q], I[(size_t)col * nT + q]); }
        const cx d = A[(size_t)col * nT + col];
        for (int q = 0; q < nT; ++q) { A[(size_t)col * nT + q] /= d; I[(size_t)col * nT + q] /= d; }
        for (int r = 0; r < nT; ++r) {
            if (r == col) continue;
            const cx f = A[(size_t)r * nT + col];
            for (int q = 0; q < nT; ++q) { A[(size_t)r * nT + q] -= f * A[(size_t)col * nT + q]; I[(size_t)r * nT + q] -= f * I[(size_t)col * nT + q]; }
        }
    }
    return true;
}

// The time shift of a phase-lagged patch side (phaseLagCyclicFvPatchField.C:
// 289-326): Dl = Re(E^-1 (M E)), M = diag(1, e^{i n IBPA} (n = 1..nH),
// e^{-i n IBPA} at nT - n), the products summed over every n as written.
int hbLagOperator(int nT, const double* omega, const double* t, double ibpa, double* Dl) {
    std::vector<cx> Einv, E;
    if (!hbFourier(nT, omega, t, Einv, E)) return -2;
    const int nH = (nT - 1) / 2;
    std::vector<cx> M((size_t)nT * nT, cx(0, 0)), T0((size_t)nT * nT, cx(0, 0));
    M[0] = cx(1, 0);
    for (int n = 1; n <= nH; ++n) {
        const cx e(std::cos(n * ibpa), std::sin(n * ibpa));
        M[(size_t)n * nT + n] = e;
        M[(size_t)(nT - n) * nT + (nT - n)] = std::conj(e);
    }
    auto mul = [](cx a, cx b) { return cx(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real()); };
    for (int l = 0; l < nT; ++l)
        for (int m2 = 0; m2 < nT; ++m2)
            for (int n = 0; n < nT; ++n) T0[(size_t)l * nT + m2] += mul(M[(size_t)l * nT + n], E[(size_t)n * nT + m2]);
    for (int l = 0; l < nT; ++l)
        for (int m2 = 0; m2 < nT; ++m2) {
            cx d(0, 0);
            for (int n = 0; n < nT; ++n) d += mul(Einv[(size_t)l * nT + n], T0[(size_t)n * nT + m2]);
            Dl[(size_t)l * nT + m2] = d.real();
        }
    return 0;
}

int hbOperator(int nT, const double* omega, int nO, const double* t, double* D) {
    if (nO != nT) return -1;
    std::vector<cx> A((size_t)nT * nT), I((size_t)nT * nT, cx(0, 0));
    for (int n = 0; n < nO; ++n)
        for (int k = 0; k < nT; ++k) A[(size_t)n * nT + k] = cx(std::cos(omega[k] * t[n]), std::sin(omega[k] * t[n]));
    const std::vector<cx> Einv = A;
    for (int i = 0; i < nT; ++i) I[(size_t)i * nT + i] = 1.0;
    for (int col = 0; col < nT; ++col) {
        int piv = col;
        for (int r = col + 1; r < nT; ++r)
            if (std::abs(A[(size_t)r * nT + col]) > std::abs(A[(size_t)piv * nT + col])) piv = r;
        if (std::abs(A[(size_t)piv * nT + col]) < 1e-13) return -2;
        for (int q = 0; q < nT; ++q) { std::swap(A[(size_t)piv * nT + q], A[(size_t)col * nT + q]); std::swap(I[(size_t)piv * nT + q], I[(size_t)col * nT + q]); }
        const cx d = A[(size_t)col * nT + col];
        for (int q = 0; q < nT; ++q) { A[(size_t)col * nT + q] /= d; I[(size_t)col * nT + q] /= d; }
        for (int r = 0; r < nT; ++r) {
            if (r == col) continue;
            const cx f = A[(size_t)r * nT + col];
            for (int q = 0; q < nT; ++q) { A[(size_t)r * nT + q] -= f * A[(size_t)col * nT + q]; I[(size_t)r * nT + q] -= f * I[(size_t)col * nT + q]; }
        }
    }
    for (int l = 0; l < nO; ++l)
        for (int m2 = 0; m2 < nO; ++m2) {
            double d0 = 0, d1 = 0;
            for (int n = 0; n < nT; ++n) {
                d0 += (Einv[(size_t)l * nT + n].imag() * omega[n]) * I[(size_t)n * nO + m2].real();
                d1 += (Einv[(size_t)l * nT + n].real() * omega[n]) * I[(size_t)n * nO + m2].imag();
            }
            D[(size_t)l * nO + m2] = -(d0 + d1);
        }
    return 0;
}

}  // namespace

// ======================================================== C interface ====
extern "C" {

void* orc_mesh_new(int nPoints, const double* pts, int nFaces, const int* fOff, const int* fV, int nInternal,
                   const int* own, const int* nei, int nPatches, const int* pStart, const int* pSize,
                   const int* pKind, const int* pNbr, const double* pT) {
    Mesh* m = new Mesh;
    m->nP = nPoints; m->nF = nFaces; m->nI = nInternal;
    m->pts.assign(pts, pts + 3 * (size_t)nPoints);
    m->fOff.assign(fOff, fOff + nFaces + 1);
    m->fV.assign(fV, fV + fOff[nFaces]);
    m->own.assign(own, own + nFaces);
    m->nei.assign(nei, nei + nInternal);
    m->pStart.assign(pStart, pStart + nPatches);
    m->pSize.assign(pSize, pSize + nPatches);
    m->pKind.assign(pKind, pKind + nPatches);
    m->pNbr.assign(nPatches, -1);
    m->pT.assign(9 * (size_t)nPatches, 0.0);
    for (int p = 0; p < nPatches; ++p)
        if (pKind[p] == PK_CYCLIC) {
            m->pNbr[p] = pNbr[p];
            for (int e = 0; e < 9; ++e) m->pT[9 * (size_t)p + e] = pT[9 * (size_t)p + e];
        }
    int nc = 0;
    for (int f = 0; f < nFaces; ++f) nc = std::max(nc, own[f] + 1);
    for (int f = 0; f < nInternal; ++f) nc = std::max(nc, nei[f] + 1);
    m->nC = nc;
    makeGeometry(*m);
    return m;
}
void orc_mesh_free(void* h) { delete (Mesh*)h; }
int orc_mesh_ncells(void* h) { return ((Mesh*)h)->nC; }

void orc_mesh_geometry(void* h, double* C, double* V, double* Sf, double* Cf, double* w, double* dc, double* nodc,
                       double* corr) {
    Mesh& m = *(Mesh*)h;
    std::copy(m.C.begin(), m.C.end(), C);
    std::copy(m.V.begin(), m.V.end(), V);
    std::copy(m.Sf.begin(), m.Sf.end(), Sf);
    std::copy(m.Cf.begin(), m.Cf.end(), Cf);
    std::copy(m.w.begin(), m.w.end(), w);
    std::copy(m.dc.begin(), m.dc.end(), dc);
    std::copy(m.nodc.begin(), m.nodc.end(), nodc);
    std::copy(m.corr.begin(), m.corr.end(), corr);
}

// meshTools::bandCompression (Cuthill-McKee from the lowest unvisited cell,
// unvisited neighbours appended by increasing connectivity, stable); the
// neighbour lists follow primitiveMesh::calcCellCells (internal-face order),
// then (decompositionMethod::calcCellCells, used by renumberMesh) the
// cells across coupled faces in boundary-face order.
static std::vector<int> cmOrder(int nC, int nI, const int* own, const int* nei, int nCp = 0,
                                const int* cOwn = nullptr, const int* cNbr = nullptr) {
    std::vector<std::vector<int>> cc(nC);
    for (int f = 0; f < nI; ++f) { cc[own[f]].push_back(nei[f]); cc[nei[f]].push_back(own[f]); }
    for (int b = 0; b < nCp; ++b) cc[cOwn[b]].push_back(cNbr[b]);
    std::vector<char> vis(nC, 0);
    std::vector<int> order;
    std::deque<int> q;
    for (int s = 0; s < nC; ++s) {
        if (vis[s]) continue;
        q.push_back(s);
        while (!q.empty()) {
            int c = q.front();
            q.pop_front();
            if (vis[c]) continue;
            vis[c] = 1;
            order.push_back(c);
            std::vector<int> cand;
            for (int x : cc[c]) if (!vis[x]) cand.push_back(x);
            std::vector<int> idx(cand.size());
            for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cc[cand[a]].size() < cc[cand[b]].size(); });
            for (int i : idx) q.push_back(cand[i]);
        }
    }
    return order;
}

void orc_rcm(int nC, int nI, const int* own, const int* nei, int* newToOld) {
    std::vector<int> order = cmOrder(nC, nI, own, nei);
    for (int i = 0; i < nC; ++i) newToOld[i] = order[nC - 1 - i];
}
void orc_rcm_coupled(int nC, int nI, const int* own, const int* nei, int nCp, const int* cOwn, const int* cNbr,
                     int* newToOld) {
    std::vector<int> order = cmOrder(nC, nI, own, nei, nCp, cOwn, cNbr);
    for (int i = 0; i < nC; ++i) newToOld[i] = order[nC - 1 - i];
}
// Cyclic-face geometry: w, dc, nodc per face, cell across each boundary
// face (-1 if not cyclic).
void orc_mesh_cyclic(void* h, double* w, double* dc, double* nodc, int* cycCell) {
    Mesh& m = *(Mesh*)h;
    std::copy(m.w.begin(), m.w.end(), w);
    std::copy(m.dc.begin(), m.dc.end(), dc);
    std::copy(m.nodc.begin(), m.nodc.end(), nodc);
    std::copy(m.cycCell.begin(), m.cycCell.end(), cycCell);
}

// Greedy colouring, cells visited in Cuthill-McKee order.
int orc_colour(int nC, int nI, const int* own, const int* nei, int* colour) {
    std::vector<std::vector<int>> cc(nC);
    for (int f = 0; f < nI; ++f) { cc[own[f]].push_back(nei[f]); cc[nei[f]].push_back(own[f]); }
    std::vector<int> order = cmOrder(nC, nI, own, nei);
    for (int c = 0; c < nC; ++c) colour[c] = -1;
    int nCol = 0;
    for (int c : order) {
        std::vector<char> used(nCol + 1, 0);
        for (int x : cc[c]) if (colour[x] >= 0) used[colour[x]] = 1;
        int k = 0;
        while (used[k]) ++k;
        colour[c] = k;
        nCol = std::max(nCol, k + 1);
    }
    return nCol;
}

// Cell renumbering with upper-triangular face re-sorting (renumberMesh).
void orc_renumber(int nC, int nF, int nI, const int* fOff, const int* fV, const int* own, const int* nei,
                  const int* newToOld, int* fVout, int* ownOut, int* neiOut, int* faceMap) {
    std::vector<int> o2n(nC);
    for (int i = 0; i < nC; ++i) o2n[newToOld[i]] = i;
    struct F { int a, b, f; bool flip; };
    std::vector<F> fs(nI);
    for (int f = 0; f < nI; ++f) {
        int a = o2n[own[f]], b = o2n[nei[f]];
        bool fl = a > b;
        if (fl) std::swap(a, b);
        fs[f] = F{a, b, f, fl};
    }
    std::stable_sort(fs.begin(), fs.end(), [](const F& x, const F& y) { return x.a < y.a || (x.a == y.a && x.b < y.b); });
    size_t pos = 0;
    for (int k = 0; k < nI; ++k) {
        const F& q = fs[k];
        ownOut[k] = q.a; neiOut[k] = q.b;
        faceMap[k] = q.flip ? -(q.f + 1) : q.f;
        int s = fOff[q.f], e = fOff[q.f + 1];
        if (!q.flip) for (int i = s; i < e; ++i) fVout[pos++] = fV[i];
        else { fVout[pos++] = fV[s]; for (int i = e - 1; i > s; --i) fVout[pos++] = fV[i]; }
    }
    for (int f = nI; f < nF; ++f) {
        ownOut[f] = o2n[own[f]];
        faceMap[f] = f;
        for (int i = fOff[f]; i < fOff[f + 1]; ++i) fVout[pos++] = fV[i];
    }
}

// iset: flux, lowMach, recon, limiter, viscous, precond, nDirs, maxIter, minIter, localTS, bounding
// rset: R, gamma, mu, Pr, entropyFix, tol, relTol, co, coMin, coMax, coIncr, coDecr, lowerBound, rhoMin, TMin, TMax
void* orc_case_new(void* mesh, const int* is, const double* rs) {
    Case* k = new Case;
    k->m = (Mesh*)mesh;
    Settings& s = k->s;
    s.flux = is[0]; s.lowMach = is[1]; s.recon = is[2]; s.limiter = is[3]; s.viscous = is[4];
    s.precond = is[5]; s.nDirs = is[6]; s.maxIter = is[7]; s.minIter = is[8]; s.localTS = is[9]; s.bounding = is[10];
    s.R = rs[0]; s.gamma = rs[1]; s.mu = rs[2]; s.Pr = rs[3]; s.entropyFix = rs[4]; s.tol = rs[5]; s.relTol = rs[6];
    s.co = rs[7]; s.coMin = rs[8]; s.coMax = rs[9]; s.coIncr = rs[10]; s.coDecr = rs[11]; s.lowerBound = rs[12];
    s.rhoMin = rs[13]; s.TMin = rs[14]; s.TMax = rs[15];
    k->Cv = s.R / (s.gamma - 1.0);
    k->bcs.assign(k->m->pStart.size(), BC());
    k->pseudoCo.assign(k->m->nC, s.co);
    return k;
}
void orc_case_free(void* h) { delete (Case*)h; }

// Frame zones: rotating zones (MRFCoupledZone.C:42-189 setMRFFaces,
// :240-296 constructor, :549-577 faceU, :580-602 correctBoundaryVelocity)
// and one translating zone (MRFTranslatingZone.C, the same face sets, faceU =
// the frame velocity).  Face types: 1 internal faces with a zone cell on
// either side and the zone cells' faces on other non-empty, non-coupled
// patches; 2 the zone cells' faces on coupled or non-rotating (non-moving)
// patches; 0 the rest.  faceU on types 1 and 2; the moving patches (type 1
// boundary faces) take U = faceU, the translating zone's after the
// rotating one's (outerLoop.H:6-7).  The face velocity the schemes see is
// MRF.faceU() & n + MRFTrans.faceU() & n (outerLoop.H:18-21); Omega (the
// Coriolis terms) comes from the rotating zones only.
// returns 1 if a moving patch's U condition is not fixedValue
int framePart(const Case& k, const int* zone, int nZone, const int* nonMov, int nNonMov,
              const std::function<V3(int)>& faceU, FramePart& out) {
    const Mesh& m = *k.m;
    const int nB = m.nF - m.nI;
    std::vector<char> zc(m.nC, 0), excl(m.pStart.size(), 0);
    for (int i = 0; i < nZone; ++i) zc[zone[i]] = 1;
    for (int i = 0; i < nNonMov; ++i) excl[nonMov[i]] = 1;
    std::vector<int> type(m.nF, 0);
    for (int f = 0; f < m.nI; ++f)
        if (zc[m.own[f]] || zc[m.nei[f]]) type[f] = 1;
    for (size_t pi = 0; pi < m.pStart.size(); ++pi)
        for (int f = m.pStart[pi]; f < m.pStart[pi] + m.pSize[pi]; ++f) {
            if (!zc[m.own[f]]) continue;
            if (m.pKind[pi] == PK_CYCLIC || excl[pi]) type[f] = 2;
            else if (m.pKind[pi] != PK_EMPTY) type[f] = 1;
        }
    for (size_t pi = 0; pi < m.pStart.size(); ++pi)
        for (int f = m.pStart[pi]; f < m.pStart[pi] + m.pSize[pi]; ++f)
            if (type[f] == 1 && k.bcs[pi].uT != BC_FIXED) return 1;
    out.on = true;
    out.un.assign(m.nF, 0.0);
    out.incl.assign(nB, 0);
    out.Ub.assign(3 * (size_t)nB, 0.0);
    out.set.assign(m.nF, 0);
    for (int f = 0; f < m.nF; ++f) {
        V3 fu{0, 0, 0};
        if (type[f]) fu = faceU(f);
        out.set[f] = type[f] != 0;
        out.un[f] = dot(fu, v3(&m.Sf[3 * f]) / m.magSf[f]);
        if (f >= m.nI && type[f] == 1) {
            const int b = f - m.nI;
            out.incl[b] = 1;
            out.Ub[3 * b] = fu.x; out.Ub[3 * b + 1] = fu.y; out.Ub[3 * b + 2] = fu.z;
        }
    }
    return 0;
}

// Several rotating zones (MRFCoupledZoneList.C:55-142): each zone in list
// order assigns faceU on its faces, Omega in its cells and U on its moving
// patch faces (MRFCoupledZone.C:307-318, 563-627), so a later zone
// overrides an earlier one where they meet.
FramePart rotCombined(const std::vector<FramePart>& zones) {
    FramePart R = zones.front();
    for (size_t z = 1; z < zones.size(); ++z) {
        const FramePart& Z = zones[z];
        for (size_t f = 0; f < R.un.size(); ++f)
            if (Z.set[f]) { R.un[f] = Z.un[f]; R.set[f] = 1; }
        for (size_t b = 0; b < R.incl.size(); ++b)
            if (Z.incl[b]) {
                R.incl[b] = 1;
                for (int d = 0; d < 3; ++d) R.Ub[3 * b + d] = Z.Ub[3 * b + d];
            }
    }
    return R;
}

// the case's frame arrays from its parts
void frameCompose(Case& k) {
    const Mesh& m = *k.m;
    const int nB = m.nF - m.nI;
    const FramePart &R = k.frot, &T = k.ftr;
    k.mrf = R.on || T.on;
    k.umrf.assign(m.nF, 0.0);
    for (int f = 0; f < m.nF; ++f)
        k.umrf[f] = R.on && T.on ? R.un[f] + T.un[f] : (R.on ? R.un[f] : T.un[f]);
    k.mrfIncl.assign(nB, 0);
    k.mrfUb.assign(3 * (size_t)nB, 0.0);
    for (int b = 0; b < nB; ++b)
        for (const FramePart* P : {&R, &T})
            if (P->on && P->incl[b]) {
                k.mrfIncl[b] = 1;
                for (int d = 0; d < 3; ++d) k.mrfUb[3 * b + d] = P->Ub[3 * b + d];
            }
    if (k.omegaC.empty()) k.omegaC.assign(3 * (size_t)m.nC, 0.0);
}

// One more rotating zone (the zone list's next entry): cells, origin,
// axis (normalised), omega [rad/s], non-rotating patches.  Returns 1 if a rotating patch's U is not
// fixedValue (the frame assigns its value, which only a fixedValue
// condition keeps), 0 on success.
int orc_set_mrf(void* h, int nZone, const int* zone, const double* origin, const double* axis, double omega,
                int nNonRot, const int* nonRot) {
    Case& k = *(Case*)h;
    const Mesh& m = *k.m;
    V3 ax = v3(axis);
    const double am = mag(ax);
    ax = am < ROOTVSMALL ? V3{0, 0, 0} : ax / am;
    const V3 Om = omega * ax, o = v3(origin);
    FramePart P;
    if (framePart(k, zone, nZone, nonRot, nNonRot, [&](int f) { return cross(Om, v3(&m.Cf[3 * f]) - o); }, P))
        return 1;
    k.rotZones.push_back(P);
    k.frot = rotCombined(k.rotZones);
    if (k.omegaC.empty()) k.omegaC.assign(3 * (size_t)m.nC, 0.0);
    for (int i = 0; i < nZone; ++i) {
        const int c = zone[i];
        k.omegaC[3 * c] = Om.x; k.omegaC[3 * c + 1] = Om.y; k.omegaC[3 * c + 2] = Om.z;
    }
    frameCompose(k);
    return 0;
}
// The translating zone (MRFTranslatingZone: cellZone, frameVelocity,
// nonMovingPatches); same return as orc_set_mrf.
int orc_set_translating(void* h, int nZone, const int* zone, const double* velocity, int nNonMov,
                        const int* nonMov) {
    Case& k = *(Case*)h;
    const V3 V = v3(velocity);
    FramePart P;
    if (framePart(k, zone, nZone, nonMov, nNonMov, [&](int) { return V; }, P)) return 1;
    k.ftr = P;
    frameCompose(k);
    return 0;
}
// The MRF set-up: umrf [nF], omegaC [3 nC], rotating boundary faces [nB]
// and their velocity [3 nB].
void orc_get_mrf(void* h, double* umrf, double* omegaC, int* incl, double* Ub) {
    const Case& k = *(const Case*)h;
    const Mesh& m = *k.m;
    const int nB = m.nF - m.nI;
    for (int f = 0; f < m.nF; ++f) umrf[f] = k.mrf ? k.umrf[f] : 0.0;
    for (int i = 0; i < 3 * m.nC; ++i) omegaC[i] = k.mrf ? k.omegaC[i] : 0.0;
    for (int b = 0; b < nB; ++b) {
        incl[b] = k.mrf ? k.mrfIncl[b] : 0;
        for (int d = 0; d < 3; ++d) Ub[3 * b + d] = k.mrf ? k.mrfUb[3 * b + d] : 0.0;
    }
}

void orc_set_bc(void* h, int patch, int pT, int uT, int tT, const double* pv, const double* uv, const double* tv) {
    Case& k = *(Case*)h;
    BC& b = k.bcs[patch];
    const int n = k.m->pSize[patch];
    b.pT = pT; b.uT = uT; b.tT = tT;
    if (pT == BC_FIXED || pT == BC_TOTAL_P) b.pv.assign(pv, pv + n);
    if (uT == BC_FIXED) b.uv.assign(uv, uv + 3 * n);
    if (uT == BC_PIOV) {
        if (uv) b.uv.assign(uv, uv + 3 * n);
        else b.uv.assign(3 * (size_t)n, 0.0);
    }
    if (tT == BC_FIXED || tT == BC_TOTAL_T) b.tv.assign(tv, tv + n);
}

// dbnsFoam's steady loop (dbnsFoam.C:105-124) under pseudotimeControl::loop
// (pseudotimeControl.C:159-241) with criteriaSatisfied (:72-101): before
// each step, corr_ advances; a converged control ends the loop; otherwise
// the criteria are tested on residuals_ (the initial residuals of the last
// solve, outerLoop.H:97), except on the first call, at corr_ == 1 and at
// corr_ == nPseudoCorr (finalIter); corr_ == 2 stores initResiduals_ and
// skips the relative test; maxInit is the largest initial residual
// (residualsIO.H:243-256) and the relative ratio divides by
// initial + ROOTVSMALL (:183-200).
int orc_pseudo_time_loop(void* h, int maxSteps, double tol, double tolRel, int nCorr, int nCorrMin, int* run,
                         int* conv) {
    Case& k = *(Case*)h;
    int corr = k.corr, n = 0;
    bool converged = false;
    while (n < maxSteps) {
        ++corr;
        if (converged) break;
        bool crit = false;
        if (k.ptFirst || corr == 1 || corr == nCorr) {
            k.ptFirst = false;
        } else {
            const double res[5] = {k.initS[0], k.initS[1], k.initV[0], k.initV[1], k.initV[2]};
            const bool store = corr == 2;
            if (store) std::copy(res, res + 5, k.ptInit);
            double mx = -1e300;
            for (double v : res) mx = std::max(mx, v);
            bool rel = false;
            if (!store) {
                double mr = -1e300;
                for (int i = 0; i < 5; ++i) mr = std::max(mr, res[i] / (k.ptInit[i] + ROOTVSMALL));
                rel = mr < tolRel;
            }
            crit = corr >= nCorrMin && (mx < tol || rel);
        }
        if (crit) converged = true;
        Res r;
        const int rc = pseudoStep(k, r);
        if (rc) return rc;
        ++n;
    }
    *run = n;
    *conv = converged ? 1 : 0;
    return 0;
}

// Boundary values, boundary mass flux and inflow switch of every boundary
// face (zero on empty and cyclic faces).
void orc_get_boundary(void* h, double* p, double* U, double* T, double* phi, int* inflow) {
    Case& k = *(Case*)h;
    const int nB = k.m->nF - k.m->nI;
    for (int b = 0; b < nB; ++b) {
        p[b] = k.pB[b]; T[b] = k.TB[b]; phi[b] = k.phiB[b]; inflow[b] = k.inflowB[b];
        for (int d = 0; d < 3; ++d) U[3 * b + d] = k.UB[3 * b + d];
    }
}

// Initial boundary values of the derived conditions (any may be NULL).
void orc_set_bc_values(void* h, int patch, const double* p, const double* U, const double* T) {
    Case& k = *(Case*)h;
    BC& b = k.bcs[patch];
    const int n = k.m->pSize[patch];
    if (p) b.pInit.assign(p, p + n);
    if (U) b.uInit.assign(U, U + 3 * n);
    if (T) b.tInit.assign(T, T + n);
}

void orc_set_fields(void* h, const double* rho, const double* U, const double* T) {
    setFields(*(Case*)h, rho, U, T);
}

void orc_get_fields(void* h, double* rho, double* U, double* T, double* p, double* W) {
    Case& k = *(Case*)h;
    const int nC = k.m->nC;
    std::copy(k.rho.begin(), k.rho.end(), rho);
    std::copy(k.U.begin(), k.U.end(), U);
    std::copy(k.T.begin(), k.T.end(), T);
    std::copy(k.p.begin(), k.p.end(), p);
    for (int c = 0; c < nC; ++c) {
        W[c] = k.rho[c];
        W[nC + c] = k.rhoE[c];
        for (int d = 0; d < 3; ++d) W[(2 + d) * (size_t)nC + c] = k.rhoU[3 * c + d];
    }
}

void orc_calc_flux(void* h, double* phi, double* phiU, double* phiE) {
    Case& k = *(Case*)h;
    vec a, b, c;
    calcFlux(k, a, b, c);
    std::copy(a.begin(), a.end(), phi);
    std::copy(b.begin(), b.end(), phiU);
    std::copy(c.begin(), c.end(), phiE);
}

void orc_residual(void* h, double* b) {
    vec r;
    residual(*(Case*)h, r);
    std::copy(r.begin(), r.end(), b);
}

void orc_local_dt(void* h, double* rDT) {
    Case& k = *(Case*)h;
    localDeltaT(k);
    std::copy(k.rDT.begin(), k.rDT.end(), rDT);
}

// Assemble the coupled Jacobian (needs orc_local_dt first) and return it
// as 5x5 blocks in [rho, rhoE, rhoU] order: diag[nC*25], upper[nI*25],
// lower[nI*25].
void orc_assemble(void* h, double* diag, double* upper, double* lower) {
    Case& k = *(Case*)h;
    assemble(k);
    const Mesh& m = *k.m;
    auto fill = [&](double* out, int idx, const vec& ss00, const vec& ss01, const vec& ss10, const vec& ss11,
                    const vec& sv0, const vec& sv1, const vec& vs0, const vec& vs1, const vec& vv) {
        double* B = out + 25 * (size_t)idx;
        B[0] = ss00[idx]; B[1] = ss01[idx]; B[5] = ss10[idx]; B[6] = ss11[idx];
        for (int d = 0; d < 3; ++d) {
            B[2 + d] = sv0[3 * idx + d];
            B[5 + 2 + d] = sv1[3 * idx + d];
            B[5 * (2 + d) + 0] = vs0[3 * idx + d];
            B[5 * (2 + d) + 1] = vs1[3 * idx + d];
            for (int e = 0; e < 3; ++e) B[5 * (2 + d) + 2 + e] = vv[9 * idx + 3 * d + e];
        }
    };
    if (diag)
        for (int c = 0; c < m.nC; ++c)
            fill(diag, c, k.SS[0][0].diag, k.SS[0][1].diag, k.SS[1][0].diag, k.SS[1][1].diag, k.SV[0].diag,
                 k.SV[1].diag, k.VS[0].diag, k.VS[1].diag, k.VV.diag);
    if (upper)
        for (int f = 0; f < m.nI; ++f)
            fill(upper, f, k.SS[0][0].upper, k.SS[0][1].upper, k.SS[1][0].upper, k.SS[1][1].upper, k.SV[0].upper,
                 k.SV[1].upper, k.VS[0].upper, k.VS[1].upper, k.VV.upper);
    if (lower)
        for (int f = 0; f < m.nI; ++f)
            fill(lower, f, k.SS[0][0].lower, k.SS[0][1].lower, k.SS[1][0].lower, k.SS[1][1].lower, k.SV[0].lower,
                 k.SV[1].lower, k.VS[0].lower, k.VS[1].lower, k.VV.lower);
}

// Interface blocks of the boundary faces ([nB][25], zero off cyclic faces),
// same layout as orc_assemble; call after orc_assemble.
void orc_get_interfaces(void* h, double* iu) {
    Case& k = *(Case*)h;
    const Mesh& m = *k.m;
    const int nB = m.nF - m.nI;
    std::fill(iu, iu + 25 * (size_t)nB, 0.0);
    auto get = [&](const LduBlock& A, int b, int q) { return A.iu.empty() ? 0.0 : A.iu[(size_t)b * A.ncomp + q]; };
    for (int b = 0; b < nB; ++b) {
        double* B = iu + 25 * (size_t)b;
        B[0] = get(k.SS[0][0], b, 0); B[1] = get(k.SS[0][1], b, 0);
        B[5] = get(k.SS[1][0], b, 0); B[6] = get(k.SS[1][1], b, 0);
        for (int d = 0; d < 3; ++d) {
            B[2 + d] = get(k.SV[0], b, d);
            B[7 + d] = get(k.SV[1], b, d);
            B[5 * (2 + d)] = get(k.VS[0], b, d);
            B[5 * (2 + d) + 1] = get(k.VS[1], b, d);
            for (int e = 0; e < 3; ++e) B[5 * (2 + d) + 2 + e] = get(k.VV, b, 3 * d + e);
        }
    }
}

// Replace the coupled matrix by given 5x5 blocks (same layout as above).
void orc_set_matrix(void* h, const double* diag, const double* upper, const double* lower) {
    Case& k = *(Case*)h;
    const Mesh& m = *k.m;
    for (int i = 0; i < 2; ++i) for (int j = 0; j < 2; ++j) k.SS[i][j].init(m.nC, m.nI, 1);
    for (int i = 0; i < 2; ++i) { k.SV[i].init(m.nC, m.nI, 3); k.VS[i].init(m.nC, m.nI, 3); }
    k.VV.init(m.nC, m.nI, 9);
    auto put = [&](const double* in, int idx, vec& ss00, vec& ss01, vec& ss10, vec& ss11, vec& sv0, vec& sv1,
                   vec& vs0, vec& vs1, vec& vv) {
        const double* B = in + 25 * (size_t)idx;
        ss00[idx] = B[0]; ss01[idx] = B[1]; ss10[idx] = B[5]; ss11[idx] = B[6];
        for (int d = 0; d < 3; ++d) {
            sv0[3 * idx + d] = B[2 + d];
            sv1[3 * idx + d] = B[7 + d];
            vs0[3 * idx + d] = B[5 * (2 + d)];
            vs1[3 * idx + d] = B[5 * (2 + d) + 1];
            for (int e = 0; e < 3; ++e) vv[9 * idx + 3 * d + e] = B[5 * (2 + d) + 2 + e];
        }
    };
    for (int c = 0; c < m.nC; ++c)
        put(diag, c, k.SS[0][0].diag, k.SS[0][1].diag, k.SS[1][0].diag, k.SS[1][1].diag, k.SV[0].diag, k.SV[1].diag,
            k.VS[0].diag, k.VS[1].diag, k.VV.diag);
    for (int f = 0; f < m.nI; ++f) {
        put(upper, f, k.SS[0][0].upper, k.SS[0][1].upper, k.SS[1][0].upper, k.SS[1][1].upper, k.SV[0].upper,
            k.SV[1].upper, k.VS[0].upper, k.VS[1].upper, k.VV.upper);
        put(lower, f, k.SS[0][0].lower, k.SS[0][1].lower, k.SS[1][0].lower, k.SS[1][1].lower, k.SV[0].lower,
            k.SV[1].lower, k.VS[0].lower, k.VS[1].lower, k.VV.lower);
    }
}

void orc_amul(void* h, const double* x, double* y) {
    Case& k = *(Case*)h;
    CVec X, Y;
    toC(x, k.m->nC, X);
    matrixMul(k, X, Y);
    fromC(Y, k.m->nC, y);
}

int orc_precondition(void* h, double* r) {
    Case& k = *(Case*)h;
    if (k.s.precond == 0) {
        int rc = lusgsSetup(k);
        if (rc) return rc;
    }
    CVec R;
    toC(r, k.m->nC, R);
    precondition(k, R);
    fromC(R, k.m->nC, r);
    return 0;
}

// solveForIncr on the current matrix: W and b in [5][nC] layout.
int orc_solve(void* h, const double* W, const double* b, double* dW, double* res /*10*/, int* nIter) {
    Case& k = *(Case*)h;
    if (k.s.precond == 0) {
        int rc = lusgsSetup(k);
        if (rc) return rc;
    }
    CVec w, bb, dw;
    toC(W, k.m->nC, w);
    toC(b, k.m->nC, bb);
    Res r = gmresSolve(k, w, bb, dw);
    fromC(dw, k.m->nC, dW);
    res[0] = r.sInit[0]; res[1] = r.sInit[1];
    for (int d = 0; d < 3; ++d) res[2 + d] = r.vInit[d];
    res[5] = r.sFinal[0]; res[6] = r.sFinal[1];
    for (int d = 0; d < 3; ++d) res[7 + d] = r.vFinal[d];
    *nIter = r.nIter;
    return 0;
}

// n pseudo-steps; res receives 10 residuals per step, nIter per step.
int orc_pseudo_steps(void* h, int n, double* res, int* nIter) {
    Case& k = *(Case*)h;
    for (int i = 0; i < n; ++i) {
        Res r;
        int rc = pseudoStep(k, r);
        if (rc) return rc;
        if (res) {
            double* o = res + 10 * (size_t)i;
            o[0] = r.sInit[0]; o[1] = r.sInit[1];
            for (int d = 0; d < 3; ++d) o[2 + d] = r.vInit[d];
            o[5] = r.sFinal[0]; o[6] = r.sFinal[1];
            for (int d = 0; d < 3; ++d) o[7 + d] = r.vFinal[d];
        }
        if (nIter) nIter[i] = r.nIter;
    }
    return 0;
}

void orc_get_pseudo_co(void* h, double* co) {
    Case& k = *(Case*)h;
    std::copy(k.pseudoCo.begin(), k.pseudoCo.end(), co);
}


// Face-level flux for pinning tests: scheme 0 AUSM+up, 1 Roe.
// in = [Sx,Sy,Sz, rhoL,pL,ULx,ULy,ULz,TL, rhoR,pR,URx,URy,URz,TR]; the
// reconstructed E, H, c* are formed from each side's state as the cell
// fields are.  out = [phi, phiUx, phiUy, phiUz, phiE].  um: the face's
// MRF normal velocity (orc_face_flux: 0).
void orc_face_flux_mrf(int scheme, int lowMach, double eFix, double R, double gamma, const double* in, double um,
                       double* out) {
    Case k;
    k.s.flux = scheme; k.s.lowMach = lowMach; k.s.entropyFix = eFix; k.s.R = R; k.s.gamma = gamma;
    k.Cv = R / (gamma - 1.0);
    V3 S{in[0], in[1], in[2]};
    const double magS = mag(S);
    double rho[2], p[2], T[2], E[2], H[2], cs[2];
    V3 U[2];
    for (int s = 0; s < 2; ++s) {
        const double* q = in + 3 + 6 * s;
        rho[s] = q[0]; p[s] = q[1]; U[s] = V3{q[2], q[3], q[4]}; T[s] = q[5];
        E[s] = k.Cv * T[s] + 0.5 * magSqr(U[s]);
        H[s] = std::max(E[s], SMALL) + std::max(p[s] / rho[s], SMALL);
        cs[s] = std::sqrt(2.0 * (gamma - 1.0) / (gamma + 1.0) * H[s]);
    }
    double ph, pe;
    V3 pu;
    if (scheme == FLUX_AUSM)
        ausmFace(k, S, magS, U[0], U[1], cs[0], cs[1], p[0], p[1], rho[0], rho[1], H[0], H[1], um, ph, pu, pe);
    else if (scheme == FLUX_ROE)
        roeFace(k, gamma, S, magS, rho[0], rho[1], p[0], p[1], U[0], U[1], E[0], E[1], H[0], H[1], um, ph, pu, pe);
    else {
        double ch[2];
        // sqrt(gamma/psi) with psi = 1/(R T)
        for (int q = 0; q < 2; ++q) ch[q] = std::max(std::sqrt(gamma / (1.0 / (R * T[q]))), VSMALL);
        hllcFace(k, gamma, S, magS, rho[0], rho[1], p[0], p[1], U[0], U[1], ch[0], ch[1], E[0], E[1], H[0], H[1], um,
                 ph, pu, pe);
    }
    out[0] = ph; out[1] = pu.x; out[2] = pu.y; out[3] = pu.z; out[4] = pe;
}
void orc_face_flux(int scheme, int lowMach, double eFix, double R, double gamma, const double* in, double* out) {
    orc_face_flux_mrf(scheme, lowMach, eFix, R, gamma, in, 0.0, out);
}

// Reconstructed face values (L, R) of one cell field of the current state:
// 0 rho, 1 p, 2 Ux, 3 Uy, 4 Uz, 5 E, 6 H, 7 c*.
void orc_reconstruct(void* h, int field, double* L, double* R) {
    Case& k = *(Case*)h;
    Mesh& m = *k.m;
    Derived d;
    derive(k, d);
    const int nB = m.nF - m.nI;
    vec vc(m.nC), vb(nB, 0.0);
    for (int c = 0; c < m.nC; ++c) {
        switch (field) {
            case 0: vc[c] = k.rho[c]; break;
            case 1: vc[c] = k.p[c]; break;
            case 2: case 3: case 4: vc[c] = k.U[3 * c + field - 2]; break;
            case 5: vc[c] = d.E[c]; break;
            case 6: vc[c] = d.H[c]; break;
            default: vc[c] = k.s.flux == FLUX_HLLC ? d.chl[c] : d.cstar[c]; break;
        }
    }
    for (int b = 0; b < nB; ++b) {
        switch (field) {
            case 0: vb[b] = k.rhoB[b]; break;
            case 1: vb[b] = k.pB[b]; break;
            case 2: case 3: case 4: vb[b] = k.UB[3 * b + field - 2]; break;
            case 5: vb[b] = d.Eb[b]; break;
            case 6: vb[b] = d.Hb[b]; break;
            default: vb[b] = k.s.flux == FLUX_HLLC ? d.chlb[b] : d.cstarb[b]; break;
        }
    }
    vec l, r;
    reconstruct(k, vc.data(), vb.data(), 1, l, r);
    std::copy(l.begin(), l.end(), L);
    std::copy(r.begin(), r.end(), R);
}

// ------------------------------------------ tangent-linear residual -----
// State from conserved W ([5][nC], [rho, rhoE, rhoU]) through updateFields
// with a zero increment.
void orc_set_state_w(void* h, const double* W) {
    Case& k = *(Case*)h;
    const int nC = k.m->nC;
    k.rho.assign(W, W + nC);
    k.rhoE.assign(W + nC, W + 2 * (size_t)nC);
    k.rhoU.assign(3 * (size_t)nC, 0.0);
    k.U.assign(3 * (size_t)nC, 0.0);
    k.e.assign(nC, 0.0); k.T.assign(nC, 0.0); k.p.assign(nC, 0.0);
    for (int c = 0; c < nC; ++c)
        for (int d = 0; d < 3; ++d) k.rhoU[3 * c + d] = W[(2 + d) * (size_t)nC + c];
    CVec z;
    z.init(nC);
    updateFields(k, z);
}

static const double* sfOr(const Case& k, const double* Sf) { return Sf ? Sf : k.m->Sf.data(); }
static const double* cfOr(const Case& k, const double* Cf) { return Cf ? Cf : k.m->Cf.data(); }

// Primal b(W, Sf, Cf) of the tangent restatement (Sf/Cf NULL: the mesh's).
void orc_residual_w(void* h, const double* W, const double* Sf, const double* Cf, double* b) {
    Case& k = *(Case*)h;
    residualT<double>(k, W, sfOr(k, Sf), cfOr(k, Cf), b);
}

// b and the Jacobian-vector product db = J_W dW + J_S dSf + J_C dCf
// (NULL tangents are zero).
void orc_residual_jvp(void* h, const double* W, const double* dW, const double* dSf, const double* dCf, double* b,
                      double* db) {
    Case& k = *(Case*)h;
    const Mesh& m = *k.m;
    const size_t n5 = 5 * (size_t)m.nC, n3 = 3 * (size_t)m.nF;
    std::vector<Dual> w(n5), s(n3), c(n3), r(n5);
    for (size_t i = 0; i < n5; ++i) w[i] = Dual(W[i], dW ? dW[i] : 0.0);
    for (size_t i = 0; i < n3; ++i) {
        s[i] = Dual(m.Sf[i], dSf ? dSf[i] : 0.0);
        c[i] = Dual(m.Cf[i], dCf ? dCf[i] : 0.0);
    }
    residualT<Dual>(k, w.data(), s.data(), c.data(), r.data());
    for (size_t i = 0; i < n5; ++i) {
        if (b) b[i] = r[i].v;
        db[i] = r[i].d;
    }
}

// Dense Jacobians by unit tangents (small meshes): JW[5nC][5nC],
// JS[5nC][3nF], JC[5nC][3nF], row = residual entry, column = input entry.
void orc_residual_jac_dense(void* h, const double* W, double* JW, double* JS, double* JC) {
    Case& k = *(Case*)h;
    const Mesh& m = *k.m;
    const size_t n5 = 5 * (size_t)m.nC, n3 = 3 * (size_t)m.nF;
    std::vector<Dual> w(n5), s(n3), c(n3), r(n5);
    for (size_t i = 0; i < n5; ++i) w[i] = Dual(W[i], 0.0);
    for (size_t i = 0; i < n3; ++i) { s[i] = Dual(m.Sf[i], 0.0); c[i] = Dual(m.Cf[i], 0.0); }
    for (size_t j = 0; j < n5; ++j) {
        w[j].d = 1.0;
        residualT<Dual>(k, w.data(), s.data(), c.data(), r.data());
        for (size_t i = 0; i < n5; ++i) JW[i * n5 + j] = r[i].d;
        w[j].d = 0.0;
    }
    for (size_t j = 0; j < n3; ++j) {
        s[j].d = 1.0;
        residualT<Dual>(k, w.data(), s.data(), c.data(), r.data());
        for (size_t i = 0; i < n5; ++i) JS[i * n3 + j] = r[i].d;
        s[j].d = 0.0;
        c[j].d = 1.0;
        residualT<Dual>(k, w.data(), s.data(), c.data(), r.data());
        for (size_t i = 0; i < n5; ++i) JC[i * n3 + j] = r[i].d;
        c[j].d = 0.0;
    }
}


// ------------------------------------------------------ harmonic balance --
void* orc_hb_new(int nO, void** cases, const double* D) {
    HB* h = new HB;
    h->nO = nO;
    for (int J = 0; J < nO; ++J) h->k.push_back((Case*)cases[J]);
    h->D.assign(D, D + (size_t)nO * nO);
    return h;
}
void orc_hb_free(void* hp) {
    HB* h = (HB*)hp;
    for (Case* k : h->k)
        if (k->pl == &h->pl) k->pl = nullptr;
    delete h;
}

// One side of a phase-lagged cyclic pair: its patch, time shift Dl
// [nO][nO] (hbLagOperator of the side's IBPA), cylCoords with the patch's
// rotation axis (normalised here, rotAxis/mag(rotAxis)) and centre.
int orc_hb_set_phase_lag(void* hp, int patch, const double* Dl, int cyl, const double* axis, const double* origin) {
    HB& h = *(HB*)hp;
    const Mesh& m = *h.k[0]->m;
    if (patch < 0 || patch >= (int)m.pStart.size() || m.pKind[patch] != PK_CYCLIC) return -1;
    if (h.k[0]->s.viscous && !cyl) return -2;  // the viscous terms take grad(U) whole
    PhaseLag& L = h.pl;
    if (L.sides.empty()) {
        L.nO = h.nO;
        L.inst = h.k;
        L.side.assign(m.nF - m.nI, -1);
    }
    PhaseLag::Side sd;
    sd.D.assign(Dl, Dl + (size_t)h.nO * h.nO);
    sd.cyl = cyl != 0;
    const V3 a = v3(axis);
    const double la = std::sqrt(magSqr(a));
    sd.axis = V3{a.x / la, a.y / la, a.z / la};
    sd.origin = v3(origin);
    L.sides.push_back(sd);
    for (int f = m.pStart[patch]; f < m.pStart[patch] + m.pSize[patch]; ++f) L.side[f - m.nI] = (int)L.sides.size() - 1;
    for (int J = 0; J < h.nO; ++J) { h.k[J]->pl = &L; h.k[J]->inst = J; }
    return 0;
}

// Far-side values of instance J's boundary faces ([nB]; zero off the
// phase-lagged faces): field PL_RHO .. PL_GAMMA (comp: velocity component;
// PL_U with comp 3..5: the velocity vector's component comp - 3, forwardT
// applied).
void orc_hb_lag_far(void* hp, int J, int field, int comp, double* out) {
    HB& h = *(HB*)hp;
    const Case& k = *h.k[J];
    const Mesh& m = *k.m;
    for (int b = 0; b < m.nF - m.nI; ++b) {
        out[b] = 0.0;
        if (!lagFace(k, b)) continue;
        if (field == PL_U && comp >= 3) {
            const V3 u = lagU(k, b);
            out[b] = (&u.x)[comp - 3];
        } else out[b] = lagScalar(k, b, field, comp);
    }
}

int orc_hb_lag_operator(int nT, const double* omega, const double* t, double ibpa, double* Dl) {
    return hbLagOperator(nT, omega, t, ibpa, Dl);
}

void orc_hb_residual(void* hp, double* b) {
    HB& h = *(HB*)hp;
    std::vector<vec> bb;
    hbResidual(h, bb);
    const size_t n5 = 5 * (size_t)h.k[0]->m->nC;
    for (int J = 0; J < h.nO; ++J) std::copy(bb[J].begin(), bb[J].end(), b + J * n5);
}

int orc_hb_assemble(void* hp) { return hbAssemble(*(HB*)hp); }

static void hbSplit(const HB& h, const double* x, std::vector<CVec>& v) {
    const int nC = h.k[0]->m->nC;
    v.resize(h.nO);
    for (int J = 0; J < h.nO; ++J) toC(x + (size_t)J * 5 * nC, nC, v[J]);
}
static void hbJoin(const HB& h, const std::vector<CVec>& v, double* x) {
    const int nC = h.k[0]->m->nC;
    for (int J = 0; J < h.nO; ++J) fromC(v[J], nC, x + (size_t)J * 5 * nC);
}

void orc_hb_amul(void* hp, const double* x, double* y) {
    HB& h = *(HB*)hp;
    std::vector<CVec> X, Y;
    hbSplit(h, x, X);
    hbAmul(h, X, Y);
    hbJoin(h, Y, y);
}

void orc_hb_precondition(void* hp, double* r) {
    HB& h = *(HB*)hp;
    std::vector<CVec> R;
    hbSplit(h, r, R);
    hbPrecondition(h, R);
    hbJoin(h, R, r);
}

static void hbResOut(const HB& h, const HBRes& r, double* res) {
    for (int J = 0; J < h.nO; ++J) {
        double* o = res + 10 * J;
        o[0] = r.sInit[2 * J]; o[1] = r.sInit[2 * J + 1];
        for (int d = 0; d < 3; ++d) o[2 + d] = r.vInit[3 * J + d];
        o[5] = r.sFinal[2 * J]; o[6] = r.sFinal[2 * J + 1];
        for (int d = 0; d < 3; ++d) o[7 + d] = r.vFinal[3 * J + d];
    }
}

// GMRES on the assembled HB system for the current state and the given b.
int orc_hb_solve(void* hp, const double* b, double* dW, double* res, int* nIter) {
    HB& h = *(HB*)hp;
    std::vector<CVec> W(h.nO), B, D;
    for (int J = 0; J < h.nO; ++J) hbConserved(*h.k[J], W[J]);
    hbSplit(h, b, B);
    HBRes r = hbGmres(h, W, B, D);
    hbJoin(h, D, dW);
    hbResOut(h, r, res);
    if (nIter) *nIter = r.nIter;
    return 0;
}

int orc_hb_pseudo_steps(void* hp, int n, double* res) {
    HB& h = *(HB*)hp;
    for (int i = 0; i < n; ++i) {
        HBRes r;
        int rc = hbPseudoStep(h, r);
        if (rc) return rc;
        hbResOut(h, r, res + (size_t)i * 10 * h.nO);
    }
    return 0;
}

int orc_hb_operator(int nT, const double* omega, int nO, const double* t, double* D) {
    return hbOperator(nT, omega, nO, t, D);
}

}  // extern "C"
