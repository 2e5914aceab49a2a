// Adjoint (vector-Jacobian product) of the coupled residual
//     b(W, Sf, Cf) = -(sum_f +-F_f / V) * V   (+ the viscous terms),
// v -> (db/dW)^T v, (db/dSf)^T v, (db/dCf)^T v, with the cell centres,
// volumes and interpolation weights held fixed.  The reference has no
// adjoint; the forward chain differentiated here is the one of
// updateFields.H:7-104 (W -> rho, U, e, T, p and the boundary values),
// ausmPlusUpFluxScheme.C:73-299 / roeFluxScheme.C:41-241 /
// hllcFluxScheme.C:76-221 (E, H, c* or the sound speed,
// reconstruction with cellLimited Gauss gradients, the face flux) and
// residualsUpdate.H:1-43 (surfaceIntegrate, laplacian(muEff, U),
// div(tauMC), div(sigmaDotU), laplacian(alphaEff, e)).
//
// The reconstruction is recomputed, not taped: the forward pass re-runs the
// boundary and gradient kernels (keeping the unlimited gradients and the
// limiter values), then the reverse kernels pull adjoints without atomics:
//   k_adj_face   one thread per face: AUSM+up or Roe reverse (each one
//                forward evaluation and one backward sweep) or HLLC (one
//                dual-number pass per face value), adjoint of
//                the face values (owner/neighbour side) and of Sf, Cf
//                (flux part)
//   k_adj_visc   one thread per face (viscous cases): adjoint of the
//                viscous face terms
//   k_adj_cell   one thread per cell: adjoint of the limited gradients, of
//                the limiter (its active face and max/min neighbour) and of
//                the Gauss gradient sum divided by V
//   k_adj_gauss  one thread per face: Gauss-gradient face values, boundary
//                values -> owner primitives, limiter pulls into Cf
//   k_adj_state  one thread per cell: cell-value adjoints -> E, H, c* ->
//                rho, U, e, T, p -> W
// Branches follow the primal value: max(a, b) takes b only if b > a,
// min(a, b) takes b only if b < a, and the limiter's min over faces keeps
// the first face that attains it.
#include "icsf_kernels.h"
#include "icsf_recon.h"
#include "icsf_roe_rev.h"
#include "icsf_views.h"

namespace icsf {

namespace {

#define DF __device__ __forceinline__

// Reconstructed fields the scheme uses: AUSM+up takes c* and not E, Roe
// takes E and not c*, HLLC all of them (its F_CS is max(sqrt(gamma/psi),
// VSMALL)).
DF bool used(int flux, int k) {
    return flux == FLUX_ROE ? k != F_CS : (flux == FLUX_HLLC ? true : k != F_E);
}

// ---- HLLC: forward-mode passes of the face flux on dual numbers (value
// and one tangent), one per input, contracted with the flux adjoint
struct Dn {
    double v, d;
};
DF Dn operator+(Dn a, Dn b) { return Dn{a.v + b.v, a.d + b.d}; }
DF Dn operator-(Dn a, Dn b) { return Dn{a.v - b.v, a.d - b.d}; }
DF Dn operator*(Dn a, Dn b) { return Dn{a.v * b.v, a.d * b.v + a.v * b.d}; }
DF Dn operator/(Dn a, Dn b) {
    const double q = a.v / b.v;
    return Dn{q, (a.d - q * b.d) / b.v};
}
DF Dn dsqrt(Dn a) {
    const double r = sqrt(a.v);
    return Dn{r, r > 0.0 ? a.d / (2.0 * r) : 0.0};
}
DF Dn dfabs(Dn a) { return a.v < 0.0 ? Dn{-a.v, -a.d} : a; }
DF Dn dmin(Dn a, Dn b) { return b.v < a.v ? b : a; }  // std::min(a, b)
DF Dn dmax(Dn a, Dn b) { return a.v < b.v ? b : a; }  // std::max(a, b)
DF Dn cst(double x) { return Dn{x, 0.0}; }
struct V3n {
    Dn x, y, z;
};
DF V3n operator+(V3n a, V3n b) { return V3n{a.x + b.x, a.y + b.y, a.z + b.z}; }
DF V3n operator*(Dn s, V3n a) { return V3n{s * a.x, s * a.y, s * a.z}; }
DF V3n operator*(V3n a, Dn s) { return V3n{a.x * s, a.y * s, a.z * s}; }
DF V3n operator/(V3n a, Dn s) { return V3n{a.x / s, a.y / s, a.z / s}; }
DF Dn dotn(V3n a, V3n b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// hllcFace() (icsf_flux.h, hllcFluxScheme.C:76-221) on dual numbers, no
// frame velocity; br: the switches (Sl, Sr choices, the signs of Sl, Sr,
// Sm, the fabs argument).
DF void hllcDual(double gIv, V3n S, Dn magS, Dn rl, Dn rr, Dn pl, Dn pr, V3n Ul, V3n Ur, Dn cl, Dn cr, Dn El,
                 Dn Er, Dn Hl, Dn Hr, Dn& phi, V3n& phiU, Dn& phiE, int* br) {
    const Dn gI = cst(gIv), one = cst(1.0);
    const V3n n = S / magS;
    const Dn coefR = dsqrt(dmax(cst(kVSMALL), rr) / dmax(cst(kVSMALL), rl));
    const V3n uAvg = (coefR * Ur + Ul) / (coefR + one);
    const Dn HAvg = (coefR * Hr + Hl) / (coefR + one);
    const Dn arg = (gI - one) * (HAvg - cst(0.5) * dotn(uAvg, uAvg));
    const Dn cAvg = dsqrt(dfabs(arg));
    const Dn ul = dotn(Ul, n), ur = dotn(Ur, n), ua = dotn(uAvg, n);
    const Dn Sl = dmin(ul - cl, ua - cAvg);
    const Dn Sr = dmax(ur + cr, ua + cAvg);
    const Dn Sm = (rr * ur * (Sr - ur) - rl * ul * (Sl - ul) + pl - pr) / (rr * (Sr - ur) - rl * (Sl - ul));
    const Dn cSl = cst(Sl.v > 0 ? 1.0 : 0.0), cSr = cst(Sr.v < 0 ? 1.0 : 0.0), cSm = cst(Sm.v > 0 ? 1.0 : 0.0);
    if (br)
        *br = (int)((ua - cAvg).v < (ul - cl).v) | (int)((ur + cr).v < (ua + cAvg).v) << 1 | (int)(Sl.v > 0) << 2 |
              (int)(Sr.v < 0) << 3 | (int)(Sm.v > 0) << 4 | (int)(arg.v < 0) << 5;
    const Dn cSlm = (one - cSl) * cSm, cSmr = (one - cSm) * (one - cSr);
    const Dn frl = Sm / (Sl - Sm) * ((Sl - ul) * rl);
    const Dn frr = Sm / (Sr - Sm) * ((Sr - ur) * rr);
    phi = (cSl * rl * ul + cSlm * frl + cSmr * frr + cSr * rr * ur) * magS;
    const Dn psl = rl * (ul - Sl) * (ul - Sm) + pl;
    const Dn psr = rr * (ur - Sr) * (ur - Sm) + pr;
    const V3n rUsl = one / (Sl - Sm) * ((Sl - ul) * rl * Ul + (psl - pl) * n);
    const V3n rUsr = one / (Sr - Sm) * ((Sr - ur) * rr * Ur + (psr - pr) * n);
    const V3n fUl = Sm * rUsl + psl * n;
    const V3n fUr = Sm * rUsr + psr * n;
    phiU = (cSl * (rl * ul * Ul + pl * n) + cSlm * fUl + cSmr * fUr + cSr * (rr * ur * Ur + pr * n)) * magS;
    const Dn rEsl = one / (Sl - Sm) * ((Sl - ul) * (rl * El) - pl * ul + psl * Sm);
    const Dn rEsr = one / (Sr - Sm) * ((Sr - ur) * (rr * Er) - pr * ur + psr * Sm);
    const Dn fEl = Sm * (rEsl + psl);
    const Dn fEr = Sm * (rEsr + psr);
    phiE = (cSl * rl * Hl * ul + cSlm * fEl + cSmr * fEr + cSr * rr * Hr * ur) * magS;
}

// Reverse of the HLLC face flux: one dual pass per face value (owner side
// L, neighbour side R, fields F_RHO..F_CS) and, with geom, per Sf
// component; each input's adjoint is the flux adjoint contracted with
// that pass's tangent.
DF void hllcFaceRev(double gI, V3 S, const double* L, const double* R, double bphi, V3 bphiU, double bphiE,
                    double* bL, double* bR, V3& bS, int* branches, bool geom) {
    const int nIn = geom ? 2 * NFLD + 3 : 2 * NFLD;
    double bs[3] = {0, 0, 0};
    for (int i = 0; i < nIn; ++i) {
        Dn l[NFLD], r[NFLD];
#pragma unroll
        for (int k = 0; k < NFLD; ++k) {
            l[k] = Dn{L[k], i == k ? 1.0 : 0.0};
            r[k] = Dn{R[k], i == NFLD + k ? 1.0 : 0.0};
        }
        const V3n Sd{Dn{S.x, i == 2 * NFLD ? 1.0 : 0.0}, Dn{S.y, i == 2 * NFLD + 1 ? 1.0 : 0.0},
                     Dn{S.z, i == 2 * NFLD + 2 ? 1.0 : 0.0}};
        const Dn magS = dsqrt(dotn(Sd, Sd));
        Dn ph, pe;
        V3n pu;
        hllcDual(gI, Sd, magS, l[F_RHO], r[F_RHO], l[F_P], r[F_P], V3n{l[F_UX], l[F_UY], l[F_UZ]},
                 V3n{r[F_UX], r[F_UY], r[F_UZ]}, l[F_CS], r[F_CS], l[F_E], r[F_E], l[F_H], r[F_H], ph, pu, pe,
                 i == 0 ? branches : nullptr);
        const double t = bphi * ph.d + (bphiU.x * pu.x.d + bphiU.y * pu.y.d + bphiU.z * pu.z.d) + bphiE * pe.d;
        if (i < NFLD) bL[i] = t;
        else if (i < 2 * NFLD) bR[i - NFLD] = t;
        else bs[i - 2 * NFLD] = t;
    }
    bS = V3{bs[0], bs[1], bs[2]};
}

// Reverse of ausmFace(): adjoints of the owner-side (L) and neighbour-side
// (R) face values, indexed by field, and of Sf.  Recomputes the forward
// quantities with the same expressions.
DF void ausmFaceAdj(bool lowMach, V3 S, double magS, const double* L, const double* R, double bphi, V3 bphiU,
                    double bphiE, double* bL, double* bR, V3& bS, int* branches) {
    const V3 UL{L[F_UX], L[F_UY], L[F_UZ]}, UR{R[F_UX], R[F_UY], R[F_UZ]};
    const double cL = L[F_CS], cR = R[F_CS], pL = L[F_P], pR = R[F_P], rL = L[F_RHO], rR = R[F_RHO];
    const double HL = L[F_H], HR = R[F_H];
    // ---- forward
    const double unL = dot(UL, S) / magS;
    const double unR = dot(UR, S) / magS;
    const bool bl = cL < unL;
    const double dl = bl ? unL : cL;
    const double cl = sqr(cL) / dl;
    const bool br = cR < -unR;
    const double dr = br ? -unR : cR;
    const double cr = sqr(cR) / dr;
    const bool bc = cr < cl;
    const double cf = bc ? cr : cl;
    const double ML = unL / cf;
    double MpL, PpL, dMpL, dPpL;
    if (fabs(ML) < 1.0) {
        const double p2 = 0.25 * sqr(ML + 1), m2 = -0.25 * sqr(ML - 1);
        const double dp2 = 0.5 * (ML + 1), dm2 = -0.5 * (ML - 1);
        MpL = p2 * (1 - 2 * m2);
        PpL = p2 * (2 - ML - 3 * ML * m2);
        dMpL = dp2 * (1 - 2 * m2) - 2 * p2 * dm2;
        dPpL = dp2 * (2 - ML - 3 * ML * m2) + p2 * (-1 - 3 * m2 - 3 * ML * dm2);
    } else {
        MpL = fmax(ML, 0.0);
        PpL = ML > 0 ? 1.0 : 0.0;
        dMpL = ML < 0.0 ? 0.0 : 1.0;
        dPpL = 0.0;
    }
    const double MR = unR / cf;
    double MmR, PmR, dMmR, dPmR;
    if (fabs(MR) < 1.0) {
        const double m2 = -0.25 * sqr(MR - 1), p2 = 0.25 * sqr(MR + 1);
        const double dm2 = -0.5 * (MR - 1), dp2 = 0.5 * (MR + 1);
        MmR = m2 * (1 + 2 * p2);
        PmR = m2 * (-2 - MR + 3 * MR * p2);
        dMmR = dm2 * (1 + 2 * p2) + 2 * m2 * dp2;
        dPmR = dm2 * (-2 - MR + 3 * MR * p2) + m2 * (-1 + 3 * p2 + 3 * MR * dp2);
    } else {
        MmR = fmin(MR, 0.0);
        PmR = MR < 0 ? 1.0 : 0.0;
        dMmR = 0.0 < MR ? 0.0 : 1.0;
        dPmR = 0.0;
    }
    const double M12 = MpL + MmR;
    double p12 = PpL * pL + PmR * pR;
    const double Mmean = 0.5 * (sqr(unL) + sqr(unR)) / sqr(cf);
    const double omM = 1.0 - Mmean;
    const double A = fmax(omM, 0.0);
    const double dA = omM < 0.0 ? 0.0 : 1.0;
    const double B = 0.5 * (rL + rR) * sqr(cf);
    const double MDiff = -0.25 * A * (pR - pL) / B;
    const double kap = ((M12 > 0.0 && M12 + MDiff <= 0.0) || (M12 < 0.0 && M12 + MDiff >= 0.0)) ? 0.2 : 1.0;
    const double M12f = kap == 0.2 ? M12 + 0.2 * MDiff : M12 + MDiff;
    const double sr = rL + rR, X = unR - unL;
    if (lowMach) p12 += -0.25 * PpL * PmR * sr * cf * X;
    const bool up = M12f >= 0;
    if (branches)  // every switch of the forward flux: where the derivative jumps
        *branches = (int)bl | (int)br << 1 | (int)bc << 2 | (int)(fabs(ML) < 1.0) << 3 |
                    (int)(fabs(ML) >= 1.0 && ML > 0) << 4 | (int)(fabs(MR) < 1.0) << 5 |
                    (int)(fabs(MR) >= 1.0 && MR < 0) << 6 | (int)(omM < 0.0) << 7 | (int)(kap == 0.2) << 8 |
                    (int)up << 9;
    const double rup = up ? rL : rR, Hup = up ? HL : HR;
    const V3 Uup = up ? UL : UR;
    const double rhoa = M12f * cf * rup;
    // ---- reverse
    const double rhoa_b = (bphi + dot(bphiU, Uup) + bphiE * Hup) * magS;
    double magS_b = bphi * rhoa + dot(bphiU, rhoa * Uup) + bphiE * rhoa * Hup;
    const V3 Uup_b = bphiU * (rhoa * magS);
    const double Hup_b = bphiE * rhoa * magS;
    const double p12_b = dot(bphiU, S);
    bS = p12 * bphiU;
    const double M12f_b = rhoa_b * cf * rup;
    double cf_b = rhoa_b * M12f * rup;
    const double rup_b = rhoa_b * M12f * cf;
    double PpL_b = 0, PmR_b = 0, rL_b = 0, rR_b = 0, unL_b = 0, unR_b = 0, pL_b = 0, pR_b = 0;
    if (lowMach) {
        const double q = -0.25 * p12_b;
        PpL_b += q * PmR * sr * cf * X;
        PmR_b += q * PpL * sr * cf * X;
        const double qs = q * PpL * PmR * cf * X;
        rL_b += qs;
        rR_b += qs;
        cf_b += q * PpL * PmR * sr * X;
        const double qx = q * PpL * PmR * sr * cf;
        unR_b += qx;
        unL_b -= qx;
    }
    const double M12_b = M12f_b;
    const double MDiff_b = kap * M12f_b;
    const double A_b = MDiff_b * (-0.25 * (pR - pL) / B);
    const double dp_b = MDiff_b * (-0.25 * A / B);
    pR_b += dp_b;
    pL_b -= dp_b;
    const double B_b = -MDiff_b * MDiff / B;
    const double Mmean_b = -A_b * dA;
    rL_b += B_b * 0.5 * sqr(cf);
    rR_b += B_b * 0.5 * sqr(cf);
    cf_b += B_b * sr * cf;
    unL_b += Mmean_b * unL / sqr(cf);
    unR_b += Mmean_b * unR / sqr(cf);
    cf_b -= Mmean_b * 2.0 * Mmean / cf;
    PpL_b += p12_b * pL;
    pL_b += p12_b * PpL;
    PmR_b += p12_b * pR;
    pR_b += p12_b * PmR;
    const double ML_b = M12_b * dMpL + PpL_b * dPpL;
    const double MR_b = M12_b * dMmR + PmR_b * dPmR;
    unR_b += MR_b / cf;
    cf_b -= MR_b * MR / cf;
    unL_b += ML_b / cf;
    cf_b -= ML_b * ML / cf;
    const double cl_b = bc ? 0.0 : cf_b, cr_b = bc ? cf_b : 0.0;
    double cL_b = cl_b * 2.0 * cL / dl;
    const double dl_b = -cl_b * cl / dl;
    if (bl) unL_b += dl_b;
    else cL_b += dl_b;
    double cR_b = cr_b * 2.0 * cR / dr;
    const double dr_b = -cr_b * cr / dr;
    if (br) unR_b -= dr_b;
    else cR_b += dr_b;
    V3 UL_b = (unL_b / magS) * S, UR_b = (unR_b / magS) * S;
    bS = bS + (unL_b / magS) * UL + (unR_b / magS) * UR;
    magS_b -= unL_b * unL / magS + unR_b * unR / magS;
    double HL_b = 0, HR_b = 0;
    if (up) { rL_b += rup_b; UL_b = UL_b + Uup_b; HL_b = Hup_b; }
    else { rR_b += rup_b; UR_b = UR_b + Uup_b; HR_b = Hup_b; }
    bS = bS + (magS_b / magS) * S;
#pragma unroll
    for (int k = 0; k < NFLD; ++k) bL[k] = bR[k] = 0.0;
    bL[F_UX] = UL_b.x; bL[F_UY] = UL_b.y; bL[F_UZ] = UL_b.z;
    bR[F_UX] = UR_b.x; bR[F_UY] = UR_b.y; bR[F_UZ] = UR_b.z;
    bL[F_CS] = cL_b; bR[F_CS] = cR_b;
    bL[F_P] = pL_b; bR[F_P] = pR_b;
    bL[F_RHO] = rL_b; bR[F_RHO] = rR_b;
    bL[F_H] = HL_b; bR[F_H] = HR_b;
}

// Face pass: adjoint of the face flux from v (b_owner gets -F, b_neighbour
// +F), AUSM+up reverse, reconstruction reverse into Cf.
template <int SCH>
__global__ void __launch_bounds__(256) k_adj_face(MeshView m, Phys ph, StateView s, AdjView a,
                                                  const double* __restrict__ v, double* __restrict__ gS,
                                                  double* __restrict__ gC, bool geom) {
    const int f = xcdBlock() * blockDim.x + threadIdx.x;
    if (f >= m.nF) return;
    double* lr = a.fLR + (size_t)f * 2 * NFLD;
    if (f >= m.nI && m.bKind[f - m.nI] == PK_EMPTY) {
#pragma unroll
        for (int k = 0; k < 2 * NFLD; ++k) lr[k] = 0.0;
        for (int d = 0; d < 3; ++d) { gS[3 * (size_t)f + d] = 0.0; gC[3 * (size_t)f + d] = 0.0; }
        if (a.fbr) a.fbr[f] = 0;
        return;
    }
    const size_t nC = m.nC;
    const int o = m.own[f];
    const bool internal = f < m.nI;
    const int n = internal ? m.nei[f] : -1;
    double bf[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) bf[q] = internal ? v[q * nC + n] - v[q * nC + o] : -v[q * nC + o];
    double L[NFLD], R[NFLD];
    if (internal) {
#pragma unroll
        for (int k = 0; k < NFLD; ++k)
            if (used(ph.flux, k)) recon(m, s, ph.recon, f, o, n, k, L[k], R[k]);
            else L[k] = R[k] = 0.0;
    } else {
        const double* bv = s.bfld + (size_t)(f - m.nI) * NFLD;
#pragma unroll
        for (int k = 0; k < NFLD; ++k) L[k] = R[k] = bv[k];
    }
    const V3 S = ld3(m.Sf + 3 * (size_t)f);
    double bL[NFLD], bR[NFLD];
    V3 bS;
    if constexpr (SCH == FLUX_HLLC) {
        hllcFaceRev(ph.gamma, S, L, R, bf[0], V3{bf[2], bf[3], bf[4]}, bf[1], bL, bR, bS,
                    a.fbr ? a.fbr + f : nullptr, geom);
        // on a boundary face both sides are the boundary state, where every
        // branch gives the physical flux: no switch of the derivative
        if (a.fbr && !internal) a.fbr[f] = 0;
    } else if constexpr (SCH == FLUX_ROE) {
        roeFaceRev(ph.gamma, ph.entropyFix, S, m.magSf[f], L, R, bf[0], V3{bf[2], bf[3], bf[4]}, bf[1], bL, bR, bS,
                   a.fbr ? a.fbr + f : nullptr, geom);
        // on a boundary face both sides are the boundary state: the
        // dissipation multiplies a zero jump, whose derivative is zero on
        // either side of every switch
        if (a.fbr && !internal) a.fbr[f] = 0;
    } else {
        ausmFaceAdj(ph.lowMach, S, m.magSf[f], L, R, bf[0], V3{bf[2], bf[3], bf[4]}, bf[1], bL, bR, bS,
                    a.fbr ? a.fbr + f : nullptr);
        // on a boundary face both sides are the boundary state: the choice
        // of the smaller c* (bit 2) and of the upwind side (bit 9) selects
        // between equal functions of the state, so it is no switch of the
        // derivative (on slip walls both follow the sign of a rounding-level
        // U.n)
        if (a.fbr && !internal) a.fbr[f] &= ~((1 << 2) | (1 << 9));
    }
    V3 bC{0, 0, 0};
    if (internal) {
#pragma unroll
        for (int k = 0; k < NFLD; ++k) { lr[k] = bL[k]; lr[NFLD + k] = bR[k]; }
        if (ph.recon == 1) {
#pragma unroll
            for (int k = 0; k < NFLD; ++k) {
                if (!used(ph.flux, k)) continue;
                const V3 go = ld3(s.grd + ((size_t)o * NFLD + k) * 3);
                const V3 gn = ld3(s.grd + ((size_t)n * NFLD + k) * 3);
                bC = bC + bL[k] * go + bR[k] * gn;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < NFLD; ++k) { lr[k] = bL[k] + bR[k]; lr[NFLD + k] = 0.0; }
    }
    gS[3 * (size_t)f] = bS.x; gS[3 * (size_t)f + 1] = bS.y; gS[3 * (size_t)f + 2] = bS.z;
    gC[3 * (size_t)f] = bC.x; gC[3 * (size_t)f + 1] = bC.y; gC[3 * (size_t)f + 2] = bC.z;
}

// Viscous face pass (viscous cases): adjoint of the face terms of
// residualsUpdate.H:16-43 - laplacian(muEff, U), div(tauMC),
// div(sigmaDotU), laplacian(alphaEff, e) - for the seed v, into a.vis.
// The terms add to b of the owner and subtract from b of the neighbour
// (the opposite sign of the convective flux).  They are linear in grad(U),
// U, e and grad(e) except div(sigmaDotU), bilinear in (grad(U), tau) and U;
// tau = mu dev2(grad(U)^T), dev2 being its own adjoint.
__global__ void __launch_bounds__(256) k_adj_visc(MeshView m, Phys ph, StateView s, AdjView a,
                                                  const double* __restrict__ v) {
    const int f = xcdBlock() * blockDim.x + threadIdx.x;
    if (f >= m.nF) return;
    double out[NVIS];
#pragma unroll
    for (int i = 0; i < NVIS; ++i) out[i] = 0.0;
    double* vs = a.vis + (size_t)f * NVIS;
    if (f >= m.nI && m.bKind[f - m.nI] == PK_EMPTY) {
        stRow(vs, out);
        return;
    }
    const size_t nC = m.nC;
    const double mu = ph.mu, alpha = ph.alpha;
    const V3 S = ld3(m.Sf + 3 * (size_t)f);
    const double ms = m.magSf[f];
    const int o = m.own[f];
    if (f < m.nI) {
        const int n = m.nei[f];
        const double w = m.w[f], nodc = m.nodc[f];
        const V3 k3 = ld3(m.corr + 3 * (size_t)f);
        const V3 lamU{v[2 * nC + o] - v[2 * nC + n], v[3 * nC + o] - v[3 * nC + n], v[4 * nC + o] - v[4 * nC + n]};
        const double lamE = v[nC + o] - v[nC + n];
        double go[NGV], gn[NGV];
        ldRow(s.gv + (size_t)o * NGV, go);
        ldRow(s.gv + (size_t)n * NGV, gn);
        T9 gUo, gUn;
#pragma unroll
        for (int i = 0; i < 9; ++i) { gUo.t[i] = go[GV_U + i]; gUn.t[i] = gn[GV_U + i]; }
        const T9 gUf = w * (gUo - gUn) + gUn;
        const T9 tauO = mu * dev2(transposeT(gUo)), tauN = mu * dev2(transposeT(gUn));
        const T9 tauf = w * (tauO - tauN) + tauN;
        const V3 Uo = cellU(s.fld, o), Un = cellU(s.fld, n);
        const V3 Uf = w * (Uo - Un) + Un;
        const double muf = w * (mu - mu) + mu, alf = w * (alpha - alpha) + alpha;
        // (muf sn) |S|, sn = nodc (Un - Uo) + k3 & gUf
        const V3 snb = (muf * ms) * lamU;
        V3 UoB = (-nodc) * snb, UnB = nodc * snb;
        T9 gUfB = outer(k3, snb);
        // Sf & tauf
        T9 taufB = outer(S, lamU);
        // ((muf gUf + tauf) & Uf) & Sf
        const T9 M = muf * gUf + tauf;
        const T9 MB = lamE * outer(S, Uf);
        const V3 UfB = lamE * vdott(S, M);
        gUfB = gUfB + muf * MB;
        taufB = taufB + MB;
        UoB = UoB + w * UfB;
        UnB = UnB + (1.0 - w) * UfB;
        // tauf interpolates mu dev2(grad(U)^T) with gUf's weights: one
        // adjoint Z of the interpolated gradient, split w / (1 - w)
        const T9 Z = gUfB + transposeT(dev2(mu * taufB));
        // (alf sne) |S|, sne = nodc (en - eo) + k3 . gef
        const double sneB = lamE * alf * ms;
#pragma unroll
        for (int i = 0; i < 9; ++i) out[VS_Z + i] = Z.t[i];
        out[VS_GE] = sneB * k3.x; out[VS_GE + 1] = sneB * k3.y; out[VS_GE + 2] = sneB * k3.z;
        out[VS_UO] = UoB.x; out[VS_UO + 1] = UoB.y; out[VS_UO + 2] = UoB.z;
        out[VS_UN] = UnB.x; out[VS_UN + 1] = UnB.y; out[VS_UN + 2] = UnB.z;
        out[VS_EO] = -nodc * sneB;
        out[VS_EN] = nodc * sneB;
    } else {
        const int b = f - m.nI;
        const V3 lamU{v[2 * nC + o], v[3 * nC + o], v[4 * nC + o]};
        const double lamE = v[nC + o];
        const V3 nn = S / ms;
        const double dc = m.dc[f];
        const double* bf = s.bfld + (size_t)b * NFLD;
        const V3 UB{bf[F_UX], bf[F_UY], bf[F_UZ]};
        const V3 Uc = cellU(s.fld, o);
        double g[NGV];
        ldRow(s.gv + (size_t)o * NGV, g);
        T9 gUc;
#pragma unroll
        for (int i = 0; i < 9; ++i) gUc.t[i] = g[GV_U + i];
        const int ut = m.bUT[b];
        V3 snU;
        if (ut == BC_ZG) snU = V3{0, 0, 0};
        else if (ut == BC_SLIP) snU = (reflect(nn, Uc) - Uc) * (dc / 2.0);
        else snU = dc * (UB - Uc);
        const T9 gUb = gUc + outer(nn, snU - vdott(nn, gUc));
        const T9 taub = mu * dev2(transposeT(gUb));
        // (mu snU) |S|
        V3 snUB = (mu * ms) * lamU;
        // Sf & taub and ((mu gUb + taub) & UB) & Sf
        const T9 M = mu * gUb + taub;
        const T9 MB = lamE * outer(S, UB);
        const V3 UBB = lamE * vdott(S, M);
        const T9 gUbB = mu * MB + transposeT(dev2(mu * (outer(S, lamU) + MB)));
        // gUb = gUc + n (snU - n & gUc)
        const V3 XB = vdott(nn, gUbB);
        snUB = snUB + XB;
        const T9 gUcB = gUbB - outer(nn, XB);
        V3 UcB{0, 0, 0};
        if (ut == BC_SLIP) UcB = (reflect(nn, snUB) - snUB) * (dc / 2.0);
        else if (ut != BC_ZG) UcB = (-dc) * snUB;  // fixedValue U: U_b is constant
        // (alpha sne) |S|, sne = dc (e_b - e_P)
        const double sneB = lamE * alpha * ms;
#pragma unroll
        for (int i = 0; i < 9; ++i) out[VS_Z + i] = gUcB.t[i];
        out[VS_UO] = UcB.x; out[VS_UO + 1] = UcB.y; out[VS_UO + 2] = UcB.z;
        out[VS_UN] = UBB.x; out[VS_UN + 1] = UBB.y; out[VS_UN + 2] = UBB.z;
        out[VS_EO] = -dc * sneB;
        out[VS_EN] = dc * sneB;
    }
    stRow(vs, out);
}

// Cell pass: the adjoint of the limited gradients and of the cell values
// (linearUpwind), then the limiter reverse (recomputing each field's
// max/min neighbour and the face that set its limiter), then the
// Gauss-gradient adjoint divided by V.  All fields are carried together,
// so the cell's faces are walked twice (once for the face adjoints and the
// max/min, once for the limiter faces) whatever the number of fields; per
// field the operations and their order are those of the forward kernels.
template <bool VISC>
__global__ void __launch_bounds__(256) k_adj_cell(MeshView m, Phys ph, StateView s, AdjView a) {
    const int c = xcdBlock() * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const int e0 = m.cfOff[c], e1 = m.cfOff[c + 1];
    const V3 Cc = ld3(m.C + 3 * (size_t)c);
    const double Vc = m.V[c];
    const bool rec = ph.recon == 1;
    const bool limit = rec && ph.limiter != 0;
    // records pass (a.fbr set): the limiter's active face and max/min
    // element are found for every field, whatever the adjoint seed
    const bool records = a.fbr != nullptr;
    double qd[NFLD], mx[NFLD], mn[NFLD], vcr[NFLD];
    V3 gl[NFLD];
    int imx[NFLD], imn[NFLD];
    ldRow(s.fld + (size_t)c * NFLD, vcr);
#pragma unroll
    for (int k = 0; k < NFLD; ++k) {
        qd[k] = 0.0;
        gl[k] = V3{0, 0, 0};
        mx[k] = mn[k] = vcr[k];
        imx[k] = imn[k] = -1;
    }
    // viscous: adjoint of the cell's Gauss grad(U) (zu, row-major) and
    // grad(e) sums from the viscous face pass
    constexpr bool visc = VISC;
    double zu[9];
    V3 ges{0, 0, 0};
#pragma unroll
    for (int i = 0; i < 9; ++i) zu[i] = 0.0;
    // internal faces first, in face order = the cell's slot order (column,
    // orientation and Cf from the slots), then the boundary faces
    const int len = m.rowLen[c], nLo = m.nLower[c];
    const size_t rslot = (size_t)m.slicePtr[c >> 6] * 64 + (c & 63);
    for (int q = e0; q < e1; ++q) {
        const int f = m.cf[q];
        const int j = q - e0;
        const size_t sk = rslot + 64 * (size_t)j;
        const V3 d = (j < len ? V3{m.sgeo[sgOff(sk, SG_CX)], m.sgeo[sgOff(sk, SG_CX + 1)],
                                   m.sgeo[sgOff(sk, SG_CX + 2)]}
                              : ld3(m.Cf + 3 * (size_t)f)) - Cc;
        const double* vo;
        int id;
        if (visc) {
            // grad(U) and grad(e) sum adjoints: the first 12 entries of the row
            double vz[VS_GE + 3];
            ldRow(a.vis + (size_t)f * NVIS, vz);
            const double wt = f < m.nI ? (j >= nLo ? m.w[f] : 1.0 - m.w[f]) : 1.0;
#pragma unroll
            for (int i = 0; i < 9; ++i) zu[i] += wt * vz[VS_Z + i];
            ges = ges + wt * V3{vz[VS_GE], vz[VS_GE + 1], vz[VS_GE + 2]};
        }
        if (f < m.nI) {
            const bool isOwn = j >= nLo;
            double lr[NFLD];
            ldRow(a.fLR + (size_t)f * 2 * NFLD + (isOwn ? 0 : NFLD), lr);
#pragma unroll
            for (int k = 0; k < NFLD; ++k) {
                if (!used(ph.flux, k)) continue;
                const double av = lr[k];
                qd[k] += av;
                if (rec) gl[k] = gl[k] + av * d;
            }
            id = m.col[sk];
            vo = s.fld + (size_t)id * NFLD;
        } else {
            const int b = f - m.nI;
            id = m.nC + b;
            vo = s.bfld + (size_t)b * NFLD;
        }
        if (limit) {
            double vr[NFLD];
            ldRow(vo, vr);
#pragma unroll
            for (int k = 0; k < NFLD; ++k) {
                if (!used(ph.flux, k)) continue;
                const double v = vr[k];
                // records pass: a candidate within rounding of the current
                // max/min does not displace it.  Such ties are two equal
                // functions of the state computed two ways (H of a cell and
                // of its zeroGradient outlet face: E + p/rho with p/rho = R T
                // on both sides), whose order follows the rounding, so that
                // any step flips them while the derivative does not change.
                const double tx = records ? 1e-12 * fabs(mx[k]) : 0.0;
                const double tn = records ? 1e-12 * fabs(mn[k]) : 0.0;
                if (mx[k] + tx < v) { mx[k] = v; imx[k] = id; }
                if (v < mn[k] - tn) { mn[k] = v; imn[k] = id; }
            }
        }
    }
    V3 g[NFLD], gb[NFLD];
    double limb[NFLD];
    bool any = false;
    double gunr[3 * NFLD], limr[NFLD];
    if (rec) {
        ldRow(a.gun + (size_t)c * NFLD * 3, gunr);
        ldRow(a.lim + (size_t)c * NFLD, limr);
    }
#pragma unroll
    for (int k = 0; k < NFLD; ++k) {
        g[k] = gb[k] = V3{0, 0, 0};
        limb[k] = 0.0;
        if (!used(ph.flux, k) || !rec) continue;
        g[k] = V3{gunr[3 * k], gunr[3 * k + 1], gunr[3 * k + 2]};
        gb[k] = limr[k] * gl[k];
        limb[k] = dot(gl[k], g[k]);
        if (limit && (limb[k] != 0.0 || records)) any = true;
    }
    int fs[NFLD], br[NFLD];
    double lm[NFLD], rs[NFLD], es[NFLD], mxD[NFLD], mnD[NFLD];
#pragma unroll
    for (int k = 0; k < NFLD; ++k) {
        fs[k] = -1; br[k] = 0; lm[k] = 1.0; rs[k] = 1.0; es[k] = 0.0;
        mxD[k] = mx[k] - vcr[k];
        mnD[k] = mn[k] - vcr[k];
    }
    if (any) {
        for (int q = e0; q < e1; ++q) {
            const int f = m.cf[q];
            const V3 d = ld3(m.Cf + 3 * (size_t)f) - Cc;
#pragma unroll
            for (int k = 0; k < NFLD; ++k) {
                if (!used(ph.flux, k) || (limb[k] == 0.0 && !records)) continue;
                const double ext = dot(d, g[k]);
                double r = 1;
                int bq = 0;
                if (ext > kSMALL) { r = mxD[k] / ext; bq = 1; }
                else if (ext < -kSMALL) { r = mnD[k] / ext; bq = -1; }
                double l;
                if (ph.limiter == 1) l = (sqr(r) + 2 * r) / (sqr(r) + r + 2);
                else l = fmin(r, 1.0);
                if (l < lm[k]) { lm[k] = l; fs[k] = f; rs[k] = r; es[k] = ext; br[k] = bq; }
            }
        }
    }
    static_assert(NFLD % 2 == 0, "fields stored in pairs");
    double pq = 0.0, pv = 0.0, px = 0.0;
    int pt = -1, pf = -1;
    V3 pG{0, 0, 0};
#pragma unroll
    for (int k = 0; k < NFLD; ++k) {
        const size_t ck = (size_t)c * NFLD + k;
        int tgt = -1, fsk = -1;
        double tv = 0.0, fx = 0.0, qk = qd[k];
        V3 G{0, 0, 0};
        // viscous: grad(U_j) is the unlimited Gauss gradient of field U_j
        const bool vk = visc && k >= F_UX && k <= F_UZ;
        if (vk && !rec) G = V3{zu[k - F_UX], zu[3 + k - F_UX], zu[6 + k - F_UX]} / Vc;
        if (used(ph.flux, k) && rec) {
            V3 gbk = gb[k];
            if (vk) gbk = gbk + V3{zu[k - F_UX], zu[3 + k - F_UX], zu[6 + k - F_UX]};
            if (limit && (limb[k] != 0.0 || records) && fs[k] >= 0 && br[k] != 0) {
                double dl;
                if (ph.limiter == 1) {
                    const double den = sqr(rs[k]) + rs[k] + 2;
                    dl = (-sqr(rs[k]) + 4 * rs[k] + 4) / sqr(den);
                } else dl = rs[k] <= 1.0 ? 1.0 : 0.0;
                const double rb = limb[k] * dl;
                const double numb = rb / es[k];
                fx = -rb * rs[k] / es[k];
                fsk = fs[k];
                gbk = gbk + fx * (ld3(m.Cf + 3 * (size_t)fsk) - Cc);
                const int id = br[k] > 0 ? imx[k] : imn[k];
                if (id >= 0) {
                    tgt = id;
                    tv = numb;
                    qk -= numb;
                }
            }
            G = gbk / Vc;
        }
        // stored a pair of fields at a time (128-bit stores; 64-bit for
        // the int rows): the even field is held until the odd one is done
        const double qv = used(ph.flux, k) ? qk : 0.0;
        if (k & 1) {
            const size_t c0 = ck - 1;
            *reinterpret_cast<double2*>(a.qc + c0) = double2{pq, qv};
            *reinterpret_cast<int2*>(a.tgt + c0) = int2{pt, tgt};
            *reinterpret_cast<double2*>(a.tval + c0) = double2{pv, tv};
            *reinterpret_cast<int2*>(a.fst + c0) = int2{pf, fsk};
            *reinterpret_cast<double2*>(a.fext + c0) = double2{px, fx};
            double2* gq = reinterpret_cast<double2*>(a.G + 3 * c0);
            gq[0] = double2{pG.x, pG.y};
            gq[1] = double2{pG.z, G.x};
            gq[2] = double2{G.y, G.z};
        } else {
            pq = qv; pt = tgt; pv = tv; pf = fsk; px = fx; pG = G;
        }
    }
    if (visc) {
        const V3 Ge = ges / Vc;
        a.Ge[3 * (size_t)c] = Ge.x; a.Ge[3 * (size_t)c + 1] = Ge.y; a.Ge[3 * (size_t)c + 2] = Ge.z;
    }
}

// Face pass 2: Gauss-gradient face values (internal faces: adjoint of
// w*(q_o - q_n) + q_n for the cell pull, Sf part), boundary values
// (flux + Gauss + limiter max/min) chained through rho_b = p_b/(R T_b),
// E_b, H_b, c*_b to the owner's p, T, U and to Sf (slip normal), and the
// limiter's Cf part.
template <bool VISC>
__global__ void __launch_bounds__(256) k_adj_gauss(MeshView m, Phys ph, StateView s, AdjView a,
                                                   double* __restrict__ gS, double* __restrict__ gC) {
    const int f = xcdBlock() * blockDim.x + threadIdx.x;
    if (f >= m.nF) return;
    double* qf = a.qf + (size_t)f * NFLD;
    if (f >= m.nI && m.bKind[f - m.nI] == PK_EMPTY) {
#pragma unroll
        for (int k = 0; k < NFLD; ++k) qf[k] = 0.0;
        double* bp = a.bprim + (size_t)(f - m.nI) * 5;
        for (int q = 0; q < 5; ++q) bp[q] = 0.0;
        return;
    }
    const V3 S = ld3(m.Sf + 3 * (size_t)f);
    const int o = m.own[f];
    V3 bS{0, 0, 0}, bC{0, 0, 0};
    const bool rec = ph.recon == 1;
    constexpr bool visc = VISC;
    // fields whose Gauss gradient carries an adjoint: the reconstructed ones
    // (second order) and, viscous, the velocity components
    const auto grads = [&](int k) { return (used(ph.flux, k) && rec) || (visc && k >= F_UX && k <= F_UZ); };
    if (f < m.nI) {
        const int n = m.nei[f];
        const double w = m.w[f];
        if (visc) a.qe[f] = dot(S, ld3(a.Ge + 3 * (size_t)o) - ld3(a.Ge + 3 * (size_t)n));
        // both cells' rows: field values, the limiter faces, and the
        // gradient adjoints two fields (six doubles) at a time
        double fo[NFLD], fn[NFLD], qv[NFLD];
        int so[NFLD], sn[NFLD];
        ldRow(s.fld + (size_t)o * NFLD, fo);
        ldRow(s.fld + (size_t)n * NFLD, fn);
        ldRowI(a.fst + (size_t)o * NFLD, so);
        ldRowI(a.fst + (size_t)n * NFLD, sn);
#pragma unroll
        for (int kp = 0; kp < NFLD; kp += 2) {
            double Go[6], Gn[6];
            if (grads(kp) || grads(kp + 1)) {
                ldRow(a.G + 3 * ((size_t)o * NFLD + kp), Go);
                ldRow(a.G + 3 * ((size_t)n * NFLD + kp), Gn);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k = kp + h;
                if (!grads(k)) { qv[k] = 0.0; continue; }
                const size_t ok = (size_t)o * NFLD + k, nk = (size_t)n * NFLD + k;
                const V3 d = V3{Go[3 * h], Go[3 * h + 1], Go[3 * h + 2]} - V3{Gn[3 * h], Gn[3 * h + 1], Gn[3 * h + 2]};
                qv[k] = dot(S, d);
                const double vo = fo[k], vn = fn[k];
                bS = bS + (w * (vo - vn) + vn) * d;
                if (so[k] == f) bC = bC + a.fext[ok] * ld3(a.gun + 3 * ok);
                if (sn[k] == f) bC = bC + a.fext[nk] * ld3(a.gun + 3 * nk);
            }
        }
        stRow(qf, qv);
    } else {
        const int b = f - m.nI;
        const double* bv = s.bfld + (size_t)b * NFLD;
        const double* lr = a.fLR + (size_t)f * 2 * NFLD;
        double qb[NFLD];
#pragma unroll
        for (int k = 0; k < NFLD; ++k) {
            qf[k] = 0.0;
            qb[k] = 0.0;
            if (used(ph.flux, k)) qb[k] = lr[k];
            if (grads(k)) {
                const size_t ok = (size_t)o * NFLD + k;
                const V3 Go = ld3(a.G + 3 * ok);
                qb[k] += dot(S, Go);
                if (a.tgt[ok] == m.nC + b) qb[k] += a.tval[ok];
                bS = bS + bv[k] * Go;
                if (a.fst[ok] == f) bC = bC + a.fext[ok] * ld3(a.gun + 3 * ok);
            }
        }
        // c*_b = sqrt(kappa H_b), H_b = max(E_b, SMALL) + max(p_b/rho_b, SMALL)
        const double g = ph.gamma;
        const double kap = 2.0 * (g - 1.0) / (g + 1.0);
        double Hb = qb[F_H];
        const bool hllc = ph.flux == FLUX_HLLC;
        if (!hllc && bv[F_CS] > 0.0) Hb += qb[F_CS] * kap / (2.0 * bv[F_CS]);
        const double TB = s.bT[b];
        const double pB = bv[F_P], rhoB = bv[F_RHO];
        const V3 UB{bv[F_UX], bv[F_UY], bv[F_UZ]};
        const double EB = ph.Cv * TB + 0.5 * magSqr(UB);
        double Eb = qb[F_E], pb = qb[F_P], rb = qb[F_RHO];
        V3 Ub{qb[F_UX], qb[F_UY], qb[F_UZ]};
        if (!(EB < kSMALL)) Eb += Hb;
        const double pr = pB / rhoB;
        if (!(pr < kSMALL)) { pb += Hb / rhoB; rb -= Hb * pr / rhoB; }
        double Tb = ph.Cv * Eb;
        // HLLC: c_b = max(sqrt(gamma/psi_b), VSMALL), sqrt(gamma R T_b)
        if (hllc && !(bv[F_CS] < kVSMALL)) Tb += qb[F_CS] * g * ph.R / (2.0 * bv[F_CS]);
        Ub = Ub + Eb * UB;
        if (visc) {
            // the viscous terms' boundary U, and the boundary e of grad(e)
            // and of laplacian(alphaEff, e): e_b = rhoE_b / rho_b - |U_b|^2 / 2
            // with rhoE_b = rho_b (Cv T_b + |U_b|^2 / 2), i.e. Cv T_b
            const double* vs = a.vis + (size_t)f * NVIS;
            Ub = Ub + V3{vs[VS_UN], vs[VS_UN + 1], vs[VS_UN + 2]};
            Tb += ph.Cv * (vs[VS_EN] + dot(S, ld3(a.Ge + 3 * (size_t)o)));
        }
        // rho_b = psi p_b, psi = 1/(R T_b)
        const double psi = 1.0 / (ph.R * TB);
        pb += rb * psi;
        Tb -= rb * pB * psi / TB;
        double* bp = a.bprim + (size_t)b * 5;
        bp[0] = m.bPT[b] == BC_FIXED ? 0.0 : pb;
        bp[1] = m.bTT[b] == BC_FIXED ? 0.0 : Tb;
        V3 Uc_b{0, 0, 0};
        const int ut = m.bUT[b];
        if (ut == BC_ZG) Uc_b = Ub;
        else if (ut == BC_SLIP) {
            // U_b = U_c - n (n.U_c), n = Sf/|Sf|
            const double ms = m.magSf[f];
            const V3 nn = S / ms;
            const V3 Uc = cellU(s.fld, o);
            const double nUb = dot(nn, Ub), nUc = dot(nn, Uc);
            Uc_b = Ub - nUb * nn;
            const V3 nb = (-nUc) * Ub - nUb * Uc;
            bS = bS + (nb - dot(nn, nb) * nn) / ms;
        }
        bp[2] = Uc_b.x; bp[3] = Uc_b.y; bp[4] = Uc_b.z;
    }
    gS[3 * (size_t)f] += bS.x; gS[3 * (size_t)f + 1] += bS.y; gS[3 * (size_t)f + 2] += bS.z;
    gC[3 * (size_t)f] += bC.x; gC[3 * (size_t)f + 1] += bC.y; gC[3 * (size_t)f + 2] += bC.z;
}

// Cell pass 2: total cell-value adjoints (direct + Gauss face values +
// limiter targets of the neighbours + boundary primitives), then
// c*, H, E -> rho, p, U, T and the updateFields chain back to W.
template <bool VISC>
__global__ void __launch_bounds__(256) k_adj_state(MeshView m, Phys ph, StateView s, AdjView a,
                                                   double* __restrict__ gW) {
    const int c = xcdBlock() * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const int e0 = m.cfOff[c], e1 = m.cfOff[c + 1];
    double q[NFLD];
    ldRow(a.qc + (size_t)c * NFLD, q);
    double pbnd = 0.0, Tbnd = 0.0;
    V3 Ubnd{0, 0, 0};
    // viscous: direct U and e adjoints of the viscous face terms, and the
    // Gauss face values of e
    constexpr bool visc = VISC;
    V3 Uv{0, 0, 0};
    double ecB = 0.0;
    const int nLo = m.nLower[c];
    const size_t rslot = (size_t)m.slicePtr[c >> 6] * 64 + (c & 63);
    for (int e = e0; e < e1; ++e) {
        const int f = m.cf[e];
        if (f < m.nI) {
            // internal faces: the cell's slots in face order
            const size_t sk = rslot + 64 * (size_t)(e - e0);
            const bool isOwn = e - e0 >= nLo;
            const int j = m.col[sk];
            const double wf = m.sgeo[sgOff(sk, SG_W)];
            const double wt = isOwn ? wf : 1.0 - wf;
            double qf[NFLD];
            int tg[NFLD];
            ldRow(a.qf + (size_t)f * NFLD, qf);
            ldRowI(a.tgt + (size_t)j * NFLD, tg);
#pragma unroll
            for (int k = 0; k < NFLD; ++k) {
                if (!used(ph.flux, k) && !(visc && k >= F_UX && k <= F_UZ)) continue;
                q[k] += wt * qf[k];
                if (tg[k] == c) q[k] += a.tval[(size_t)j * NFLD + k];
            }
            if (visc) {
                // the state adjoints of the viscous face terms: entries
                // VS_UO .. VS_EN of the row
                double vu[NVIS - VS_UO];
                ldRow(a.vis + (size_t)f * NVIS + VS_UO, vu);
                const int u0 = (isOwn ? VS_UO : VS_UN) - VS_UO;
                Uv = Uv + V3{vu[u0], vu[u0 + 1], vu[u0 + 2]};
                ecB += (isOwn ? vu[VS_EO - VS_UO] : vu[VS_EN - VS_UO]) + wt * a.qe[f];
            }
        } else {
            const double* bp = a.bprim + (size_t)(f - m.nI) * 5;
            pbnd += bp[0];
            Tbnd += bp[1];
            Ubnd = Ubnd + V3{bp[2], bp[3], bp[4]};
            if (visc) {
                const double* vs = a.vis + (size_t)f * NVIS;
                Uv = Uv + V3{vs[VS_UO], vs[VS_UO + 1], vs[VS_UO + 2]};
                ecB += vs[VS_EO];
            }
        }
    }
    const size_t nC = m.nC;
    const double* fl = s.fld + (size_t)c * NFLD;
    const double rho = fl[F_RHO], p = fl[F_P], H = fl[F_H], cs = fl[F_CS], E = fl[F_E];
    const V3 U{fl[F_UX], fl[F_UY], fl[F_UZ]};
    const double T = s.T[c];
    const double g = ph.gamma;
    const double kap = 2.0 * (g - 1.0) / (g + 1.0);
    double Hb = q[F_H];
    const bool hllc = ph.flux == FLUX_HLLC;
    if (!hllc && cs > 0.0) Hb += q[F_CS] * kap / (2.0 * cs);
    double rb = q[F_RHO], pb = q[F_P] + pbnd, Eb = q[F_E];
    V3 Ub = V3{q[F_UX], q[F_UY], q[F_UZ]} + Ubnd + Uv;
    if (!(E < kSMALL)) Eb += Hb;
    const double pr = p / rho;
    if (!(pr < kSMALL)) { pb += Hb / rho; rb -= Hb * pr / rho; }
    double Tb = ph.Cv * Eb + Tbnd;
    // HLLC: c = max(sqrt(gamma/psi), VSMALL), sqrt(gamma R T)
    if (hllc && !(cs < kVSMALL)) Tb += q[F_CS] * g * ph.R / (2.0 * cs);
    Ub = Ub + Eb * U;
    // p = rho/psi = rho R T
    const double psi = 1.0 / (ph.R * T);
    rb += pb / psi;
    Tb += pb * p / T;
    // updateFields: rho = max(W0, rhoMin), U = rhoU/rho,
    // e = clip(rhoE/rho - |U|^2/2), T = e/Cv
    const double W0 = a.W0[c], W1 = a.W0[nC + c];
    const V3 rU{a.W0[2 * nC + c], a.W0[3 * nC + c], a.W0[4 * nC + c]};
    const bool rhoFree = !(W0 < ph.rhoMin);
    const double r0 = rhoFree ? W0 : ph.rhoMin;
    const V3 U0 = rU / r0;
    const double eU = W1 / r0 - 0.5 * magSqr(U0);
    bool eFree = !(eU < ph.Cv * ph.TMin);
    if (ph.TMax < kGREAT && ph.Cv * ph.TMax < eU) eFree = false;
    // e_calc = rhoE/rho - |U|^2/2 with rhoE rebuilt from the clipped e: e
    const double eb = eFree ? Tb / ph.Cv + ecB : 0.0;
    double rEb = eb / r0;
    rb -= eb * (W1 / r0) / r0;
    Ub = Ub - eb * U0;
    const V3 rUb = Ub / r0;
    rb -= dot(Ub, U0) / r0;
    gW[c] = rhoFree ? rb : 0.0;
    gW[nC + c] = rEb;
    gW[2 * nC + c] = rUb.x;
    gW[3 * nC + c] = rUb.y;
    gW[4 * nC + c] = rUb.z;
}

#undef DF

}  // namespace

void launchResidualVjp(const MeshView& m, const Phys& ph, const StateView& s, const AdjView& a, const double* v,
                       double* gW, double* gS, double* gC, bool geom, hipStream_t st) {
    const dim3 gf(xcdGrid(m.nF, 256)), gc(xcdGrid(m.nC, 256));
    if (ph.flux == FLUX_ROE) k_adj_face<FLUX_ROE><<<gf, 256, 0, st>>>(m, ph, s, a, v, gS, gC, geom);
    else if (ph.flux == FLUX_HLLC) k_adj_face<FLUX_HLLC><<<gf, 256, 0, st>>>(m, ph, s, a, v, gS, gC, geom);
    else k_adj_face<FLUX_AUSM><<<gf, 256, 0, st>>>(m, ph, s, a, v, gS, gC, geom);
    if (ph.viscous) {
        k_adj_visc<<<gf, 256, 0, st>>>(m, ph, s, a, v);
        k_adj_cell<true><<<gc, 256, 0, st>>>(m, ph, s, a);
        k_adj_gauss<true><<<gf, 256, 0, st>>>(m, ph, s, a, gS, gC);
        k_adj_state<true><<<gc, 256, 0, st>>>(m, ph, s, a, gW);
    } else {
        k_adj_cell<false><<<gc, 256, 0, st>>>(m, ph, s, a);
        k_adj_gauss<false><<<gf, 256, 0, st>>>(m, ph, s, a, gS, gC);
        k_adj_state<false><<<gc, 256, 0, st>>>(m, ph, s, a, gW);
    }
}

}  // namespace icsf
