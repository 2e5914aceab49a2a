// Face-flux functions of the convective schemes, device side.
//   AUSM+up : ausmPlusUpFluxScheme::calcFlux (ausmPlusUpFluxScheme.C:73-299)
//   Roe     : roeFluxScheme::calcFlux + getRoeDissipation (roeFluxScheme.C:41-241,276-409)
// Expressions keep the reference's association order.
#pragma once
#include "icsf_math.h"

namespace icsf {

#define DF __device__ __forceinline__

// um: the face's normal frame velocity of a rotating zone (0 outside one):
// AUSM+up transports with the normal velocities relative to the frame and
// adds p_1/2 um |S| to the energy flux (ausmPlusUpFluxScheme.C:101-102,293).
DF void ausmFace(bool lowMach, V3 S, double magS, V3 UL, V3 UR, double cL, double cR, double pL, double pR,
                 double rL, double rR, double HL, double HR, double um, double& phi, V3& phiU, double& phiE) {
    const double unL = dot(UL, S) / magS - um;
    const double unR = dot(UR, S) / magS - um;
    const double cl = sqr(cL) / fmax(cL, unL);
    const double cr = sqr(cR) / fmax(cR, -unR);
    const double cf = fmin(cl, cr);
    const double ML = unL / cf;
    double MpL, PpL;
    if (fabs(ML) < 1.0) {
        const double p2 = 0.25 * sqr(ML + 1), m2 = -0.25 * sqr(ML - 1);
        MpL = p2 * (1 - 2 * m2);
        PpL = p2 * (2 - ML - 3 * ML * m2);
    } else {
        MpL = fmax(ML, 0.0);
        PpL = ML > 0 ? 1.0 : 0.0;
    }
    const double MR = unR / cf;
    double MmR, PmR;
    if (fabs(MR) < 1.0) {
        const double m2 = -0.25 * sqr(MR - 1), p2 = 0.25 * sqr(MR + 1);
        MmR = m2 * (1 + 2 * p2);
        PmR = m2 * (-2 - MR + 3 * MR * p2);
    } else {
        MmR = fmin(MR, 0.0);
        PmR = MR < 0 ? 1.0 : 0.0;
    }
    double M12 = MpL + MmR;
    double p12 = PpL * pL + PmR * pR;
    const double Mmean = 0.5 * (sqr(unL) + sqr(unR)) / sqr(cf);
    const double MDiff = -0.25 * fmax(1.0 - Mmean, 0.0) * (pR - pL) / (0.5 * (rL + rR) * sqr(cf));
    if ((M12 > 0.0 && M12 + MDiff <= 0.0) || (M12 < 0.0 && M12 + MDiff >= 0.0)) M12 += 0.2 * MDiff;
    else M12 += MDiff;
    if (lowMach) p12 += -0.25 * PpL * PmR * (rL + rR) * cf * (unR - unL);
    const bool up = M12 >= 0;
    const V3 Uf = up ? UL : UR;
    const double rhoa = M12 * cf * (up ? rL : rR);
    const V3 rhoaU = rhoa * Uf;
    const double rhoah = rhoa * (up ? HL : HR);
    phi = rhoa * magS;
    phiU = rhoaU * magS + p12 * S;
    phiE = rhoah * magS + p12 * um * magS;
}

// Roe flux on one face; the P|Lambda|P^-1 dissipation matrix is formed
// explicitly exactly as the reference does and applied to the jump of the
// reconstructed conserved state.
// um: the face's normal frame velocity: the dissipation takes the
// Roe-averaged normal velocity relative to the frame, and -1/2 |S| um
// (Q_l + Q_r) follows the central flux (roeFluxScheme.C:361-363,404-408).
DF void roeFace(double gI, double eFix, V3 S, double magS, double rl, double rr, double pl, double pr, V3 Ul,
                V3 Ur, double El, double Er, double Hl, double Hr, double um, double& phi, V3& phiU, double& phiE) {
    const V3 n = S / magS;
    const double coefR = sqrt(fmax(kVSMALL, rr) / fmax(kVSMALL, rl));
    const double rho = coefR * rl;
    const V3 Vel = (coefR * Ur + Ul) / (coefR + 1.0);
    const double HRoe = (coefR * Hr + Hl) / (coefR + 1.0);
    const double c = sqrt(fabs((gI - 1.0) * (HRoe - 0.5 * magSqr(Vel))));
    const double uMl = dot(Ul, n), uMr = dot(Ur, n);
    const double u = dot(Vel, n) - um;
    const V3 rUl = rl * Ul, rUr = rr * Ur;
    const double rEl = rl * El, rEr = rr * Er;

    const double a1 = gI - 1;
    const double theta = 0.5 * a1 * magSqr(Vel);
    const double c2 = sqr(c);
    const double a2 = 1 / (rho * c * sqrt(2.0));
    const double a3 = rho / (c * sqrt(2.0));
    const double a4 = (theta + c2) / a1;
    const double a5 = 1 - theta / c2;
    const double a6 = theta / a1;
    const V3 iP11 = (n * a5) - cross(Vel, n) / rho;
    T9 iP12 = outer(a1 / c2 * n, Vel);
    iP12.t[1] += n.z / rho; iP12.t[2] -= n.y / rho; iP12.t[3] -= n.z / rho;
    iP12.t[5] += n.x / rho; iP12.t[6] += n.y / rho; iP12.t[7] -= n.x / rho;
    const V3 iP13 = (-a1) / c2 * n;
    const double iP21 = a2 * (theta - c * u);
    const V3 iP22 = (-a2) * (a1 * Vel - c * n);
    const double iP23 = a1 * a2;
    const double iP31 = a2 * (theta + c * u);
    const V3 iP32 = (-a2) * (a1 * Vel + c * n);
    const double iP33 = a1 * a2;
    double L1 = fabs(u), L2 = fabs(u + c), L3 = fabs(u - c);
    const double eps = eFix * fmax(L2, L3);
    if (L1 < eps) L1 = (sqr(L1) + sqr(eps)) / (2.0 * eps);
    if (L2 < eps) L2 = (sqr(L2) + sqr(eps)) / (2.0 * eps);
    if (L3 < eps) L3 = (sqr(L3) + sqr(eps)) / (2.0 * eps);
    const double P12 = a3, P13 = a3;
    T9 P21 = outer(Vel, n);
    P21.t[1] -= n.z * rho; P21.t[2] += n.y * rho; P21.t[3] += n.z * rho;
    P21.t[5] -= n.x * rho; P21.t[6] -= n.y * rho; P21.t[7] += n.x * rho;
    const V3 P22 = a3 * (Vel + c * n);
    const V3 P23 = a3 * (Vel - c * n);
    const V3 P31 = n * a6 + rho * cross(Vel, n);
    const double P32 = a3 * (a4 + c * u);
    const double P33 = a3 * (a4 - c * u);
    const V3 P11L = n * L1;
    const T9 P21L = P21 * L1;
    const V3 P31L = P31 * L1;
    const double CR = dot(P11L, iP11) + L2 * P12 * iP21 + L3 * P13 * iP31;
    const V3 CU = vdott(P11L, iP12) + (L2 * P12) * iP22 + (L3 * P13) * iP32;
    const double CE = dot(P11L, iP13) + L2 * P12 * iP23 + L3 * P13 * iP33;
    const V3 MRv = tdotv(P21L, iP11) + (L2 * P22) * iP21 + (L3 * P23) * iP31;
    const T9 MU = tdott(P21L, iP12) + outer(L2 * P22, iP22) + outer(L3 * P23, iP32);
    const V3 ME = tdotv(P21L, iP13) + (L2 * P22) * iP23 + (L3 * P23) * iP33;
    const double ER = dot(P31L, iP11) + L2 * P32 * iP21 + L3 * P33 * iP31;
    const V3 EU = vdott(P31L, iP12) + (L2 * P32) * iP22 + (L3 * P33) * iP32;
    const double EE = dot(P31L, iP13) + L2 * P32 * iP23 + L3 * P33 * iP33;

    const double dR = rr - rl;
    const V3 dRU = rUr - rUl;
    const double dRE = rEr - rEl;
    phi = -0.5 * magS * (CR * dR + dot(CU, dRU) + CE * dRE);
    phiU = -0.5 * magS * (MRv * dR + tdotv(MU, dRU) + ME * dRE);
    phiE = -0.5 * magS * (ER * dR + dot(EU, dRU) + EE * dRE);
    const double rnl = rl * uMl, rnr = rr * uMr;
    phi += 0.5 * magS * (rnl + rnr);
    phiU = phiU + 0.5 * magS * (rnl * Ul + rnr * Ur + n * (pl + pr));
    phiE += 0.5 * magS * (rnl * Hl + rnr * Hr);
    const double hm = 0.5 * magS * um;
    phi -= hm * (rl + rr);
    phiU = phiU - hm * (rUl + rUr);
    phiE -= hm * (rEl + rEr);
}

// HLLC flux on one face (hllcFluxScheme.C:70-240): Roe-averaged wave
// speeds, pos(x) = x > 0 and neg(x) = x < 0 switches, all four branch terms
// evaluated and weighted as in the reference.
// um: the face's normal frame velocity: relative normal velocities, p* um
// in the star energy fluxes (hllcFluxScheme.C:157-161,217-218).
DF void hllcFace(double gI, V3 S, double magS, double rl, double rr, double pl, double pr, V3 Ul, V3 Ur, double cl,
                 double cr, double El, double Er, double Hl, double Hr, double um, double& phi, V3& phiU,
                 double& phiE) {
    const V3 n = S / magS;
    const double coefR = sqrt(fmax(kVSMALL, rr) / fmax(kVSMALL, rl));
    const V3 uAvg = (coefR * Ur + Ul) / (coefR + 1.0);
    const double HAvg = (coefR * Hr + Hl) / (coefR + 1.0);
    const double cAvg = sqrt(fabs((gI - 1.0) * (HAvg - 0.5 * magSqr(uAvg))));
    const double ul = dot(Ul, n) - um, ur = dot(Ur, n) - um, ua = dot(uAvg, n) - um;
    const double Sl = fmin(ul - cl, ua - cAvg);
    const double Sr = fmax(ur + cr, ua + cAvg);
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

// Convective flux Jacobian of one side of a face, rows/cols ordered
// [rho, rhoE, rhoUx, rhoUy, rhoUz] (convectiveFluxScheme.C:402-464).
// Entries with no insertBlock in the reference are left at zero.
DF void fluxJacobian(double gI, V3 U, double E, V3 n, double A[25]) {
    const double th = 0.5 * (gI - 1) * magSqr(U);
    const double a1 = gI * E - th;
    const double a2 = gI - 1;
    const double pu = dot(U, n);
    A[0] = 0.0; A[1] = 0.0; A[2] = n.x; A[3] = n.y; A[4] = n.z;
    A[5] = pu * (th - a1);
    A[6] = gI * pu;
    const V3 eu = (n * a1) - ((a2 * U) * pu);
    A[7] = eu.x; A[8] = eu.y; A[9] = eu.z;
    const V3 mr = (n * th) - (U * pu);
    const V3 me = n * a2;
    T9 mu = outer(U, n) - outer(a2 * n, U);
    mu.t[0] += pu; mu.t[4] += pu; mu.t[8] += pu;
    A[10] = mr.x; A[11] = me.x; A[12] = mu.t[0]; A[13] = mu.t[1]; A[14] = mu.t[2];
    A[15] = mr.y; A[16] = me.y; A[17] = mu.t[3]; A[18] = mu.t[4]; A[19] = mu.t[5];
    A[20] = mr.z; A[21] = me.z; A[22] = mu.t[6]; A[23] = mu.t[7]; A[24] = mu.t[8];
}

// The same in precision C from scalar arguments, operation for operation
// (C = double gives fluxJacobian's bits).
template <typename C>
DF void fluxJacobianT(C gI, C ux, C uy, C uz, C E, C nx, C ny, C nz, C A[25]) {
    const C th = (C)0.5 * (gI - (C)1) * (ux * ux + uy * uy + uz * uz);
    const C a1 = gI * E - th;
    const C a2 = gI - (C)1;
    const C pu = ux * nx + uy * ny + uz * nz;
    A[0] = (C)0; A[1] = (C)0; A[2] = nx; A[3] = ny; A[4] = nz;
    A[5] = pu * (th - a1);
    A[6] = gI * pu;
    A[7] = nx * a1 - (a2 * ux) * pu;
    A[8] = ny * a1 - (a2 * uy) * pu;
    A[9] = nz * a1 - (a2 * uz) * pu;
    A[10] = nx * th - ux * pu; A[11] = nx * a2;
    A[15] = ny * th - uy * pu; A[16] = ny * a2;
    A[20] = nz * th - uz * pu; A[21] = nz * a2;
    const C an0 = a2 * nx, an1 = a2 * ny, an2 = a2 * nz;
    A[12] = ux * nx - an0 * ux; A[13] = ux * ny - an0 * uy; A[14] = ux * nz - an0 * uz;
    A[17] = uy * nx - an1 * ux; A[18] = uy * ny - an1 * uy; A[19] = uy * nz - an1 * uz;
    A[22] = uz * nx - an2 * ux; A[23] = uz * ny - an2 * uy; A[24] = uz * nz - an2 * uz;
    A[12] += pu; A[18] += pu; A[24] += pu;
}

// Entries of one off-diagonal block from its generators (genOff layout):
// h A(U, E, n), diagonal entries then minus the frame divergence term mo
// (hasMo: a frame zone), minus v and minus sv, as the Jacobian kernel forms
// the stored block.
template <typename C>
DF void blockFromGen(const C g[10], C gamma, C b[25], C mo = (C)0, bool hasMo = false) {
    fluxJacobianT<C>(gamma, g[0], g[1], g[2], g[3], g[4], g[5], g[6], b);
#pragma unroll
    for (int e = 0; e < 25; ++e) {
        C x = g[7] * b[e];
        if (e == 0 || e == 6 || e == 12 || e == 18 || e == 24) {
            if (hasMo) x = x - mo;
            x = x - g[8];
            x = x - g[9];
        }
        b[e] = x;
    }
}

#undef DF
}  // namespace icsf
