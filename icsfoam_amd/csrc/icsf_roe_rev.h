// Reverse mode of roeFace() (icsf_flux.h; roeFluxScheme.C:41-241) for the
// residual adjoint: one forward evaluation of the Roe-averaged state, then
// one backward sweep.
//
// The dissipation -|S|/2 M dQ, M = P |L| P^-1, is the sum over the three
// wave groups of the eigen-decomposition,
//   M = L1 P_1 iP_1 + L2 p2 ip2 + L3 p3 ip3
// (P_1: the 5x3 columns [n; P21; P31], iP_1 the 3x5 rows [iP11 iP12 iP13];
// p2, p3 the acoustic columns, ip2, ip3 the acoustic rows).  With the flux
// adjoint g and gd = -|S|/2 g, the adjoint of M is the rank-one gd dQ^T,
// so with the wave strengths a_k = iP_k dQ and the projections
// b_k = P_k^T gd:
//   dQ-bar   = L1 iP_1^T b_1 + L2 b_2 ip2 + L3 b_3 ip3
//   L_k-bar  = b_k . a_k
//   P_k-bar  = L_k gd (x) a_k,   iP_k-bar = L_k b_k (x) dQ,
// and no 5x5 matrix is formed.  The entries of P and P^-1 are then pulled
// back onto rho, the Roe velocity, c, u, theta and the a2-a6 factors, and
// those onto the face values.  Branches (|x|, the max of the entropy fix,
// the fix itself, the VSMALL clips of rho) follow the value the forward
// takes; ties take the first argument, the convention of the switch
// records.
#pragma once
#include "icsf_math.h"
#include "icsf_views.h"

namespace icsf {

#define DFR __device__ __forceinline__

// cross-product rule: y = a x b, y-bar given -> a-bar += b x y-bar,
// b-bar += y-bar x a

// L, R: owner- and neighbour-side face values [NFLD]; (bphi, bphiU,
// bphiE): adjoint of the face flux (mass, momentum, energy); out: adjoints
// of the face values (bL, bR, fields rho p U E H; F_CS gets zero) and, when
// geom, of Sf (|Sf| taken as |Sf|(Sf)); branches (optional): the switch
// bits of the value (0 u < 0, 1 u + c < 0, 2 u - c < 0, 3-5 the entropy fix
// of L1, L2, L3, 6 max(L2, L3) = L3, 7 the argument of c's square root
// below zero, 8 rr below VSMALL, 9 rl below VSMALL).
DFR void roeFaceRev(double gI, double eFix, V3 S, double magS, const double* L, const double* R, double bphi,
                    V3 bphiU, double bphiE, double* bL, double* bR, V3& bS, int* branches, bool geom) {
    const double rl = L[F_RHO], rr = R[F_RHO], pl = L[F_P], pr = R[F_P];
    const double El = L[F_E], Er = R[F_E], Hl = L[F_H], Hr = R[F_H];
    const V3 Ul{L[F_UX], L[F_UY], L[F_UZ]}, Ur{R[F_UX], R[F_UY], R[F_UZ]};

    // ---- forward (the values roeFace forms, in its order)
    const V3 n = S / magS;
    const bool clipR = !(kVSMALL < rr), clipL = !(kVSMALL < rl);
    const double coefR = sqrt(fmax(kVSMALL, rr) / fmax(kVSMALL, rl));
    const double rho = coefR * rl;
    const double cp1 = coefR + 1.0;
    const V3 Vel = (coefR * Ur + Ul) / cp1;
    const double HRoe = (coefR * Hr + Hl) / cp1;
    const double carg = (gI - 1.0) * (HRoe - 0.5 * magSqr(Vel));
    const double c = sqrt(fabs(carg));
    const double uMl = dot(Ul, n), uMr = dot(Ur, n);
    const double u = dot(Vel, n);
    const double a1 = gI - 1;
    const double theta = 0.5 * a1 * magSqr(Vel);
    const double c2 = sqr(c);
    const double a2 = 1 / (rho * c * sqrt(2.0));
    const double a3 = rho / (c * sqrt(2.0));
    const double a4 = (theta + c2) / a1;
    const double a5 = 1 - theta / c2;
    const double a6 = theta / a1;
    const double q = a1 / c2;
    const double L1r = fabs(u), L2r = fabs(u + c), L3r = fabs(u - c);
    const bool m3 = L2r < L3r;
    const double eps = eFix * fmax(L2r, L3r);
    const bool f1 = L1r < eps, f2 = L2r < eps, f3 = L3r < eps;
    if (branches)
        *branches = (int)(u < 0) | (int)(u + c < 0) << 1 | (int)(u - c < 0) << 2 | (int)f1 << 3 | (int)f2 << 4 |
                    (int)f3 << 5 | (int)m3 << 6 | (int)(carg < 0) << 7 | (int)clipR << 8 | (int)clipL << 9;
    const double L1 = f1 ? (sqr(L1r) + sqr(eps)) / (2.0 * eps) : L1r;
    const double L2 = f2 ? (sqr(L2r) + sqr(eps)) / (2.0 * eps) : L2r;
    const double L3 = f3 ? (sqr(L3r) + sqr(eps)) / (2.0 * eps) : L3r;
    const V3 Vxn = cross(Vel, n);

    // conservative jump and the flux adjoint of the dissipation
    const double dQ0 = rr - rl;
    const V3 dQU = rr * Ur - rl * Ul;
    const double dQ4 = rr * Er - rl * El;
    const double h = 0.5 * magS;
    const double gd0 = -h * bphi, gd4 = -h * bphiE;
    const V3 gdU = (-h) * bphiU;

    // wave rows of P^-1 and columns of P
    const V3 iP11 = n * a5 - Vxn / rho;
    const double iP21 = a2 * (theta - c * u), iP31 = a2 * (theta + c * u), iP23 = a1 * a2;
    const V3 iP22 = (-a2) * (a1 * Vel - c * n), iP32 = (-a2) * (a1 * Vel + c * n);
    const V3 P31 = n * a6 + rho * Vxn;
    const V3 P22 = a3 * (Vel + c * n), P23 = a3 * (Vel - c * n);
    const double P32 = a3 * (a4 + c * u), P33 = a3 * (a4 - c * u);

    // wave strengths a_k = iP_k dQ, projections b_k = P_k^T gd
    const double VdQ = dot(Vel, dQU);
    const V3 al1 = iP11 * dQ0 + (q * VdQ) * n - cross(n, dQU) / rho - (q * dQ4) * n;
    const double al2 = iP21 * dQ0 + dot(iP22, dQU) + iP23 * dQ4;
    const double al3 = iP31 * dQ0 + dot(iP32, dQU) + iP23 * dQ4;
    const double VgU = dot(Vel, gdU);
    const V3 be1 = (gd0 + VgU) * n - rho * cross(n, gdU) + gd4 * P31;
    const double be2 = a3 * gd0 + dot(P22, gdU) + P32 * gd4;
    const double be3 = a3 * gd0 + dot(P23, gdU) + P33 * gd4;

    // ---- backward
    const double L1b = dot(be1, al1), L2b = be2 * al2, L3b = be3 * al3;
    // dQ-bar
    const double nb1 = dot(n, be1);
    const double dQ0b = L1 * dot(iP11, be1) + L2 * be2 * iP21 + L3 * be3 * iP31;
    const V3 dQUb = L1 * ((q * nb1) * Vel + cross(n, be1) / rho) + (L2 * be2) * iP22 + (L3 * be3) * iP32;
    const double dQ4b = L1 * (-q * nb1) + L2 * be2 * iP23 + L3 * be3 * iP23;

    V3 nb{0, 0, 0}, Velb{0, 0, 0};
    double rhob = 0, cb = 0, ub = 0, thb = 0, c2b = 0, a2b = 0, a3b = 0, a4b = 0, a5b = 0, a6b = 0, qb = 0;
    // P_1-bar = L1 gd (x) al1: row 0 (n), rows 1-3 (P21 = Vel n^T + rho [n]x), row 4 (P31)
    nb = nb + (L1 * gd0) * al1;
    Velb = Velb + (L1 * dot(al1, n)) * gdU;
    nb = nb + (L1 * dot(gdU, Vel)) * al1;
    rhob += L1 * dot(gdU, cross(n, al1));
    nb = nb + (rho * L1) * cross(al1, gdU);
    {
        const V3 P31b = (L1 * gd4) * al1;
        nb = nb + a6 * P31b;
        a6b += dot(P31b, n);
        rhob += dot(P31b, Vxn);
        const V3 yb = rho * P31b;  // Vxn-bar
        Velb = Velb + cross(n, yb);
        nb = nb + cross(yb, Vel);
    }
    // p2-bar = L2 al2 gd, p3-bar = L3 al3 gd
    {
        const double s2 = L2 * al2, s3 = L3 * al3;
        const V3 p2U = s2 * gdU, p3U = s3 * gdU;
        a3b += s2 * gd0 + dot(p2U, Vel + c * n) + s2 * gd4 * (a4 + c * u);
        a3b += s3 * gd0 + dot(p3U, Vel - c * n) + s3 * gd4 * (a4 - c * u);
        Velb = Velb + a3 * (p2U + p3U);
        cb += a3 * (dot(p2U, n) + s2 * gd4 * u) - a3 * (dot(p3U, n) + s3 * gd4 * u);
        nb = nb + (a3 * c) * (p2U - p3U);
        a4b += a3 * (s2 + s3) * gd4;
        ub += a3 * c * (s2 - s3) * gd4;
    }
    // iP_1-bar = L1 be1 (x) dQ: column 0 (iP11), columns 1-3 (iP12), column 4 (iP13)
    {
        const V3 iP11b = (L1 * dQ0) * be1;
        nb = nb + a5 * iP11b;
        a5b += dot(iP11b, n);
        const V3 yb = (-1.0 / rho) * iP11b;  // Vxn-bar
        Velb = Velb + cross(n, yb);
        nb = nb + cross(yb, Vel);
        rhob += dot(iP11b, Vxn) / (rho * rho);
        qb += L1 * nb1 * VdQ;
        nb = nb + (q * L1 * VdQ) * be1;
        Velb = Velb + (q * L1 * nb1) * dQU;
        const V3 nxdQ = cross(n, dQU);
        rhob += L1 / (rho * rho) * dot(be1, nxdQ);
        nb = nb - (L1 / rho) * cross(dQU, be1);
        const V3 iP13b = (L1 * dQ4) * be1;
        qb -= dot(iP13b, n);
        nb = nb - q * iP13b;
    }
    // ip2-bar = L2 be2 dQ, ip3-bar = L3 be3 dQ
    {
        const double t2 = L2 * be2, t3 = L3 * be3;
        const double i20 = t2 * dQ0, i30 = t3 * dQ0;
        const V3 i2U = t2 * dQU, i3U = t3 * dQU;
        a2b += i20 * (theta - c * u) + i30 * (theta + c * u);
        thb += a2 * (i20 + i30);
        cb += a2 * u * (i30 - i20);
        ub += a2 * c * (i30 - i20);
        a2b -= dot(i2U, a1 * Vel - c * n) + dot(i3U, a1 * Vel + c * n);
        Velb = Velb - (a2 * a1) * (i2U + i3U);
        cb += a2 * (dot(i2U, n) - dot(i3U, n));
        nb = nb + (a2 * c) * (i2U - i3U);
        a2b += a1 * (t2 * dQ4 + t3 * dQ4);
    }
    // eigenvalues: entropy fix, max, |x|
    {
        double L1rb = L1b, L2rb = L2b, L3rb = L3b, epsb = 0;
        const double e2 = 2.0 * eps * eps;
        if (f1) { L1rb = L1b * L1r / eps; epsb += L1b * (eps * eps - L1r * L1r) / e2; }
        if (f2) { L2rb = L2b * L2r / eps; epsb += L2b * (eps * eps - L2r * L2r) / e2; }
        if (f3) { L3rb = L3b * L3r / eps; epsb += L3b * (eps * eps - L3r * L3r) / e2; }
        if (m3) L3rb += eFix * epsb;
        else L2rb += eFix * epsb;
        const double s1 = u < 0 ? -L1rb : L1rb;
        const double s2 = u + c < 0 ? -L2rb : L2rb;
        const double s3 = u - c < 0 ? -L3rb : L3rb;
        ub += s1 + s2 + s3;
        cb += s2 - s3;
    }
    // scalar factors
    c2b -= qb * q / c2;
    thb += a6b / a1;
    thb -= a5b / c2;
    c2b += a5b * theta / (c2 * c2);
    thb += a4b / a1;
    c2b += a4b / a1;
    rhob += a3b * a3 / rho;
    cb -= a3b * a3 / c;
    rhob -= a2b * a2 / rho;
    cb -= a2b * a2 / c;
    cb += 2.0 * c * c2b;
    Velb = Velb + (a1 * thb) * Vel;
    Velb = Velb + ub * n;
    nb = nb + ub * Vel;
    const double cargb = (carg < 0 ? -cb : cb) / (2.0 * c);
    const double HRoeb = a1 * cargb;
    Velb = Velb - (a1 * cargb) * Vel;

    // central flux 0.5|S| (F_l + F_r) and |S|
    const double rnl = rl * uMl, rnr = rr * uMr;
    const double gc0 = h * bphi, gcE = h * bphiE;
    const V3 gcU = h * bphiU;
    const double rnlb = gc0 + dot(gcU, Ul) + gcE * Hl;
    const double rnrb = gc0 + dot(gcU, Ur) + gcE * Hr;
    V3 Ulb = rnl * gcU, Urb = rnr * gcU;
    nb = nb + (pl + pr) * gcU;
    const double pnb = dot(gcU, n);
    double Hlb = rnl * gcE, Hrb = rnr * gcE;
    double rlb = rnlb * uMl, rrb = rnrb * uMr;
    Ulb = Ulb + (rnlb * rl) * n;
    Urb = Urb + (rnrb * rr) * n;
    nb = nb + (rnlb * rl) * Ul + (rnrb * rr) * Ur;

    // conservative jump -> face values
    rrb += dQ0b + dot(dQUb, Ur) + dQ4b * Er;
    Urb = Urb + rr * dQUb;
    const double Erb = rr * dQ4b;
    rlb -= dQ0b + dot(dQUb, Ul) + dQ4b * El;
    Ulb = Ulb - rl * dQUb;
    const double Elb = -rl * dQ4b;

    // Roe average
    double cRb = HRoeb * (Hr - HRoe) / cp1 + dot(Velb, Ur - Vel) / cp1 + rhob * rl;
    Hrb += HRoeb * coefR / cp1;
    Hlb += HRoeb / cp1;
    Urb = Urb + (coefR / cp1) * Velb;
    Ulb = Ulb + (1.0 / cp1) * Velb;
    rlb += rhob * coefR;
    if (!clipR) rrb += cRb * coefR / (2.0 * rr);
    if (!clipL) rlb -= cRb * coefR / (2.0 * rl);

#pragma unroll
    for (int k = 0; k < NFLD; ++k) bL[k] = bR[k] = 0.0;
    bL[F_RHO] = rlb; bR[F_RHO] = rrb;
    bL[F_P] = pnb; bR[F_P] = pnb;
    bL[F_UX] = Ulb.x; bL[F_UY] = Ulb.y; bL[F_UZ] = Ulb.z;
    bR[F_UX] = Urb.x; bR[F_UY] = Urb.y; bR[F_UZ] = Urb.z;
    bL[F_E] = Elb; bR[F_E] = Erb;
    bL[F_H] = Hlb; bR[F_H] = Hrb;

    bS = V3{0, 0, 0};
    if (geom) {
        // flux = |S| (...): |S|-bar is g . flux / |S|
        const V3 Fu = rnl * Ul + rnr * Ur + n * (pl + pr);
        const double magSb = 0.5 * (bphi * (rnl + rnr) + dot(bphiU, Fu) + bphiE * (rnl * Hl + rnr * Hr)) +
                             (L1 * L1b + L2 * L2b + L3 * L3b) / magS;
        // n = S / |S|, |S| = |S|(S)
        bS = (nb - dot(nb, n) * n) / magS + magSb * n;
    }
}

#undef DFR
}  // namespace icsf
