// Forward-mode dual numbers (value, one tangent) for the residual adjoint's
// face pass: a flux evaluated on duals seeded in one input gives that
// column of the flux Jacobian.  Values are formed by the same operations,
// in the same order, as the double-precision kernels (icsf_math.h), so
// every branch (max/min, |x|, the entropy fix) follows the value the
// residual itself takes.
#pragma once
#include "icsf_math.h"

namespace icsf {

#define DF __device__ __forceinline__

struct Dd {
    double v, d;
};
DF Dd dcon(double v) { return Dd{v, 0.0}; }
DF Dd operator+(Dd a, Dd b) { return Dd{a.v + b.v, a.d + b.d}; }
DF Dd operator-(Dd a, Dd b) { return Dd{a.v - b.v, a.d - b.d}; }
DF Dd operator-(Dd a) { return Dd{-a.v, -a.d}; }
DF Dd operator*(Dd a, Dd b) { return Dd{a.v * b.v, a.d * b.v + a.v * b.d}; }
DF Dd operator/(Dd a, Dd b) {
    const double q = a.v / b.v;
    return Dd{q, (a.d - q * b.d) / b.v};
}
DF Dd operator*(double a, Dd b) { return Dd{a * b.v, a * b.d}; }
DF Dd operator*(Dd a, double b) { return Dd{a.v * b, a.d * b}; }
DF Dd operator+(Dd a, double b) { return Dd{a.v + b, a.d}; }
DF Dd operator-(Dd a, double b) { return Dd{a.v - b, a.d}; }
DF Dd operator-(double a, Dd b) { return Dd{a - b.v, -b.d}; }
DF Dd operator/(Dd a, double b) { return Dd{a.v / b, a.d / b}; }
DF Dd operator/(double a, Dd b) {
    const double q = a / b.v;
    return Dd{q, -q * b.d / b.v};
}
DF Dd sqrtD(Dd a) {
    const double s = sqrt(a.v);
    return Dd{s, a.d / (2.0 * s)};
}
// |a|, fmax, fmin: the branch of the value (ties: the first argument)
DF Dd fabsD(Dd a) { return Dd{fabs(a.v), a.v < 0 ? -a.d : a.d}; }
DF Dd fmaxD(Dd a, Dd b) { return a.v < b.v ? b : a; }
DF Dd sqrD(Dd a) { return a * a; }

struct V3D {
    Dd x, y, z;
};
DF V3D operator+(V3D a, V3D b) { return V3D{a.x + b.x, a.y + b.y, a.z + b.z}; }
DF V3D operator-(V3D a, V3D b) { return V3D{a.x - b.x, a.y - b.y, a.z - b.z}; }
DF V3D operator*(Dd s, V3D a) { return V3D{s * a.x, s * a.y, s * a.z}; }
DF V3D operator*(V3D a, Dd s) { return V3D{a.x * s, a.y * s, a.z * s}; }
DF V3D operator*(double s, V3D a) { return V3D{s * a.x, s * a.y, s * a.z}; }
DF V3D operator/(V3D a, Dd s) { return V3D{a.x / s, a.y / s, a.z / s}; }
DF Dd dotD(V3D a, V3D b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DF Dd magSqrD(V3D a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
DF V3D crossD(V3D a, V3D b) {
    return V3D{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct T9D {
    Dd t[9];
};
DF T9D outerD(V3D a, V3D b) {
    return T9D{{a.x * b.x, a.x * b.y, a.x * b.z, a.y * b.x, a.y * b.y, a.y * b.z, a.z * b.x, a.z * b.y, a.z * b.z}};
}
DF T9D operator+(const T9D& a, const T9D& b) {
    T9D r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.t[i] = a.t[i] + b.t[i];
    return r;
}
DF T9D operator*(const T9D& a, Dd s) {
    T9D r;
#pragma unroll
    for (int i = 0; i < 9; ++i) r.t[i] = a.t[i] * s;
    return r;
}
DF V3D tdotvD(const T9D& A, V3D v) {
    return V3D{A.t[0] * v.x + A.t[1] * v.y + A.t[2] * v.z, A.t[3] * v.x + A.t[4] * v.y + A.t[5] * v.z,
               A.t[6] * v.x + A.t[7] * v.y + A.t[8] * v.z};
}
DF V3D vdottD(V3D v, const T9D& A) {
    return V3D{v.x * A.t[0] + v.y * A.t[3] + v.z * A.t[6], v.x * A.t[1] + v.y * A.t[4] + v.z * A.t[7],
               v.x * A.t[2] + v.y * A.t[5] + v.z * A.t[8]};
}
DF T9D tdottD(const T9D& A, const T9D& B) {
    T9D r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            r.t[3 * i + j] = A.t[3 * i] * B.t[j] + A.t[3 * i + 1] * B.t[3 + j] + A.t[3 * i + 2] * B.t[6 + j];
    return r;
}

// roeFace() of icsf_flux.h on duals, expression for expression; bits
// (optional) receives the switches of the value: 0 u < 0, 1 u + c < 0,
// 2 u - c < 0, 3-5 the entropy fix of L1, L2, L3, 6 max(L2, L3) = L3,
// 7 the argument of c's square root below zero, 8 rr below VSMALL, 9 rl
// below VSMALL.
DF void roeFaceD(double gI, double eFix, V3D S, Dd magS, Dd rl, Dd rr, Dd pl, Dd pr, V3D Ul, V3D Ur, Dd El, Dd Er,
                 Dd Hl, Dd Hr, Dd& phi, V3D& phiU, Dd& phiE, int* bits) {
    const V3D n = S / magS;
    const Dd coefR = sqrtD(fmaxD(dcon(kVSMALL), rr) / fmaxD(dcon(kVSMALL), rl));
    const Dd rho = coefR * rl;
    const V3D Vel = (coefR * Ur + Ul) / (coefR + 1.0);
    const Dd HRoe = (coefR * Hr + Hl) / (coefR + 1.0);
    const Dd carg = (gI - 1.0) * (HRoe - 0.5 * magSqrD(Vel));
    const Dd c = sqrtD(fabsD(carg));
    const Dd uMl = dotD(Ul, n), uMr = dotD(Ur, n);
    const Dd u = dotD(Vel, n);
    const V3D rUl = rl * Ul, rUr = rr * Ur;
    const Dd rEl = rl * El, rEr = rr * Er;

    const double a1 = gI - 1;
    const Dd theta = (0.5 * a1) * magSqrD(Vel);
    const Dd c2 = sqrD(c);
    const Dd a2 = 1 / (rho * c * sqrt(2.0));
    const Dd a3 = rho / (c * sqrt(2.0));
    const Dd a4 = (theta + c2) / a1;
    const Dd a5 = 1 - theta / c2;
    const Dd a6 = theta / a1;
    const V3D iP11 = (n * a5) - crossD(Vel, n) / rho;
    T9D iP12 = outerD(a1 / c2 * n, Vel);
    iP12.t[1] = iP12.t[1] + n.z / rho; iP12.t[2] = iP12.t[2] - n.y / rho; iP12.t[3] = iP12.t[3] - n.z / rho;
    iP12.t[5] = iP12.t[5] + n.x / rho; iP12.t[6] = iP12.t[6] + n.y / rho; iP12.t[7] = iP12.t[7] - n.x / rho;
    const V3D iP13 = (-a1) / c2 * n;
    const Dd iP21 = a2 * (theta - c * u);
    const V3D iP22 = (-a2) * (a1 * Vel - c * n);
    const Dd iP23 = a1 * a2;
    const Dd iP31 = a2 * (theta + c * u);
    const V3D iP32 = (-a2) * (a1 * Vel + c * n);
    const Dd iP33 = a1 * a2;
    Dd L1 = fabsD(u), L2 = fabsD(u + c), L3 = fabsD(u - c);
    const Dd eps = eFix * fmaxD(L2, L3);
    const bool f1 = L1.v < eps.v, f2 = L2.v < eps.v, f3 = L3.v < eps.v;
    if (bits)
        *bits = (int)(u.v < 0) | (int)((u + c).v < 0) << 1 | (int)((u - c).v < 0) << 2 | (int)f1 << 3 |
                (int)f2 << 4 | (int)f3 << 5 | (int)(L2.v < L3.v) << 6 | (int)(carg.v < 0) << 7 |
                (int)(!(kVSMALL < rr.v)) << 8 | (int)(!(kVSMALL < rl.v)) << 9;
    if (f1) L1 = (sqrD(L1) + sqrD(eps)) / (2.0 * eps);
    if (f2) L2 = (sqrD(L2) + sqrD(eps)) / (2.0 * eps);
    if (f3) L3 = (sqrD(L3) + sqrD(eps)) / (2.0 * eps);
    const Dd P12 = a3, P13 = a3;
    T9D P21 = outerD(Vel, n);
    P21.t[1] = P21.t[1] - n.z * rho; P21.t[2] = P21.t[2] + n.y * rho; P21.t[3] = P21.t[3] + n.z * rho;
    P21.t[5] = P21.t[5] - n.x * rho; P21.t[6] = P21.t[6] - n.y * rho; P21.t[7] = P21.t[7] + n.x * rho;
    const V3D P22 = a3 * (Vel + c * n);
    const V3D P23 = a3 * (Vel - c * n);
    const V3D P31 = n * a6 + rho * crossD(Vel, n);
    const Dd P32 = a3 * (a4 + c * u);
    const Dd P33 = a3 * (a4 - c * u);
    const V3D P11L = n * L1;
    const T9D P21L = P21 * L1;
    const V3D P31L = P31 * L1;
    const Dd CR = dotD(P11L, iP11) + L2 * P12 * iP21 + L3 * P13 * iP31;
    const V3D CU = vdottD(P11L, iP12) + (L2 * P12) * iP22 + (L3 * P13) * iP32;
    const Dd CE = dotD(P11L, iP13) + L2 * P12 * iP23 + L3 * P13 * iP33;
    const V3D MRv = tdotvD(P21L, iP11) + (L2 * P22) * iP21 + (L3 * P23) * iP31;
    const T9D MU = tdottD(P21L, iP12) + outerD(L2 * P22, iP22) + outerD(L3 * P23, iP32);
    const V3D ME = tdotvD(P21L, iP13) + (L2 * P22) * iP23 + (L3 * P23) * iP33;
    const Dd ER = dotD(P31L, iP11) + L2 * P32 * iP21 + L3 * P33 * iP31;
    const V3D EU = vdottD(P31L, iP12) + (L2 * P32) * iP22 + (L3 * P33) * iP32;
    const Dd EE = dotD(P31L, iP13) + L2 * P32 * iP23 + L3 * P33 * iP33;

    const Dd dR = rr - rl;
    const V3D dRU = rUr - rUl;
    const Dd dRE = rEr - rEl;
    phi = -0.5 * magS * (CR * dR + dotD(CU, dRU) + CE * dRE);
    phiU = -0.5 * magS * (MRv * dR + tdotvD(MU, dRU) + ME * dRE);
    phiE = -0.5 * magS * (ER * dR + dotD(EU, dRU) + EE * dRE);
    const Dd rnl = rl * uMl, rnr = rr * uMr;
    phi = phi + 0.5 * magS * (rnl + rnr);
    phiU = phiU + 0.5 * magS * (rnl * Ul + rnr * Ur + n * (pl + pr));
    phiE = phiE + 0.5 * magS * (rnl * Hl + rnr * Hr);
}

#undef DF
}  // namespace icsf
