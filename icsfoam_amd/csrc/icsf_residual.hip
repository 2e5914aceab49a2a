// Residual and Jacobian assembly kernels: one thread per cell, pulling the
// cell's faces from a face-ordered cell->face list, so every cell sums its
// face contributions in the same order as OpenFOAM's face loops
// (surfaceIntegrate, gaussGrad, LduMatrix::negSumDiag) and no atomics are
// needed.  Faces are evaluated once per adjacent cell with identical code
// and inputs, hence identical bits on both sides.
#include "icsf_flux.h"
#include "icsf_kernels.h"
#include "icsf_recon.h"
#include "icsf_views.h"

namespace icsf {

namespace {

// Primitive/derived cell fields from conserved W after an increment
// (updateFields.H:7-78), or from (rho, U, T) at initialisation.
__global__ void k_update(MeshView m, Phys ph, StateView s, const double* __restrict__ dW, int mode,
                         const double* __restrict__ rho0, const double* __restrict__ U0,
                         const double* __restrict__ T0) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    if (s.halt && *s.halt) return;
    const int nC = m.nC;
    double rho, e, T, p;
    V3 U;
    if (mode == 0) {
        // initialisation from rho, U, T (createFields.H)
        rho = rho0[c];
        U = V3{U0[3 * c], U0[3 * c + 1], U0[3 * c + 2]};
        T = T0[c];
        e = ph.Cv * T;
        const double psi = 1.0 / (ph.R * T);
        p = rho / psi;
    } else {
        rho = s.W[c];
        V3 rU = V3{s.W[2 * (size_t)nC + c], s.W[3 * (size_t)nC + c], s.W[4 * (size_t)nC + c]};
        double rE = s.W[(size_t)nC + c];
        if (dW) {
            rho += dW[c];
            rU = V3{rU.x + dW[2 * (size_t)nC + c], rU.y + dW[3 * (size_t)nC + c], rU.z + dW[4 * (size_t)nC + c]};
            rE += dW[(size_t)nC + c];
        }
        // a non-finite conserved state (before the bounds, which would mask
        // a NaN) is reported through the error flags
        if (s.err && !(isfinite(rho) && isfinite(rE) && isfinite(rU.x) && isfinite(rU.y) && isfinite(rU.z)))
            atomicOr(s.err, ERR_FLAG_NONFINITE);
        rho = fmax(rho, ph.rhoMin);
        U = rU / rho;
        e = rE / rho - 0.5 * magSqr(U);
        e = fmax(e, ph.Cv * ph.TMin);
        if (ph.TMax < kGREAT) e = fmin(e, ph.Cv * ph.TMax);
        T = e / ph.Cv;
        const double psi = 1.0 / (ph.R * T);
        p = rho / psi;
    }
    const double rE = rho * (e + 0.5 * magSqr(U));
    s.W[c] = rho;
    s.W[(size_t)nC + c] = rE;
    s.W[2 * (size_t)nC + c] = rho * U.x;
    s.W[3 * (size_t)nC + c] = rho * U.y;
    s.W[4 * (size_t)nC + c] = rho * U.z;
    const double E = ph.Cv * T + 0.5 * magSqr(U);
    const double H = fmax(E, kSMALL) + fmax(p / rho, kSMALL);
    const double g = ph.gamma;
    double* f = s.fld + (size_t)c * NFLD;
    f[F_RHO] = rho;
    f[F_P] = p;
    f[F_UX] = U.x;
    f[F_UY] = U.y;
    f[F_UZ] = U.z;
    f[F_E] = E;
    f[F_H] = H;
    if (ph.flux == FLUX_HLLC) {
        const double psi = 1.0 / (ph.R * T);
        f[F_CS] = fmax(sqrt(g / psi), kVSMALL);
    } else {
        f[F_CS] = sqrt(2.0 * (g - 1.0) / (g + 1.0) * H);
    }
    s.T[c] = T;
    s.ecalc[c] = rE / rho - 0.5 * magSqr(U);
    if (s.e) s.e[c] = e;
}

// Boundary values of p, U and T after a state update, on every non-empty,
// non-coupled boundary face, in the order updateFields.H:90-105 corrects
// them (p, then U, then T), into StateView::bcV.  fixedValue, zeroGradient
// and slip from the updated cells; the derived inflow conditions
// (OpenFOAM-v2112 finiteVolume, as ICSFoam's tutorials use them):
//   totalPressure (psi, gamma > 1): p0 / (1 + 0.5 psi (gamma-1)/gamma
//     (1 - pos0(phi)) |U_b|^2)^(gamma/(gamma-1)), psi = 1/(R T_b) and U_b
//     the values before this update (thermo.correct() precedes the p
//     correction, U is corrected after p);
//   pressureInletOutletVelocity: directionMixed, valueFraction
//     neg(phi) (I - n n), refValue the tangential part of
//     tangentialVelocity: U_b = vf & ref + (I - vf) & U_P;
//   totalTemperature: T0 / (1 + 0.5 psi (gamma-1)/gamma (1 - pos0(phi))
//     |U_b|^2) with the corrected U_b;
// phi = bcPhi, the boundary mass flux of the last face-terms pass.
// init: the createFields state (createFields.H:51-145): the derived
// conditions take their initial values from bcInit, valueFraction starts
// at zero, and phi = Sf & (rho_b U_b).
__global__ void k_bc_update(MeshView m, Phys ph, StateView s, int init) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nB) return;
    if (s.halt && *s.halt) return;
    if (m.bKind[b] == PK_EMPTY || m.bKind[b] == PK_CYCLIC) return;
    const int f = m.nI + b;
    const int c = m.own[f];
    double* v = s.bcV + 5 * (size_t)b;
    const V3 UOld{v[1], v[2], v[3]};
    const double TOld = v[4];
    const double gM1ByG = (ph.gamma - 1.0) / ph.gamma;
    const double phi = init ? 0.0 : s.bcPhi[b];
    const double npos = 1.0 - (phi >= 0 ? 1.0 : 0.0);
    const double* q = s.bcInit + 5 * (size_t)b;
    const int pt = m.bPT[b], ut = m.bUT[b], tt = m.bTT[b];
    double pB;
    if (pt == BC_FIXED) pB = m.bPv[b];
    else if (pt == BC_TOTAL_P) {
        if (init) pB = q[0];
        else {
            const double psip = 1.0 / (ph.R * TOld);
            pB = m.bPv[b] / pow(1.0 + 0.5 * psip * gM1ByG * npos * magSqr(UOld), 1.0 / gM1ByG);
        }
    } else pB = s.fld[(size_t)c * NFLD + F_P];
    const V3 Uc = cellU(s.fld, c);
    V3 UB;
    if (ut == BC_FIXED) UB = ld3(m.bUv + 3 * (size_t)b);
    else if (ut == BC_SLIP) {
        const V3 n = ld3(m.Sf + 3 * (size_t)f) / m.magSf[f];
        UB = (Uc + reflect(n, Uc)) / 2.0;
    } else if (ut == BC_PIOV) {
        if (init) {
            UB = V3{q[1], q[2], q[3]};
            s.bcIn[b] = 0;
        } else if (phi < 0) {
            s.bcIn[b] = 1;
            const V3 n = ld3(m.Sf + 3 * (size_t)f) / m.magSf[f];
            const V3 tv = ld3(m.bUv + 3 * (size_t)b);
            const V3 ref = tv - n * dot(n, tv);
            const T9 vf{{1 - n.x * n.x, -n.x * n.y, -n.x * n.z,
                         -n.y * n.x, 1 - n.y * n.y, -n.y * n.z,
                         -n.z * n.x, -n.z * n.y, 1 - n.z * n.z}};
            const T9 I{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
            UB = tdotv(vf, ref) + tdotv(I - vf, Uc);
        } else {
            s.bcIn[b] = 0;
            UB = Uc;
        }
    } else UB = Uc;
    double TB;
    if (tt == BC_FIXED) TB = m.bTv[b];
    else if (tt == BC_TOTAL_T) {
        if (init) TB = q[4];
        else {
            const double psip = 1.0 / (ph.R * TOld);
            TB = m.bTv[b] / (1.0 + 0.5 * psip * gM1ByG * npos * magSqr(UB));
        }
    } else TB = s.T[c];
    v[0] = pB; v[1] = UB.x; v[2] = UB.y; v[3] = UB.z; v[4] = TB;
    if (init) {
        const double psi = 1.0 / (ph.R * TB);
        const double rhoB = psi * pB;
        const V3 rUB{rhoB * UB.x, rhoB * UB.y, rhoB * UB.z};
        s.bcPhi[b] = dot(ld3(m.Sf + 3 * (size_t)f), rUB);
    }
}

// Boundary values of p, U, T and the derived reconstruction fields on
// every non-empty boundary face: fixedValue / zeroGradient / slip from the
// current cells, the derived inflow conditions from bcV.
__device__ __forceinline__ void boundaryBody(const MeshView& m, const Phys& ph, const StateView& s) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nB) return;
    if (m.bKind[b] == PK_EMPTY || m.bKind[b] == PK_CYCLIC) return;
    const int f = m.nI + b;
    const int c = m.own[f];
    const double* fc = s.fld + (size_t)c * NFLD;
    const double* bv = s.bcV + 5 * (size_t)b;
    const int pt = m.bPT[b], tt = m.bTT[b];
    const double pB = pt == BC_FIXED ? m.bPv[b] : pt == BC_TOTAL_P ? bv[0] : fc[F_P];
    const double TB = tt == BC_FIXED ? m.bTv[b] : tt == BC_TOTAL_T ? bv[4] : s.T[c];
    const V3 Uc = cellU(s.fld, c);
    V3 UB;
    if (m.bUT[b] == BC_FIXED) UB = ld3(m.bUv + 3 * (size_t)b);
    else if (m.bUT[b] == BC_PIOV) UB = V3{bv[1], bv[2], bv[3]};
    else if (m.bUT[b] == BC_SLIP) {
        const V3 n = ld3(m.Sf + 3 * (size_t)f) / m.magSf[f];
        UB = (Uc + reflect(n, Uc)) / 2.0;
    } else UB = Uc;
    const double psi = 1.0 / (ph.R * TB);
    const double rhoB = psi * pB;
    const double E = ph.Cv * TB + 0.5 * magSqr(UB);
    const double H = fmax(E, kSMALL) + fmax(pB / rhoB, kSMALL);
    const double g = ph.gamma;
    double* o = s.bfld + (size_t)b * NFLD;
    o[F_RHO] = rhoB;
    o[F_P] = pB;
    o[F_UX] = UB.x;
    o[F_UY] = UB.y;
    o[F_UZ] = UB.z;
    o[F_E] = E;
    o[F_H] = H;
    if (ph.flux == FLUX_HLLC) o[F_CS] = fmax(sqrt(g / psi), kVSMALL);
    else o[F_CS] = sqrt(2.0 * (g - 1.0) / (g + 1.0) * H);
    s.bT[b] = TB;
}

__global__ void k_boundary(MeshView m, Phys ph, StateView s) { boundaryBody(m, ph, s); }
__global__ void k_boundary_hb(const HBInst* __restrict__ I, Phys ph) {
    const HBInst q = I[blockIdx.y];
    boundaryBody(q.m, ph, q.s);
}

// Gauss linear gradients with cellLimited<limiter> (k = 1) of the NG
// reconstruction fields; unlimited grad(U) and grad(ecalc) when viscous.
// gun/limo (optional, [nC][NFLD][3] and [nC][NFLD]) receive the unlimited
// gradients and the limiter values for the adjoint pass.
template <int NG>
__device__ __forceinline__ void gradBody(const MeshView& m, const Phys& ph, const StateView& s,
                                         double* __restrict__ gun, double* __restrict__ limo) {
    const int c = xcdBlock() * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    double gx[NG], gy[NG], gz[NG], mx[NG], mn[NG], vc[NFLD];
    ldRow(s.fld + (size_t)c * NFLD, vc);
#pragma unroll
    for (int k = 0; k < NG; ++k) {
        gx[k] = gy[k] = gz[k] = 0.0;
        mx[k] = mn[k] = vc[k];
    }
    const bool visc = ph.viscous;
    V3 gex{0, 0, 0};
    const double ecc = visc ? s.ecalc[c] : 0.0;
    // internal faces in face order from the cell's slots (column, slot
    // geometry), then its boundary faces from the cell-face list
    const int e0 = m.cfOff[c], e1 = m.cfOff[c + 1];
    const int len = m.rowLen[c], nLo = m.nLower[c];
    const size_t rs = (size_t)m.slicePtr[c >> 6] * 64 + (c & 63);
    for (int j = 0; j < len; ++j) {
        const size_t sk = rs + 64 * (size_t)j;
        const int other = ldnt(m.col + sk);
        const bool isOwn = j >= nLo;
        const V3 S{ldnt(m.sgeo + sgOff(sk, SG_SX)), ldnt(m.sgeo + sgOff(sk, SG_SX + 1)),
                   ldnt(m.sgeo + sgOff(sk, SG_SX + 2))};
        const double w = ldnt(m.sgeo + sgOff(sk, SG_W));
        // the face's owner and neighbour values are this cell's row and the
        // other cell's, gathered once
        double vO[NFLD];
        ldRow(s.fld + (size_t)other * NFLD, vO);
#pragma unroll
        for (int k = 0; k < NG; ++k) {
            const double vo = isOwn ? vc[k] : vO[k], vn = isOwn ? vO[k] : vc[k];
            const double vf = w * (vo - vn) + vn;
            const V3 t = S * vf;
            if (isOwn) { gx[k] = gx[k] + t.x; gy[k] = gy[k] + t.y; gz[k] = gz[k] + t.z; }
            else { gx[k] = gx[k] - t.x; gy[k] = gy[k] - t.y; gz[k] = gz[k] - t.z; }
            mx[k] = fmax(mx[k], vO[k]);
            mn[k] = fmin(mn[k], vO[k]);
        }
        if (visc) {
            const double eO = s.ecalc[other];
            const double eo = isOwn ? ecc : eO, en = isOwn ? eO : ecc;
            const double vf = w * (eo - en) + en;
            const V3 t = S * vf;
            gex = isOwn ? gex + t : gex - t;
        }
    }
    for (int q = e0 + len; q < e1; ++q) {
        const int f = m.cf[q];
        const V3 S = ld3(m.Sf + 3 * (size_t)f);
        if (m.bKind[f - m.nI] == PK_CYCLIC) {
            // coupled linear interpolation w v_P + (1 - w) v_far; max/min
            // over the far-side value (cellLimitedGrad, coupled patches)
            const int kc = m.cycOf[f - m.nI];
            const double w = m.w[f];
#pragma unroll
            for (int k = 0; k < NG; ++k) {
                const double vn = cycNbrValue(m, s, kc, k);
                const double vf = w * vc[k] + (1.0 - w) * vn;
                const V3 t = S * vf;
                gx[k] = gx[k] + t.x; gy[k] = gy[k] + t.y; gz[k] = gz[k] + t.z;
                mx[k] = fmax(mx[k], vn);
                mn[k] = fmin(mn[k], vn);
            }
            if (visc) {
                // grad(ecalc): the far cell's value (eCalc is not lagged)
                const double ef = w * ecc + (1.0 - w) * s.ecalc[m.cycCell[kc]];
                gex = gex + S * ef;
            }
        } else {
            const int b = f - m.nI;
            double bf[NFLD];
            ldRow(s.bfld + (size_t)b * NFLD, bf);
#pragma unroll
            for (int k = 0; k < NG; ++k) {
                const double vb = bf[k];
                const V3 t = S * vb;
                gx[k] = gx[k] + t.x; gy[k] = gy[k] + t.y; gz[k] = gz[k] + t.z;
                mx[k] = fmax(mx[k], vb);
                mn[k] = fmin(mn[k], vb);
            }
            if (visc) {
                const V3 UB = V3{bf[F_UX], bf[F_UY], bf[F_UZ]};
                const double rhoB = bf[F_RHO];
                const double eB = ph.Cv * s.bT[b];
                const double rhoEB = rhoB * (eB + 0.5 * magSqr(UB));
                const double ebv = rhoEB / rhoB - 0.5 * magSqr(UB);
                gex = gex + S * ebv;
            }
        }
    }
    const double Vc = m.V[c];
#pragma unroll
    for (int k = 0; k < NG; ++k) { gx[k] = gx[k] / Vc; gy[k] = gy[k] / Vc; gz[k] = gz[k] / Vc; }
    if (gun) {
        double o[3 * NG];
#pragma unroll
        for (int k = 0; k < NG; ++k) { o[3 * k] = gx[k]; o[3 * k + 1] = gy[k]; o[3 * k + 2] = gz[k]; }
        stRow(gun + (size_t)c * NFLD * 3, o);
    }
    if (visc) {
        // grad(U)_ij = d U_j / d x_i; column j is the gradient of U_j
        gex = gex / Vc;
        const double o[NGV] = {gx[F_UX], gx[F_UY], gx[F_UZ], gy[F_UX], gy[F_UY], gy[F_UZ],
                               gz[F_UX], gz[F_UY], gz[F_UZ], gex.x,    gex.y,    gex.z};
        stRow(s.gv + (size_t)c * NGV, o);
    }
    if (ph.recon == 1 && ph.limiter != 0) {
        double lim[NG];
#pragma unroll
        for (int k = 0; k < NG; ++k) { mx[k] -= vc[k]; mn[k] -= vc[k]; lim[k] = 1.0; }
        const V3 Cc = ld3(m.C + 3 * (size_t)c);
        for (int j = 0; j < len; ++j) {
            const size_t sk = rs + 64 * (size_t)j;
            const V3 d = V3{ldnt(m.sgeo + sgOff(sk, SG_CX)), ldnt(m.sgeo + sgOff(sk, SG_CX + 1)),
                            ldnt(m.sgeo + sgOff(sk, SG_CX + 2))} - Cc;
#pragma unroll
            for (int k = 0; k < NG; ++k)
                limitFace(lim[k], mx[k], mn[k], dot(d, V3{gx[k], gy[k], gz[k]}), ph.limiter);
        }
        for (int q = e0 + len; q < e1; ++q) {
            const int f = m.cf[q];
            const V3 d = ld3(m.Cf + 3 * (size_t)f) - Cc;
#pragma unroll
            for (int k = 0; k < NG; ++k)
                limitFace(lim[k], mx[k], mn[k], dot(d, V3{gx[k], gy[k], gz[k]}), ph.limiter);
        }
#pragma unroll
        for (int k = 0; k < NG; ++k) { gx[k] = gx[k] * lim[k]; gy[k] = gy[k] * lim[k]; gz[k] = gz[k] * lim[k]; }
        if (limo) stRow(limo + (size_t)c * NFLD, lim);
    } else if (limo) {
        double one[NG];
#pragma unroll
        for (int k = 0; k < NG; ++k) one[k] = 1.0;
        stRow(limo + (size_t)c * NFLD, one);
    }
    double o[3 * NG];
#pragma unroll
    for (int k = 0; k < NG; ++k) { o[3 * k] = gx[k]; o[3 * k + 1] = gy[k]; o[3 * k + 2] = gz[k]; }
    stRow(s.grd + (size_t)c * NFLD * 3, o);
}

// W: minimum waves per SIMD the register allocation must allow (1: the
// compiler's choice; 4: at most 128 VGPRs); the autotuner picks.
template <int NG, int W = 1>
__global__ void __launch_bounds__(256, W) k_grad(MeshView m, Phys ph, StateView s, double* __restrict__ gun,
                                                 double* __restrict__ limo) {
    gradBody<NG>(m, ph, s, gun, limo);
}
template <int NG>
__global__ void __launch_bounds__(256) k_grad_hb(const HBInst* __restrict__ I, Phys ph) {
    const HBInst q = I[blockIdx.y];
    gradBody<NG>(q.m, ph, q.s, nullptr, nullptr);
}

__device__ __forceinline__ double soundSpeed(const Phys& ph, double T) {
    const double psi = 1.0 / (ph.R * T);
    return sqrt(ph.gamma / psi);
}

// Convective flux of face f (internal or boundary) -> (rho, rhoE, rhoU).
// rec (internal and cyclic faces): the reconstructed Ux, Uy, Uz, E of the
// owner side, then of the neighbour side, for the Jacobian (the same
// values the flux used; E reconstructed once more for AUSM+up, whose flux
// uses c* instead).
// Internal faces take the field rows of their owner o and neighbour n
// (fo, fn), which the caller gathers once for the flux and its own terms.
// um: the face's normal frame velocity (rotating zone; 0 otherwise).
__device__ __forceinline__ void faceFlux(const MeshView& m, const Phys& ph, const StateView& s, int f, int o, int n,
                                         const double (&fo)[NFLD], const double (&fn)[NFLD], double um, double& fr,
                                         double& fe, V3& fu, double* rec) {
    const V3 S{ldnt(m.Sf + 3 * (size_t)f), ldnt(m.Sf + 3 * (size_t)f + 1), ldnt(m.Sf + 3 * (size_t)f + 2)};
    const double ms = ldnt(m.magSf + f);
    double rL, rR, pL, pR, uxL, uxR, uyL, uyR, uzL, uzR, HL, HR, xL, xR, eL = 0, eR = 0;
    double gI = ph.gamma;  // fvc::interpolate(gamma): exact on internal faces
    if (f >= m.nI && m.bKind[f - m.nI] == PK_CYCLIC) {
        const int kc = m.cycOf[f - m.nI];
        reconCyclic(m, s, ph.recon, f, o, kc, F_UX, uxL, uxR);
        reconCyclic(m, s, ph.recon, f, o, kc, F_UY, uyL, uyR);
        reconCyclic(m, s, ph.recon, f, o, kc, F_UZ, uzL, uzR);
        reconCyclic(m, s, ph.recon, f, o, kc, F_P, pL, pR);
        reconCyclic(m, s, ph.recon, f, o, kc, F_RHO, rL, rR);
        reconCyclic(m, s, ph.recon, f, o, kc, F_H, HL, HR);
        reconCyclic(m, s, ph.recon, f, o, kc, ph.flux == FLUX_ROE ? F_E : F_CS, xL, xR);
        if (ph.flux == FLUX_HLLC) reconCyclic(m, s, ph.recon, f, o, kc, F_E, eL, eR);
        else if (ph.flux == FLUX_ROE) { eL = xL; eR = xR; }
        else reconCyclic(m, s, ph.recon, f, o, kc, F_E, eL, eR);
        gI = m.w[f] * ph.gamma + (1.0 - m.w[f]) * (plFace(m, s, kc) ? s.plFar[(size_t)kc * PLW + PL_G] : ph.gamma);
    } else if (f < m.nI) {
        double L[NFLD], R[NFLD];
        reconRows(m, s, ph.recon, ph.flux == FLUX_ROE ? NFLD - 1 : NFLD, f, o, n, fo, fn, L, R);
        uxL = L[F_UX]; uxR = R[F_UX];
        uyL = L[F_UY]; uyR = R[F_UY];
        uzL = L[F_UZ]; uzR = R[F_UZ];
        pL = L[F_P]; pR = R[F_P];
        rL = L[F_RHO]; rR = R[F_RHO];
        HL = L[F_H]; HR = R[F_H];
        const int kx = ph.flux == FLUX_ROE ? F_E : F_CS;
        xL = L[kx]; xR = R[kx];
        eL = L[F_E]; eR = R[F_E];
    } else {
        const double* bf = s.bfld + (size_t)(f - m.nI) * NFLD;
        rL = rR = bf[F_RHO];
        pL = pR = bf[F_P];
        uxL = uxR = bf[F_UX];
        uyL = uyR = bf[F_UY];
        uzL = uzR = bf[F_UZ];
        HL = HR = bf[F_H];
        xL = xR = bf[ph.flux == FLUX_ROE ? F_E : F_CS];
        eL = eR = bf[F_E];
    }
    const V3 UL{uxL, uyL, uzL}, UR{uxR, uyR, uzR};
    rec[0] = uxL; rec[1] = uyL; rec[2] = uzL; rec[3] = eL;
    rec[4] = uxR; rec[5] = uyR; rec[6] = uzR; rec[7] = eR;
    if (ph.flux == FLUX_AUSM)
        ausmFace(ph.lowMach, S, ms, UL, UR, xL, xR, pL, pR, rL, rR, HL, HR, um, fr, fu, fe);
    else if (ph.flux == FLUX_ROE)
        roeFace(gI, ph.entropyFix, S, ms, rL, rR, pL, pR, UL, UR, xL, xR, HL, HR, um, fr, fu, fe);
    else
        hllcFace(gI, S, ms, rL, rR, pL, pR, UL, UR, xL, xR, eL, eR, HL, HR, um, fr, fu, fe);
}

// Face terms, one thread per face (each face evaluated once):
//   FT_CONV..+4  convective flux (rho, rhoE, rhoUx, rhoUy, rhoUz)
//   FT_LAP..+2   |S| mu_f snGrad(U)            (laplacian(muEff,U))
//   FT_TAU..+2   Sf & tauMC_f                  (div(tauMC))
//   FT_SIG       (mu_f grad(U)_f + tauMC_f) & U_f & Sf (div(sigmaDotU & Sf))
//   FT_LAPE      |S| alpha_f snGrad(e)         (laplacian(alphaEff,e))
//   FT_LAM       lambda_f = interp(c) + |interp(U).n - um| (internal) or
//                c_b + |U_b.n - um| (boundary); um the normal frame
//                velocity of a rotating zone, else 0
//   FT_FRDT      nonOrthDeltaCoeffs*lambda (internal), wall-face
//                0.5*nonOrthDeltaCoeffs*(c_P + |U_P.n - um|), else -1
//   FT_REC..+7   reconstructed Ux,Uy,Uz,E on the owner side, then the
//                neighbour side (internal faces)
//   FT_SV        0.5*lambdaVisc*|S|*deltaCoeffs (internal, viscous)
// Cell kernels then sum these in face order (the reference's face loops).
// o: the face's NFT-entry record (entries a face kind does not use are
// left as they are).
__device__ __forceinline__ void faceTermsBody(const MeshView& m, const Phys& ph, const StateView& s, int f,
                                              double* o) {
    double fr, fe;
    V3 fu;
    double rec[8];
    const bool internal = f < m.nI;
    const int ow = internal ? ldnt(m.own + f) : m.own[f], ne = internal ? ldnt(m.nei + f) : -1;
    double fo[NFLD], fn[NFLD];
    if (internal) {
        ldRow(s.fld + (size_t)ow * NFLD, fo);
        ldRow(s.fld + (size_t)ne * NFLD, fn);
    }
    const double um = m.umrf ? m.umrf[f] : 0.0;
    faceFlux(m, ph, s, f, ow, ne, fo, fn, um, fr, fe, fu, rec);
    o[FT_CONV] = fr;
    o[FT_CONV + 1] = fe;
    o[FT_CONV + 2] = fu.x;
    o[FT_CONV + 3] = fu.y;
    o[FT_CONV + 4] = fu.z;
    const V3 S{ldnt(m.Sf + 3 * (size_t)f), ldnt(m.Sf + 3 * (size_t)f + 1), ldnt(m.Sf + 3 * (size_t)f + 2)};
    const double ms = ldnt(m.magSf + f);
    const bool visc = ph.viscous;
    const double mu = ph.mu, alpha = ph.alpha;
    if (internal) {
        const double w = ldnt(m.w + f);
        const double co = soundSpeed(ph, s.T[ow]), cn = soundSpeed(ph, s.T[ne]);
        const double cf = w * (co - cn) + cn;
        const V3 Uo{fo[F_UX], fo[F_UY], fo[F_UZ]}, Un{fn[F_UX], fn[F_UY], fn[F_UZ]};
        const V3 Uf = w * (Uo - Un) + Un;
        const double lam = cf + fabs(dot(Uf, S / ms) - um);
        o[FT_LAM] = lam;
        o[FT_FRDT] = m.nodc[f] * lam;
#pragma unroll
        for (int k = 0; k < 8; ++k) o[FT_REC + k] = rec[k];
        if (visc) {
            double vo[NGV], vn[NGV];
            ldRow(s.gv + (size_t)ow * NGV, vo);
            ldRow(s.gv + (size_t)ne * NGV, vn);
            T9 gUo, gUn;
#pragma unroll
            for (int i = 0; i < 9; ++i) { gUo.t[i] = vo[GV_U + i]; gUn.t[i] = vn[GV_U + i]; }
            const T9 gUf = w * (gUo - gUn) + gUn;
            const T9 tauO = mu * dev2(transposeT(gUo));
            const T9 tauN = mu * dev2(transposeT(gUn));
            const T9 tauf = w * (tauO - tauN) + tauN;
            const double muf = w * (mu - mu) + mu;
            const double alf = w * (alpha - alpha) + alpha;
            const V3 k3{ldnt(m.corr + 3 * (size_t)f), ldnt(m.corr + 3 * (size_t)f + 1), ldnt(m.corr + 3 * (size_t)f + 2)};
            const V3 sn = m.nodc[f] * (Un - Uo) + vdott(k3, gUf);
            const V3 lap = (muf * sn) * ms;
            const V3 dt = vdott(S, tauf);
            const V3 sig = tdotv(muf * gUf + tauf, Uf);
            const V3 geO{vo[GV_E], vo[GV_E + 1], vo[GV_E + 2]}, geN{vn[GV_E], vn[GV_E + 1], vn[GV_E + 2]};
            const V3 gef = w * (geO - geN) + geN;
            const double sne = m.nodc[f] * (s.ecalc[ne] - s.ecalc[ow]) + dot(k3, gef);
            o[FT_LAP] = lap.x; o[FT_LAP + 1] = lap.y; o[FT_LAP + 2] = lap.z;
            o[FT_TAU] = dt.x; o[FT_TAU + 1] = dt.y; o[FT_TAU + 2] = dt.z;
            o[FT_SIG] = dot(sig, S);
            o[FT_LAPE] = (alf * sne) * ms;
            const double rf = w * (fo[F_RHO] - fn[F_RHO]) + fn[F_RHO];
            const double lv = (muf + alf) / rf;
            o[FT_SV] = 0.5 * lv * ms * m.dc[f];
        }
    } else if (m.bKind[f - m.nI] == PK_CYCLIC) {
        // coupled interpolation of c and U (setCoAndDeltaT.H:44-56,78-89,
        // convectiveFluxScheme.C:509-524); reconstructed U, E for the
        // interface block (owner side, far side)
        // (the far side lagged on a phase-lagged face)
        const int kc = m.cycOf[f - m.nI];
        const int nc = m.cycCell[kc];
        const double w = m.w[f];
        const bool pl = plFace(m, s, kc);
        const double* pf = pl ? s.plFar + (size_t)kc * PLW : nullptr;
        const double cf = w * soundSpeed(ph, s.T[ow]) + (1.0 - w) * (pl ? pf[PL_C] : soundSpeed(ph, s.T[nc]));
        const V3 Un = pl ? V3{pf[PL_UV], pf[PL_UV + 1], pf[PL_UV + 2]} : rot3(m.cycR + 9 * (size_t)kc, cellU(s.fld, nc));
        const V3 Uf = w * cellU(s.fld, ow) + (1.0 - w) * Un;
        const double lam = cf + fabs(dot(Uf, S / ms) - um);
        o[FT_LAM] = lam;
        o[FT_FRDT] = m.nodc[f] * lam;
#pragma unroll
        for (int k = 0; k < 8; ++k) o[FT_REC + k] = rec[k];
        if (visc) {
            // coupled viscous terms: every interpolation w v_P + (1 - w)
            // v_far with the far value through patchNeighbourField: U
            // (lagged on a phase-lagged face), grad(U) and tauMC as tensors
            // (forwardT & T & forwardT^T), grad(eCalc) as a vector; the
            // snGrad correction interpolates the component gradients
            // grad(U.component(i)), each rotated as a vector (forwardT &
            // grad(U)); the coupled face's nonOrthDeltaCoeffs, correction
            // vector and deltaCoeffs
            const double* R = m.cycR + 9 * (size_t)kc;
            const T9 Rt{{R[0], R[1], R[2], R[3], R[4], R[5], R[6], R[7], R[8]}};
            double vo[NGV], vn[NGV];
            ldRow(s.gv + (size_t)ow * NGV, vo);
            ldRow(s.gv + (size_t)nc * NGV, vn);
            T9 gUo, gUn;
#pragma unroll
            for (int i = 0; i < 9; ++i) { gUo.t[i] = vo[GV_U + i]; gUn.t[i] = vn[GV_U + i]; }
            const T9 tauO = mu * dev2(transposeT(gUo));
            const T9 tauN = mu * dev2(transposeT(gUn));
            const T9 gUc = w * gUo + (1.0 - w) * tdott(Rt, gUn);
            const T9 gUf = w * gUo + (1.0 - w) * tdott(tdott(Rt, gUn), transposeT(Rt));
            const T9 tauf = w * tauO + (1.0 - w) * tdott(tdott(Rt, tauN), transposeT(Rt));
            const V3 UP = cellU(s.fld, ow);
            const double muf = w * mu + (1.0 - w) * mu;
            const double alf = w * alpha + (1.0 - w) * alpha;
            const V3 k3 = ld3(m.corr + 3 * (size_t)f);
            const V3 sn = m.nodc[f] * (Un - UP) + vdott(k3, gUc);
            const V3 lap = (muf * sn) * ms;
            const V3 dt = vdott(S, tauf);
            const V3 sig = tdotv(muf * gUf + tauf, Uf);
            const V3 geO{vo[GV_E], vo[GV_E + 1], vo[GV_E + 2]}, geN{vn[GV_E], vn[GV_E + 1], vn[GV_E + 2]};
            const V3 gef = w * geO + (1.0 - w) * rot3(R, geN);
            const double sne = m.nodc[f] * (s.ecalc[nc] - s.ecalc[ow]) + dot(k3, gef);
            o[FT_LAP] = lap.x; o[FT_LAP + 1] = lap.y; o[FT_LAP + 2] = lap.z;
            o[FT_TAU] = dt.x; o[FT_TAU + 1] = dt.y; o[FT_TAU + 2] = dt.z;
            o[FT_SIG] = dot(sig, S);
            o[FT_LAPE] = (alf * sne) * ms;
            const double rN = pl ? pf[F_RHO] : s.fld[(size_t)nc * NFLD + F_RHO];
            const double rf = w * s.fld[(size_t)ow * NFLD + F_RHO] + (1.0 - w) * rN;
            const double lv = (muf + alf) / rf;
            o[FT_SV] = 0.5 * lv * ms * m.dc[f];
        }
    } else {
        const int b = f - m.nI;
        const int c = m.own[f];
        if (!(s.halt && *s.halt)) s.bcPhi[b] = fr;  // the registered phi the derived conditions read
        const double* bf = s.bfld + (size_t)b * NFLD;
        const V3 UB{bf[F_UX], bf[F_UY], bf[F_UZ]};
        o[FT_LAM] = soundSpeed(ph, s.bT[b]) + fabs(dot(UB, S / ms) - um);
        o[FT_FRDT] = m.bKind[b] == PK_WALL
                         ? 0.5 * m.nodc[f] * (soundSpeed(ph, s.T[c]) + fabs(dot(cellU(s.fld, c), S / ms) - um))
                         : -1.0;
        if (visc) {
            const V3 n = S / ms;
            const V3 Uc = cellU(s.fld, c);
            V3 snU;
            const int ut = m.bUT[b];
            if (ut == BC_ZG) snU = V3{0, 0, 0};
            else if (ut == BC_SLIP) snU = (reflect(n, Uc) - Uc) * (m.dc[f] / 2.0);
            else snU = m.dc[f] * (UB - Uc);
            T9 gUc;
#pragma unroll
            for (int i = 0; i < 9; ++i) gUc.t[i] = s.gv[(size_t)c * NGV + GV_U + i];
            const T9 gUb = gUc + outer(n, snU - vdott(n, gUc));
            const T9 taub = mu * dev2(transposeT(gUb));
            const V3 lap = (mu * snU) * ms;
            const V3 dt = vdott(S, taub);
            const V3 sig = tdotv(mu * gUb + taub, UB);
            const double rhoB = bf[F_RHO];
            const double rhoEB = rhoB * (ph.Cv * s.bT[b] + 0.5 * magSqr(UB));
            const double eb = rhoEB / rhoB - 0.5 * magSqr(UB);
            o[FT_LAP] = lap.x; o[FT_LAP + 1] = lap.y; o[FT_LAP + 2] = lap.z;
            o[FT_TAU] = dt.x; o[FT_TAU + 1] = dt.y; o[FT_TAU + 2] = dt.z;
            o[FT_SIG] = dot(sig, S);
            o[FT_LAPE] = (alpha * (m.dc[f] * (eb - s.ecalc[c]))) * ms;
        }
    }
}

// Face records of workgroup blk: assembled in LDS (row stride NFT + 1
// doubles, zero where a face kind writes nothing, as the buffer holds from
// its allocation), then stored as one contiguous range of whole records.
// Each thread storing its own record directly issues NFT 8-byte stores at a
// 192-byte stride across the wave, which reach memory as partial cache
// lines, about twice the record bytes.
__device__ __forceinline__ void faceTermsBlock(const MeshView& m, const Phys& ph, const StateView& s, int blk) {
    extern __shared__ double shFT[];
    const int bs = blockDim.x;
    const int f0 = blk * bs;
    double* o = shFT + (size_t)threadIdx.x * (NFT + 1);
#pragma unroll
    for (int k = 0; k < NFT; ++k) o[k] = 0.0;
    const int f = f0 + threadIdx.x;
    if (f < m.nF && !(f >= m.nI && m.bKind[f - m.nI] == PK_EMPTY)) faceTermsBody(m, ph, s, f, o);
    __syncthreads();
    const int nRec = min(bs, m.nF - f0);
    double* g = s.ft + (size_t)f0 * NFT;
    for (int i = threadIdx.x; i < nRec * NFT; i += bs) {
        const int r = i / NFT, k = i - r * NFT;
        stnt(g + i, shFT[r * (NFT + 1) + k]);
    }
}

template <int W>
__global__ void __launch_bounds__(256, W) k_face_terms(MeshView m, Phys ph, StateView s) {
    faceTermsBlock(m, ph, s, xcdBlock());
}
__global__ void __launch_bounds__(256) k_face_terms_hb(const HBInst* __restrict__ I, Phys ph) {
    const HBInst q = I[blockIdx.y];
    faceTermsBlock(q.m, ph, q.s, xcdBlock());
}

// Residual b = R*V (residualsUpdate.H) and rPseudoDeltaT
// (setCoAndDeltaT.H:57-141) of a cell from the face terms: add() each face
// record in face order (plus: the cell owns the face, or it is a boundary
// face), then store().
struct CellSums {
    double ar = 0, ae = 0;
    V3 au{0, 0, 0}, aLap{0, 0, 0}, aTau{0, 0, 0};
    double aSig = 0, aLapE = 0, rdt = 0.0;
    __device__ __forceinline__ void add(const double* t, bool plus, bool visc) {
        const double fr = t[FT_CONV], fe = t[FT_CONV + 1];
        const V3 fu{t[FT_CONV + 2], t[FT_CONV + 3], t[FT_CONV + 4]};
        if (plus) { ar += fr; ae += fe; au = au + fu; }
        else { ar -= fr; ae -= fe; au = au - fu; }
        if (t[FT_FRDT] >= 0) rdt = fmax(rdt, t[FT_FRDT]);
        if (visc) {
            const V3 lap{t[FT_LAP], t[FT_LAP + 1], t[FT_LAP + 2]};
            const V3 dt{t[FT_TAU], t[FT_TAU + 1], t[FT_TAU + 2]};
            if (plus) { aLap = aLap + lap; aTau = aTau + dt; aSig += t[FT_SIG]; aLapE += t[FT_LAPE]; }
            else { aLap = aLap - lap; aTau = aTau - dt; aSig -= t[FT_SIG]; aLapE -= t[FT_LAPE]; }
        }
    }
    // writes b and rDT of cell c; returns rDT
    __device__ __forceinline__ double store(const MeshView& m, const StateView& s, int c, bool visc) const {
        const double Vc = m.V[c];
        const int nC = m.nC;
        double Rr = -(ar / Vc);
        double Re = -(ae / Vc);
        V3 Ru = V3{-(au.x / Vc), -(au.y / Vc), -(au.z / Vc)};
        if (visc) {
            Ru = Ru + aLap / Vc;
            Ru = Ru + aTau / Vc;
            Re += aSig / Vc;
            Re += aLapE / Vc;
        }
        s.b[c] = Rr * Vc;
        s.b[(size_t)nC + c] = Re * Vc;
        s.b[2 * (size_t)nC + c] = Ru.x * Vc;
        s.b[3 * (size_t)nC + c] = Ru.y * Vc;
        s.b[4 * (size_t)nC + c] = Ru.z * Vc;
        if (m.omC) {
            // rotating zone: -(Omega ^ rho U) V, which addMRFSource adds to
            // the coupled system's source (convectiveFluxScheme.C:123-129)
            const double rho = s.fld[(size_t)c * NFLD + F_RHO];
            const V3 U = cellU(s.fld, c);
            const V3 sc = cross(ld3(m.omC + 3 * (size_t)c), V3{rho * U.x, rho * U.y, rho * U.z}) * Vc;
            s.b[2 * (size_t)nC + c] -= sc.x;
            s.b[3 * (size_t)nC + c] -= sc.y;
            s.b[4 * (size_t)nC + c] -= sc.z;
        }
        const double r = rdt / s.pseudoCo[c];
        s.rDT[c] = r;
        return r;
    }
};

__device__ __forceinline__ void cellResidualBody(const MeshView& m, const Phys& ph, const StateView& s) {
    const int c = xcdBlock() * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const bool visc = ph.viscous;
    const int e0 = m.cfOff[c], e1 = m.cfOff[c + 1];
    const int nLo = m.nLower[c];
    CellSums cs;
    for (int q = e0; q < e1; ++q) {
        const int f = m.cf[q];
        cs.add(s.ft + (size_t)f * NFT, q - e0 >= nLo, visc);  // boundary and upper faces follow the lower faces
    }
    cs.store(m, s, c, visc);
}

__global__ void __launch_bounds__(256) k_cell_residual(MeshView m, Phys ph, StateView s) {
    cellResidualBody(m, ph, s);
}
__global__ void __launch_bounds__(256) k_cell_residual_hb(const HBInst* __restrict__ I, Phys ph) {
    const HBInst q = I[blockIdx.y];
    cellResidualBody(q.m, ph, q.s);
}

// Face fluxes of every face (convectiveFluxScheme::calcFlux output) from
// the face terms; fluxes of empty faces are zero.
__global__ void k_face_flux(MeshView m, StateView s, double* phi, double* phiU, double* phiE) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.nF) return;
    const bool empty = f >= m.nI && m.bKind[f - m.nI] == PK_EMPTY;
    const double* t = s.ft + (size_t)f * NFT;
    phi[f] = empty ? 0.0 : t[FT_CONV];
    phiE[f] = empty ? 0.0 : t[FT_CONV + 1];
    for (int d = 0; d < 3; ++d) phiU[3 * (size_t)f + d] = empty ? 0.0 : t[FT_CONV + 2 + d];
}

// Row assembly of the coupled Jacobian: off-diagonal 5x5 blocks of cell c
// into its sliced-ELL slots, the diagonal block, and the LU-SGS scalar
// reciprocal diagonal.  Reference: convectiveFluxScheme::addFluxTerms,
// addDissipationJacobian, addBoundaryTerms, addTemporalTerms
// (convectiveFluxScheme.C:219-546), viscousFluxScheme::addFluxTerms
// (viscousFluxScheme.C:220-247), blockFvMatrix::insertBlock /
// insertDissipationBlock (blockFvMatrix.C:212-326), lusgs::lusgs
// (lusgs.C:50-125).  The diagonal accumulates the convective part, the
// dissipation and the viscous parts separately and combines them in the
// order the reference adds its temporary matrices.
// RES: the cell's residual sums (cellResidualBody) are formed in the same
// face walk, from the same face-record loads, and rPseudoDeltaT is taken
// from them (one pass over the records instead of two; same bits).
// HBC (harmonic balance, instance J of the instances I): with RES the
// spectral source -V sum_K D_JK W_K is added to b into the coupled
// right-hand side (HBZone::addSource, HBZoneTemplates.C:39-93), and V D_JJ
// to the diagonal block as stored (HBZone::addBlock, HBZone.C:435-518);
// the instance's own scalar diagonal is not formed (the LU-SGS sweeps of
// the coupled system take the shared one, k_hb_shared_diag).
template <typename TB, bool RES, bool HBC = false>
__device__ __forceinline__ void cellJacobianBody(const MeshView& m, const Phys& ph, const StateView& s,
                                                 TB* __restrict__ blk, TB* __restrict__ dg,
                                                 double* __restrict__ rD, int* __restrict__ err,
                                                 const HBInst* __restrict__ I = nullptr, int J = 0,
                                                 const HBCouple* hc = nullptr) {
    const int c = xcdBlock() * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const double g = ph.gamma;
    double D[NB];
#pragma unroll
    for (int e = 0; e < NB; ++e) D[e] = 0.0;
    double diss = 0.0, vis = 0.0;
    const bool visc = ph.viscous;
    const int e0 = m.cfOff[c], e1 = m.cfOff[c + 1];
    const int sl = c >> 6, lane = c & 63;
    const size_t rowSlot0 = (size_t)m.slicePtr[sl];
    const size_t rs = rowSlot0 * 64 + lane;
    const int nLo = m.nLower[c];
    CellSums cs;
    int slot = 0;
    // rotating zone: fvj::div(w, um |S|) (upper sf (1 - w), lower -sf w,
    // its own negSumDiag mrfD, coupled faces diag -= -sf w), subtracted
    // from the rho, rhoU and rhoE diagonal sub-blocks after the central part
    // (convectiveFluxScheme.C:477-481, blockFvOperatorsTemplates.C:148-200)
    const bool mrf = m.umrf != nullptr;
    double mrfD = 0.0;
    for (int q = e0; q < e1; ++q) {
        const int f = m.cf[q];
        // the face record, gathered whole (12 128-bit loads)
        double t[NFT];
        if (RES || f < m.nI || m.nCyc) ldRow(s.ft + (size_t)f * NFT, t);
        if (RES) cs.add(t, q - e0 >= nLo, visc);
        if (f >= m.nI) {
            // cyclic face: interface block (interfacesUpper: 1/2|S| A_R -
            // 1/2|S| lambda I) and diag -= interfacesLower (+1/2|S| A_L),
            // after all internal faces as in insertBlock
            // (blockFvMatrix.C:241-265, :297-323); gamma interpolated
            if (m.nCyc == 0 || m.bKind[f - m.nI] != PK_CYCLIC) continue;
            const int kc = m.cycOf[f - m.nI];
            const double ms = m.magSf[f];
            const V3 nrm = ld3(m.Sf + 3 * (size_t)f) / ms;
            const double gf = m.w[f] * g + (1.0 - m.w[f]) * (plFace(m, s, kc) ? s.plFar[(size_t)kc * PLW + PL_G] : g);
            double A[NB];
            fluxJacobian(gf, V3{t[FT_REC + 4], t[FT_REC + 5], t[FT_REC + 6]}, t[FT_REC + 7], nrm, A);
            const double hb = 0.5 * ms;
            const double v = 0.5 * ms * t[FT_LAM];
            const double sf = mrf ? m.umrf[f] * ms : 0.0;
            // the viscous LF coefficient on the coupled face: interface
            // diagonal - sv, owner diagonal + sv (after the internal faces)
            const double sv = visc ? t[FT_SV] : 0.0;
            TB* ib = blk + m.cycBase + (size_t)kc * NB;
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                double x = hb * A[e];
                if (e == 0 || e == 6 || e == 12 || e == 18 || e == 24) {
                    if (mrf) x = x - sf * (1 - m.w[f]);
                    x = x - v;
                    if (visc) x = x - sv;
                }
                ib[e] = (TB)x;
            }
            if (visc) vis -= sv;
            fluxJacobian(gf, V3{t[FT_REC], t[FT_REC + 1], t[FT_REC + 2]}, t[FT_REC + 3], nrm, A);
            const double hd = -0.5 * ms;
#pragma unroll
            for (int e = 0; e < NB; ++e) D[e] -= hd * A[e];
            if (mrf) mrfD -= -sf * m.w[f];
            continue;
        }
        const size_t sk = rs + 64 * (size_t)slot;
        const bool isOwn = slot >= nLo;
        const V3 S{ldnt(m.sgeo + sgOff(sk, SG_SX)), ldnt(m.sgeo + sgOff(sk, SG_SX + 1)),
                   ldnt(m.sgeo + sgOff(sk, SG_SX + 2))};
        const double ms = ldnt(m.sgeo + sgOff(sk, SG_MAG));
        const V3 nrm = S / ms;
        // own side: block = upper (A_R), diagonal -= lower (A_L);
        // neighbour side: block = lower (A_L), diagonal -= upper (A_R)
        const int bs = isOwn ? FT_REC + 4 : FT_REC;
        const int ds = isOwn ? FT_REC : FT_REC + 4;
        double A[NB];
        const double hb = isOwn ? 0.5 * ms : -0.5 * ms;
        const double v = 0.5 * ms * t[FT_LAM];
        const double sv = visc ? t[FT_SV] : 0.0;
        // the frame divergence: upper sf (1 - w) in the owner's row, lower
        // -sf w in the neighbour's; negSumDiag gives the owner -lower and
        // the neighbour -upper
        double mo = 0.0, md = 0.0;
        if (mrf) {
            const double sf = m.umrf[f] * ms, w = ldnt(m.sgeo + sgOff(sk, SG_W));
            const double up = sf * (1 - w), lo = -sf * w;
            mo = isOwn ? up : lo;
            md = isOwn ? lo : up;
        }
        if (m.blkGen) {
            // generators: the solver kernels rebuild h A(U, E, n) - v I - sv I
            const size_t gb = ((rowSlot0 + slot) * NGEN) * 64 + lane;
            const double gv[NGEN] = {t[bs], t[bs + 1], t[bs + 2], t[bs + 3], nrm.x, nrm.y, nrm.z, hb, v, sv};
#pragma unroll
            for (int e = 0; e < NGEN; ++e) stnt(blk + gb + (size_t)e * 64, (TB)gv[e]);
            // the frame divergence term beside them (a frame zone)
            if (mrf && m.genMo) stnt((TB*)m.genMo + (rowSlot0 + slot) * 64 + lane, (TB)mo);
        } else {
            fluxJacobian(g, V3{t[bs], t[bs + 1], t[bs + 2]}, t[bs + 3], nrm, A);
            const size_t base = ((rowSlot0 + slot) * NB) * 64 + lane;
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                double x = hb * A[e];
                if (e == 0 || e == 6 || e == 12 || e == 18 || e == 24) {
                    if (mrf) x = x - mo;
                    x = x - v;
                    if (visc) x = x - sv;
                }
                blk[base + (size_t)e * 64] = (TB)x;
            }
        }
        fluxJacobian(g, V3{t[ds], t[ds + 1], t[ds + 2]}, t[ds + 3], nrm, A);
        const double hd = isOwn ? -0.5 * ms : 0.5 * ms;
#pragma unroll
        for (int e = 0; e < NB; ++e) D[e] -= hd * A[e];
        // negSumDiag of the frame divergence, in face order
        if (mrf) mrfD -= md;
        diss -= v;
        if (visc) vis -= sv;
        ++slot;
    }
    if (mrf) { D[0] = D[0] - mrfD; D[6] = D[6] - mrfD; D[12] = D[12] - mrfD; D[18] = D[18] - mrfD; D[24] = D[24] - mrfD; }
    double rDTc = 0.0;
    if (RES) rDTc = cs.store(m, s, c, visc);
    if constexpr (HBC && RES) {
        const size_t nC = m.nC, vs = 5 * nC;
        const int nO = hc->nO;
        for (int q = 0; q < 5; ++q) {
            double src = 0.0;
            for (int K = 0; K < nO; ++K) src -= m.V[c] * hc->D[J * nO + K] * I[K].s.W[(size_t)q * nC + c];
            hc->ball[J * vs + (size_t)q * nC + c] = s.b[(size_t)q * nC + c] + src;
            hc->Wall[J * vs + (size_t)q * nC + c] = s.W[(size_t)q * nC + c];
        }
    }
    // boundary dissipation (all non-empty patches) follows the internal faces
    for (int q = e0; q < e1; ++q) {
        const int f = m.cf[q];
        if (f < m.nI) continue;
        diss -= 0.5 * m.magSf[f] * s.ft[(size_t)f * NFT + FT_LAM];
    }
    D[0] = D[0] - diss; D[6] = D[6] - diss; D[12] = D[12] - diss; D[18] = D[18] - diss; D[24] = D[24] - diss;
    // boundary terms (addBoundaryTerms), patch/face order
    const double Cv = ph.Cv;
    const double rhoI = s.fld[(size_t)c * NFLD + F_RHO];
    const V3 UI = cellU(s.fld, c);
    for (int q = e0; q < e1; ++q) {
        const int f = m.cf[q];
        if (f < m.nI) continue;
        const int b = f - m.nI;
        if (m.bKind[b] == PK_CYCLIC) continue;  // coupled patches carry no boundary terms
        const V3 Sf = ld3(m.Sf + 3 * (size_t)f);
        const V3 n = Sf / m.magSf[f];
        // valueInternalCoeffs: 0 for fixedValue, totalPressure and
        // totalTemperature, 1 for zeroGradient, 1 - |n| for slip, and
        // 1 - sqrt(|diag(valueFraction)|) for pressureInletOutletVelocity
        // (valueFraction I - n n after an inflow update, else 0)
        const int pt = m.bPT[b], tt = m.bTT[b], ut = m.bUT[b];
        const double pVIC = (pt == BC_FIXED || pt == BC_TOTAL_P) ? 0.0 : 1.0;
        const double tVIC = (tt == BC_FIXED || tt == BC_TOTAL_T) ? 0.0 : 1.0;
        V3 uVIC;
        if (ut == BC_FIXED) uVIC = V3{0, 0, 0};
        else if (ut == BC_SLIP) uVIC = V3{1 - fabs(n.x), 1 - fabs(n.y), 1 - fabs(n.z)};
        else if (ut == BC_PIOV && s.bcIn[b])
            uVIC = V3{1 - sqrt(fabs(1 - n.x * n.x)), 1 - sqrt(fabs(1 - n.y * n.y)), 1 - sqrt(fabs(1 - n.z * n.z))};
        else uVIC = V3{1, 1, 1};
        const double* bf = s.bfld + (size_t)b * NFLD;
        const double rhoB = bf[F_RHO], pB = bf[F_P], TB_ = s.bT[b];
        const V3 UB{bf[F_UX], bf[F_UY], bf[F_UZ]};
        double UrelSf = dot(UB, Sf);
        if (mrf) UrelSf -= m.umrf[f] * m.magSf[f];  // convectiveFluxScheme.C:94
        const double rhoEB = rhoB * (Cv * TB_ + 0.5 * magSqr(UB));
        const V3 SfV{Sf.x * uVIC.x, Sf.y * uVIC.y, Sf.z * uVIC.z};
        const double dCdp = rhoB / pB * UrelSf * pVIC;
        const V3 dCdU = rhoB * SfV;
        const double dCdT = -rhoB / TB_ * UrelSf * tVIC;
        const V3 dMdp = (rhoB / pB * UB * UrelSf + Sf) * pVIC;
        T9 dMdU = outer(rhoB * UB, SfV);
        const V3 dMdUDiag = rhoB * UrelSf * uVIC;
        const V3 dMdT = -rhoB / TB_ * UB * UrelSf * tVIC;
        const double dEdp = (rhoEB / pB * UrelSf + dot(UB, Sf)) * pVIC;
        const V3 UBV{UB.x * uVIC.x, UB.y * uVIC.y, UB.z * uVIC.z};
        const V3 dEdU = SfV * (rhoEB + pB) + rhoB * UrelSf * UBV;
        const double dEdT = UrelSf * (rhoB * Cv - rhoEB / TB_) * tVIC;
        dMdU.t[0] += dMdUDiag.x; dMdU.t[4] += dMdUDiag.y; dMdU.t[8] += dMdUDiag.z;
        const double rhoEI = rhoI * (Cv * s.T[c] + 0.5 * magSqr(UI));
        const double dPdRho = 0.5 * (g - 1) * magSqr(UI);
        const V3 dUdRho{-UI.x / rhoI, -UI.y / rhoI, -UI.z / rhoI};
        const double dTdRho = -1.0 / (Cv * rhoI) * (rhoEI / rhoI - magSqr(UI));
        const V3 dPdRhoU = -(g - 1) * UI;
        const double dUdRhoU = 1.0 / rhoI;
        const V3 dTdRhoU{-UI.x / (Cv * rhoI), -UI.y / (Cv * rhoI), -UI.z / (Cv * rhoI)};
        const double dPdRhoE = g - 1;
        const double dTdRhoE = 1.0 / (Cv * rhoI);
        // continuity row
        D[0] += dCdp * dPdRho; D[0] += dot(dCdU, dUdRho); D[0] += dCdT * dTdRho;
        {
            const V3 t1 = dCdp * dPdRhoU, t2 = dCdU * dUdRhoU, t3 = dCdT * dTdRhoU;
            D[2] += t1.x; D[3] += t1.y; D[4] += t1.z;
            D[2] += t2.x; D[3] += t2.y; D[4] += t2.z;
            D[2] += t3.x; D[3] += t3.y; D[4] += t3.z;
        }
        D[1] += dCdp * dPdRhoE; D[1] += 0.0; D[1] += dCdT * dTdRhoE;
        // momentum rows
        {
            const V3 a = dMdp * dPdRho, bb = tdotv(dMdU, dUdRho), cc = dMdT * dTdRho;
            D[10] += a.x; D[15] += a.y; D[20] += a.z;
            D[10] += bb.x; D[15] += bb.y; D[20] += bb.z;
            D[10] += cc.x; D[15] += cc.y; D[20] += cc.z;
            const T9 A1 = outer(dMdp, dPdRhoU), B1 = dMdU * dUdRhoU, C1 = outer(dMdT, dTdRhoU);
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) D[5 * (2 + i) + 2 + j] += A1.t[3 * i + j];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) D[5 * (2 + i) + 2 + j] += B1.t[3 * i + j];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) D[5 * (2 + i) + 2 + j] += C1.t[3 * i + j];
            const V3 e1v = dMdp * dPdRhoE, e3v = dMdT * dTdRhoE;
            D[11] += e1v.x; D[16] += e1v.y; D[21] += e1v.z;
            D[11] += 0.0; D[16] += 0.0; D[21] += 0.0;
            D[11] += e3v.x; D[16] += e3v.y; D[21] += e3v.z;
        }
        // energy row
        D[5] += dEdp * dPdRho; D[5] += dot(dEdU, dUdRho); D[5] += dEdT * dTdRho;
        {
            const V3 f1 = dEdp * dPdRhoU, f2 = dEdU * dUdRhoU, f3 = dEdT * dTdRhoU;
            D[7] += f1.x; D[8] += f1.y; D[9] += f1.z;
            D[7] += f2.x; D[8] += f2.y; D[9] += f2.z;
            D[7] += f3.x; D[8] += f3.y; D[9] += f3.z;
        }
        D[6] += dEdp * dPdRhoE; D[6] += 0.0; D[6] += dEdT * dTdRhoE;
    }
    // temporal term: ddtCoeff = (rDT*V)/V, diag += ddtCoeff*V
    {
        const double Vc = m.V[c];
        const double ddtCoeff = ((RES ? rDTc : s.rDT[c]) * Vc) / Vc;
        const double dcv = ddtCoeff * Vc;
        D[0] += dcv; D[12] += dcv; D[18] += dcv; D[24] += dcv; D[6] += dcv;
    }
    if (m.omC) {
        // addMRFSource (convectiveFluxScheme.C:130-137): the Coriolis term's
        // derivative on the rhoU block, entries (row, col) xy xz yx yz zx zy
        const double Vc = m.V[c];
        const V3 om = ld3(m.omC + 3 * (size_t)c);
        D[13] -= om.z * Vc; D[14] += om.y * Vc; D[17] += om.z * Vc;
        D[19] -= om.x * Vc; D[22] -= om.y * Vc; D[23] += om.x * Vc;
    }
    if (visc) { D[0] = D[0] - vis; D[12] = D[12] - vis; D[18] = D[18] - vis; D[24] = D[24] - vis; D[6] = D[6] - vis; }
    if constexpr (HBC) {
        // added to the diagonal as stored (FP32 with mixed precision)
        const double VD = m.V[c] * hc->D[J * hc->nO + J];
        for (int e = 0; e < NB; e += 6) D[e] = (double)(TB)D[e] + VD;
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) stnt(dg + diagOff(c, e), (TB)D[e]);
    if constexpr (HBC) return;
    // LU-SGS scalar diagonal
    double r = kGREAT;
    r = 1.0 / fmax(1.0 / r, fabs(D[0]));
    r = 1.0 / fmax(1.0 / r, fabs(D[6]));
    r = 1.0 / fmax(1.0 / r, fabs(D[12]));
    r = 1.0 / fmax(1.0 / r, fabs(D[18]));
    r = 1.0 / fmax(1.0 / r, fabs(D[24]));
    rD[c] = r;
    if (!(r >= kVSMALL)) atomicOr(err, ERR_FLAG_DIAG);
}

template <typename TB, int W = 1, bool RES = false>
__global__ void __launch_bounds__(256, W) k_cell_jacobian(MeshView m, Phys ph, StateView s, TB* __restrict__ blk,
                                                          TB* __restrict__ dg, double* __restrict__ rD,
                                                          int* __restrict__ err) {
    cellJacobianBody<TB, RES>(m, ph, s, blk, dg, rD, err);
}
// The shared LU-SGS diagonal of the coupled harmonic-balance system
// (lusgs::lusgs, lusgs.C:50-125, over the 2 nO scalar and nO vector
// diagonals in coupledMatrix order), one thread per cell, after the
// instances' Jacobian launch has stored their diagonal blocks.
template <typename TB>
__global__ void __launch_bounds__(256) k_hb_shared_diag(const HBInst* __restrict__ I, HBCouple hc, int nC) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    const auto d = [&](int K, int e) { return fabs((double)((const TB*)I[K].dg)[diagOff(c, e)]); };
    double r = kGREAT;
    for (int K = 0; K < hc.nO; ++K) {
        r = 1.0 / fmax(1.0 / r, d(K, 0));
        r = 1.0 / fmax(1.0 / r, d(K, 6));
    }
    for (int K = 0; K < hc.nO; ++K) {
        r = 1.0 / fmax(1.0 / r, d(K, 12));
        r = 1.0 / fmax(1.0 / r, d(K, 18));
        r = 1.0 / fmax(1.0 / r, d(K, 24));
    }
    hc.rD[c] = r;
    if (!(r >= kVSMALL)) atomicOr(hc.err, ERR_FLAG_DIAG);
}

template <typename TB, bool RES = false>
__global__ void __launch_bounds__(256) k_cell_jacobian_hb(const HBInst* __restrict__ I, Phys ph, HBCouple hc) {
    const int J = blockIdx.y;
    const HBInst q = I[J];
    cellJacobianBody<TB, RES, true>(q.m, ph, q.s, (TB*)q.blk, (TB*)q.dg, q.rD, q.err, I, J, &hc);
}

}  // namespace

// Fourier reconstruction of the instances at one time (HBZone::reconstruct,
// HBZoneTemplates.C:328-386): each primitive field of a cell is
// sum_l w_l field_l, summed from zero in instance order.  out [nC][6]:
// rho, Ux, Uy, Uz, T, p.  (For a static mesh the reference's cylindrical
// variant for U gives the same field: the rotation to cylindrical
// components is the same for every instance and commutes with the sum.)
namespace {
__global__ void k_hb_reconstruct(int nC, int nO, const HBInst* __restrict__ I, const double* __restrict__ w,
                                 double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    double r = 0, ux = 0, uy = 0, uz = 0, T = 0, p = 0;
    for (int l = 0; l < nO; ++l) {
        const double* f = I[l].s.fld + (size_t)c * NFLD;
        const double wl = w[l];
        r += wl * f[F_RHO];
        ux += wl * f[F_UX];
        uy += wl * f[F_UY];
        uz += wl * f[F_UZ];
        T += wl * I[l].s.T[c];
        p += wl * f[F_P];
    }
    double* o = out + 6 * (size_t)c;
    o[0] = r; o[1] = ux; o[2] = uy; o[3] = uz; o[4] = T; o[5] = p;
}
}  // namespace

// ------------------------------------------------------------ launchers --
void launchHBReconstruct(int nC, int nO, const HBInst* inst, const double* w, double* out, hipStream_t st) {
    k_hb_reconstruct<<<(nC + 255) / 256, 256, 0, st>>>(nC, nO, inst, w, out);
}


static inline int nblk(int n, int b) { return (n + b - 1) / b; }

void launchUpdate(const MeshView& m, const Phys& ph, const StateView& s, const double* dW, int mode,
                  const double* rho0, const double* U0, const double* T0, hipStream_t st) {
    k_update<<<nblk(m.nC, 256), 256, 0, st>>>(m, ph, s, dW, mode, rho0, U0, T0);
    // boundary conditions follow every state update (updateFields.H:90-105)
    if (m.nB > 0) k_bc_update<<<nblk(m.nB, 256), 256, 0, st>>>(m, ph, s, mode == 0 ? 1 : 0);
}

void launchResidual(const MeshView& m, const Phys& ph, const StateView& s, hipStream_t st, bool cellSums) {
    if (m.nB > 0) k_boundary<<<nblk(m.nB, 256), 256, 0, st>>>(m, ph, s);
    const int bc = m.bsCell, bf = m.bsFace;
    const dim3 gc(xcdGrid(m.nC, bc)), gf(xcdGrid(m.nF, bf));
    // register budget (occ*): 0 the compiler's choice, 1 at most 128 VGPRs
    // (4 waves per SIMD), 2 at most 168 (3 waves)
    if (ph.flux == FLUX_ROE) {
        if (m.occCell == 1) k_grad<7, 4><<<gc, bc, 0, st>>>(m, ph, s, nullptr, nullptr);
        else if (m.occCell == 2) k_grad<7, 3><<<gc, bc, 0, st>>>(m, ph, s, nullptr, nullptr);
        else k_grad<7><<<gc, bc, 0, st>>>(m, ph, s, nullptr, nullptr);
    } else {
        if (m.occCell == 1) k_grad<8, 4><<<gc, bc, 0, st>>>(m, ph, s, nullptr, nullptr);
        else if (m.occCell == 2) k_grad<8, 3><<<gc, bc, 0, st>>>(m, ph, s, nullptr, nullptr);
        else k_grad<8><<<gc, bc, 0, st>>>(m, ph, s, nullptr, nullptr);
    }
    const size_t lds = (size_t)bf * (NFT + 1) * sizeof(double);
    if (m.occFace == 1) k_face_terms<4><<<gf, bf, lds, st>>>(m, ph, s);
    else if (m.occFace == 2) k_face_terms<3><<<gf, bf, lds, st>>>(m, ph, s);
    else k_face_terms<1><<<gf, bf, lds, st>>>(m, ph, s);
    if (cellSums) k_cell_residual<<<gc, bc, 0, st>>>(m, ph, s);
}

void launchResidualHB(const HBInst* inst, int nO, const MeshView& m, const Phys& ph, hipStream_t st,
                      bool cellSums) {
    if (m.nB > 0) k_boundary_hb<<<dim3(nblk(m.nB, 256), nO), 256, 0, st>>>(inst, ph);
    const int bc = m.bsCell, bf = m.bsFace;
    if (ph.flux == FLUX_ROE) k_grad_hb<7><<<dim3(xcdGrid(m.nC, bc), nO), bc, 0, st>>>(inst, ph);
    else k_grad_hb<8><<<dim3(xcdGrid(m.nC, bc), nO), bc, 0, st>>>(inst, ph);
    k_face_terms_hb<<<dim3(xcdGrid(m.nF, bf), nO), bf, (size_t)bf * (NFT + 1) * sizeof(double), st>>>(inst, ph);
    if (cellSums) k_cell_residual_hb<<<dim3(xcdGrid(m.nC, bc), nO), bc, 0, st>>>(inst, ph);
}

void launchJacobianHB(const HBInst* inst, int nO, int fp32, const MeshView& m, const Phys& ph, hipStream_t st,
                      bool cellSums, const HBCouple& hc) {
    const dim3 g(xcdGrid(m.nC, m.bsJac), nO);
    if (cellSums) {
        if (fp32) k_cell_jacobian_hb<float, true><<<g, m.bsJac, 0, st>>>(inst, ph, hc);
        else k_cell_jacobian_hb<double, true><<<g, m.bsJac, 0, st>>>(inst, ph, hc);
    } else {
        if (fp32) k_cell_jacobian_hb<float><<<g, m.bsJac, 0, st>>>(inst, ph, hc);
        else k_cell_jacobian_hb<double><<<g, m.bsJac, 0, st>>>(inst, ph, hc);
    }
    const int nb = (m.nC + 255) / 256;
    if (fp32) k_hb_shared_diag<float><<<nb, 256, 0, st>>>(inst, hc, m.nC);
    else k_hb_shared_diag<double><<<nb, 256, 0, st>>>(inst, hc, m.nC);
}

void launchAdjointForward(const MeshView& m, const Phys& ph, const StateView& s, double* gun, double* limo,
                          hipStream_t st) {
    if (m.nB > 0) k_boundary<<<nblk(m.nB, 256), 256, 0, st>>>(m, ph, s);
    // limited gradients (second order) and the viscous Gauss gradients
    if (ph.recon == 1 || ph.viscous) {
        if (ph.flux == FLUX_ROE) k_grad<7><<<xcdGrid(m.nC, 256), 256, 0, st>>>(m, ph, s, gun, limo);
        else k_grad<8><<<xcdGrid(m.nC, 256), 256, 0, st>>>(m, ph, s, gun, limo);
    }
}

void launchFaceFlux(const MeshView& m, const Phys& ph, const StateView& s, double* phi, double* phiU, double* phiE,
                    hipStream_t st) {
    (void)ph;
    k_face_flux<<<nblk(m.nF, 256), 256, 0, st>>>(m, s, phi, phiU, phiE);
}

template <bool RES>
static void launchCellJacobian(const MeshView& m, const Phys& ph, const StateView& s, void* blk, void* dg, int fp32,
                               double* rD, int* err, hipStream_t st) {
    const int bj = m.bsJac;
    const dim3 g(xcdGrid(m.nC, bj));
    if (fp32) {
        float *b = (float*)blk, *d = (float*)dg;
        if (m.occJac == 1) k_cell_jacobian<float, 4, RES><<<g, bj, 0, st>>>(m, ph, s, b, d, rD, err);
        else if (m.occJac == 2) k_cell_jacobian<float, 3, RES><<<g, bj, 0, st>>>(m, ph, s, b, d, rD, err);
        else k_cell_jacobian<float, 1, RES><<<g, bj, 0, st>>>(m, ph, s, b, d, rD, err);
    } else {
        double *b = (double*)blk, *d = (double*)dg;
        if (m.occJac == 1) k_cell_jacobian<double, 4, RES><<<g, bj, 0, st>>>(m, ph, s, b, d, rD, err);
        else if (m.occJac == 2) k_cell_jacobian<double, 3, RES><<<g, bj, 0, st>>>(m, ph, s, b, d, rD, err);
        else k_cell_jacobian<double, 1, RES><<<g, bj, 0, st>>>(m, ph, s, b, d, rD, err);
    }
}
void launchJacobian(const MeshView& m, const Phys& ph, const StateView& s, void* blk, void* dg, int fp32,
                    double* rD, int* err, hipStream_t st) {
    launchCellJacobian<false>(m, ph, s, blk, dg, fp32, rD, err, st);
}
void launchCellSumsJacobian(const MeshView& m, const Phys& ph, const StateView& s, void* blk, void* dg, int fp32,
                            double* rD, int* err, hipStream_t st) {
    launchCellJacobian<true>(m, ph, s, blk, dg, fp32, rD, err, st);
}

}  // namespace icsf
