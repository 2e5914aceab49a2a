// Device helpers shared by the residual, Jacobian and adjoint kernels:
// cell velocity, the cellLimited face test and the linearUpwind face values.
#pragma once
#include "icsf_math.h"
#include "icsf_views.h"

namespace icsf {

__device__ __forceinline__ V3 cellU(const double* fld, int c) {
    return V3{fld[c * NFLD + F_UX], fld[c * NFLD + F_UY], fld[c * NFLD + F_UZ]};
}

// One face's contribution to a cellLimited limiter (k = 1): the face
// extrapolation ext = (Cf - C).grad, the cell's max/min neighbour
// differences, Venkatakrishnan (limiter 1) or minmod (limiter 2).
__device__ __forceinline__ void limitFace(double& lim, double maxD, double minD, double ext, int limiter) {
    double r = 1;
    if (ext > kSMALL) r = maxD / ext;
    else if (ext < -kSMALL) r = minD / ext;
    double l;
    if (limiter == 1) l = (sqr(r) + 2 * r) / (sqr(r) + r + 2);
    else l = fmin(r, 1.0);
    lim = fmin(lim, l);
}

// L (owner-side) and R (neighbour-side) reconstructed value of field k on
// internal face f (linearUpwind with pos_/neg_ weights).
__device__ __forceinline__ void recon(const MeshView& m, const StateView& s, int recon, int f, int o, int n, int k,
                                      double& L, double& R) {
    const double vo = s.fld[(size_t)o * NFLD + k], vn = s.fld[(size_t)n * NFLD + k];
    double l = 1.0 * (vo - vn) + vn;
    double r = 0.0 * (vo - vn) + vn;
    if (recon == 1) {
        const V3 fc = ld3(m.Cf + 3 * (size_t)f);
        const double* go = s.grd + ((size_t)o * NFLD + k) * 3;
        const double* gn = s.grd + ((size_t)n * NFLD + k) * 3;
        l += dot(fc - ld3(m.C + 3 * (size_t)o), V3{go[0], go[1], go[2]});
        r += dot(fc - ld3(m.C + 3 * (size_t)n), V3{gn[0], gn[1], gn[2]});
    }
    L = l;
    R = r;
}

// recon() of the first nk fields at once from the owner's and the
// neighbour's field rows (fo, fn), each cell's gradient row gathered whole.
__device__ __forceinline__ void reconRows(const MeshView& m, const StateView& s, int recon, int nk, int f, int o,
                                          int n, const double (&fo)[NFLD], const double (&fn)[NFLD],
                                          double (&L)[NFLD], double (&R)[NFLD]) {
#pragma unroll
    for (int k = 0; k < NFLD; ++k) {
        L[k] = 1.0 * (fo[k] - fn[k]) + fn[k];
        R[k] = 0.0 * (fo[k] - fn[k]) + fn[k];
    }
    if (recon == 1) {
        const V3 fc = ld3(m.Cf + 3 * (size_t)f);
        double g[3 * NFLD];
        ldRow(s.grd + (size_t)o * NFLD * 3, g);
        const V3 dO = fc - ld3(m.C + 3 * (size_t)o);
#pragma unroll
        for (int k = 0; k < NFLD; ++k)
            if (k < nk) L[k] += dot(dO, V3{g[3 * k], g[3 * k + 1], g[3 * k + 2]});
        ldRow(s.grd + (size_t)n * NFLD * 3, g);
        const V3 dN = fc - ld3(m.C + 3 * (size_t)n);
#pragma unroll
        for (int k = 0; k < NFLD; ++k)
            if (k < nk) R[k] += dot(dN, V3{g[3 * k], g[3 * k + 1], g[3 * k + 2]});
    }
}

// Whether cyclic face kc is phase-lagged (its far values in s.plFar).
__device__ __forceinline__ bool plFace(const MeshView& m, const StateView& s, int kc) {
    return s.plFar != nullptr && m.plSide[kc] >= 0;
}

// Far-side value of field k on the kc-th cyclic face (cyclicFvPatchField::
// patchNeighbourField as patched by the reference,
// originalOFFiles/constraintFvPatchFields/cyclic/cyclicFvPatchField.C:127-188):
// the far cell's value, and for the velocity components ("U.component(i)")
// component i of forwardT & U of the far cell; on a phase-lagged face the
// lagged record (phaseLagCyclicFvPatchField.C:343-396).
__device__ __forceinline__ double cycNbrValue(const MeshView& m, const StateView& s, int kc, int k) {
    if (plFace(m, s, kc)) return s.plFar[(size_t)kc * PLW + k];
    const double* R = m.cycR + 9 * (size_t)kc;
    const int n = m.cycCell[kc];
    if (k >= F_UX && k <= F_UZ) {
        const V3 u = rot3(R, cellU(s.fld, n));
        return k == F_UX ? u.x : (k == F_UY ? u.y : u.z);
    }
    return s.fld[(size_t)n * NFLD + k];
}

// L (owner side) and R (far side) reconstructed value of field k on cyclic
// face f (linearUpwind::correction, coupled branch): L = v_P + (Cf - C_P) &
// grad_P, R = vn + ((Cf - delta) - C_P) & (forwardT & grad_N).
__device__ __forceinline__ void reconCyclic(const MeshView& m, const StateView& s, int recon, int f, int o, int kc,
                                            int k, double& L, double& R) {
    const double* Rt = m.cycR + 9 * (size_t)kc;
    const int n = m.cycCell[kc];
    double l = s.fld[(size_t)o * NFLD + k];
    double r = cycNbrValue(m, s, kc, k);
    if (recon == 1) {
        const V3 fc = ld3(m.Cf + 3 * (size_t)f);
        const double* go = s.grd + ((size_t)o * NFLD + k) * 3;
        const double* gn = s.grd + ((size_t)n * NFLD + k) * 3;
        l += dot(fc - ld3(m.C + 3 * (size_t)o), V3{go[0], go[1], go[2]});
        const V3 gR = rot3(Rt, V3{gn[0], gn[1], gn[2]});
        r += dot((fc - ld3(m.cycDelta + 3 * (size_t)kc)) - ld3(m.C + 3 * (size_t)o), gR);
    }
    L = l;
    R = r;
}

}  // namespace icsf
