// Linear-algebra kernels of the coupled solve:
//   block SpMV        coupledMatrix::matrixMul -> blockFvMatrix::Amul
//                     (coupledMatrix.C:66-123, blockFvMatrix.C:331-381)
//   LU-SGS            lusgs::precondition (lusgs.C:220-382), level-scheduled
//   block-Jacobi      Jacobi::precondition / JacobiSmoother (Jacobi.C:55-132)
//   GMRES pieces      gmres::solveDelta (gmres.C:772-1108): reductions,
//                     orthogonalisation, Givens, back-substitution, update.
//
// Vectors are SoA [5][nC] in [rho, rhoE, rhoUx, rhoUy, rhoUz] order.
// Off-diagonal 5x5 blocks live in a sliced-ELL layout (64-row slices, one
// wavefront per slice, entry-major inside a slot) so every block load of a
// wavefront is one contiguous 64-element run.
#include "icsf_flux.h"
#include "icsf_kernels.h"
#include "icsf_math.h"
#include "icsf_views.h"

namespace icsf {

namespace {

template <typename T>
__device__ __forceinline__ double ld(const T* p, size_t i) { return (double)p[i]; }

// ---------------------------------------------------------------- SpMV ----
// Per output component three partial sums are kept (scalar-scalar,
// scalar-scalar', scalar/vector-vector parts) and combined in the order
// coupledMatrix::matrixMul adds its sub-block products.
// One block row of y = A (x/scal): the diagonal block, then the
// off-diagonal slots in column order, three partial sums per row (scalar
// rho column, scalar rhoE column, vector rhoU columns) combined as the
// blockFvMatrix sub-block sums (coupledMatrix.C:66-123).  With FWD, the
// slots are all lower neighbours p of a colour-1 row and the forward
// LU-SGS terms L_cp (rD_p y_p) are accumulated from the same block loads
// (three partial sums per row as well); the row result is then
// rD_c (y_c - fwd), i.e. the forward and the first backward level of the
// LU-SGS sweep of a two-colour level structure.
// Off-diagonal block of sliced-ELL slot sk for the row in lane `lane`, in
// arithmetic type A: read from full storage, or generated from the slot's
// NGEN generators in the storage precision (the same bits as the block the
// Jacobian kernel would have stored, for FP64).
// NT: non-temporal loads (the SpMV rows: each slot is read once per pass
// and a full L2 pass lies between reads; keeping these lines out of L2
// leaves it to the gathered vectors).
template <typename TB, typename A, bool GEN, bool NT = false>
__device__ __forceinline__ void slotBlock(const MeshView& m, const TB* __restrict__ blk, size_t sk, int lane,
                                          A b[NB]) {
    if (GEN) {
        const size_t gb = sk * NGEN * 64 + lane;
        TB g[NGEN], t[NB];
#pragma unroll
        for (int e = 0; e < NGEN; ++e)
            g[e] = NT ? __builtin_nontemporal_load(blk + gb + (size_t)e * 64) : blk[gb + (size_t)e * 64];
        const TB* gm = (const TB*)m.genMo;
        TB mo = (TB)0;
        if (gm) mo = NT ? __builtin_nontemporal_load(gm + sk * 64 + lane) : gm[sk * 64 + lane];
        blockFromGen<TB>(g, (TB)m.gamma, t, mo, gm != nullptr);
#pragma unroll
        for (int e = 0; e < NB; ++e) b[e] = (A)t[e];
    } else {
        const size_t bb = sk * NB * 64 + lane;
#pragma unroll
        for (int e = 0; e < NB; ++e) b[e] = (A)blk[bb + (size_t)e * 64];
    }
}

// Cyclic faces of row c: interfacesUpper & (x_far with its momentum part
// transformed by forwardT), added to the three partial sums after the
// internal faces (blockFvMatrix::Amul, blockFvMatrix.C:398-420).  Not part
// of the LU-SGS sweeps (lusgs.C:143-186 visits internal faces only).
template <typename TB, typename A, typename XF>
__device__ __forceinline__ void spmvCyclic(const MeshView& m, const TB* __restrict__ blk, int c, XF X,
                                           A acc[5][3]) {
    const int q1 = m.cycRowOff[c + 1];
    for (int q = m.cycRowOff[c]; q < q1; ++q) {
        const int kc = m.cycRowK[q];
        const int j = m.cycCell[kc];
        const double* R = m.cycR + 9 * (size_t)kc;
        const A x0 = X(0, j), x1 = X(1, j), x2 = X(2, j), x3 = X(3, j), x4 = X(4, j);
        const A v2 = (A)R[0] * x2 + (A)R[1] * x3 + (A)R[2] * x4;
        const A v3 = (A)R[3] * x2 + (A)R[4] * x3 + (A)R[5] * x4;
        const A v4 = (A)R[6] * x2 + (A)R[7] * x3 + (A)R[8] * x4;
        const TB* ib = blk + m.cycBase + (size_t)kc * NB;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const A b0 = (A)ib[5 * i], b1 = (A)ib[5 * i + 1], b2 = (A)ib[5 * i + 2], b3 = (A)ib[5 * i + 3],
                    b4 = (A)ib[5 * i + 4];
            acc[i][0] += b0 * x0;
            acc[i][1] += b1 * x1;
            const A dd = b2 * v2 + b3 * v3 + b4 * v4;
            acc[i][2] += dd;
        }
    }
}

template <typename TB, typename TV, bool FWD, bool GEN>
__device__ __forceinline__ void spmvRow(const MeshView& m, const TB* __restrict__ blk, const TB* __restrict__ dg,
                                        const TV* __restrict__ x, const double* __restrict__ scal,
                                        TV* __restrict__ vout, TV* __restrict__ y, const double* __restrict__ rD,
                                        int c) {
    const size_t nC = m.nC;
    typedef TV A;
    const A beta = scal ? (A)(*scal) : (A)1;
    auto X = [&](int k, int j) -> A {
        const A v = x[(size_t)k * nC + j];
        return scal ? v / beta : v;
    };
    A xc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) xc[k] = X(k, c);
    if (vout)
#pragma unroll
        for (int k = 0; k < 5; ++k) stnt(vout + (size_t)k * nC + c, xc[k]);
    A acc[5][3];
    A fw[5][3];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const A b0 = (A)__builtin_nontemporal_load(dg + diagOff(c, 5 * i)),
                b1 = (A)__builtin_nontemporal_load(dg + diagOff(c, 5 * i + 1));
        const A b2 = (A)__builtin_nontemporal_load(dg + diagOff(c, 5 * i + 2)),
                b3 = (A)__builtin_nontemporal_load(dg + diagOff(c, 5 * i + 3)),
                b4 = (A)__builtin_nontemporal_load(dg + diagOff(c, 5 * i + 4));
        acc[i][0] = b0 * xc[0];
        acc[i][1] = b1 * xc[1];
        acc[i][2] = b2 * xc[2] + b3 * xc[3] + b4 * xc[4];
        if (FWD) fw[i][0] = fw[i][1] = fw[i][2] = (A)0;
    }
    const int sl = c >> 6, lane = c & 63;
    const int len = m.rowLen[c];
    const size_t base = (size_t)m.slicePtr[sl];
    for (int k = 0; k < len; ++k) {
        const size_t slot = (base + k) * 64 + lane;
        const int j = ldnt(m.col + slot);
        A xj[5], d[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) xj[q] = X(q, j);
        if (FWD) {
            const A rd = (A)rD[j];
#pragma unroll
            for (int q = 0; q < 5; ++q) d[q] = rd * y[(size_t)q * nC + j];
        }
        A bk[NB];
        slotBlock<TB, A, GEN, true>(m, blk, base + k, lane, bk);
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const A b0 = bk[5 * i], b1 = bk[5 * i + 1];
            const A b2 = bk[5 * i + 2], b3 = bk[5 * i + 3], b4 = bk[5 * i + 4];
            acc[i][0] += b0 * xj[0];
            acc[i][1] += b1 * xj[1];
            const A dd = b2 * xj[2] + b3 * xj[3] + b4 * xj[4];
            acc[i][2] += dd;
            if (FWD) {
                fw[i][0] += b0 * d[0];
                fw[i][1] += b1 * d[1];
                const A de = b2 * d[2] + b3 * d[3] + b4 * d[4];
                fw[i][2] += de;
            }
        }
    }
    if (m.nCyc > 0) spmvCyclic<TB, A>(m, blk, c, X, acc);
    A rdc = (A)1;
    if (FWD) rdc = (A)rD[c];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        A r = (A)0 + acc[i][0];
        r = r + acc[i][1];
        r = r + acc[i][2];
        if (FWD) {
            r = r - fw[i][0];
            r = r - fw[i][1];
            r = r - fw[i][2];
            r = rdc * r;
        }
        y[(size_t)i * nC + c] = r;
    }
}

template <typename TB, typename TV, bool GEN, int W = 1>
__global__ void __launch_bounds__(256, W) k_spmv(MeshView m, const TB* __restrict__ blk, const TB* __restrict__ dg,
                                              const TV* __restrict__ x, const double* __restrict__ scal,
                                              TV* __restrict__ vout, TV* __restrict__ y) {
    const int c = xcdBlock() * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    spmvRow<TB, TV, false, GEN>(m, blk, dg, x, scal, vout, y, nullptr, c);
}

// SpMV of the rows in a cell list; FWD: fused forward sweep (see spmvRow).
template <typename TB, typename TV, bool FWD, bool GEN, int W = 1>
__global__ void __launch_bounds__(256, W) k_spmv_cells(MeshView m, const TB* __restrict__ blk,
                                                    const TB* __restrict__ dg, const TV* __restrict__ x,
                                                    const double* __restrict__ scal, TV* __restrict__ vout,
                                                    TV* __restrict__ y, const double* __restrict__ rD,
                                                    const int* __restrict__ cells, int n) {
    const int t = xcdBlock() * blockDim.x + threadIdx.x;
    if (t >= n) return;
    spmvRow<TB, TV, FWD, GEN>(m, blk, dg, x, scal, vout, y, rD, ldnt(cells + t));
}

// Harmonic-balance SpMV: row cell c of instance J = blockIdx.y.  The own
// instance's partial sums are those of k_spmv; the cross-instance terms
// (V_c D_JK) x_K are added in coupledMatrix::matrixMul order
// (coupledMatrix.C:66-123): scalar rows add the scalar columns of
// instances 0..nO-1 (own rho, rhoE blocks at K = J) and then the own
// vector column; vector rows add the own scalar columns and then the vector
// columns of instances 0..nO-1 (own block at K = J).
// With FWD (colour-1 rows of a two-colour numbering): the forward LU-SGS
// terms of the own instance are accumulated from the same block loads and
// the row result is rD_c (y_c - fwd), as in spmvRow.
template <typename TB, typename TV, bool FWD, bool GEN>
__device__ __forceinline__ void spmvRowHB(const MeshView& m, const HBBlocks& hb, int J, const TV* __restrict__ x,
                                          const double* __restrict__ scal, TV* __restrict__ vout,
                                          TV* __restrict__ y, const double* __restrict__ rD, int c) {
    typedef TV A;
    const size_t nC = m.nC;
    const size_t vs = 5 * nC;
    const TB* blk = (const TB*)hb.blk[J];
    const TB* dg = (const TB*)hb.dg[J];
    const A beta = scal ? (A)(*scal) : (A)1;
    auto X = [&](int K, int k, int j) -> A {
        const A v = x[(size_t)K * vs + (size_t)k * nC + j];
        return scal ? v / beta : v;
    };
    const TV* yJ = y + (size_t)J * vs;
    A xc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) xc[k] = X(J, k, c);
    if (vout)
#pragma unroll
        for (int k = 0; k < 5; ++k) vout[(size_t)J * vs + (size_t)k * nC + c] = xc[k];
    A acc[5][3], fw[5][3];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const A b0 = (A)__builtin_nontemporal_load(dg + diagOff(c, 5 * i)),
                b1 = (A)__builtin_nontemporal_load(dg + diagOff(c, 5 * i + 1));
        const A b2 = (A)__builtin_nontemporal_load(dg + diagOff(c, 5 * i + 2)),
                b3 = (A)__builtin_nontemporal_load(dg + diagOff(c, 5 * i + 3)),
                b4 = (A)__builtin_nontemporal_load(dg + diagOff(c, 5 * i + 4));
        acc[i][0] = b0 * xc[0];
        acc[i][1] = b1 * xc[1];
        acc[i][2] = b2 * xc[2] + b3 * xc[3] + b4 * xc[4];
        if (FWD) fw[i][0] = fw[i][1] = fw[i][2] = (A)0;
    }
    const int sl = c >> 6, lane = c & 63;
    const int len = m.rowLen[c];
    const size_t base = (size_t)m.slicePtr[sl];
    for (int k = 0; k < len; ++k) {
        const size_t slot = (base + k) * 64 + lane;
        const int j = ldnt(m.col + slot);
        A xj[5], d[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) xj[q] = X(J, q, j);
        if (FWD) {
            const A rd = (A)rD[j];
#pragma unroll
            for (int q = 0; q < 5; ++q) d[q] = rd * yJ[(size_t)q * nC + j];
        }
        A bk[NB];
        slotBlock<TB, A, GEN, true>(m, blk, base + k, lane, bk);
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const A b0 = bk[5 * i], b1 = bk[5 * i + 1];
            const A b2 = bk[5 * i + 2], b3 = bk[5 * i + 3], b4 = bk[5 * i + 4];
            acc[i][0] += b0 * xj[0];
            acc[i][1] += b1 * xj[1];
            const A dd = b2 * xj[2] + b3 * xj[3] + b4 * xj[4];
            acc[i][2] += dd;
            if (FWD) {
                fw[i][0] += b0 * d[0];
                fw[i][1] += b1 * d[1];
                const A de = b2 * d[2] + b3 * d[3] + b4 * d[4];
                fw[i][2] += de;
            }
        }
    }
    if (m.nCyc > 0) spmvCyclic<TB, A>(m, blk, c, [&](int k, int j) { return X(J, k, j); }, acc);
    const double Vc = m.V[c];
    const int nO = hb.nO;
    A out[5];
    // scalar rows
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        A r = (A)0;
        for (int K = 0; K < nO; ++K) {
            if (K == J) {
                r = r + acc[i][0];
                r = r + acc[i][1];
            } else {
                r = r + (A)(Vc * hb.D[J * nO + K]) * X(K, i, c);
            }
        }
        out[i] = r + acc[i][2];
    }
    // vector rows
#pragma unroll
    for (int i = 2; i < 5; ++i) {
        A r = (A)0 + acc[i][0];
        r = r + acc[i][1];
        for (int K = 0; K < nO; ++K) {
            if (K == J) r = r + acc[i][2];
            else r = r + (A)(Vc * hb.D[J * nO + K]) * X(K, i, c);
        }
        out[i] = r;
    }
    A rdc = (A)1;
    if (FWD) rdc = (A)rD[c];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        A r = out[i];
        if (FWD) {
            r = r - fw[i][0];
            r = r - fw[i][1];
            r = r - fw[i][2];
            r = rdc * r;
        }
        y[(size_t)J * vs + (size_t)i * nC + c] = r;
    }
}

template <typename TB, typename TV, bool GEN>
__global__ void __launch_bounds__(256) k_spmv_hb(MeshView m, HBBlocks hb, const TV* __restrict__ x,
                                                 const double* __restrict__ scal, TV* __restrict__ vout,
                                                 TV* __restrict__ y) {
    const int c = xcdBlock() * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    spmvRowHB<TB, TV, false, GEN>(m, hb, blockIdx.y, x, scal, vout, y, nullptr, c);
}

// HB rows of a cell list (blockIdx.y = instance); FWD: fused forward sweep.
template <typename TB, typename TV, bool FWD, bool GEN>
__global__ void __launch_bounds__(256) k_spmv_hb_cells(MeshView m, HBBlocks hb, const TV* __restrict__ x,
                                                       const double* __restrict__ scal, TV* __restrict__ vout,
                                                       TV* __restrict__ y, const double* __restrict__ rD,
                                                       const int* __restrict__ cells, int n) {
    const int t = xcdBlock() * blockDim.x + threadIdx.x;
    if (t >= n) return;
    spmvRowHB<TB, TV, FWD, GEN>(m, hb, blockIdx.y, x, scal, vout, y, rD, cells[t]);
}

// ------------------------------------------------------------- LU-SGS -----
// Forward sweep of one level: r[c] -= sum over lower neighbours p (in
// increasing p) of L_cp * (rD[p] r[p]).
template <typename TB, typename TV, bool GEN>
__device__ __forceinline__ void lusgsFwdRow(const MeshView& m, const TB* __restrict__ blk,
                                            const double* __restrict__ rD, TV* __restrict__ r, int c) {
    const int nl = m.nLower[c];
    if (nl == 0) return;
    const size_t nC = m.nC;
    typedef TV A;
    A acc[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) acc[i] = r[(size_t)i * nC + c];
    const int sl = c >> 6, lane = c & 63;
    const size_t base = (size_t)m.slicePtr[sl];
    for (int k = 0; k < nl; ++k) {
        const size_t slot = (base + k) * 64 + lane;
        const int p = m.col[slot];
        const A rd = (A)rD[p];
        A d[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) d[q] = rd * r[(size_t)q * nC + p];
        A bk[NB];
        slotBlock<TB, A, GEN>(m, blk, base + k, lane, bk);
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const A b0 = bk[5 * i], b1 = bk[5 * i + 1];
            const A b2 = bk[5 * i + 2], b3 = bk[5 * i + 3], b4 = bk[5 * i + 4];
            acc[i] -= b0 * d[0];
            acc[i] -= b1 * d[1];
            const A dd = b2 * d[2] + b3 * d[3] + b4 * d[4];
            acc[i] -= dd;
        }
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) r[(size_t)i * nC + c] = acc[i];
}

// Backward sweep of one level: r[c] = rD[c] (r[c] - sum over upper
// neighbours q (decreasing q) of U_cq r[q]), r[q] already final.
template <typename TB, typename TV, bool GEN>
__device__ __forceinline__ void lusgsBwdRow(const MeshView& m, const TB* __restrict__ blk,
                                            const double* __restrict__ rD, TV* __restrict__ r, int c) {
    const size_t nC = m.nC;
    typedef TV A;
    A acc[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) acc[i] = r[(size_t)i * nC + c];
    const int sl = c >> 6, lane = c & 63;
    const size_t base = (size_t)m.slicePtr[sl];
    const int nl = m.nLower[c], len = m.rowLen[c];
    for (int k = len - 1; k >= nl; --k) {
        const size_t slot = (base + k) * 64 + lane;
        const int q = m.col[slot];
        A d[5];
#pragma unroll
        for (int z = 0; z < 5; ++z) d[z] = r[(size_t)z * nC + q];
        A bk[NB];
        slotBlock<TB, A, GEN>(m, blk, base + k, lane, bk);
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const A b0 = bk[5 * i], b1 = bk[5 * i + 1];
            const A b2 = bk[5 * i + 2], b3 = bk[5 * i + 3], b4 = bk[5 * i + 4];
            acc[i] -= b0 * d[0];
            acc[i] -= b1 * d[1];
            const A dd = b2 * d[2] + b3 * d[3] + b4 * d[4];
            acc[i] -= dd;
        }
    }
    const A rd = (A)rD[c];
#pragma unroll
    for (int i = 0; i < 5; ++i) r[(size_t)i * nC + c] = rd * acc[i];
}

template <typename TB, typename TV, bool GEN, int W = 1>
__global__ void __launch_bounds__(256, W) k_lusgs_fwd(MeshView m, const TB* __restrict__ blk,
                                                   const double* __restrict__ rD, TV* __restrict__ r,
                                                   const int* __restrict__ cells, int n) {
    const int t = xcdBlock() * blockDim.x + threadIdx.x;
    if (t >= n) return;
    lusgsFwdRow<TB, TV, GEN>(m, blk, rD, r, cells[t]);
}

template <typename TB, typename TV, bool GEN, int W = 1>
__global__ void __launch_bounds__(256, W) k_lusgs_bwd(MeshView m, const TB* __restrict__ blk,
                                                   const double* __restrict__ rD, TV* __restrict__ r,
                                                   const int* __restrict__ cells, int n) {
    const int t = xcdBlock() * blockDim.x + threadIdx.x;
    if (t >= n) return;
    lusgsBwdRow<TB, TV, GEN>(m, blk, rD, r, cells[t]);
}

// Harmonic balance: one launch per level for all instances (blockIdx.y =
// instance; the cross-instance blocks have no faces, so the sweeps are
// independent per instance and share the diagonal rD).
template <typename TB, typename TV, bool FWD, bool GEN>
__global__ void __launch_bounds__(256) k_lusgs_hb(MeshView m, HBBlocks hb, const double* __restrict__ rD,
                                                  TV* __restrict__ r, const int* __restrict__ cells, int n) {
    const int t = xcdBlock() * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int J = blockIdx.y;
    TV* rJ = r + (size_t)J * 5 * m.nC;
    if (FWD) lusgsFwdRow<TB, TV, GEN>(m, (const TB*)hb.blk[J], rD, rJ, cells[t]);
    else lusgsBwdRow<TB, TV, GEN>(m, (const TB*)hb.blk[J], rD, rJ, cells[t]);
}

// ------------------------------------------------------- block-Jacobi -----
// 5x5 inverse of each diagonal block by LU with partial pivoting.
template <typename TB>
__global__ void k_jacobi_setup(int nC, const TB* __restrict__ dg, TB* __restrict__ dinv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    double a[5][5];
    int piv[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        piv[i] = i;
#pragma unroll
        for (int j = 0; j < 5; ++j) a[i][j] = (double)dg[diagOff(c, 5 * i + j)];
    }
    for (int col = 0; col < 5; ++col) {
        int p = col;
        for (int r = col + 1; r < 5; ++r)
            if (fabs(a[r][col]) > fabs(a[p][col])) p = r;
        if (p != col) {
            for (int q = 0; q < 5; ++q) { double t = a[p][q]; a[p][q] = a[col][q]; a[col][q] = t; }
            int t = piv[p]; piv[p] = piv[col]; piv[col] = t;
        }
        for (int r = col + 1; r < 5; ++r) {
            a[r][col] /= a[col][col];
            for (int q = col + 1; q < 5; ++q) a[r][q] -= a[r][col] * a[col][q];
        }
    }
    for (int e = 0; e < 5; ++e) {
        double yv[5];
        for (int i = 0; i < 5; ++i) {
            double s = piv[i] == e ? 1.0 : 0.0;
            for (int q = 0; q < i; ++q) s -= a[i][q] * yv[q];
            yv[i] = s;
        }
        for (int i = 4; i >= 0; --i) {
            double s = yv[i];
            for (int q = i + 1; q < 5; ++q) s -= a[i][q] * yv[q];
            yv[i] = s / a[i][i];
        }
        for (int i = 0; i < 5; ++i) dinv[diagOff(c, 5 * i + e)] = (TB)yv[i];
    }
}

template <typename TB, typename TV>
__global__ void k_jacobi_apply(int nC, const TB* __restrict__ dinv, TV* __restrict__ r) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    typedef TV A;
    A v[5], o[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) { v[i] = r[(size_t)i * nC + c]; o[i] = 0; }
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) o[i] += (A)dinv[diagOff(c, 5 * i + j)] * v[j];
#pragma unroll
    for (int i = 0; i < 5; ++i) r[(size_t)i * nC + c] = o[i];
}

// ---------------------------------------------------------- reductions ----
__device__ __forceinline__ double waveSum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Block-level sum of NV values per thread -> out[blockIdx.x*NV + k].
template <int NV>
__device__ __forceinline__ void blockReduceStore(double (&v)[NV], double* out) {
    __shared__ double sh[4][NV > 0 ? NV : 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double s = waveSum(v[k]);
        if (lane == 0) sh[wv][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int k = threadIdx.x;
        out[(size_t)blockIdx.x * NV + k] = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
    }
}

// Sum of value k over nb partial rows of `stride` values, in a fixed order
// (thread t adds rows t, t+256, ...; then the wave sums; then the four
// waves in order).  Called by all threads of a 256-thread block; every
// thread gets the sum.
__device__ __forceinline__ double sumPartials(const double* __restrict__ part, int nb, int stride, int k) {
    __shared__ double sh[4];
    double v = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) v += part[(size_t)b * stride + k];
    v = waveSum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const double s = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    __syncthreads();
    return s;
}

// Final reduction of [nb][nv] partials: value k = blockIdx.x (grid nv,
// 256 threads); out[k].
__global__ void __launch_bounds__(256) k_reduce_final(const double* __restrict__ part, int nb, int nv,
                                                      double* __restrict__ out) {
    const int k = blockIdx.x;
    const double s = sumPartials(part, nb, nv, k);
    if (threadIdx.x == 0) out[k] = s;
}

// sums of the 5 components of x (for the average) -> partials
template <typename TV>
__global__ void __launch_bounds__(256) k_sum5(int nC, const TV* __restrict__ x, double* __restrict__ part) {
    double v[5] = {0, 0, 0, 0, 0};
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nC; c += gridDim.x * blockDim.x)
#pragma unroll
        for (int k = 0; k < 5; ++k) v[k] += (double)x[(size_t)k * nC + c];
    blockReduceStore<5>(v, part);
}

// norm-factor partials (gmres.C:839-850):
// [0..1] sum |t_s|+|b_s|, [2] sum |t_v|+|b_v|, [3..4] sum |b_s|, [5..7] sum |b_v,d|
template <typename TV>
__global__ void __launch_bounds__(256) k_normfactor(int nC, const TV* __restrict__ t, const TV* __restrict__ b,
                                                    double* __restrict__ part) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nC; c += gridDim.x * blockDim.x) {
        double tv[5], bv[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) { tv[k] = (double)t[(size_t)k * nC + c]; bv[k] = (double)b[(size_t)k * nC + c]; }
        v[0] += fabs(tv[0]) + fabs(bv[0]);
        v[1] += fabs(tv[1]) + fabs(bv[1]);
        v[2] += sqrt(tv[2] * tv[2] + tv[3] * tv[3] + tv[4] * tv[4]) + sqrt(bv[2] * bv[2] + bv[3] * bv[3] + bv[4] * bv[4]);
        v[3] += fabs(bv[0]);
        v[4] += fabs(bv[1]);
        v[5] += fabs(bv[2]);
        v[6] += fabs(bv[3]);
        v[7] += fabs(bv[4]);
    }
    blockReduceStore<8>(v, part);
}

// final residual partials of r = b - A dW: [0..4] sum |r_k|
template <typename TV>
__global__ void __launch_bounds__(256) k_resid_l1(int nC, const TV* __restrict__ b, TV* __restrict__ Ax,
                                                  double* __restrict__ part) {
    double v[5] = {0, 0, 0, 0, 0};
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nC; c += gridDim.x * blockDim.x)
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const TV r = b[(size_t)k * nC + c] - Ax[(size_t)k * nC + c];
            Ax[(size_t)k * nC + c] = r;
            v[k] += fabs((double)r);
        }
    blockReduceStore<5>(v, part);
}

template <typename TV>
__global__ void __launch_bounds__(256) k_norm2(int nC, const TV* __restrict__ x, double* __restrict__ part) {
    double v[1] = {0};
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nC; c += gridDim.x * blockDim.x)
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const double q = (double)x[(size_t)k * nC + c];
            v[0] += q * q;
        }
    blockReduceStore<1>(v, part);
}

// CGS projection partials: <w, V_j> for j < nv (nv <= 64, chunked by 16).
// The basis vectors are streamed with non-temporal loads here and in the
// update and combine kernels: each is read once per pass and the passes
// between two reads of one vector move far more than L2 holds.
template <typename TV, int CH>
__global__ void __launch_bounds__(256) k_multidot(int nC, const TV* __restrict__ w, const TV* __restrict__ Vb,
                                                  size_t vstride, int j0, int nv, double* __restrict__ part) {
    double v[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) v[k] = 0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nC; c += gridDim.x * blockDim.x) {
        TV wv[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) wv[q] = w[(size_t)q * nC + c];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            if (k < nv) {
                const TV* vj = Vb + (size_t)(j0 + k) * vstride;
                double s = 0;
#pragma unroll
                for (int q = 0; q < 5; ++q) s += (double)(wv[q] * __builtin_nontemporal_load(vj + (size_t)q * nC + c));
                v[k] += s;
            }
        }
    }
    blockReduceStore<CH>(v, part);
}

// CGS projections of one pass over w: ch (8, 16 or 32) basis vectors.
template <typename TV>
void launchMultidot(int ch, int nbk, int nC, const TV* w, const TV* V, size_t vs, int j0, int nv, double* part,
                    hipStream_t st) {
    if (ch == 8) k_multidot<TV, 8><<<nbk, 256, 0, st>>>(nC, w, V, vs, j0, nv, part);
    else if (ch == 32) k_multidot<TV, 32><<<nbk, 256, 0, st>>>(nC, w, V, vs, j0, nv, part);
    else k_multidot<TV, 16><<<nbk, 256, 0, st>>>(nC, w, V, vs, j0, nv, part);
}

// w -= sum_j h_j V_j (j < nv), then ||w||^2 partials
template <typename TV>
__global__ void __launch_bounds__(256) k_cgs_update(int nC, TV* __restrict__ w, const TV* __restrict__ Vb,
                                                    size_t vstride, int nv, const double* __restrict__ h,
                                                    double* __restrict__ part) {
    double v[1] = {0};
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nC; c += gridDim.x * blockDim.x) {
        // five independent chains (one per component), basis vectors in
        // order j = 0..nv-1 on each; four basis vectors' loads in flight
        TV a[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) a[q] = w[(size_t)q * nC + c];
        int j = 0;
        for (; j + 4 <= nv; j += 4) {
            TV x[4][5];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < 5; ++q)
                    x[u][q] = __builtin_nontemporal_load(Vb + (size_t)(j + u) * vstride + (size_t)q * nC + c);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const TV hj = (TV)h[j + u];
#pragma unroll
                for (int q = 0; q < 5; ++q) a[q] -= hj * x[u][q];
            }
        }
        for (; j < nv; ++j) {
            const TV hj = (TV)h[j];
#pragma unroll
            for (int q = 0; q < 5; ++q) a[q] -= hj * Vb[(size_t)j * vstride + (size_t)q * nC + c];
        }
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            w[(size_t)q * nC + c] = a[q];
            v[0] += (double)a[q] * (double)a[q];
        }
    }
    blockReduceStore<1>(v, part);
}

// MGS step: (optionally) w -= h * V_prev, then partials of <w, V_next>
// (5 per-variable partial sums, combined in the reference's order on reduce)
template <typename TV>
__global__ void __launch_bounds__(256) k_mgs_step(int nC, TV* __restrict__ w, const TV* __restrict__ Vprev,
                                                  const double* __restrict__ hprev, const TV* __restrict__ Vnext,
                                                  int wantNorm, double* __restrict__ part) {
    double v[3] = {0, 0, 0};
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nC; c += gridDim.x * blockDim.x) {
        TV a[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            a[q] = w[(size_t)q * nC + c];
            if (Vprev) {
                a[q] -= (TV)(*hprev) * Vprev[(size_t)q * nC + c];
                w[(size_t)q * nC + c] = a[q];
            }
        }
        if (Vnext) {
            v[0] += (double)(a[0] * Vnext[c]);
            v[1] += (double)(a[1] * Vnext[(size_t)nC + c]);
            v[2] += (double)(a[2] * Vnext[2 * (size_t)nC + c] + a[3] * Vnext[3 * (size_t)nC + c] +
                             a[4] * Vnext[4 * (size_t)nC + c]);
        }
        if (wantNorm) {
            v[0] += (double)(a[0] * a[0]);
            v[1] += (double)(a[1] * a[1]);
            v[2] += (double)(a[2] * a[2] + a[3] * a[3] + a[4] * a[4]);
        }
    }
    blockReduceStore<3>(v, part);
}

// x = y - avg (component-wise); dW = 0 afterwards is done by memset.
template <typename TV>
__global__ void k_sub_avg(int nC, const TV* __restrict__ W, const double* __restrict__ sums, TV* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
#pragma unroll
    for (int k = 0; k < 5; ++k) out[(size_t)k * nC + c] = W[(size_t)k * nC + c] - (TV)(sums[k] / nC);
}

template <typename TV>
__global__ void k_copy(size_t n, const TV* __restrict__ a, TV* __restrict__ b) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

template <typename TA, typename TB2>
__global__ void k_convert(size_t n, const TA* __restrict__ a, TB2* __restrict__ b) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (TB2)a[i];
}

// dW += sum_i y_i V_i (gmres.C:1050-1059), sequential in i per entry
template <typename TV, typename TX>
__global__ void __launch_bounds__(256) k_update_sol(int nC, TX* __restrict__ dW, const TV* __restrict__ Vb,
                                                    size_t vstride, int nv, const double* __restrict__ yh) {
    const size_t n = (size_t)5 * nC;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    TX a = dW[i];
    // four basis vectors' loads in flight, added in order j = 0..nv-1
    int j = 0;
    for (; j + 4 <= nv; j += 4) {
        TV x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = __builtin_nontemporal_load(Vb + (size_t)(j + u) * vstride + i);
#pragma unroll
        for (int u = 0; u < 4; ++u) a += (TX)yh[j + u] * (TX)x[u];
    }
    for (; j < nv; ++j) a += (TX)yh[j] * (TX)Vb[(size_t)j * vstride + i];
    dW[i] = a;
}

}  // namespace

// ------------------------------------------------------------ GMRES state --
// Scalars of one GMRES(m) cycle kept on the device (no host round trips).
struct GmresScalars {
    double H[KMAX][KMAX];
    double cs[KMAX], sn[KMAX], bh[KMAX + 1], yh[KMAX];
    double beta;
    double sNF[2], vNF;
    double res[10];   // sInit[2] vInit[3] sFinal[2] vFinal[3]
    double red[64];   // scratch reduction results
    double hcol[KMAX];  // contiguous copy of the current Hessenberg column
};

namespace {

__device__ void givensRot(double h, double beta, double& c, double& s) {
    if (beta == 0) { c = 1; s = 0; }
    else if (fabs(beta) > fabs(h)) {
        const double tau = -h / beta;
        s = 1.0 / sqrt(1.0 + tau * tau);
        c = s * tau;
    } else {
        const double tau = -beta / h;
        c = 1.0 / sqrt(1.0 + tau * tau);
        s = c * tau;
    }
}

// Norm factors and initial residuals from the k_normfactor reduction.
__global__ void k_gm_init(GmresScalars* g) {
    if (threadIdx.x != 0) return;
    const double* r = g->red;
    g->sNF[0] = r[0] + kVSMALL;
    g->sNF[1] = r[1] + kVSMALL;
    g->vNF = r[2] + kVSMALL;
    g->res[0] = r[3] / g->sNF[0];
    g->res[1] = r[4] / g->sNF[1];
    for (int d = 0; d < 3; ++d) g->res[2 + d] = r[5 + d] / g->vNF;
    for (int k = 0; k < 5; ++k) g->res[5 + k] = g->res[k];
}

// beta = ||r|| at the start of a cycle
__global__ void k_gm_beta0(GmresScalars* g, int nDirs) {
    if (threadIdx.x != 0) return;
    g->beta = sqrt(g->red[0]);
    for (int i = 0; i <= nDirs; ++i) g->bh[i] = 0;
    g->bh[0] = g->beta;
}

// Apply the previous rotations to column i of H, form the new rotation and
// update bh (gmres.C:999-1027); beta2 = ||w||^2 after the orthogonalisation.
// The block's threads stage column i and the previous rotations in LDS, so
// that the chain of dependent updates run by one thread waits on LDS, not
// on device memory.  Called by all threads of the block.
__device__ void givensBlock(GmresScalars* g, int i, double beta2) {
    __shared__ double hc[KMAX], cc[KMAX], ss[KMAX];
    for (int t = threadIdx.x; t <= i; t += blockDim.x) hc[t] = g->H[t][i];
    for (int t = threadIdx.x; t < i; t += blockDim.x) {
        cc[t] = g->cs[t];
        ss[t] = g->sn[t];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double beta = sqrt(beta2);
        for (int j = 0; j < i; ++j) {
            const double Hji = hc[j];
            hc[j] = cc[j] * Hji - ss[j] * hc[j + 1];
            hc[j + 1] = ss[j] * Hji + cc[j] * hc[j + 1];
        }
        double ci, si;
        givensRot(hc[i], beta, ci, si);
        g->cs[i] = ci;
        g->sn[i] = si;
        const double bhi = g->bh[i];
        const double bhn = g->bh[i + 1];
        g->bh[i] = ci * bhi - si * bhn;
        g->bh[i + 1] = si * bhi + ci * bhn;
        hc[i] = ci * hc[i] - si * beta;
        g->beta = beta;
    }
    __syncthreads();
    for (int t = threadIdx.x; t <= i; t += blockDim.x) g->H[t][i] = hc[t];
}

// CGS projections of one multidot pass (partial rows of ch values): value
// k = blockIdx.x (grid nv) is stored as H[j0+k][i] and hcol[j0+k].
__global__ void __launch_bounds__(256) k_reduce_col(const double* __restrict__ part, int nb, int ch,
                                                    GmresScalars* g, int i, int j0) {
    const int k = blockIdx.x;
    const double s = sumPartials(part, nb, ch, k);
    if (threadIdx.x == 0) {
        g->red[k] = s;
        g->H[j0 + k][i] = s;
        g->hcol[j0 + k] = s;
    }
}

// ||w||^2 of the CGS update pass, then the Givens step (one block).
__global__ void __launch_bounds__(256) k_reduce_givens(const double* __restrict__ part, int nb, GmresScalars* g,
                                                       int i) {
    const double s = sumPartials(part, nb, 1, 0);
    if (threadIdx.x == 0) g->red[0] = s;
    givensBlock(g, i, s);
}

// MGS step j of column i: the three per-variable partial sums of <w, V_j>
// (j <= i, stored as H[j][i]) or of ||w||^2 (j = i+1, then the Givens
// step), combined in the reference's order (one block).
__global__ void __launch_bounds__(256) k_reduce_mgs(const double* __restrict__ part, int nb, GmresScalars* g, int i,
                                                    int j) {
    double s[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = sumPartials(part, nb, 3, k);
    const double hs = (0.0 + s[0] + s[1]) + s[2];
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k) g->red[k] = s[k];
        if (j <= i) g->H[j][i] = hs;
    }
    if (j > i) givensBlock(g, i, hs);
}

__global__ void k_gm_backsub(GmresScalars* g, int n) {
    if (threadIdx.x != 0) return;
    for (int i = n - 1; i >= 0; --i) {
        double s = g->bh[i];
        for (int j = i + 1; j < n; ++j) s -= g->H[i][j] * g->yh[j];
        double hd = g->H[i][i];
        hd = hd >= 0 ? hd + kVSMALL : hd - kVSMALL;
        g->yh[i] = s / hd;
    }
}

__global__ void k_gm_final(GmresScalars* g, int emptyDir) {
    if (threadIdx.x != 0) return;
    const double* r = g->red;
    g->res[5] = r[0] / g->sNF[0];
    g->res[6] = r[1] / g->sNF[1];
    for (int d = 0; d < 3; ++d) g->res[7 + d] = (d == emptyDir) ? 0.0 : r[2 + d] / g->vNF;
}

}  // namespace

// --------------------------------------------------------------- host side --
static inline int nb(size_t n, int b) { return (int)((n + b - 1) / b); }

// The CGS projection passes of one Arnoldi step against nv basis vectors,
// ch vectors per pass (autotuner timing).
void launchProjectionPasses(const MeshView& m, int vecFp32, int ch, const void* w, const void* V, int nv,
                            double* part, hipStream_t st) {
    const int nC = m.nC, nbk = reduceBlocks(nC);
    const size_t vs = (size_t)5 * nC;
    for (int j0 = 0; j0 < nv; j0 += ch) {
        const int n = (nv - j0) < ch ? (nv - j0) : ch;
        if (vecFp32) launchMultidot<float>(ch, nbk, nC, (const float*)w, (const float*)V, vs, j0, n, part, st);
        else launchMultidot<double>(ch, nbk, nC, (const double*)w, (const double*)V, vs, j0, n, part, st);
    }
}

int reduceBlocks(int nC) {
    int n = nb(nC, 256);
    return n < 1024 ? (n < 1 ? 1 : n) : 1024;
}

// Full off-diagonal blocks from generator storage (read-back of the
// assembled matrix; padding slots are left untouched).
template <typename TB>
__global__ void k_expand_blocks(MeshView m, const TB* __restrict__ gen, TB* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC) return;
    const int sl = c >> 6, lane = c & 63;
    const size_t base = (size_t)m.slicePtr[sl];
    for (int k = 0; k < m.rowLen[c]; ++k) {
        TB b[NB];
        slotBlock<TB, TB, true>(m, gen, base + k, lane, b);
#pragma unroll
        for (int e = 0; e < NB; ++e) out[((base + k) * NB + e) * 64 + lane] = b[e];
    }
}

void launchExpandBlocks(const MeshView& m, int fp32, const void* gen, void* out, hipStream_t st) {
    if (fp32) k_expand_blocks<float><<<nb(m.nC, 256), 256, 0, st>>>(m, (const float*)gen, (float*)out);
    else k_expand_blocks<double><<<nb(m.nC, 256), 256, 0, st>>>(m, (const double*)gen, (double*)out);
}

template <typename TB, typename TV>
void spmv(const MeshView& m, const void* blk, const void* dg, const void* x, const double* scal, void* vout, void* y,
          hipStream_t st) {
    const dim3 g(xcdGrid(m.nC, m.bsOp));
    const TB *B = (const TB*)blk, *G = (const TB*)dg;
    if (m.blkGen && m.occOp)
        k_spmv<TB, TV, true, 4><<<g, m.bsOp, 0, st>>>(m, B, G, (const TV*)x, scal, (TV*)vout, (TV*)y);
    else if (m.blkGen)
        k_spmv<TB, TV, true><<<g, m.bsOp, 0, st>>>(m, B, G, (const TV*)x, scal, (TV*)vout, (TV*)y);
    else
        k_spmv<TB, TV, false><<<g, m.bsOp, 0, st>>>(m, B, G, (const TV*)x, scal, (TV*)vout, (TV*)y);
}

void launchSpmv(const MeshView& m, int blkFp32, int vecFp32, const void* blk, const void* dg, const void* x,
                const double* scal, void* vout, void* y, hipStream_t st) {
    if (blkFp32 && vecFp32) spmv<float, float>(m, blk, dg, x, scal, vout, y, st);
    else if (blkFp32) spmv<float, double>(m, blk, dg, x, scal, vout, y, st);
    else if (vecFp32) spmv<double, float>(m, blk, dg, x, scal, vout, y, st);
    else spmv<double, double>(m, blk, dg, x, scal, vout, y, st);
}

template <typename TB, typename TV, bool GEN, int W>
static void spmvLusgsG(const MeshView& m, const Levels& L, const void* blk, const void* dg, const double* rD,
                       const void* x, const double* scal, void* vout, void* y, hipStream_t st) {
    const int n0 = L.fwdPtrHost[1], n1 = L.fwdPtrHost[2] - n0;
    const TB* B = (const TB*)blk;
    const TB* G = (const TB*)dg;
    const int bs = m.bsOp;
    k_spmv_cells<TB, TV, false, GEN, W><<<xcdGrid(n0, bs), bs, 0, st>>>(m, B, G, (const TV*)x, scal, (TV*)vout,
                                                                        (TV*)y, rD, L.fwdCells, n0);
    k_spmv_cells<TB, TV, true, GEN, W><<<xcdGrid(n1, bs), bs, 0, st>>>(m, B, G, (const TV*)x, scal, (TV*)vout,
                                                                       (TV*)y, rD, L.fwdCells + n0, n1);
    const int s0 = L.bwdPtrHost[1], nbw = L.bwdPtrHost[2] - s0;
    k_lusgs_bwd<TB, TV, GEN, W><<<xcdGrid(nbw, bs), bs, 0, st>>>(m, B, rD, (TV*)y, L.bwdCells + s0, nbw);
}

template <typename TB, typename TV>
static void spmvLusgsT(const MeshView& m, const Levels& L, const void* blk, const void* dg, const double* rD,
                       const void* x, const double* scal, void* vout, void* y, hipStream_t st) {
    if (m.blkGen && m.occOp) spmvLusgsG<TB, TV, true, 4>(m, L, blk, dg, rD, x, scal, vout, y, st);
    else if (m.blkGen) spmvLusgsG<TB, TV, true, 1>(m, L, blk, dg, rD, x, scal, vout, y, st);
    else spmvLusgsG<TB, TV, false, 1>(m, L, blk, dg, rD, x, scal, vout, y, st);
}

bool launchSpmvLusgs(const MeshView& m, const Levels& L, int blkFp32, int vecFp32, const void* blk, const void* dg,
                     const double* rD, const void* x, const double* scal, void* vout, void* y, hipStream_t st) {
    if (!L.twoColour) return false;
    if (blkFp32 && vecFp32) spmvLusgsT<float, float>(m, L, blk, dg, rD, x, scal, vout, y, st);
    else if (blkFp32) spmvLusgsT<float, double>(m, L, blk, dg, rD, x, scal, vout, y, st);
    else if (vecFp32) spmvLusgsT<double, float>(m, L, blk, dg, rD, x, scal, vout, y, st);
    else spmvLusgsT<double, double>(m, L, blk, dg, rD, x, scal, vout, y, st);
    return true;
}

template <typename TB, typename TV, bool GEN, int W>
static void lusgsG(const MeshView& m, const Levels& L, const void* blk, const double* rD, void* r, hipStream_t st) {
    const int bs = m.bsOp;
    for (int l = 0; l < L.nFwd; ++l) {
        const int s0 = L.fwdPtrHost[l], n = L.fwdPtrHost[l + 1] - s0;
        if (L.fwdSkip && l == 0) continue;  // first level has no lower neighbours
        k_lusgs_fwd<TB, TV, GEN, W><<<xcdGrid(n, bs), bs, 0, st>>>(m, (const TB*)blk, rD, (TV*)r, L.fwdCells + s0,
                                                                   n);
    }
    for (int l = 0; l < L.nBwd; ++l) {
        const int s0 = L.bwdPtrHost[l], n = L.bwdPtrHost[l + 1] - s0;
        k_lusgs_bwd<TB, TV, GEN, W><<<xcdGrid(n, bs), bs, 0, st>>>(m, (const TB*)blk, rD, (TV*)r, L.bwdCells + s0,
                                                                   n);
    }
}

template <typename TB, typename TV>
static void lusgsT(const MeshView& m, const Levels& L, const void* blk, const double* rD, void* r, hipStream_t st) {
    if (m.blkGen && m.occOp) lusgsG<TB, TV, true, 4>(m, L, blk, rD, r, st);
    else if (m.blkGen) lusgsG<TB, TV, true, 1>(m, L, blk, rD, r, st);
    else lusgsG<TB, TV, false, 1>(m, L, blk, rD, r, st);
}

void launchLusgs(const MeshView& m, const Levels& L, int blkFp32, int vecFp32, const void* blk, const double* rD,
                 void* r, hipStream_t st) {
    if (blkFp32 && vecFp32) lusgsT<float, float>(m, L, blk, rD, r, st);
    else if (blkFp32) lusgsT<float, double>(m, L, blk, rD, r, st);
    else if (vecFp32) lusgsT<double, float>(m, L, blk, rD, r, st);
    else lusgsT<double, double>(m, L, blk, rD, r, st);
}

void launchJacobiSetup(int nC, int blkFp32, const void* dg, void* dinv, hipStream_t st) {
    if (blkFp32) k_jacobi_setup<float><<<nb(nC, 256), 256, 0, st>>>(nC, (const float*)dg, (float*)dinv);
    else k_jacobi_setup<double><<<nb(nC, 256), 256, 0, st>>>(nC, (const double*)dg, (double*)dinv);
}

void launchJacobiApply(int nC, int blkFp32, int vecFp32, const void* dinv, void* r, hipStream_t st) {
    if (blkFp32 && vecFp32) k_jacobi_apply<float, float><<<nb(nC, 256), 256, 0, st>>>(nC, (const float*)dinv, (float*)r);
    else if (blkFp32) k_jacobi_apply<float, double><<<nb(nC, 256), 256, 0, st>>>(nC, (const float*)dinv, (double*)r);
    else if (vecFp32) k_jacobi_apply<double, float><<<nb(nC, 256), 256, 0, st>>>(nC, (const double*)dinv, (float*)r);
    else k_jacobi_apply<double, double><<<nb(nC, 256), 256, 0, st>>>(nC, (const double*)dinv, (double*)r);
}

// --------------------------------------------------------- GMRES driver --
template <typename TV>
void gmresCycleT(const GmresArgs& a, hipStream_t st) {
    const MeshView& m = *a.m;
    const int nC = a.nVec > 0 ? a.nVec : m.nC;
    const int nbk = reduceBlocks(nC);
    const size_t vs = (size_t)5 * nC;
    TV* r = (TV*)a.r;    // current w
    TV* r2 = (TV*)a.r2;  // SpMV output (ping-pong with r)
    TV* V = (TV*)a.V;
    GmresScalars* g = a.g;
    double* red = g->red;
    auto mark = [&](int k, int b) {
        if (a.mark) a.mark(k, b);
    };
    auto precond = [&](TV* v) {
        mark(TK_PRECOND, 1);
        if (a.pre) a.pre(v);
        else if (a.precond == 0) launchLusgs(m, *a.lev, a.blkFp32, sizeof(TV) == 4, a.blk, a.rD, v, st);
        else if (a.precond == 1) launchJacobiApply(nC, a.blkFp32, sizeof(TV) == 4, a.dinv, v, st);
        mark(TK_PRECOND, 0);
    };
    // r <- P^-1 r; beta = ||r||
    precond(r);
    k_norm2<TV><<<nbk, 256, 0, st>>>(nC, r, a.wk->part);
    k_reduce_final<<<1, 256, 0, st>>>(a.wk->part, nbk, 1, red);
    k_gm_beta0<<<1, 64, 0, st>>>(g, a.nDirs);
    for (int i = 0; i < a.nDirs; ++i) {
        TV* vi = V + (size_t)i * vs;
        // v_i = r / beta ; w = A v_i (into the other buffer: the SpMV gathers
        // neighbour entries of its input, so it never runs in place)
        // fused SpMV + LU-SGS on two-colour level structures (the colour-1
        // blocks are read once for both; the SPMV timing class then covers
        // the sweeps too); otherwise SpMV, then the preconditioner
        mark(TK_SPMV, 1);
        bool fused = false;
        if (a.opPre) fused = a.opPre(r, &g->beta, vi, r2);
        else if (!a.op && !a.pre && a.precond == 0 && a.fuse)
            fused = launchSpmvLusgs(m, *a.lev, a.blkFp32, sizeof(TV) == 4, a.blk, a.dg, a.rD, r, &g->beta, vi, r2,
                                    st);
        if (!fused) {
            if (a.op) a.op(r, &g->beta, vi, r2);
            else launchSpmv(m, a.blkFp32, sizeof(TV) == 4, a.blk, a.dg, r, &g->beta, vi, r2, st);
        }
        mark(TK_SPMV, 0);
        {
            TV* t = r;
            r = r2;
            r2 = t;
        }
        if (!fused) precond(r);
        mark(TK_ORTH, 1);
        if (a.orth == 0) {
            // modified Gram-Schmidt, reference order
            for (int j = 0; j <= i + 1; ++j) {
                const TV* vprev = j > 0 ? V + (size_t)(j - 1) * vs : nullptr;
                const TV* vnext = j <= i ? V + (size_t)j * vs : nullptr;
                const double* hprev = j > 0 ? &g->H[j - 1][i] : nullptr;
                k_mgs_step<TV><<<nbk, 256, 0, st>>>(nC, r, vprev, hprev, vnext, j == i + 1, a.wk->part);
                k_reduce_mgs<<<1, 256, 0, st>>>(a.wk->part, nbk, g, i, j);
            }
        } else {
            // classical Gram-Schmidt: one projection pass + one update pass
            const int ch = m.mdCh;
            for (int j0 = 0; j0 <= i; j0 += ch) {
                const int nv = (i + 1 - j0) < ch ? (i + 1 - j0) : ch;
                launchMultidot<TV>(ch, nbk, nC, r, V, vs, j0, nv, a.wk->part, st);
                k_reduce_col<<<nv, 256, 0, st>>>(a.wk->part, nbk, ch, g, i, j0);
            }
            k_cgs_update<TV><<<nbk, 256, 0, st>>>(nC, r, V, vs, i + 1, g->hcol, a.wk->part);
            k_reduce_givens<<<1, 256, 0, st>>>(a.wk->part, nbk, g, i);
        }
        mark(TK_ORTH, 0);
    }
    k_gm_backsub<<<1, 64, 0, st>>>(g, a.nDirs);
}

}  // namespace icsf

namespace icsf {

namespace {
template <typename TV>
__global__ void k_zero_comp(int nC, int comp, TV* x) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nC) x[(size_t)comp * nC + c] = (TV)0;
}
}  // namespace

size_t gmresScalarsBytes() { return sizeof(GmresScalars); }

void gmresCycle(const GmresArgs& a, int vecFp32, hipStream_t st) {
    if (vecFp32) gmresCycleT<float>(a, st);
    else gmresCycleT<double>(a, st);
}

template <typename TV>
static void prepareT(const MeshView& m, Work& wk, GmresScalars* g, const void* W, const void* b, void* tmp1,
                     void* tmp2, int blkFp32, const void* blk, const void* dg, hipStream_t st) {
    const int nC = m.nC;
    const int nbk = reduceBlocks(nC);
    k_sum5<TV><<<nbk, 256, 0, st>>>(nC, (const TV*)W, wk.part);
    k_reduce_final<<<5, 256, 0, st>>>(wk.part, nbk, 5, g->red);
    k_sub_avg<TV><<<nb(nC, 256), 256, 0, st>>>(nC, (const TV*)W, g->red, (TV*)tmp1);
    launchSpmv(m, blkFp32, sizeof(TV) == 4, blk, dg, tmp1, nullptr, nullptr, tmp2, st);
    k_normfactor<TV><<<nbk, 256, 0, st>>>(nC, (const TV*)tmp2, (const TV*)b, wk.part);
    k_reduce_final<<<8, 256, 0, st>>>(wk.part, nbk, 8, g->red);
    k_gm_init<<<1, 64, 0, st>>>(g);
}

void gmresPrepare(const MeshView& m, Work& wk, GmresScalars* g, int vecFp32, const void* W, const void* b,
                  void* tmp1, void* tmp2, int blkFp32, const void* blk, const void* dg, hipStream_t st) {
    if (vecFp32) prepareT<float>(m, wk, g, W, b, tmp1, tmp2, blkFp32, blk, dg, st);
    else prepareT<double>(m, wk, g, W, b, tmp1, tmp2, blkFp32, blk, dg, st);
}

template <typename TV>
static void updResT(const MeshView& m, Work& wk, GmresScalars* g, int blkFp32, const void* blk, const void* dg,
                    int nDirs, const void* V, void* dW, const void* b, void* r, hipStream_t st) {
    const int nC = m.nC;
    const int nbk = reduceBlocks(nC);
    const size_t vs = (size_t)5 * nC;
    k_update_sol<TV, TV><<<nb(vs, 256), 256, 0, st>>>(nC, (TV*)dW, (const TV*)V, vs, nDirs, g->yh);
    launchSpmv(m, blkFp32, sizeof(TV) == 4, blk, dg, dW, nullptr, nullptr, r, st);
    k_resid_l1<TV><<<nbk, 256, 0, st>>>(nC, (const TV*)b, (TV*)r, wk.part);
    k_reduce_final<<<5, 256, 0, st>>>(wk.part, nbk, 5, g->red);
    k_gm_final<<<1, 64, 0, st>>>(g, m.emptyDir);
}

void gmresUpdateAndResidual(const MeshView& m, Work& wk, GmresScalars* g, int vecFp32, int blkFp32,
                            const void* blk, const void* dg, int nDirs, const void* V, void* dW, const void* b,
                            void* r, hipStream_t st) {
    if (vecFp32) updResT<float>(m, wk, g, blkFp32, blk, dg, nDirs, V, dW, b, r, st);
    else updResT<double>(m, wk, g, blkFp32, blk, dg, nDirs, V, dW, b, r, st);
}

void gmresReadResiduals(const GmresScalars* g, double* res10, hipStream_t st) {
    (void)hipMemcpyAsync(res10, g->res, sizeof(double) * 10, hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
}

const double* gmresResidualsDevice(const GmresScalars* g) { return g->res; }

void launchCopy(size_t n, int fp32, const void* a, void* b, hipStream_t st) {
    if (fp32) k_copy<float><<<nb(n, 256), 256, 0, st>>>(n, (const float*)a, (float*)b);
    else k_copy<double><<<nb(n, 256), 256, 0, st>>>(n, (const double*)a, (double*)b);
}

void launchConvert(size_t n, int fromFp32, const void* a, void* b, hipStream_t st) {
    if (fromFp32) k_convert<float, double><<<nb(n, 256), 256, 0, st>>>(n, (const float*)a, (double*)b);
    else k_convert<double, float><<<nb(n, 256), 256, 0, st>>>(n, (const double*)a, (float*)b);
}

namespace {
template <typename T>
__global__ void k_zero(size_t n, T* x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (T)0;
}
}  // namespace

// Zero fill by a kernel (no memset nodes, so the sequence can be captured).
void launchZero(size_t n, int fp32, void* x, hipStream_t st) {
    if (fp32) k_zero<float><<<nb(n, 256), 256, 0, st>>>(n, (float*)x);
    else k_zero<double><<<nb(n, 256), 256, 0, st>>>(n, (double*)x);
}

void launchZeroComponent(int nC, int comp, int fp32, void* x, hipStream_t st) {
    if (fp32) k_zero_comp<float><<<nb(nC, 256), 256, 0, st>>>(nC, comp, (float*)x);
    else k_zero_comp<double><<<nb(nC, 256), 256, 0, st>>>(nC, comp, (double*)x);
}

}  // namespace icsf

// ------------------------------------------------- harmonic balance -----
namespace icsf {

template <typename TB, typename TV>
void spmvHBT(const MeshView& m, const HBBlocks& hb, const void* x, const double* scal, void* vout, void* y,
             hipStream_t st) {
    dim3 grid(xcdGrid(m.nC, 256), hb.nO);
    if (m.blkGen) k_spmv_hb<TB, TV, true><<<grid, 256, 0, st>>>(m, hb, (const TV*)x, scal, (TV*)vout, (TV*)y);
    else k_spmv_hb<TB, TV, false><<<grid, 256, 0, st>>>(m, hb, (const TV*)x, scal, (TV*)vout, (TV*)y);
}

void launchSpmvHB(const MeshView& m, const HBBlocks& hb, int vecFp32, const void* x, const double* scal, void* vout,
                  void* y, hipStream_t st) {
    if (hb.fp32 && vecFp32) spmvHBT<float, float>(m, hb, x, scal, vout, y, st);
    else if (hb.fp32) spmvHBT<float, double>(m, hb, x, scal, vout, y, st);
    else if (vecFp32) spmvHBT<double, float>(m, hb, x, scal, vout, y, st);
    else spmvHBT<double, double>(m, hb, x, scal, vout, y, st);
}

template <typename TB, typename TV, bool GEN>
void lusgsHBG(const MeshView& m, const Levels& L, const HBBlocks& hb, const double* rD, void* r, hipStream_t st) {
    for (int l = 0; l < L.nFwd; ++l) {
        const int s0 = L.fwdPtrHost[l], n = L.fwdPtrHost[l + 1] - s0;
        if (L.fwdSkip && l == 0) continue;
        dim3 grid(xcdGrid(n, 256), hb.nO);
        k_lusgs_hb<TB, TV, true, GEN><<<grid, 256, 0, st>>>(m, hb, rD, (TV*)r, L.fwdCells + s0, n);
    }
    for (int l = 0; l < L.nBwd; ++l) {
        const int s0 = L.bwdPtrHost[l], n = L.bwdPtrHost[l + 1] - s0;
        dim3 grid(xcdGrid(n, 256), hb.nO);
        k_lusgs_hb<TB, TV, false, GEN><<<grid, 256, 0, st>>>(m, hb, rD, (TV*)r, L.bwdCells + s0, n);
    }
}

template <typename TB, typename TV>
void lusgsHBT(const MeshView& m, const Levels& L, const HBBlocks& hb, const double* rD, void* r, hipStream_t st) {
    if (m.blkGen) lusgsHBG<TB, TV, true>(m, L, hb, rD, r, st);
    else lusgsHBG<TB, TV, false>(m, L, hb, rD, r, st);
}

void launchLusgsHB(const MeshView& m, const Levels& L, const HBBlocks& hb, int vecFp32, const double* rD, void* r,
                   hipStream_t st) {
    if (hb.fp32 && vecFp32) lusgsHBT<float, float>(m, L, hb, rD, r, st);
    else if (hb.fp32) lusgsHBT<float, double>(m, L, hb, rD, r, st);
    else if (vecFp32) lusgsHBT<double, float>(m, L, hb, rD, r, st);
    else lusgsHBT<double, double>(m, L, hb, rD, r, st);
}

template <typename TB, typename TV, bool GEN>
void spmvLusgsHBG(const MeshView& m, const Levels& L, const HBBlocks& hb, const double* rD, const void* x,
                  const double* scal, void* vout, void* y, hipStream_t st) {
    const int n0 = L.fwdPtrHost[1], n1 = L.fwdPtrHost[2] - n0;
    k_spmv_hb_cells<TB, TV, false, GEN><<<dim3(xcdGrid(n0, 256), hb.nO), 256, 0, st>>>(
        m, hb, (const TV*)x, scal, (TV*)vout, (TV*)y, rD, L.fwdCells, n0);
    k_spmv_hb_cells<TB, TV, true, GEN><<<dim3(xcdGrid(n1, 256), hb.nO), 256, 0, st>>>(
        m, hb, (const TV*)x, scal, (TV*)vout, (TV*)y, rD, L.fwdCells + n0, n1);
    const int s0 = L.bwdPtrHost[1], nbw = L.bwdPtrHost[2] - s0;
    k_lusgs_hb<TB, TV, false, GEN><<<dim3(xcdGrid(nbw, 256), hb.nO), 256, 0, st>>>(m, hb, rD, (TV*)y,
                                                                                  L.bwdCells + s0, nbw);
}

template <typename TB, typename TV>
void spmvLusgsHBT(const MeshView& m, const Levels& L, const HBBlocks& hb, const double* rD, const void* x,
                  const double* scal, void* vout, void* y, hipStream_t st) {
    if (m.blkGen) spmvLusgsHBG<TB, TV, true>(m, L, hb, rD, x, scal, vout, y, st);
    else spmvLusgsHBG<TB, TV, false>(m, L, hb, rD, x, scal, vout, y, st);
}

bool launchSpmvLusgsHB(const MeshView& m, const Levels& L, const HBBlocks& hb, int vecFp32, const double* rD,
                       const void* x, const double* scal, void* vout, void* y, hipStream_t st) {
    if (!L.twoColour) return false;
    if (hb.fp32 && vecFp32) spmvLusgsHBT<float, float>(m, L, hb, rD, x, scal, vout, y, st);
    else if (hb.fp32) spmvLusgsHBT<float, double>(m, L, hb, rD, x, scal, vout, y, st);
    else if (vecFp32) spmvLusgsHBT<double, float>(m, L, hb, rD, x, scal, vout, y, st);
    else spmvLusgsHBT<double, double>(m, L, hb, rD, x, scal, vout, y, st);
    return true;
}

void sliceSubAverage(const MeshView& m, Work& wk, GmresScalars* g, const double* W, double* out, hipStream_t st) {
    const int nC = m.nC;
    const int nbk = reduceBlocks(nC);
    k_sum5<double><<<nbk, 256, 0, st>>>(nC, W, wk.part);
    k_reduce_final<<<5, 256, 0, st>>>(wk.part, nbk, 5, g->red);
    k_sub_avg<double><<<nb(nC, 256), 256, 0, st>>>(nC, W, g->red, out);
}

void sliceNormFactors(const MeshView& m, Work& wk, GmresScalars* g, const double* t, const double* b,
                      hipStream_t st) {
    const int nC = m.nC;
    const int nbk = reduceBlocks(nC);
    k_normfactor<double><<<nbk, 256, 0, st>>>(nC, t, b, wk.part);
    k_reduce_final<<<8, 256, 0, st>>>(wk.part, nbk, 8, g->red);
    k_gm_init<<<1, 64, 0, st>>>(g);
}

void sliceFinalResidual(const MeshView& m, Work& wk, GmresScalars* g, const double* b, double* r, hipStream_t st) {
    const int nC = m.nC;
    const int nbk = reduceBlocks(nC);
    k_resid_l1<double><<<nbk, 256, 0, st>>>(nC, b, r, wk.part);
    k_reduce_final<<<5, 256, 0, st>>>(wk.part, nbk, 5, g->red);
    k_gm_final<<<1, 64, 0, st>>>(g, m.emptyDir);
}

void launchUpdateSolution(size_t nVec, double* dW, const double* V, size_t vstride, int nv, const GmresScalars* g,
                          hipStream_t st) {
    k_update_sol<double, double><<<nb(5 * nVec, 256), 256, 0, st>>>((int)nVec, dW, V, vstride, nv, g->yh);
}

}  // namespace icsf

// ------------------------------------------------------- flexible GMRES ---
namespace icsf {

namespace {
template <typename TV>
__global__ void k_scale_by(size_t n, const TV* __restrict__ a, const double* __restrict__ beta, TV* __restrict__ b) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i] / (TV)(*beta);
}
}  // namespace

// One FGMRES(m) cycle in FP64 (Saad's right-preconditioned flexible GMRES):
// v_0 = r/||r||; z_j = M_j^-1 v_j (callback); w = A z_j; Gram-Schmidt
// against v_0..v_j; Givens; back-substitution into g->yh.  Z holds z_j.
void fgmresCycle(const FgmresArgs& a, hipStream_t st) {
    const MeshView& m = *a.m;
    const int nC = a.nVec ? a.nVec : m.nC;
    const int nbk = reduceBlocks(nC);
    const size_t vs = (size_t)5 * nC;
    double* r = (double*)a.r;
    double* V = (double*)a.V;
    double* Z = (double*)a.Z;
    GmresScalars* g = a.g;
    double* red = g->red;
    k_norm2<double><<<nbk, 256, 0, st>>>(nC, r, a.wk->part);
    k_reduce_final<<<1, 256, 0, st>>>(a.wk->part, nbk, 1, red);
    k_gm_beta0<<<1, 64, 0, st>>>(g, a.nDirs);
    for (int i = 0; i < a.nDirs; ++i) {
        double* vi = V + (size_t)i * vs;
        double* zi = Z + (size_t)i * vs;
        k_scale_by<double><<<nb(vs, 256), 256, 0, st>>>(vs, r, &g->beta, vi);
        a.precond(vi, zi);
        // the outer product is timed with the other work, so that the
        // operator class holds the Arnoldi operator groups only
        if (a.mark) a.mark(TK_OTHER, 1);
        if (a.op) a.op(zi, r);
        else launchSpmv(m, a.blkFp32, 0, a.blk, a.dg, zi, nullptr, nullptr, r, st);
        if (a.mark) a.mark(TK_OTHER, 0);
        if (a.mark) a.mark(TK_ORTH, 1);
        if (a.orth == 0) {
            for (int j = 0; j <= i + 1; ++j) {
                const double* vprev = j > 0 ? V + (size_t)(j - 1) * vs : nullptr;
                const double* vnext = j <= i ? V + (size_t)j * vs : nullptr;
                const double* hprev = j > 0 ? &g->H[j - 1][i] : nullptr;
                k_mgs_step<double><<<nbk, 256, 0, st>>>(nC, r, vprev, hprev, vnext, j == i + 1, a.wk->part);
                k_reduce_mgs<<<1, 256, 0, st>>>(a.wk->part, nbk, g, i, j);
            }
        } else {
            const int ch = m.mdCh;
            for (int j0 = 0; j0 <= i; j0 += ch) {
                const int nv = (i + 1 - j0) < ch ? (i + 1 - j0) : ch;
                launchMultidot<double>(ch, nbk, nC, r, V, vs, j0, nv, a.wk->part, st);
                k_reduce_col<<<nv, 256, 0, st>>>(a.wk->part, nbk, ch, g, i, j0);
            }
            k_cgs_update<double><<<nbk, 256, 0, st>>>(nC, r, V, vs, i + 1, g->hcol, a.wk->part);
            k_reduce_givens<<<1, 256, 0, st>>>(a.wk->part, nbk, g, i);
        }
        if (a.mark) a.mark(TK_ORTH, 0);
    }
    k_gm_backsub<<<1, 64, 0, st>>>(g, a.nDirs);
}

// z = sum_i y_i V_i from zero (inner solve result), FP32 or FP64 vectors.
void launchCombine(int nC, int vecFp32, const void* V, int nv, const GmresScalars* g, void* z, hipStream_t st) {
    const size_t vs = (size_t)5 * nC;
    launchZero(vs, vecFp32, z, st);
    if (vecFp32)
        k_update_sol<float, float><<<nb(vs, 256), 256, 0, st>>>(nC, (float*)z, (const float*)V, vs, nv, g->yh);
    else
        k_update_sol<double, double><<<nb(vs, 256), 256, 0, st>>>(nC, (double*)z, (const double*)V, vs, nv, g->yh);
}

void fgmresUpdateAndResidual(const MeshView& m, Work& wk, GmresScalars* g, int blkFp32, const void* blk,
                             const void* dg, int nDirs, const void* Z, void* dW, const void* b, void* r,
                             hipStream_t st) {
    updResT<double>(m, wk, g, blkFp32, blk, dg, nDirs, Z, dW, b, r, st);
}

}  // namespace icsf
