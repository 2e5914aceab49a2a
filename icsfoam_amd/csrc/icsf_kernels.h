// Launch interfaces of the HIP kernels (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <functional>
#include <vector>

#include "icsf_views.h"

namespace icsf {

constexpr int KMAX = 64;  // largest nDirections supported by the device GMRES

// Sweep levels of the LU-SGS dependency graph (cells grouped by level).
struct Levels {
    int nFwd = 0, nBwd = 0;
    std::vector<int> fwdPtrHost, bwdPtrHost;  // [n+1] offsets into the cell lists
    int* fwdCells = nullptr;                  // device, [nC]
    int* bwdCells = nullptr;                  // device, [nC]
    bool fwdSkip = true;                      // level 0 has no lower neighbours
    // two levels each way, forward level 1 == backward level 0 (a two-colour
    // numbering: every neighbour of a level-1 cell is a level-0 cell)
    bool twoColour = false;
};

// Scratch for two-stage reductions.
struct Work {
    double* part = nullptr;  // [1024][64]
};

struct GmresScalars;

struct GmresArgs {
    const MeshView* m;
    const Levels* lev;
    Work* wk;
    GmresScalars* g;  // device
    int nDirs;
    int precond;      // 0 LUSGS, 1 Jacobi, 2 none
    int orth;         // 0 MGS (reference order), 1 CGS
    int blkFp32;
    const void* blk;
    const void* dg;
    const double* rD;
    const void* dinv;
    void* r;          // residual / work vector, [5][nC]
    void* r2;         // second work vector (SpMV output), [5][nC]
    void* V;          // Krylov basis, nDirs vectors of [5][nC]
    // optional timing marks: mark(kernelClass, begin?1:0)
    std::function<void(int, int)> mark;
    // Batched (harmonic-balance) systems: vectors of 5*nVec entries and
    // caller-supplied operator y = A (x/scal) (vout = x/scal when given)
    // and preconditioner v <- M^-1 v.  nVec 0 / empty hooks: the single
    // mesh SpMV and LU-SGS/Jacobi above.
    int nVec = 0;
    std::function<void(const void*, const double*, void*, void*)> op;
    std::function<void(void*)> pre;
    // fuse the Arnoldi SpMV with the LU-SGS sweeps when the level structure
    // allows it (same algorithm; the forward-sweep terms of a row are summed
    // in three partial sums instead of one running difference)
    bool fuse = false;
    // batched systems: fused operator + preconditioner (returns false when
    // not applicable; op and pre are then used)
    std::function<bool(const void*, const double*, void*, void*)> opPre;
};

// Harmonic balance, one launch per stage for all instances
// (blockIdx.y = instance): each instance's mesh view (its own boundary
// conditions), state and FP64 Jacobian storage, in device memory.
struct HBInst {
    MeshView m;
    StateView s;
    void* blk;    // FP64, or FP32 with mixed precision
    void* dg;
    double* rD;
    int* err;
};
// The harmonic-balance terms formed in the instances' Jacobian launch
// (HBZone.C:359-518): with the residual cell sums, the spectral source into
// the coupled right-hand side; V D_JJ on each instance's diagonal blocks;
// then (one small launch) the shared LU-SGS diagonal over all instances.
struct HBCouple {
    int nO;
    const double* D;  // [nO][nO] spectral operator
    double* Wall;     // [nO][5][nC] coupled state (written with the source)
    double* ball;     // [nO][5][nC] coupled right-hand side b + source
    double* rD;       // [nC] shared LU-SGS reciprocal diagonal
    int* err;         // zero-diagonal flag of the coupled system
};
// cellSums: as launchResidual / launchCellSumsJacobian, for all instances
void launchResidualHB(const HBInst* inst, int nO, const MeshView& m, const Phys& ph, hipStream_t st,
                      bool cellSums = true);
// Fourier reconstruction of the instances' fields at one time, out [nC][6]
void launchHBReconstruct(int nC, int nO, const HBInst* inst, const double* w, double* out, hipStream_t st);
void launchJacobianHB(const HBInst* inst, int nO, int fp32, const MeshView& m, const Phys& ph, hipStream_t st,
                      bool cellSums, const HBCouple& hc);

// Harmonic-balance coupling: per-instance block storage (FP64) and the
// spectral operator D [nO][nO] (device).
constexpr int HB_MAXO = 16;
struct HBBlocks {
    int nO;
    int fp32;                     // blocks stored in FP32 (mixed precision)
    const void* blk[HB_MAXO];
    const void* dg[HB_MAXO];
    const double* D;
};

// kernel classes of the timing marks
enum { TK_RESIDUAL = 0, TK_JACOBIAN = 1, TK_SPMV = 2, TK_PRECOND = 3, TK_ORTH = 4, TK_OTHER = 5 };

// Flexible GMRES (FP64 outer); precond(v, z) writes z = M^-1 v.
struct FgmresArgs {
    const MeshView* m;
    Work* wk;
    GmresScalars* g;
    int nDirs;
    int orth;
    int blkFp32;
    const void* blk;
    const void* dg;
    void* r;  // [5][nC] FP64
    void* V;  // nDirs+1 FP64 vectors
    void* Z;  // nDirs FP64 vectors
    std::function<void(const double*, double*)> precond;
    std::function<void(int, int)> mark;
    // batched (harmonic-balance) systems: 5*nVec entries per vector and a
    // caller-supplied y = A z (FP64 vectors); nVec 0 / empty: the mesh SpMV
    int nVec = 0;
    std::function<void(const double*, double*)> op;
};

int reduceBlocks(int nC);
// Full 5x5 off-diagonal blocks (sliced layout) from generator storage.
void launchExpandBlocks(const MeshView& m, int fp32, const void* gen, void* out, hipStream_t st);
void launchProjectionPasses(const MeshView& m, int vecFp32, int ch, const void* w, const void* V, int nv,
                            double* part, hipStream_t st);

void launchUpdate(const MeshView& m, const Phys& ph, const StateView& s, const double* dW, int mode,
                  const double* rho0, const double* U0, const double* T0, hipStream_t st);
// Forward part of the residual adjoint: boundary values and gradients,
// keeping the unlimited gradients and limiter values.
void launchAdjointForward(const MeshView& m, const Phys& ph, const StateView& s, double* gun, double* limo,
                          hipStream_t st);
// Reverse pass: gW = (db/dW)^T v, gS = (db/dSf)^T v, gC = (db/dCf)^T v
// (geom false: gS and gC are work space only, and the Roe reverse skips
// its Sf passes).
void launchResidualVjp(const MeshView& m, const Phys& ph, const StateView& s, const AdjView& a, const double* v,
                       double* gW, double* gS, double* gC, bool geom, hipStream_t st);
// cellSums false: everything up to the face terms; the cell sums (b,
// rPseudoDeltaT) are then formed by launchCellSumsJacobian.
void launchResidual(const MeshView& m, const Phys& ph, const StateView& s, hipStream_t st, bool cellSums = true);
void launchFaceFlux(const MeshView& m, const Phys& ph, const StateView& s, double* phi, double* phiU, double* phiE,
                    hipStream_t st);
void launchJacobian(const MeshView& m, const Phys& ph, const StateView& s, void* blk, void* dg, int fp32,
                    double* rD, int* err, hipStream_t st);
// Cell sums of the residual (b, rPseudoDeltaT) and the Jacobian rows in one
// kernel, after launchResidual(..., cellSums = false); same results as
// k_cell_residual followed by launchJacobian.
void launchCellSumsJacobian(const MeshView& m, const Phys& ph, const StateView& s, void* blk, void* dg, int fp32,
                            double* rD, int* err, hipStream_t st);

void launchSpmv(const MeshView& m, int blkFp32, int vecFp32, const void* blk, const void* dg, const void* x,
                const double* scal, void* vout, void* y, hipStream_t st);
void launchLusgs(const MeshView& m, const Levels& L, int blkFp32, int vecFp32, const void* blk, const double* rD,
                 void* r, hipStream_t st);
// y = LU-SGS(A (x/scal)) in three passes on a two-colour level structure
// (false: not applicable, nothing launched).
bool launchSpmvLusgs(const MeshView& m, const Levels& L, int blkFp32, int vecFp32, const void* blk, const void* dg,
                     const double* rD, const void* x, const double* scal, void* vout, void* y, hipStream_t st);
void launchJacobiSetup(int nC, int blkFp32, const void* dg, void* dinv, hipStream_t st);
void launchJacobiApply(int nC, int blkFp32, int vecFp32, const void* dinv, void* r, hipStream_t st);

// GMRES building blocks (icsf_krylov.hip)
size_t gmresScalarsBytes();
void gmresPrepare(const MeshView& m, Work& wk, GmresScalars* g, int vecFp32, const void* W, const void* b,
                  void* tmp1, void* tmp2, int blkFp32, const void* blk, const void* dg, hipStream_t st);
void gmresCycle(const GmresArgs& a, int vecFp32, hipStream_t st);
void gmresUpdateAndResidual(const MeshView& m, Work& wk, GmresScalars* g, int vecFp32, int blkFp32,
                            const void* blk, const void* dg, int nDirs, const void* V, void* dW, const void* b,
                            void* r, hipStream_t st);
void fgmresCycle(const FgmresArgs& a, hipStream_t st);
void fgmresUpdateAndResidual(const MeshView& m, Work& wk, GmresScalars* g, int blkFp32, const void* blk,
                             const void* dg, int nDirs, const void* Z, void* dW, const void* b, void* r,
                             hipStream_t st);
void launchCombine(int nC, int vecFp32, const void* V, int nv, const GmresScalars* g, void* z, hipStream_t st);

// Harmonic balance (icsf_krylov.hip): y = A_HB (x/scal) over [nO][5][nC]
// vectors in coupledMatrix order, with vout = x/scal when given.
void launchSpmvHB(const MeshView& m, const HBBlocks& hb, int vecFp32, const void* x, const double* scal, void* vout,
                  void* y, hipStream_t st);
// Per-variable bookkeeping of one instance slice ([5][nC]) into g: W minus
// its per-component average; norm factors and initial residuals from
// t = A (W - avg) and b; final residuals from r = b - A dW (r holds A dW on
// entry).
// y = LU-SGS(A_HB (x/scal)) in three passes on a two-colour level
// structure (false: not applicable, nothing launched).
bool launchSpmvLusgsHB(const MeshView& m, const Levels& L, const HBBlocks& hb, int vecFp32, const double* rD,
                       const void* x, const double* scal, void* vout, void* y, hipStream_t st);
// LU-SGS of all HB instances (per-instance sweeps, shared diagonal rD).
void launchLusgsHB(const MeshView& m, const Levels& L, const HBBlocks& hb, int vecFp32, const double* rD, void* r,
                   hipStream_t st);
void sliceSubAverage(const MeshView& m, Work& wk, GmresScalars* g, const double* W, double* out, hipStream_t st);
void sliceNormFactors(const MeshView& m, Work& wk, GmresScalars* g, const double* t, const double* b,
                      hipStream_t st);
void sliceFinalResidual(const MeshView& m, Work& wk, GmresScalars* g, const double* b, double* r, hipStream_t st);
// dW += sum_i y_i V_i over 5*nVec entries
void launchUpdateSolution(size_t nVec, double* dW, const double* V, size_t vstride, int nv, const GmresScalars* g,
                          hipStream_t st);
void gmresReadResiduals(const GmresScalars* g, double* res10, hipStream_t st);
const double* gmresResidualsDevice(const GmresScalars* g);

void launchCopy(size_t n, int fp32, const void* a, void* b, hipStream_t st);
void launchConvert(size_t n, int fromFp32, const void* a, void* b, hipStream_t st);
void launchZero(size_t n, int fp32, void* x, hipStream_t st);
void launchZeroComponent(int nC, int comp, int fp32, void* x, hipStream_t st);

}  // namespace icsf
