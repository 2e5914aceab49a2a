// Device-side views of the mesh, physical settings and per-case buffers.
// All pointers are device pointers owned by the library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace icsf {

// reconstructed cell fields, AoS [nCells][NFLD]
// F_CS holds the scheme's reconstructed speed: c* (AUSM+up) or the sound
// speed max(sqrt(gamma/psi), VSMALL) (HLLC); unused by Roe.
enum { F_RHO = 0, F_P = 1, F_UX = 2, F_UY = 3, F_UZ = 4, F_E = 5, F_H = 6, F_CS = 7, NFLD = 8 };
enum { FLUX_AUSM = 0, FLUX_ROE = 1, FLUX_HLLC = 2 };
// boundary-condition kinds: p zeroGradient / fixedValue / totalPressure;
// U zeroGradient / fixedValue / slip / pressureInletOutletVelocity;
// T zeroGradient / fixedValue / totalTemperature
enum { BC_ZG = 0, BC_FIXED = 1, BC_SLIP = 2, BC_TOTAL_P = 3, BC_PIOV = 4, BC_TOTAL_T = 5 };
enum { PK_PATCH = 0, PK_WALL = 1, PK_EMPTY = 2, PK_CYCLIC = 3 };
constexpr int NB = 25;  // entries of a 5x5 block

// viscous gradient row of a cell (StateView::gv): 12 doubles (96 B), so a
// row is six 16-byte accesses
enum { GV_U = 0, GV_E = 9, NGV = 12 };

// per-face terms written by the face kernel, AoS [nFaces][NFT]
enum { FT_CONV = 0, FT_LAP = 5, FT_TAU = 8, FT_SIG = 11, FT_LAPE = 12, FT_LAM = 13, FT_FRDT = 14, FT_REC = 15,
       FT_SV = 23, NFT = 24 };

struct MeshView {
    int nC, nF, nI, nB;
    const int* own;       // [nF]
    const int* nei;       // [nI]
    const double* Sf;     // [nF][3]
    const double* magSf;  // [nF]
    const double* Cf;     // [nF][3]
    const double* w;      // [nF]
    const double* dc;     // [nF]
    const double* nodc;   // [nF]
    const double* corr;   // [nF][3]
    const double* C;      // [nC][3]
    const double* V;      // [nC]
    const int* cfOff;     // [nC+1] cell -> faces (sorted by face, no empty faces)
    const int* cf;
    const int* bKind;     // [nB] patch kind of each boundary face
    const int* bPT;       // [nB] p boundary-condition type
    const int* bUT;       // [nB] U boundary-condition type
    const int* bTT;       // [nB] T boundary-condition type
    const double* bPv;    // [nB] fixed p, or totalPressure p0
    const double* bUv;    // [nB][3] fixed U, or the tangentialVelocity of
                          //   pressureInletOutletVelocity
    const double* bTv;    // [nB] fixed T, or totalTemperature T0
    // sliced-ELL off-diagonal layout
    int nSlices;
    const int* slicePtr;  // [nSlices+1]
    const int* col;       // [slicePtr[nSlices]*64]
    const int* rowLen;    // [nC]
    const int* nLower;    // [nC]
    // geometry of each slot's internal face, same sliced layout (sgOff):
    // Sf (3), weight, Cf (3), magSf; a cell's slots follow its internal
    // faces in face order (upper-triangular numbering: lower faces, then
    // upper faces, both by column), so cell kernels read it contiguously
    const double* sgeo;
    int emptyDir;
    // cyclic faces k = 0..nCyc-1 (face order); their 5x5 interface blocks
    // follow the sliced-ELL blocks in the same allocation, at
    // cycBase + k*25 + e
    int nCyc;
    size_t cycBase;
    const int* cycOf;        // [nB] k, or -1
    const int* cycCell;      // [nCyc] cell on the far side
    const double* cycR;      // [nCyc][9] forwardT (far side -> this side)
    const double* cycDelta;  // [nCyc][3] cyclicFvPatch::delta()
    const int* cycRowOff;    // [nC+1] cell -> its cyclic faces
    const int* cycRowK;      // [nCyc]
    // launch parameters (host side; chosen by the autotuner, else these
    // defaults): workgroup sizes of the Arnoldi-operator row kernels (SpMV,
    // LU-SGS), the face-terms kernel, the gradient and cell-residual
    // kernels and the Jacobian rows; basis vectors per CGS projection pass.
    // None of them changes a result bit: row kernels compute each row in
    // one thread, and the projection sums are per vector.
    int bsOp = 256, bsFace = 256, bsCell = 256, bsJac = 256, mdCh = 16;
    // register budget of the same kernel classes: 0 the compiler's choice,
    // 1 at most 128 VGPRs (4 waves per SIMD, spilling what does not fit),
    // 2 at most 168 (3 waves; residual and Jacobian kernels)
    int occOp = 0, occFace = 0, occCell = 0, occJac = 0;
    // off-diagonal blocks of the internal faces stored as generators (1) or
    // as full 5x5 blocks (0); gamma for generating them
    int blkGen = 1;
    double gamma = 1.4;
    // rotating zone (MRF): normal frame velocity of every face
    // (faceU & Sf/|Sf|, zero outside the zone) and Omega of every cell (zero
    // outside); null without one
    const double* umrf = nullptr;  // [nF]
    const double* omC = nullptr;   // [nC][3]
    // harmonic balance with phase-lagged cyclic patches: the lagged side of
    // every cyclic face (-1: a plain cyclic face); null without any
    const int* plSide = nullptr;  // [nCyc]
    // rotating or translating zone with generator storage: the frame
    // divergence's diagonal term of every off-diagonal block (one value per
    // sliced-ELL slot, in the generators' precision, slot * 64 + lane);
    // the same for every harmonic-balance instance (geometry and frame
    // only); null without a zone
    void* genMo = nullptr;
};

// Lagged far-side record of a phase-lagged cyclic face (PLW doubles per
// face and instance): slots 0..NFLD-1 the values cycNbrValue gives field k
// (the velocity components unrotated unless cylCoords), PL_UV..+2 the
// lagged velocity vector with forwardT applied, PL_C the lagged
// sqrt(gamma/psi), PL_G the lagged gamma.
enum { PL_UV = 8, PL_C = 11, PL_G = 12, PLW = 16 };

struct Phys {
    double R, gamma, Cv, mu, alpha, entropyFix;
    int flux, lowMach, recon, limiter, viscous;
    double rhoMin, TMin, TMax;
};

struct StateView {
    double* W;        // [5][nC] conserved, order rho, rhoE, rhoUx, rhoUy, rhoUz
    double* fld;      // [nC][NFLD]
    double* T;        // [nC]
    double* ecalc;    // [nC]  rhoE/rho - |U|^2/2
    double* bfld;     // [nB][NFLD]
    double* bT;       // [nB]
    double* grd;      // [nC][NFLD][3] limited gradients
    double* gv;       // [nC][NGV] viscous gradients: Gauss grad(U) (GV_U..+8,
                      //   row-major, grad(U)_ij = d U_j / d x_i), then Gauss
                      //   grad(ecalc) (GV_E..+2)
    double* pseudoCo; // [nC]
    double* rDT;      // [nC]
    double* b;        // [5][nC]
    double* ft;       // [nF][NFT] face terms
    int* err;         // device error flags: ERR_FLAG_DIAG | ERR_FLAG_NONFINITE
    // boundary state of the derived conditions (totalPressure,
    // pressureInletOutletVelocity, totalTemperature), which depend on
    // lagged values: bcV [nB][5] = p, Ux, Uy, Uz, T of every boundary face
    // as of the last state update (updateFields.H:90-105); bcPhi [nB] the
    // boundary mass flux of the last face-terms pass (the registered phi);
    // bcIn [nB] the valueFraction switch of pressureInletOutletVelocity (1:
    // inflow at its last update); bcInit [nB][5] the initial (`value`)
    // boundary values used by the initialisation
    double* bcV;
    double* bcPhi;
    int* bcIn;
    const double* bcInit;
    // end flag of the device-controlled pseudo-time loop (nonzero: the loop
    // has ended, and steps still queued leave the state as it is: the
    // state update, the boundary correction and the boundary mass flux
    // return at once)
    const int* halt;
    // phase-lagged cyclic patches (harmonic balance): e [nC] the cells'
    // internal energy as the state update leaves it (the registered e),
    // written when non-null; plFar [nCyc][PLW] this instance's lagged
    // far-side records (MeshView::plSide)
    double* e = nullptr;
    const double* plFar = nullptr;
};

// Device control block of the pseudo-time loop (pseudotimeControl):
// ints [NCTL] and the stored initial residuals [5].
enum { CTL_CORR = 0, CTL_CONV = 1, CTL_FIN = 2, CTL_RUN = 3, CTL_EN = 4, CTL_FIRST = 5, NCTL = 8 };

// bits of the device error flags
enum { ERR_FLAG_DIAG = 1, ERR_FLAG_NONFINITE = 2 };

// Buffers of the residual adjoint (vector-Jacobian product).
struct AdjView {
    double* W0;     // [5][nC] conserved state the residual is differentiated at
    double* gun;    // [nC][NFLD][3] unlimited Gauss gradients
    double* lim;    // [nC][NFLD] limiter values
    double* fLR;    // [nF][2*NFLD] adjoint of owner-/neighbour-side face values
                    //   (boundary faces: slot k holds the boundary value's adjoint)
    double* G;      // [nC][NFLD][3] adjoint of the unlimited gradient / V
    double* qc;     // [nC][NFLD] partial cell-value adjoint
    int* tgt;       // [nC][NFLD] limiter max/min element: cell, nC + boundary face, -1 none
    double* tval;   // [nC][NFLD] adjoint sent to that element
    int* fst;       // [nC][NFLD] face that set the limiter, -1 none
    double* fext;   // [nC][NFLD] adjoint of that face's extrapolation
    double* qf;     // [nF][NFLD] adjoint of the Gauss face values (internal faces)
    double* bprim;  // [nB][5] boundary adjoint forwarded to the owner's p, T, U
    int* fbr;       // [nF] branch bits of each face's flux (records for derivative
                    //   checks; null: not written)
    // viscous cases (null otherwise):
    double* vis;    // [nF][NVIS] face adjoints of the viscous terms: internal faces
                    //   Z (9, the interpolated grad(U), tau path included), grad(e)
                    //   (3), U of the owner (3) and of the neighbour (3), e of the
                    //   owner and of the neighbour; boundary faces grad(U) of the
                    //   owner (9), 0 (3), U of the owner (3), the boundary U (3), e
                    //   of the owner, the boundary e
    double* qe;     // [nF] adjoint of the Gauss face value of e (internal faces)
    double* Ge;     // [nC][3] adjoint of the Gauss sum of grad(e), divided by V
};
enum { VS_Z = 0, VS_GE = 9, VS_UO = 12, VS_UN = 15, VS_EO = 18, VS_EN = 19, NVIS = 20 };

// Workgroups of a launch are dispatched round-robin over the 8 XCDs, each
// with its own L2.  Kernels that gather neighbour data launch a multiple of
// 8 workgroups (xcdGrid) and remap with xcdBlock(), so that XCD x runs the
// contiguous workgroup range [x g/8, (x+1) g/8): consecutive rows, whose
// neighbour gathers overlap, then share one L2.
constexpr int kXCD = 8;
inline int xcdGrid(size_t n, int b) {
    const int g = (int)((n + b - 1) / b);
    return (g + kXCD - 1) / kXCD * kXCD;
}
__device__ __forceinline__ int xcdBlock() {
    const int per = gridDim.x / kXCD;
    return (blockIdx.x % kXCD) * per + blockIdx.x / kXCD;
}

// Block storage offset in the sliced-ELL layout: slot index (as stored in
// col[]) and block entry e.
__host__ __device__ __forceinline__ size_t sellOff(size_t slot, int e) {
    return ((slot >> 6) * NB + (size_t)e) * 64 + (slot & 63);
}
// Generator storage of an off-diagonal LF block (NGEN values per slot,
// same sliced layout): Ux, Uy, Uz, E of the face side the block belongs
// to, the unit normal, the signed scale h (+|S|/2 upper, -|S|/2 lower) and
// the diagonal shifts v = |S| lambda / 2 and sv (viscous).
constexpr int NGEN = 10;
__host__ __device__ __forceinline__ size_t genOff(size_t slot, int e) {
    return ((slot >> 6) * NGEN + (size_t)e) * 64 + (slot & 63);
}
// Non-temporal loads and stores for data streamed once per pass (mesh
// geometry in face or slot order, face records, matrix output): they keep
// L2 for the neighbour gathers.  Values are unchanged.
template <typename T>
__device__ __forceinline__ T ldnt(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void stnt(T* p, T v) { __builtin_nontemporal_store(v, p); }

// Rows read or written whole (cell fields, gradients, face records): N
// doubles at a 16-byte aligned address as 128-bit accesses, one
// global_load_dwordx4 / global_store_dwordx4 per pair.  A gather of a
// neighbour's row then takes N/2 memory instructions instead of N; in the
// cell and face kernels, where every lane of a gather touches its own cache
// line, the number of gather instructions is what bounds the kernel.
template <int N>
__device__ __forceinline__ void ldRow(const double* p, double (&v)[N]) {
    static_assert(N % 2 == 0, "rows of an even number of doubles");
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
        const double2 t = q[i];
        v[2 * i] = t.x;
        v[2 * i + 1] = t.y;
    }
}
template <int N>
__device__ __forceinline__ void stRow(double* p, const double (&v)[N]) {
    double2* q = reinterpret_cast<double2*>(p);
#pragma unroll
    for (int i = 0; i < N / 2; ++i) q[i] = double2{v[2 * i], v[2 * i + 1]};
    if (N % 2) p[N - 1] = v[N - 1];
}

// Integer rows (N a multiple of 4, 16-byte aligned): int4 accesses.
template <int N>
__device__ __forceinline__ void ldRowI(const int* p, int (&v)[N]) {
    static_assert(N % 4 == 0, "rows of a multiple of four ints");
    const int4* q = reinterpret_cast<const int4*>(p);
#pragma unroll
    for (int i = 0; i < N / 4; ++i) {
        const int4 t = q[i];
        v[4 * i] = t.x; v[4 * i + 1] = t.y; v[4 * i + 2] = t.z; v[4 * i + 3] = t.w;
    }
}

// Slot geometry (MeshView::sgeo) of flat slot index `slot`, entry e.
constexpr int NSG = 8;
enum { SG_SX = 0, SG_W = 3, SG_CX = 4, SG_MAG = 7 };
__host__ __device__ __forceinline__ size_t sgOff(size_t slot, int e) {
    return ((slot >> 6) * NSG + (size_t)e) * 64 + (slot & 63);
}
// Diagonal block of row r, entry e.
__host__ __device__ __forceinline__ size_t diagOff(int r, int e) {
    return ((size_t)(r >> 6) * NB + (size_t)e) * 64 + (r & 63);
}

}  // namespace icsf
