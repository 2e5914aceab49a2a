// Host driver and C ABI of the implicit coupled pseudo-time step.
//
// Mirrors dbnsFoam's per-step sequence (applications/solvers/dbnsFoam:
// outerLoop.H, setCoAndDeltaT.H, boundLocalTimeStep.H, updateFields.H) with
// all fields, the block Jacobian and the Krylov basis resident on the GPU;
// the only host<->device traffic inside a step is kernel launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <functional>
#include <vector>

#include "icsf.h"
#include "icsf_dict.h"
#include "icsf_hb.h"
#include "icsf_kernels.h"
#include "icsf_math.h"
#include "icsf_mesh.h"

using namespace icsf;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(ICSF_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

int checkDevice() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        return fail(ICSF_ERR_NODEVICE, "no HIP device is available; the coupled solver runs only on the GPU");
    return ICSF_OK;
}

struct DevBuf {
    std::vector<void*> ptrs;
    template <class T>
    int alloc(T*& p, size_t n) {
        p = nullptr;
        if (n == 0) n = 1;
        hipError_t e = hipMalloc((void**)&p, n * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(ICSF_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        ptrs.push_back(p);
        return ICSF_OK;
    }
    template <class T>
    int upload(T*& p, const T* h, size_t n, hipStream_t st = nullptr) {
        int rc = alloc(p, n);
        if (rc) return rc;
        if (n) HIPCHK(hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice));
        (void)st;
        return ICSF_OK;
    }
    // frees one allocation of this buffer set (after the stream is idle)
    void release(void* p) {
        for (size_t i = 0; i < ptrs.size(); ++i)
            if (ptrs[i] == p) {
                (void)hipFree(p);
                ptrs.erase(ptrs.begin() + (long)i);
                return;
            }
    }
    ~DevBuf() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

// Fixed work space of icsf_residual_vjp_device_noalloc, for callers that
// cannot allocate device memory (code under graph capture, serving loops):
// one array per GPU, held until the process ends.  In host code the name
// `scratch` is only a stand-in; its device address on each GPU comes from
// hipGetSymbolAddress (scratchAddress), taken once per device.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
std::mutex g_scratchMu;
char* g_scratchDev[64] = {};

// device address of `scratch` on device dev (the current device)
int scratchAddress(int dev, char** out) {
    if (dev < 0 || dev >= 64) return fail(ICSF_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchDev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ICSF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchDev[dev] = (char*)p;
    }
    *out = g_scratchDev[dev];
    return ICSF_OK;
}

// Consecutive 256-byte aligned pieces of the scratch array (base nullptr:
// sizes only).
struct ScratchCarve {
    char* base;
    size_t off;
    template <class T>
    T* take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T* p = base ? (T*)(base + off) : nullptr;
        off += (n ? n : 1) * sizeof(T);
        return p;
    }
};

#define RC(x)                   \
    do {                        \
        int rc_ = (x);          \
        if (rc_) return rc_;    \
    } while (0)

}  // namespace

// ------------------------------------------------------------------ mesh --
struct icsf_mesh {
    HostMesh h;
    SellLayout sell;
    Levels lev;
    DevBuf buf;
    MeshView view{};
};

// ------------------------------------------------------------------ case --
enum { NTIMER = 6 };

// One frame zone's contribution (rotating or translating): faceU & Sf/|Sf|
// per face, the moving boundary faces and their U.
struct FramePart {
    bool on = false;
    std::vector<double> un;
    std::vector<char> incl;
    std::vector<double> Ub;
    std::vector<char> set;  // faces the zone assigns (face types 1 and 2)
};

struct icsf_case {
    icsf_mesh* mesh = nullptr;
    hipStream_t st = nullptr;
    Phys ph{};
    // solution controls
    int solver = 0;      // 0 GMRES (left, reference), 1 FGMRES (right, outer)
    int precond = 0;     // 0 LUSGS, 1 Jacobi, 2 none
    int nDirs = 30, maxIter = 1, minIter = 0, innerDirs = 30, orth = 0, mixed = 0;
    bool fuse = true;  // SpMV fused with the LU-SGS sweeps on two-colour numberings
    bool fuseCells = true;  // pseudo-step: residual cell sums formed in the Jacobian's face walk
    bool autotune = true, tuned = false;
    bool blkGen = true;  // off-diagonal blocks as generators (jacobianStorage)
    double tol = 1e-12, relTol = 1e-2;
    double co = 1, coMin = 0.1, coMax = 25, coIncr = 1.25, coDecr = 0.1, lowerBound = 0.95;
    int localTS = 1, bounding = 1;
    // boundary conditions (per boundary face)
    std::vector<int> bPT, bUT, bTT;
    std::vector<double> bPv, bUv, bTv;
    int *d_bPT = nullptr, *d_bUT = nullptr, *d_bTT = nullptr;
    double *d_bPv = nullptr, *d_bUv = nullptr, *d_bTv = nullptr;
    // initial values of the derived conditions [nB][5] (p, U, T; NaN: the
    // default p0, the owner cell's U, T0), uploaded by icsf_set_fields
    std::vector<double> bInit;
    double* d_bInit = nullptr;
    bool bcDirty = true;
    // frame zones (icsf_set_mrf, icsf_set_translating_frame): their parts,
    // the boundary faces that move with them and the frame velocity there
    // (their fixedValue U is set to it)
    bool mrf = false;
    // rotating zones in list order (MRFCoupledZoneList), their combination
    // frot and the translating zone ftr; Omega of every cell
    std::vector<FramePart> rotZones;
    FramePart frot, ftr;
    std::vector<double> omcHost;
    std::vector<char> mrfIncl;
    std::vector<double> mrfUb;
    double *d_umrf = nullptr, *d_omC = nullptr;
    void* d_genMo = nullptr;  // the frame divergence term per slot (generator storage)
    MeshView mv{};
    StateView sv{};
    // residual adjoint (allocated on first use)
    AdjView adj{};
    double *adjGS = nullptr, *adjGC = nullptr;
    bool adjReady = false;
    // Jacobian
    void *blk = nullptr, *dg = nullptr, *dinv = nullptr;
    double* rD = nullptr;
    int* derr = nullptr;
    // pseudo-time loop control (pseudotimeControl): device block [NCTL],
    // stored initial residuals [5]; criteria from fvSolution pseudoTime
    // (pseudoTol / pseudoTolRel NaN when absent)
    int* dctl = nullptr;
    double* dctlInit = nullptr;
    double pseudoTol = 0, pseudoTolRel = 0;
    int nPseudoCorr = 20, nPseudoCorrMin = 1;
    int blkFp32 = 0;
    bool haveResidual = false, assembled = false, haveFields = false;
    // Krylov work (FP64)
    double *V = nullptr, *r = nullptr, *t1 = nullptr, *t2 = nullptr, *dW = nullptr, *Z = nullptr;
    double* dres = nullptr;  // [2][5] init / previous initial residuals
    double* fac = nullptr;
    // FP32 inner solve
    float *V32 = nullptr, *r32 = nullptr, *r32b = nullptr;
    GmresScalars* g = nullptr;
    GmresScalars* g2 = nullptr;
    Work wk;
    // pseudo-time
    int corr = 0;
    bool haveInit = false, havePrev = false;
    // timing
    bool timing = false;
    int tmask = 0;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> events;
    std::vector<hipEvent_t> open;
    double tms[NTIMER] = {0};
    long long tcnt[NTIMER] = {0};
    DevBuf buf;
    bool ownStream = true;
    int device = 0;  // HIP device the case's buffers and stream live on
    // one steady-state pseudo-step captured as a HIP graph and replayed
    bool useGraph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    ~icsf_case() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& e : events) { (void)hipEventDestroy(e.second.first); (void)hipEventDestroy(e.second.second); }
        for (auto& e : pool) (void)hipEventDestroy(e);
        for (auto& e : open)
            if (e) (void)hipEventDestroy(e);
        if (st && ownStream) (void)hipStreamDestroy(st);
    }
};

namespace {

hipEvent_t takeEvent(icsf_case* c) {
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// begin/end marks of a kernel class on the case stream (nested per class)
void markClass(icsf_case* c, int cls, int begin) {
    if (!c->timing || !((c->tmask >> cls) & 1)) return;
    if (c->open.size() < NTIMER) c->open.assign(NTIMER, nullptr);
    if (begin) {
        hipEvent_t a = takeEvent(c);
        (void)hipEventRecord(a, c->st);
        c->open[cls] = a;
    } else if (c->open[cls]) {
        hipEvent_t b = takeEvent(c);
        (void)hipEventRecord(b, c->st);
        c->events.push_back({cls, {c->open[cls], b}});
        c->open[cls] = nullptr;
    }
}

struct Timed {
    icsf_case* c;
    int cls;
    Timed(icsf_case* cc, int k) : c(cc), cls(k) { markClass(c, cls, 1); }
    ~Timed() { markClass(c, cls, 0); }
};

void resolveEvents(icsf_case* c) {
    if (c->events.empty()) return;
    (void)hipStreamSynchronize(c->st);
    for (auto& e : c->events) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e.second.first, e.second.second);
        c->tms[e.first] += ms;
        c->tcnt[e.first] += 1;
        c->pool.push_back(e.second.first);
        c->pool.push_back(e.second.second);
    }
    c->events.clear();
}

// ---- device kernels local to the driver --------------------------------
__global__ void k_ser(int nC, double* co, const double* dres, double coIncr, double coDecr, double coMin,
                      double coMax, const int* halt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC || *halt) return;
    const double* ini = dres;      // latest initial residuals
    const double* prv = dres + 5;  // previous
    const double nI = sqrt(ini[0] * ini[0] + ini[1] * ini[1] + (ini[2] * ini[2] + ini[3] * ini[3] + ini[4] * ini[4]));
    const double nP = sqrt(prv[0] * prv[0] + prv[1] * prv[1] + (prv[2] * prv[2] + prv[3] * prv[3] + prv[4] * prv[4]));
    double ratio = nP / nI;
    ratio = fmax(fmin(ratio, coIncr), coDecr);
    double v = co[c] * ratio;
    v = fmax(fmin(v, coMax), coMin);
    co[c] = v;
}

__global__ void k_copy5(const double* a, double* b) {
    if (threadIdx.x < 5) b[threadIdx.x] = a[threadIdx.x];
}
// the same inside a pseudo-step (nothing once the loop has ended)
__global__ void k_copy5_step(const double* a, double* b, const int* halt) {
    if (threadIdx.x < 5 && !*halt) b[threadIdx.x] = a[threadIdx.x];
}

// pseudotimeControl::loop and criteriaSatisfied (pseudotimeControl.C:72-101,
// 159-241; steady state) at the start of a pseudo-step, on the device so
// that replayed steps evaluate it without the host: ctl[CTL_CORR] is the
// outer-iteration counter corr_, dres the initial residuals of the
// previous step's solve (residuals_, outerLoop.H:97), init the residuals
// stored at corr_ == 2.  A step started after convergence ends the loop
// (CTL_FIN, the halt flag of the state kernels); the step in which the
// criteria first hold is the final iteration and runs.  maxInit takes the
// largest of the five initial residuals; the relative check divides by
// init + ROOTVSMALL (residualsIO operator/).  Disabled (CTL_EN 0) outside
// icsf_pseudo_time_loop.
__global__ void k_pseudo_ctl(int* ctl, double* init, const double* dres, double tol, double tolRel, int nCorr,
                             int nCorrMin) {
    if (threadIdx.x != 0 || !ctl[CTL_EN] || ctl[CTL_FIN]) return;
    const int corr = ++ctl[CTL_CORR];
    if (ctl[CTL_CONV]) {
        ctl[CTL_FIN] = 1;
        return;
    }
    bool crit = false;
    if (ctl[CTL_FIRST] || corr == 1 || corr == nCorr) {
        ctl[CTL_FIRST] = 0;
    } else {
        const bool store = corr == 2;
        if (store)
            for (int i = 0; i < 5; ++i) init[i] = dres[i];
        double mx = -1.0e300;
        for (int i = 0; i < 5; ++i) mx = fmax(mx, dres[i]);
        const bool absC = mx < tol;
        bool relC = false;
        if (!store) {
            double mr = -1.0e300;
            for (int i = 0; i < 5; ++i) mr = fmax(mr, dres[i] / (init[i] + 1.0e-150));
            relC = mr < tolRel;
        }
        crit = corr >= nCorrMin && (absC || relC);
    }
    if (crit) ctl[CTL_CONV] = 1;
    ctl[CTL_RUN]++;
}

// boundLocalTimeStep.H:174-271 evaluated before the state update (as in
// outerLoop.H:99), so the state compared is the pre-solve state.
__global__ void k_bound_flag(int nC, double lb, const double* W, double* bad) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    const double rho = W[c], rE = W[(size_t)nC + c];
    const double ux = W[2 * (size_t)nC + c] / rho, uy = W[3 * (size_t)nC + c] / rho, uz = W[4 * (size_t)nC + c] / rho;
    const double rhoMin = lb * rho;
    const double eMin = lb * (rE / rho - 0.5 * (ux * ux + uy * uy + uz * uz));
    const double eT = rE / rho - 0.5 * (ux * ux + uy * uy + uz * uz);
    // bit 1: the density test, bit 2: the energy tests
    bad[c] = (rho < rhoMin ? 1.0 : 0.0) + ((eT < eMin || eT < 1.0e-15) ? 2.0 : 0.0);
}

// The harmonic-balance instance J's phase-lagged far test: rho_ is lagged
// over the instances as they stand (those before J already updated,
// outerLoop.H:221-229), rhoMin, eTemp and eMin are not (their far cell's
// bits).
struct BoundLag {
    const HBInst* I = nullptr;
    int nO = 0, J = 0;
    const int* side = nullptr;
    const double* Dl = nullptr;
    double lb = 0;
};

// boundLocalTimeStep.H:28-97: internal faces, then cyclic faces (the far
// cell through patchNeighbourField).
// Over the cell's internal and cyclic faces (the factor is a minimum of
// 0.5 and 0.75, independent of the face order): the neighbours come from
// the cell's matrix slots and cyclic-face list, read contiguously.
__global__ void k_bound_apply(MeshView m, const double* bad, double* co, const int* halt, BoundLag bl = {}) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.nC || (halt && *halt)) return;
    double f = 1.0;
    const bool badC = bad[c] != 0.0;
    const int len = m.rowLen[c];
    const size_t rs = (size_t)m.slicePtr[c >> 6] * 64 + (c & 63);
    for (int k = 0; k < len; ++k) {
        const int other = m.col[rs + 64 * (size_t)k];
        if (badC) f = fmin(0.5, f);
        if (bad[other] != 0.0) f = fmin(0.75, f);
    }
    if (m.nCyc > 0)
        for (int q = m.cycRowOff[c]; q < m.cycRowOff[c + 1]; ++q) {
            const int kc = m.cycRowK[q];
            const int other = m.cycCell[kc];
            if (badC) f = fmin(0.5, f);
            bool farBad = bad[other] != 0.0;
            if (bl.side && bl.side[kc] >= 0) {
                const double* D = bl.Dl + ((size_t)bl.side[kc] * bl.nO + bl.J) * bl.nO;
                double r = 0.0;
                for (int K = 0; K < bl.nO; ++K) r = r + D[K] * bl.I[K].s.W[other];
                farBad = r < bl.lb * bl.I[bl.J].s.W[other] || bad[other] >= 2.0;
            }
            if (farBad) f = fmin(0.75, f);
        }
    co[c] *= f;
}

__global__ void k_fill(int n, double v, double* x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

int nblk(size_t n, int b) { return (int)((n + b - 1) / b); }

// Read the device error flags of the given flag words (after the work on
// `st`), clear them, and turn them into a return code: a non-finite state
// first, then a diagonal whose LU-SGS coefficient fell below VSMALL
// (lusgs.C:119-123).
int takeDeviceErrors(hipStream_t st, const std::vector<int*>& flags) {
    int any = 0;
    for (int* f : flags) {
        if (!f) continue;
        int e = 0;
        HIPCHK(hipMemcpyAsync(&e, f, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (e) HIPCHK(hipMemset(f, 0, sizeof(int)));
        any |= e;
    }
    if (any & ERR_FLAG_NONFINITE) return fail(ICSF_ERR_NONFINITE, "non-finite conserved state after the update");
    if (any & ERR_FLAG_DIAG) return fail(ICSF_ERR_ZERO_DIAG, "All diagonals of coupledMatrix are zero.");
    return ICSF_OK;
}

// ---- settings -----------------------------------------------------------
int readSettings(icsf_case* c, const char* thermo, const char* schemes, const char* solution) {
    Dict td, sd, fd;
    std::string err;
    if (!(err = parseDict(thermo ? thermo : "", td)).empty()) return fail(ICSF_ERR_ARG, "thermophysicalProperties: " + err);
    if (!(err = parseDict(schemes ? schemes : "", sd)).empty()) return fail(ICSF_ERR_ARG, "fvSchemes: " + err);
    if (!(err = parseDict(solution ? solution : "", fd)).empty()) return fail(ICSF_ERR_ARG, "fvSolution: " + err);
    Phys& ph = c->ph;
    // thermophysical model: hePsiThermo / hConst / perfectGas, e = Cv T
    double R = td.scalar("R", -1);
    const double molW = td.scalar("mixture/specie/molWeight", 28.966);
    if (R <= 0) R = 8314.47 / molW;
    double gamma = td.scalar("gamma", -1);
    const double Cp = td.scalar("mixture/thermodynamics/Cp", 1005.0);
    if (gamma <= 0) gamma = Cp / (Cp - R);
    ph.R = R;
    ph.gamma = gamma;
    ph.Cv = R / (gamma - 1.0);
    ph.mu = td.scalar("mixture/transport/mu", 0.0);
    const double Pr = td.scalar("mixture/transport/Pr", 1.0);
    ph.alpha = gamma * ph.mu / Pr;  // alphahe = Cp/Cv * kappa/Cp
    ph.viscous = ph.mu > 0 ? 1 : 0;
    // fvSchemes
    const std::string fs = sd.word("convectiveFluxScheme/fluxScheme", "");
    if (fs == "AUSMPlusUp") ph.flux = FLUX_AUSM;
    else if (fs == "ROE") ph.flux = FLUX_ROE;
    else if (fs == "HLLC") ph.flux = FLUX_HLLC;
    else
        return fail(ICSF_ERR_KEYWORD, "Unknown convectiveFluxScheme type '" + fs +
                                          "'; valid types are AUSMPlusUp HLLC ROE");
    ph.lowMach = sd.sw("convectiveFluxScheme/lowMachAusm", true) ? 1 : 0;
    ph.entropyFix = sd.scalar("convectiveFluxScheme/entropyFixCoeff", 0.05);
    const auto* ddt = sd.find("ddtSchemes/default");
    if (!ddt || ddt->empty() || ddt->back() != "steadyState" || (*ddt)[0] != "dualTime")
        return fail(ICSF_ERR_KEYWORD, "ddtSchemes default must be 'dualTime rPseudoDeltaT steadyState'");
    // reconstruction: all three reconstruct(...) entries must agree
    std::string recon;
    for (const char* k : {"reconstruct(rho)", "reconstruct(U)", "reconstruct(T)"}) {
        const auto* v = sd.find(std::string("interpolationSchemes/") + k);
        std::string s;
        if (v) for (size_t i = 0; i < v->size(); ++i) s += (i ? " " : "") + (*v)[i];
        if (recon.empty()) recon = s;
        else if (s != recon) return fail(ICSF_ERR_KEYWORD, "reconstruct(rho|U|T) schemes must be identical");
    }
    ph.limiter = 0;
    if (recon == "upwind") ph.recon = 0;
    else if (recon.rfind("linearUpwind", 0) == 0) {
        ph.recon = 1;
        std::string gname = recon.size() > 13 ? recon.substr(13) : "default";
        const auto* gs = sd.find("gradSchemes/" + gname);
        if (!gs) gs = sd.find("gradSchemes/default");
        std::string g;
        if (gs) for (size_t i = 0; i < gs->size(); ++i) g += (i ? " " : "") + (*gs)[i];
        if (g == "Gauss linear") ph.limiter = 0;
        else if (g == "cellLimited<Venkatakrishnan> Gauss linear 1") ph.limiter = 1;
        else if (g == "cellLimited<minmod> Gauss linear 1" || g == "cellLimited Gauss linear 1") ph.limiter = 2;
        else return fail(ICSF_ERR_KEYWORD, "unsupported gradScheme '" + g + "'");
    } else return fail(ICSF_ERR_KEYWORD, "unsupported reconstruction scheme '" + recon + "'");
    // fvSolution
    const std::string sv = fd.word("flowSolver/solver", "GMRES");
    if (sv == "GMRES") c->solver = 0;
    else if (sv == "FGMRES") c->solver = 1;
    else return fail(ICSF_ERR_KEYWORD, "Unknown coupled solver '" + sv + "'; valid solvers are GMRES FGMRES");
    const std::string base = "flowSolver/" + sv + "/";
    const std::string pc = fd.word(base + "preconditioner", "LUSGS");
    if (pc == "LUSGS") c->precond = 0;
    else if (pc == "Jacobi") c->precond = 1;
    else if (pc == "none") c->precond = 2;
    else return fail(ICSF_ERR_KEYWORD, "Unknown preconditioner '" + pc + "'; valid are LUSGS Jacobi none");
    c->nDirs = (int)fd.scalar(base + "nDirections", 30);
    c->maxIter = (int)fd.scalar(base + "maxIter", 1000);
    c->minIter = (int)fd.scalar(base + "minIter", 0);
    c->tol = fd.scalar(base + "tolerance", 1e-12);
    c->relTol = fd.scalar(base + "relTol", 1e-2);
    c->innerDirs = (int)fd.scalar(base + "innerDirections", 30);
    const std::string prec = fd.word(base + "precision", "double");
    if (prec == "double") c->mixed = 0;
    else if (prec == "mixed") c->mixed = 1;
    else return fail(ICSF_ERR_KEYWORD, "precision must be double or mixed");
    const std::string orth = fd.word(base + "orthogonalisation", "MGS");
    if (orth == "MGS") c->orth = 0;
    else if (orth == "CGS") c->orth = 1;
    else return fail(ICSF_ERR_KEYWORD, "orthogonalisation must be MGS or CGS");
    const std::string gcw = fd.word(base + "graphCapture", "true");
    if (gcw == "true" || gcw == "on" || gcw == "yes") c->useGraph = true;
    else if (gcw == "false" || gcw == "off" || gcw == "no") c->useGraph = false;
    else return fail(ICSF_ERR_KEYWORD, "graphCapture must be true or false");
    const std::string jsw = fd.word(base + "jacobianStorage", "generators");
    if (jsw == "generators") c->blkGen = true;
    else if (jsw == "blocks") c->blkGen = false;
    else return fail(ICSF_ERR_KEYWORD, "jacobianStorage must be generators or blocks");
    const std::string atw = fd.word(base + "autotune", "true");
    if (atw == "true" || atw == "on" || atw == "yes") c->autotune = true;
    else if (atw == "false" || atw == "off" || atw == "no") c->autotune = false;
    else return fail(ICSF_ERR_KEYWORD, "autotune must be true or false");
    const std::string fsw = fd.word(base + "fusedSweeps", "true");
    if (fsw == "true" || fsw == "on" || fsw == "yes") c->fuse = true;
    else if (fsw == "false" || fsw == "off" || fsw == "no") c->fuse = false;
    else return fail(ICSF_ERR_KEYWORD, "fusedSweeps must be true or false");
    const std::string fcw = fd.word(base + "fusedCellSums", "true");
    if (fcw == "true" || fcw == "on" || fcw == "yes") c->fuseCells = true;
    else if (fcw == "false" || fcw == "off" || fcw == "no") c->fuseCells = false;
    else return fail(ICSF_ERR_KEYWORD, "fusedCellSums must be true or false");
    if (c->nDirs < 1 || c->nDirs > KMAX) return fail(ICSF_ERR_ARG, "nDirections must be in [1, 64]");
    // innerDirections 0: the mixed FGMRES is right-preconditioned by the
    // FP32 preconditioner alone (no inner Krylov iteration)
    if (c->innerDirs < 0 || c->innerDirs > KMAX) return fail(ICSF_ERR_ARG, "innerDirections must be in [0, 64]");
    if (c->mixed && c->solver != 1)
        return fail(ICSF_ERR_KEYWORD, "precision mixed requires solver FGMRES (FP64 outer, FP32 inner)");
    c->co = fd.scalar("pseudoTime/pseudoCoNum", 1.0);
    c->coMin = fd.scalar("pseudoTime/pseudoCoNumMin", 0.1);
    c->coMax = fd.scalar("pseudoTime/pseudoCoNumMax", 25.0);
    c->coIncr = fd.scalar("pseudoTime/pseudoCoNumMaxIncreaseFactor", 1.25);
    c->coDecr = fd.scalar("pseudoTime/pseudoCoNumMinDecreaseFactor", 0.1);
    c->localTS = fd.sw("pseudoTime/localTimestepping", true) ? 1 : 0;
    if (!c->localTS) return fail(ICSF_ERR_KEYWORD, "only localTimestepping true is supported");
    c->bounding = fd.sw("pseudoTime/localTimesteppingBounding", true) ? 1 : 0;
    double lb = fd.scalar("pseudoTime/localTimesteppingLowerBound", 0.95);
    lb = lb > 0 ? lb : 0.0;
    lb = lb < 0.99 ? lb : 0.99;
    c->lowerBound = lb;
    c->pseudoTol = fd.scalar("pseudoTime/pseudoTol", std::nan(""));
    c->pseudoTolRel = fd.scalar("pseudoTime/pseudoTolRel", std::nan(""));
    c->nPseudoCorr = (int)fd.scalar("pseudoTime/nPseudoCorr", 20);
    c->nPseudoCorrMin = (int)fd.scalar("pseudoTime/nPseudoCorrMin", 1);
    ph.rhoMin = fd.scalar("pseudoTime/rhoMin", -kGREAT);
    ph.TMin = fd.scalar("pseudoTime/TMin", kSMALL);
    ph.TMax = fd.scalar("pseudoTime/TMax", kGREAT);
    c->blkFp32 = c->mixed;
    return ICSF_OK;
}

int uploadBC(icsf_case* c) {
    if (!c->bcDirty) return ICSF_OK;
    const int nB = c->mesh->h.nFaces - c->mesh->h.nInternal;
    const size_t n = (size_t)std::max(nB, 1);
    HIPCHK(hipMemcpy(c->d_bPT, c->bPT.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_bUT, c->bUT.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_bTT, c->bTT.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_bPv, c->bPv.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_bUv, c->bUv.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_bTv, c->bTv.data(), n * sizeof(double), hipMemcpyHostToDevice));
    c->bcDirty = false;
    return ICSF_OK;
}

// cellSums false (pseudo-step): the residual's cell sums are left to
// jacobianStage(c, true), which forms them in the Jacobian's face walk.
int residualStage(icsf_case* c, bool cellSums = true) {
    RC(uploadBC(c));
    {
        Timed t(c, TK_RESIDUAL);
        launchResidual(c->mv, c->ph, c->sv, c->st, cellSums);
    }
    // without the cell sums b and rDT are complete only after
    // jacobianStage(c, true)
    c->haveResidual = cellSums;
    return ICSF_OK;
}

int jacobianStage(icsf_case* c, bool cellSums = false) {
    {
        Timed t(c, TK_JACOBIAN);
        if (cellSums) launchCellSumsJacobian(c->mv, c->ph, c->sv, c->blk, c->dg, c->blkFp32, c->rD, c->derr, c->st);
        else launchJacobian(c->mv, c->ph, c->sv, c->blk, c->dg, c->blkFp32, c->rD, c->derr, c->st);
    }
    if (cellSums) c->haveResidual = true;
    if (c->precond == 1) launchJacobiSetup(c->mv.nC, c->blkFp32, c->dg, c->dinv, c->st);
    c->assembled = true;
    return ICSF_OK;
}

void applyPrecond(icsf_case* c, int vecFp32, void* v) {
    Timed t(c, TK_PRECOND);
    if (c->precond == 0) launchLusgs(c->mv, c->mesh->lev, c->blkFp32, vecFp32, c->blk, c->rD, v, c->st);
    else if (c->precond == 1) launchJacobiApply(c->mv.nC, c->blkFp32, vecFp32, c->dinv, v, c->st);
}

// Left-preconditioned GMRES(m) cycles of gmres::solveDelta on the FP64
// vectors (reference algorithm); returns cycles run.
int gmresSolve(icsf_case* c, const double* W, const double* b, int& nIter) {
    const MeshView& m = c->mv;
    const size_t vs = (size_t)5 * m.nC;
    {
        Timed t(c, 5);
        gmresPrepare(m, c->wk, c->g, 0, W, b, c->t1, c->t2, c->blkFp32, c->blk, c->dg, c->st);
        launchZero(vs, 0, c->dW, c->st);
        launchCopy(vs, 0, b, c->r, c->st);
    }
    nIter = 0;
    for (;;) {
        GmresArgs a{&c->mv, &c->mesh->lev, &c->wk, c->g, c->nDirs, c->precond, c->orth, c->blkFp32,
                    c->blk, c->dg, c->rD, c->dinv, c->r, c->t2, c->V};
        a.mark = [c](int k, int b) { markClass(c, k, b); };
        a.fuse = c->fuse;
        gmresCycle(a, 0, c->st);
        {
            Timed t(c, 5);
            gmresUpdateAndResidual(m, c->wk, c->g, 0, c->blkFp32, c->blk, c->dg, c->nDirs, c->V, c->dW, b, c->r,
                                   c->st);
        }
        ++nIter;
        if (nIter < c->minIter) continue;
        if (nIter >= c->maxIter) break;
        double res[10];
        gmresReadResiduals(c->g, res, c->st);
        double mx = -1e300, rel = -1e300;
        for (int k = 5; k < 10; ++k) mx = std::max(mx, res[k]);
        for (int k = 0; k < 5; ++k) {
            if (k >= 2 && (k - 2) == m.emptyDir) continue;
            rel = std::max(rel, res[5 + k] / (res[k] + kROOTVSMALL));
        }
        if (mx < c->tol || rel < c->relTol) break;
    }
    if (m.emptyDir >= 0) launchZeroComponent(m.nC, 2 + m.emptyDir, 0, c->dW, c->st);
    return ICSF_OK;
}

// Flexible GMRES: FP64 outer Arnoldi on A with right preconditioning by an
// FP32 inner solve (left-preconditioned GMRES(innerDirections) with FP32
// LU-SGS on the FP32 Jacobian), or by the preconditioner alone when
// innerDirections == 0.  The true residual b - A dW is evaluated in FP64.
int fgmresSolve(icsf_case* c, const double* W, const double* b, int& nIter);

int solveStage(icsf_case* c, const double* W, const double* b, int& nIter) {
    if (c->solver == 1) return fgmresSolve(c, W, b, nIter);
    return gmresSolve(c, W, b, nIter);
}

// Autotuner: on the first pseudo-step, after the residual and Jacobian
// (every buffer then holds the case's real data), each launch parameter is
// timed with HIP events on the case stream over its candidates and the
// fastest kept: workgroup sizes 64/128/256 of the face-terms, gradient,
// cell-sums + Jacobian-row and Arnoldi-operator kernels, and 8/16/32
// basis vectors per CGS projection pass.  Only scratch vectors are written
// (or state is recomputed identically); no candidate changes a result bit.
int autotune(icsf_case* c) {
    {
        const hipError_t pend = hipGetLastError();
        if (pend != hipSuccess)
            return fail(ICSF_ERR_HIP, std::string("autotune: HIP error pending on entry: ") + hipGetErrorString(pend));
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    hipError_t terr = hipSuccess;
    const char* twhere = "";
    auto note = [&](hipError_t e, const char* where) {
        if (e != hipSuccess && terr == hipSuccess) { terr = e; twhere = where; }
    };
    auto timeIt = [&](const std::function<void()>& fn) -> float {
        fn();
        note(hipGetLastError(), "launch");
        note(hipEventRecord(e0, c->st), "hipEventRecord");
        for (int r = 0; r < 3; ++r) fn();
        note(hipGetLastError(), "launch");
        note(hipEventRecord(e1, c->st), "hipEventRecord");
        note(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0;
        note(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        return ms;
    };
    auto pick = [&](int& param, const std::vector<int>& cand, const std::function<void()>& fn) {
        int best = param;
        float bt = 1e30f;
        for (int v : cand) {
            param = v;
            const float t = timeIt(fn);
            if (t < bt) { bt = t; best = v; }
        }
        param = best;
    };
    // workgroup size and register budget of one kernel class, jointly
    // (budgets: 0 the compiler's choice, 1 at most 128 VGPRs, 2 at most 168)
    auto pick2 = [&](int& bsP, int& occP, const std::function<void()>& fn, int nOcc = 3) {
        int bb = bsP, bo = occP;
        float bt = 1e30f;
        for (int o = 0; o < nOcc; ++o)
            for (int b : {64, 128, 256}) {
                bsP = b;
                occP = o;
                const float t = timeIt(fn);
                if (t < bt) { bt = t; bb = b; bo = o; }
            }
        bsP = bb;
        occP = bo;
    };
    MeshView& m = c->mv;
    const std::vector<int> bs{64, 128, 256};
    // the pseudo-step's launches: residual up to the face terms, then the
    // cell sums with the Jacobian rows (the last candidate leaves b, rDT and
    // the matrix as the step computed them)
    const bool fc = c->fuseCells;
    pick2(m.bsFace, m.occFace, [&] { launchResidual(m, c->ph, c->sv, c->st, !fc); });
    pick2(m.bsCell, m.occCell, [&] { launchResidual(m, c->ph, c->sv, c->st, !fc); });
    pick2(m.bsJac, m.occJac, [&] {
        if (fc) launchCellSumsJacobian(m, c->ph, c->sv, c->blk, c->dg, c->blkFp32, c->rD, c->derr, c->st);
        else launchJacobian(m, c->ph, c->sv, c->blk, c->dg, c->blkFp32, c->rD, c->derr, c->st);
    });
    // the Arnoldi operator in the precision of the inner iteration
    const int v32 = c->mixed ? 1 : 0;
    void* x = v32 ? (void*)c->r32 : (void*)c->t1;
    void* y = v32 ? (void*)c->r32b : (void*)c->t2;
    if (x && y && c->precond == 0) {
        launchZero((size_t)5 * m.nC, v32, x, c->st);
        int occDummy = 0;
        pick2(m.bsOp, m.blkGen ? m.occOp : occDummy, [&] {
            if (!(c->fuse && launchSpmvLusgs(m, c->mesh->lev, c->blkFp32, v32, c->blk, c->dg, c->rD, x, nullptr,
                                             nullptr, y, c->st))) {
                launchSpmv(m, c->blkFp32, v32, c->blk, c->dg, x, nullptr, nullptr, y, c->st);
                launchLusgs(m, c->mesh->lev, c->blkFp32, v32, c->blk, c->rD, y, c->st);
            }
        }, 2);
        const void* V = v32 ? (const void*)c->V32 : (const void*)c->V;
        const int nv = v32 ? c->innerDirs : c->nDirs;
        if (c->orth == 1 && V && nv > 0)
            pick(m.mdCh, std::vector<int>{8, 16, 32},
                 [&] { launchProjectionPasses(m, v32, m.mdCh, x, V, nv, c->wk.part, c->st); });
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (terr != hipSuccess)
        return fail(ICSF_ERR_HIP, std::string("autotune: ") + twhere + ": " + hipGetErrorString(terr));
    c->tuned = true;
    return ICSF_OK;
}

int pseudoStep(icsf_case* c, int& nIterOut) {
    const MeshView& m = c->mv;
    const int* halt = c->sv.halt;
    c->corr++;
    k_pseudo_ctl<<<1, 64, 0, c->st>>>(c->dctl, c->dctlInit, c->dres, c->pseudoTol, c->pseudoTolRel,
                                       c->nPseudoCorr, c->nPseudoCorrMin);
    if (c->haveInit) {
        if (c->corr != 1 && c->havePrev)
            k_ser<<<nblk(m.nC, 256), 256, 0, c->st>>>(m.nC, c->sv.pseudoCo, c->dres, c->coIncr, c->coDecr, c->coMin,
                                                      c->coMax, halt);
        k_copy5_step<<<1, 64, 0, c->st>>>(c->dres, c->dres + 5, halt);
        c->havePrev = true;
    }
    RC(residualStage(c, !c->fuseCells));
    RC(jacobianStage(c, c->fuseCells));
    if (c->autotune && !c->tuned) RC(autotune(c));
    int nIter = 0;
    RC(solveStage(c, c->sv.W, c->sv.b, nIter));
    nIterOut = nIter;
    k_copy5_step<<<1, 64, 0, c->st>>>(gmresResidualsDevice(c->g), c->dres, halt);
    c->haveInit = true;
    if (c->localTS && c->bounding) {
        k_bound_flag<<<nblk(m.nC, 256), 256, 0, c->st>>>(m.nC, c->lowerBound, c->sv.W, c->fac);
        k_bound_apply<<<nblk(m.nC, 256), 256, 0, c->st>>>(c->mv, c->fac, c->sv.pseudoCo, halt);
    }
    {
        Timed t(c, 5);
        launchUpdate(m, c->ph, c->sv, c->dW, 1, nullptr, nullptr, nullptr, c->st);
    }
    // the residual and the matrix belong to the state before the update
    c->haveResidual = c->assembled = false;
    return ICSF_OK;
}

// From the third step on (SER active, initial/previous residuals held on
// the device) a pseudo-step launches the same kernels with the same
// arguments every time and, with one GMRES cycle (maxIter 1), never waits
// on the host: it is captured once as a HIP graph and replayed.  Not with
// event timing (its events are created per launch) or pending boundary
// uploads.
bool graphable(const icsf_case* c) {
    return c->useGraph && !c->timing && c->maxIter == 1 && c->minIter <= 1 && c->haveInit && c->havePrev &&
           c->corr >= 2 && !c->bcDirty;
}

int graphStep(icsf_case* c) {
    if (!c->gexec) {
        HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        int it = 0;
        const int rc = pseudoStep(c, it);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->st, &g);
        c->corr--;  // the capture enqueued nothing; the replay below is the step
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (rc) return rc;
            HIPCHK(e);
        }
        const hipError_t ie = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(g);
            c->gexec = nullptr;
            HIPCHK(ie);
        }
        c->graph = g;
    }
    c->corr++;
    HIPCHK(hipGraphLaunch(c->gexec, c->st));
    return ICSF_OK;
}

}  // namespace

// FGMRES (outer FP64, inner FP32) -----------------------------------------
namespace {
int fgmresSolve(icsf_case* c, const double* W, const double* b, int& nIter) {
    const MeshView& m = c->mv;
    const size_t vs = (size_t)5 * m.nC;
    {
        Timed t(c, 5);
        gmresPrepare(m, c->wk, c->g, 0, W, b, c->t1, c->t2, c->blkFp32, c->blk, c->dg, c->st);
        launchZero(vs, 0, c->dW, c->st);
        launchCopy(vs, 0, b, c->r, c->st);
    }
    // z = M^-1 v
    auto precond = [c, vs](const double* v, double* z) {
        if (c->mixed) {
            launchConvert(vs, 0, v, c->r32, c->st);
            if (c->innerDirs > 0) {
                GmresArgs a{&c->mv, &c->mesh->lev, &c->wk, c->g2, c->innerDirs, c->precond, c->orth, c->blkFp32,
                            c->blk, c->dg, c->rD, c->dinv, c->r32, c->r32b, c->V32};
                a.mark = [c](int k, int b) { markClass(c, k, b); };
                a.fuse = c->fuse;
                gmresCycle(a, 1, c->st);
                Timed t(c, 5);
                launchCombine(c->mv.nC, 1, c->V32, c->innerDirs, c->g2, c->r32, c->st);
            } else {
                applyPrecond(c, 1, c->r32);
            }
            launchConvert(vs, 1, c->r32, z, c->st);
        } else {
            launchCopy(vs, 0, v, z, c->st);
            applyPrecond(c, 0, z);
        }
    };
    nIter = 0;
    for (;;) {
        FgmresArgs a{&c->mv, &c->wk, c->g, c->nDirs, c->orth, c->blkFp32, c->blk, c->dg, c->r, c->V, c->Z, precond,
                     [c](int k, int b) { markClass(c, k, b); }};
        fgmresCycle(a, c->st);
        {
            Timed t(c, 5);
            fgmresUpdateAndResidual(m, c->wk, c->g, c->blkFp32, c->blk, c->dg, c->nDirs, c->Z, c->dW, b, c->r,
                                    c->st);
        }
        ++nIter;
        if (nIter < c->minIter) continue;
        if (nIter >= c->maxIter) break;
        double res[10];
        gmresReadResiduals(c->g, res, c->st);
        double mx = -1e300, rel = -1e300;
        for (int k = 5; k < 10; ++k) mx = std::max(mx, res[k]);
        for (int k = 0; k < 5; ++k) {
            if (k >= 2 && (k - 2) == m.emptyDir) continue;
            rel = std::max(rel, res[5 + k] / (res[k] + kROOTVSMALL));
        }
        if (mx < c->tol || rel < c->relTol) break;
    }
    if (m.emptyDir >= 0) launchZeroComponent(m.nC, 2 + m.emptyDir, 0, c->dW, c->st);
    return ICSF_OK;
}
}  // namespace

// ================================================================ C ABI ===
namespace {

int adjointSupported(const icsf_case* c) {
    if (c->mrf) return fail(ICSF_ERR_ARG, "residual VJP: cases with an MRF zone are not supported");
    for (size_t b = 0; b < c->bPT.size(); ++b)
        if (c->bPT[b] == BC_TOTAL_P || c->bUT[b] == BC_PIOV || c->bTT[b] == BC_TOTAL_T)
            return fail(ICSF_ERR_ARG, "residual VJP: totalPressure, pressureInletOutletVelocity and totalTemperature "
                                      "conditions are not supported");
    if (c->mv.nCyc > 0) return fail(ICSF_ERR_ARG, "residual VJP: meshes with cyclic patches are not supported");
    return ICSF_OK;
}

int adjointPrepare(icsf_case* c) {
    RC(adjointSupported(c));
    if (c->adjReady) return ICSF_OK;
    const size_t nC = c->mv.nC, nF = c->mv.nF, nB = std::max(c->mv.nB, 1);
    AdjView& a = c->adj;
    if (!a.W0) RC(c->buf.alloc(a.W0, 5 * nC));
    RC(c->buf.alloc(a.gun, nC * NFLD * 3));
    RC(c->buf.alloc(a.lim, nC * NFLD));
    RC(c->buf.alloc(a.fLR, nF * 2 * NFLD));
    RC(c->buf.alloc(a.G, nC * NFLD * 3));
    RC(c->buf.alloc(a.qc, nC * NFLD));
    RC(c->buf.alloc(a.tgt, nC * NFLD));
    RC(c->buf.alloc(a.tval, nC * NFLD));
    RC(c->buf.alloc(a.fst, nC * NFLD));
    RC(c->buf.alloc(a.fext, nC * NFLD));
    RC(c->buf.alloc(a.qf, nF * NFLD));
    RC(c->buf.alloc(a.bprim, nB * 5));
    RC(c->buf.alloc(c->adjGS, 3 * nF));
    RC(c->buf.alloc(c->adjGC, 3 * nF));
    if (c->ph.viscous) {
        RC(c->buf.alloc(a.vis, nF * NVIS));
        RC(c->buf.alloc(a.qe, nF));
        RC(c->buf.alloc(a.Ge, nC * 3));
    }
    c->adjReady = true;
    return ICSF_OK;
}

// State from W (device, or the current W when NULL) through updateFields
// with a zero increment; W is kept in adj.W0.
int stateFromW(icsf_case* c, const double* W) {
    const size_t n5 = 5 * (size_t)c->mv.nC;
    if (!c->adj.W0) RC(c->buf.alloc(c->adj.W0, n5));
    if (W) HIPCHK(hipMemcpyAsync(c->sv.W, W, n5 * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->adj.W0, c->sv.W, n5 * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    launchUpdate(c->mv, c->ph, c->sv, nullptr, 1, nullptr, nullptr, nullptr, c->st);
    c->haveFields = true;
    c->haveResidual = c->assembled = false;
    return ICSF_OK;
}

int vjpDevice(icsf_case* c, const double* W, const double* v, double* gW, double* gSf, double* gCf, bool geom) {
    if (!W && !c->haveFields) return fail(ICSF_ERR_STATE, "fields not set");
    RC(adjointPrepare(c));
    RC(uploadBC(c));
    RC(stateFromW(c, W));
    launchAdjointForward(c->mv, c->ph, c->sv, c->adj.gun, c->adj.lim, c->st);
    launchResidualVjp(c->mv, c->ph, c->sv, c->adj, v, gW, gSf ? gSf : c->adjGS, gCf ? gCf : c->adjGC, geom, c->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    return takeDeviceErrors(c->st, {c->derr});
}

}  // namespace


// ====================================================== harmonic balance ==
// dbnsFullyImplicitHBFoam (outerLoop.H:1-230): nO instances of one case on
// one mesh, coupled by the spectral source -V sum_K D_JK W_K
// (HBZoneTemplates.C:39-93) and the diagonal blocks V D_JK I
// (HBZone.C:435-518).  Coupled vectors are instance-major [nO][5][nC].
struct icsf_hb {
    icsf_mesh* mesh = nullptr;
    int nO = 0;
    std::vector<icsf_case*> inst;  // owned
    hipStream_t st = nullptr;
    std::vector<double> D;
    double* dD = nullptr;
    HBBlocks hbb{};
    double *Wall = nullptr, *ball = nullptr, *r = nullptr, *r2 = nullptr, *V = nullptr, *dW = nullptr;
    double *t1 = nullptr, *t2 = nullptr, *rD = nullptr, *dres = nullptr;
    // mixed precision: FP64 outer FGMRES (Z), FP32 inner GMRES over all
    // instances (r32, r32b, V32, g2)
    int solver = 0, mixed = 0, innerDirs = 30;
    double *Z = nullptr;
    float *r32 = nullptr, *r32b = nullptr, *V32 = nullptr;
    GmresScalars* g2 = nullptr;
    int* derr = nullptr;
    HBInst* dInst = nullptr;  // per-instance views, blockIdx.y of the batched launches
    GmresScalars* g = nullptr;
    std::vector<GmresScalars*> gJ;
    Work wk;
    int nDirs = 30, maxIter = 1, minIter = 0, orth = 0;
    bool fuse = true;
    double tol = 1e-12, relTol = 1e-2;
    int corr = 0;
    bool haveInit = false, havePrev = false, assembled = false;
    bool useGraph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    DevBuf buf;
    // phase-lagged cyclic patch sides: the side of every cyclic face (-1
    // none), the sides' time shifts [nSides][nO][nO] and geometry
    // [nSides][8] (unit rotation axis, rotation centre, cylCoords), the
    // paired faces' centres [nCyc][3] and the lagged far-side records
    // [nO][nCyc][PLW]
    std::vector<int> plSide;
    std::vector<double> plD, plGeo;
    int* dplSide = nullptr;
    double *dplD = nullptr, *dplGeo = nullptr, *dplCf = nullptr, *dplFar = nullptr;
    ~icsf_hb() {
        if (st) (void)hipStreamSynchronize(st);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (icsf_case* c : inst) delete c;
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {

// Instance J = blockIdx.y: out_J = b_J + source_J with
// source_J = ((0 - V D_J0 W_0) - V D_J1 W_1) - ..., and W_J copied into the
// coupled state vector Wall (written only, never read here).
__global__ void k_hb_source(int nC, int nO, const HBInst* __restrict__ I, const double* __restrict__ D,
                            const double* __restrict__ Vol, double* __restrict__ Wall, double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y;
    if (c >= nC) return;
    const size_t vs = (size_t)5 * nC;
    const double* WJ = I[J].s.W;
    const double* bJ = I[J].s.b;
    for (int q = 0; q < 5; ++q) {
        double src = 0.0;
        for (int K = 0; K < nO; ++K) src -= Vol[c] * D[J * nO + K] * I[K].s.W[(size_t)q * nC + c];
        out[J * vs + (size_t)q * nC + c] = bJ[(size_t)q * nC + c] + src;
        Wall[J * vs + (size_t)q * nC + c] = WJ[(size_t)q * nC + c];
    }
}

// Phase-lagged cyclic patches (phaseLagCyclicFvPatchField.C:179-396,
// phaseLagCyclicFvPatchFields.C:41-702): instance J = blockIdx.y, one
// thread per cyclic face of a lagged side: its far-side record, each slot
// sum_K Dl[J][K] x_K(far cell) accumulated from zero in K order, x_K formed
// from instance K's registered rho, p, U, e and psi = 1/(R T):
// E = e + 0.5|U|^2, H = E + p/rho, c = sqrt(gamma/psi); the velocity
// components unrotated, or with cylCoords combined as (r, theta, axis)
// components at the paired face's centre and forwardT applied.  HLLC's
// reconstructed speed keeps the plain far value.
__global__ void k_phase_lag_hb(const HBInst* __restrict__ I, int nO, int nCyc, const int* __restrict__ side,
                               const double* __restrict__ Dl, const double* __restrict__ geo,
                               const double* __restrict__ cfPair, Phys ph, double* __restrict__ far) {
    const int kc = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y;
    if (kc >= nCyc) return;
    const int sd = side[kc];
    if (sd < 0) return;
    const int n = I[J].m.cycCell[kc];
    const double* D = Dl + ((size_t)sd * nO + J) * nO;
    const double* g = geo + 8 * (size_t)sd;
    const bool cyl = g[6] != 0.0;
    const V3 ah{g[0], g[1], g[2]};
    V3 rh{0, 0, 0}, th{0, 0, 0};
    if (cyl) {
        V3 r = ld3(cfPair + 3 * (size_t)kc) - V3{g[3], g[4], g[5]};
        const double ar = dot(ah, r);
        r = r - ar * ah;
        const double mr = sqrt(magSqr(r));
        rh = V3{r.x / mr, r.y / mr, r.z / mr};
        th = cross(ah, rh);
    }
    double aR = 0, aP = 0, aE = 0, aH = 0, aC = 0, aG = 0;
    V3 aU{0, 0, 0}, sc{0, 0, 0};
    for (int K = 0; K < nO; ++K) {
        const StateView& q = I[K].s;
        const double* f = q.fld + (size_t)n * NFLD;
        const V3 u{f[F_UX], f[F_UY], f[F_UZ]};
        const double e = q.e[n];
        const double dk = D[K];
        aR = aR + dk * f[F_RHO];
        aP = aP + dk * f[F_P];
        aU = aU + dk * u;
        if (cyl) sc = sc + dk * V3{dot(u, rh), dot(u, th), dot(u, ah)};
        aE = aE + dk * (e + 0.5 * magSqr(u));
        aH = aH + dk * ((e + 0.5 * magSqr(u)) + f[F_P] / f[F_RHO]);
        aC = aC + dk * sqrt(ph.gamma / (1.0 / (ph.R * q.T[n])));
        aG = aG + dk * ph.gamma;
    }
    const double* Rt = I[J].m.cycR + 9 * (size_t)kc;
    const V3 uv = cyl ? rot3(Rt, (sc.x * rh + sc.y * th) + sc.z * ah) : rot3(Rt, aU);
    const V3 uc = cyl ? uv : aU;
    double* o = far + ((size_t)J * nCyc + kc) * PLW;
    o[F_RHO] = aR;
    o[F_P] = aP;
    o[F_UX] = uc.x;
    o[F_UY] = uc.y;
    o[F_UZ] = uc.z;
    o[F_E] = aE;
    o[F_H] = aH;
    o[F_CS] = ph.flux == FLUX_HLLC ? I[J].s.fld[(size_t)n * NFLD + F_CS] : aC;
    o[PL_UV] = uv.x;
    o[PL_UV + 1] = uv.y;
    o[PL_UV + 2] = uv.z;
    o[PL_C] = aC;
    o[PL_G] = aG;
}

// SER over all instances (outerLoop.H:30-66): dres = [init nO*5][prev nO*5]
__global__ void k_hb_ser(int nC, int nO, double* co, const double* dres, double coIncr, double coDecr, double coMin,
                         double coMax) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    double nI = 0, nP = 0;
    for (int J = 0; J < nO; ++J) {
        const double* a = dres + 5 * J;
        const double* b = dres + 5 * (nO + J);
        nI += a[0] * a[0] + a[1] * a[1] + (a[2] * a[2] + a[3] * a[3] + a[4] * a[4]);
        nP += b[0] * b[0] + b[1] * b[1] + (b[2] * b[2] + b[3] * b[3] + b[4] * b[4]);
    }
    double ratio = sqrt(nP) / sqrt(nI);
    ratio = fmax(fmin(ratio, coIncr), coDecr);
    double v = co[c] * ratio;
    v = fmax(fmin(v, coMax), coMin);
    co[c] = v;
}

__global__ void k_copyn(int n, const double* a, double* b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

void hbLusgs(icsf_hb* h, void* v, int vecFp32 = 0) {
    launchLusgsHB(h->inst[0]->mv, h->mesh->lev, h->hbb, vecFp32, h->rD, v, h->st);
}

int hbUploadBC(icsf_hb* h) {
    for (icsf_case* c : h->inst) RC(uploadBC(c));
    return ICSF_OK;
}

// Residuals of all instances, each kernel launched once over the
// instances (blockIdx.y), then the spectral source.
// cellSums false (pseudo-step): the instances' cell sums and the spectral
// source are left to hbJacobianStage(h, true), which forms the cell sums in
// the Jacobian kernel's face walk and then adds the source.
int hbResidualStage(icsf_hb* h, bool cellSums = true) {
    const MeshView& m = h->inst[0]->mv;
    RC(hbUploadBC(h));
    if (h->dplFar)
        k_phase_lag_hb<<<dim3(nblk(m.nCyc, 64), h->nO), 64, 0, h->st>>>(
            h->dInst, h->nO, m.nCyc, h->dplSide, h->dplD, h->dplGeo, h->dplCf, h->inst[0]->ph, h->dplFar);
    launchResidualHB(h->dInst, h->nO, m, h->inst[0]->ph, h->st, cellSums);
    for (icsf_case* c : h->inst) c->haveResidual = cellSums;
    if (cellSums)
        k_hb_source<<<dim3(nblk(m.nC, 256), h->nO), 256, 0, h->st>>>(m.nC, h->nO, h->dInst, h->dD, m.V, h->Wall,
                                                                      h->ball);
    return ICSF_OK;
}

// All instances' Jacobian rows in one launch, with V D_JJ on their
// diagonals (then the shared LU-SGS diagonal); cellSums: also the
// instances' residual cell sums and the spectral source into the coupled
// right-hand side (the pseudo-step: gradients, face terms, then this).
int hbJacobianStage(icsf_hb* h, bool cellSums = false) {
    const MeshView& m = h->inst[0]->mv;
    const HBCouple hc{h->nO, h->dD, h->Wall, h->ball, h->rD, h->derr};
    launchJacobianHB(h->dInst, h->nO, h->hbb.fp32, m, h->inst[0]->ph, h->st, cellSums, hc);
    if (cellSums)
        for (icsf_case* c : h->inst) c->haveResidual = true;
    for (icsf_case* c : h->inst) c->assembled = true;
    h->assembled = true;
    return ICSF_OK;
}

void hbReadResiduals(icsf_hb* h, std::vector<double>& res) {
    res.assign(10 * (size_t)h->nO, 0.0);
    (void)hipStreamSynchronize(h->st);
    for (int J = 0; J < h->nO; ++J) gmresReadResiduals(h->gJ[J], &res[10 * (size_t)J], h->st);
}

// gmres::solveDelta on the HB system (left LU-SGS, reference algorithm)
int hbSolve(icsf_hb* h, const double* b, int& nIter) {
    const MeshView& m = h->inst[0]->mv;
    const size_t vs = (size_t)5 * m.nC, vA = vs * h->nO;
    for (int J = 0; J < h->nO; ++J) sliceSubAverage(m, h->wk, h->gJ[J], h->Wall + J * vs, h->t1 + J * vs, h->st);
    launchSpmvHB(m, h->hbb, 0, h->t1, nullptr, nullptr, h->t2, h->st);
    for (int J = 0; J < h->nO; ++J) sliceNormFactors(m, h->wk, h->gJ[J], h->t2 + J * vs, b + J * vs, h->st);
    launchZero(vA, 0, h->dW, h->st);
    launchCopy(vA, 0, b, h->r, h->st);
    nIter = 0;
    for (;;) {
        GmresArgs a{&m, &h->mesh->lev, &h->wk, h->g, h->nDirs, 0, h->orth, 0, nullptr, nullptr, h->rD, nullptr,
                    h->r, h->r2, h->V};
        a.nVec = m.nC * h->nO;
        a.op = [h, &m](const void* x, const double* scal, void* vout, void* y) {
            launchSpmvHB(m, h->hbb, 0, x, scal, vout, y, h->st);
        };
        a.pre = [h](void* v) { hbLusgs(h, v); };
        if (h->fuse)
            a.opPre = [h, &m](const void* x, const double* scal, void* vout, void* y) {
                return launchSpmvLusgsHB(m, h->mesh->lev, h->hbb, 0, h->rD, x, scal, vout, y, h->st);
            };
        gmresCycle(a, 0, h->st);
        launchUpdateSolution((size_t)m.nC * h->nO, h->dW, h->V, vA, h->nDirs, h->g, h->st);
        launchSpmvHB(m, h->hbb, 0, h->dW, nullptr, nullptr, h->r, h->st);
        for (int J = 0; J < h->nO; ++J) sliceFinalResidual(m, h->wk, h->gJ[J], b + J * vs, h->r + J * vs, h->st);
        ++nIter;
        if (nIter < h->minIter) continue;
        if (nIter >= h->maxIter) break;
        std::vector<double> res;
        hbReadResiduals(h, res);
        double mx = -1e300, rel = -1e300;
        for (int J = 0; J < h->nO; ++J) {
            const double* q = &res[10 * (size_t)J];
            for (int k = 5; k < 10; ++k) mx = std::max(mx, q[k]);
            for (int k = 0; k < 5; ++k) {
                if (k >= 2 && (k - 2) == m.emptyDir) continue;
                rel = std::max(rel, q[5 + k] / (q[k] + kROOTVSMALL));
            }
        }
        if (mx < h->tol || rel < h->relTol) break;
    }
    if (m.emptyDir >= 0)
        for (int J = 0; J < h->nO; ++J) launchZeroComponent(m.nC, 2 + m.emptyDir, 0, h->dW + J * vs, h->st);
    return ICSF_OK;
}

// Mixed precision (precision mixed, solver FGMRES): outer FGMRES on the FP64
// coupled vectors with the FP32-stored HB operator, each preconditioner
// application an inner left-preconditioned GMRES(innerDirections) in FP32
// over all instances (FP32 HB SpMV, shared-diagonal LU-SGS, fused when the
// numbering allows), as the single-instance mixed form.
int hbSolveMixed(icsf_hb* h, const double* b, int& nIter) {
    const MeshView& m = h->inst[0]->mv;
    const size_t vs = (size_t)5 * m.nC, vA = vs * h->nO;
    const int nVec = m.nC * h->nO;
    for (int J = 0; J < h->nO; ++J) sliceSubAverage(m, h->wk, h->gJ[J], h->Wall + J * vs, h->t1 + J * vs, h->st);
    launchSpmvHB(m, h->hbb, 0, h->t1, nullptr, nullptr, h->t2, h->st);
    for (int J = 0; J < h->nO; ++J) sliceNormFactors(m, h->wk, h->gJ[J], h->t2 + J * vs, b + J * vs, h->st);
    launchZero(vA, 0, h->dW, h->st);
    launchCopy(vA, 0, b, h->r, h->st);
    auto precond = [h, &m, vA, nVec](const double* v, double* z) {
        launchConvert(vA, 0, v, h->r32, h->st);
        if (h->innerDirs == 0) {
            hbLusgs(h, h->r32, 1);
            launchConvert(vA, 1, h->r32, z, h->st);
            return;
        }
        GmresArgs a{&m, &h->mesh->lev, &h->wk, h->g2, h->innerDirs, 0, h->orth, 1, nullptr, nullptr, h->rD,
                    nullptr, h->r32, h->r32b, h->V32};
        a.nVec = nVec;
        a.op = [h, &m](const void* x, const double* scal, void* vout, void* y) {
            launchSpmvHB(m, h->hbb, 1, x, scal, vout, y, h->st);
        };
        a.pre = [h](void* v) { hbLusgs(h, v, 1); };
        if (h->fuse)
            a.opPre = [h, &m](const void* x, const double* scal, void* vout, void* y) {
                return launchSpmvLusgsHB(m, h->mesh->lev, h->hbb, 1, h->rD, x, scal, vout, y, h->st);
            };
        gmresCycle(a, 1, h->st);
        launchCombine(nVec, 1, h->V32, h->innerDirs, h->g2, h->r32, h->st);
        launchConvert(vA, 1, h->r32, z, h->st);
    };
    nIter = 0;
    for (;;) {
        FgmresArgs a{&m, &h->wk, h->g, h->nDirs, h->orth, 1, nullptr, nullptr, h->r, h->V, h->Z, precond, nullptr};
        a.nVec = nVec;
        a.op = [h, &m](const double* z, double* y) { launchSpmvHB(m, h->hbb, 0, z, nullptr, nullptr, y, h->st); };
        fgmresCycle(a, h->st);
        launchUpdateSolution((size_t)nVec, h->dW, h->Z, vA, h->nDirs, h->g, h->st);
        launchSpmvHB(m, h->hbb, 0, h->dW, nullptr, nullptr, h->r, h->st);
        for (int J = 0; J < h->nO; ++J) sliceFinalResidual(m, h->wk, h->gJ[J], b + J * vs, h->r + J * vs, h->st);
        ++nIter;
        if (nIter < h->minIter) continue;
        if (nIter >= h->maxIter) break;
        std::vector<double> res;
        hbReadResiduals(h, res);
        double mx = -1e300, rel = -1e300;
        for (int J = 0; J < h->nO; ++J) {
            const double* q = &res[10 * (size_t)J];
            for (int k = 5; k < 10; ++k) mx = std::max(mx, q[k]);
            for (int k = 0; k < 5; ++k) {
                if (k >= 2 && (k - 2) == m.emptyDir) continue;
                rel = std::max(rel, q[5 + k] / (q[k] + kROOTVSMALL));
            }
        }
        if (mx < h->tol || rel < h->relTol) break;
    }
    if (m.emptyDir >= 0)
        for (int J = 0; J < h->nO; ++J) launchZeroComponent(m.nC, 2 + m.emptyDir, 0, h->dW + J * vs, h->st);
    return ICSF_OK;
}

int hbPseudoStep(icsf_hb* h) {
    const MeshView& m = h->inst[0]->mv;
    const size_t vs = (size_t)5 * m.nC;
    const icsf_case* c0 = h->inst[0];
    h->corr++;
    if (h->haveInit) {
        if (h->corr != 1 && h->havePrev)
            for (int J = 0; J < h->nO; ++J)
                k_hb_ser<<<nblk(m.nC, 256), 256, 0, h->st>>>(m.nC, h->nO, h->inst[J]->sv.pseudoCo, h->dres,
                                                             c0->coIncr, c0->coDecr, c0->coMin, c0->coMax);
        k_copyn<<<1, 5 * HB_MAXO, 0, h->st>>>(5 * h->nO, h->dres, h->dres + 5 * h->nO);
        h->havePrev = true;
    }
    const bool fc = c0->fuseCells;
    RC(hbResidualStage(h, !fc));
    RC(hbJacobianStage(h, fc));
    int nIter = 0;
    if (h->solver == 1) RC(hbSolveMixed(h, h->ball, nIter));
    else RC(hbSolve(h, h->ball, nIter));
    for (int J = 0; J < h->nO; ++J) k_copy5<<<1, 64, 0, h->st>>>(gmresResidualsDevice(h->gJ[J]), h->dres + 5 * J);
    h->haveInit = true;
    for (int J = 0; J < h->nO; ++J) {
        icsf_case* c = h->inst[J];
        if (c->localTS && c->bounding) {
            BoundLag bl;
            if (h->dplFar) bl = BoundLag{h->dInst, h->nO, J, h->dplSide, h->dplD, c->lowerBound};
            k_bound_flag<<<nblk(m.nC, 256), 256, 0, h->st>>>(m.nC, c->lowerBound, c->sv.W, c->fac);
            k_bound_apply<<<nblk(m.nC, 256), 256, 0, h->st>>>(c->mv, c->fac, c->sv.pseudoCo, nullptr, bl);
        }
        // the instance's own view: its boundary conditions and values (the
        // instants' inflow total pressures differ)
        launchUpdate(c->mv, c->ph, c->sv, h->dW + J * vs, 1, nullptr, nullptr, nullptr, h->st);
        // the residual and the matrix belong to the state before the update
        c->haveResidual = c->assembled = false;
    }
    h->assembled = false;
    return ICSF_OK;
}

// As graphStep: from the third step on an HB pseudo-step with one GMRES
// cycle is a fixed launch sequence without host waits; it is captured once
// and replayed.
bool hbGraphable(const icsf_hb* h) {
    if (!h->useGraph || h->maxIter != 1 || h->minIter > 1 || !h->haveInit || !h->havePrev || h->corr < 2)
        return false;
    for (const icsf_case* c : h->inst)
        if (c->bcDirty) return false;
    return true;
}

int hbGraphStep(icsf_hb* h) {
    if (!h->gexec) {
        HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
        const int rc = hbPseudoStep(h);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(h->st, &g);
        h->corr--;  // the capture enqueued nothing; the replay below is the step
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (rc) return rc;
            HIPCHK(e);
        }
        const hipError_t ie = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(g);
            h->gexec = nullptr;
            HIPCHK(ie);
        }
        h->graph = g;
    }
    h->corr++;
    HIPCHK(hipGraphLaunch(h->gexec, h->st));
    return ICSF_OK;
}

int hbCheckInstance(const icsf_hb* h, int J) {
    if (!h) return fail(ICSF_ERR_ARG, "NULL HB case");
    if (J < 0 || J >= h->nO) return fail(ICSF_ERR_ARG, "instant index out of range");
    return ICSF_OK;
}

}  // namespace

namespace icsf {
// error text for the other host translation units (case file I/O)
int reportError(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace icsf

extern "C" {

const char* icsf_last_error(void) { return g_err.c_str(); }
int icsf_version(void) { return 100; }

int icsf_rcm_order(int nCells, int nInternalFaces, const int* owner, const int* neighbour, int* newToOld) {
    if (nCells <= 0 || nInternalFaces < 0 || !newToOld) return fail(ICSF_ERR_ARG, "icsf_rcm_order: bad arguments");
    std::vector<int> o;
    rcmOrder(nCells, nInternalFaces, owner, neighbour, o);
    std::copy(o.begin(), o.end(), newToOld);
    return ICSF_OK;
}

int icsf_rcm_order_coupled(int nCells, int nInternalFaces, const int* owner, const int* neighbour, int nCoupled,
                           const int* coupledOwner, const int* coupledNeighbour, int* newToOld) {
    if (nCells <= 0 || nInternalFaces < 0 || nCoupled < 0 || !newToOld || (nCoupled > 0 && (!coupledOwner ||
                                                                                          !coupledNeighbour)))
        return fail(ICSF_ERR_ARG, "icsf_rcm_order_coupled: bad arguments");
    for (int b = 0; b < nCoupled; ++b)
        if (coupledOwner[b] < 0 || coupledOwner[b] >= nCells || coupledNeighbour[b] < 0 ||
            coupledNeighbour[b] >= nCells)
            return fail(ICSF_ERR_ARG, "icsf_rcm_order_coupled: coupled cell out of range");
    std::vector<int> o;
    rcmOrder(nCells, nInternalFaces, owner, neighbour, o, nCoupled, coupledOwner, coupledNeighbour);
    std::copy(o.begin(), o.end(), newToOld);
    return ICSF_OK;
}

int icsf_colour_cells(int nCells, int nInternalFaces, const int* owner, const int* neighbour, int* colour,
                      int* nColours) {
    if (nCells <= 0 || !colour) return fail(ICSF_ERR_ARG, "icsf_colour_cells: bad arguments");
    std::vector<int> col;
    const int n = greedyColour(nCells, nInternalFaces, owner, neighbour, col);
    std::copy(col.begin(), col.end(), colour);
    if (nColours) *nColours = n;
    return ICSF_OK;
}

int icsf_colour_major_order(int nCells, const int* colour, int nColours, int* newToOld) {
    if (nCells <= 0 || !colour || !newToOld) return fail(ICSF_ERR_ARG, "icsf_colour_major_order: bad arguments");
    std::vector<int> col(colour, colour + nCells), o;
    colourMajorOrder(col, nColours, o);
    if ((int)o.size() != nCells) return fail(ICSF_ERR_ARG, "colours out of range");
    std::copy(o.begin(), o.end(), newToOld);
    return ICSF_OK;
}

int icsf_renumber_mesh(int nCells, int nFaces, int nInternalFaces, const int* faceOffsets, const int* faceVerts,
                       const int* owner, const int* neighbour, const int* newToOld, int* faceVertsOut,
                       int* ownerOut, int* neighbourOut, int* faceNewToOld) {
    if (nCells <= 0 || nFaces <= 0 || nInternalFaces < 0) return fail(ICSF_ERR_ARG, "icsf_renumber_mesh: bad sizes");
    std::vector<char> seen(nCells, 0);
    for (int i = 0; i < nCells; ++i) {
        if (newToOld[i] < 0 || newToOld[i] >= nCells || seen[newToOld[i]])
            return fail(ICSF_ERR_ARG, "newToOld is not a permutation");
        seen[newToOld[i]] = 1;
    }
    std::vector<int> fv, ow, nb, fm;
    renumberMesh(nCells, nFaces, nInternalFaces, faceOffsets, faceVerts, owner, neighbour, newToOld, fv, ow, nb, fm);
    std::copy(fv.begin(), fv.end(), faceVertsOut);
    std::copy(ow.begin(), ow.end(), ownerOut);
    std::copy(nb.begin(), nb.end(), neighbourOut);
    if (faceNewToOld) std::copy(fm.begin(), fm.end(), faceNewToOld);
    return ICSF_OK;
}

int icsf_block_mesh_hex(int nx, int ny, int nz, const int* sidePatch, int nPatches, int* faceVerts, int* owner,
                        int* neighbour, int* patchSize) {
    if (nx < 1 || ny < 1 || nz < 1 || nPatches < 1 || !sidePatch || !faceVerts || !owner || !neighbour || !patchSize)
        return fail(ICSF_ERR_ARG, "icsf_block_mesh_hex: bad arguments");
    for (int s = 0; s < 6; ++s)
        if (sidePatch[s] < 0 || sidePatch[s] >= nPatches) return fail(ICSF_ERR_ARG, "sidePatch out of range");
    hexBlockTopology(nx, ny, nz, sidePatch, nPatches, faceVerts, owner, neighbour, patchSize);
    return ICSF_OK;
}

int icsf_sweep_levels(int nCells, int nInternalFaces, const int* owner, const int* neighbour, int* fwdLevel,
                      int* bwdLevel, int* nFwd, int* nBwd) {
    if (nCells <= 0) return fail(ICSF_ERR_ARG, "icsf_sweep_levels: bad sizes");
    std::vector<int> f, b;
    int nf = 0, nbw = 0;
    sweepLevels(nCells, nInternalFaces, owner, neighbour, f, b, nf, nbw);
    if (fwdLevel) std::copy(f.begin(), f.end(), fwdLevel);
    if (bwdLevel) std::copy(b.begin(), b.end(), bwdLevel);
    if (nFwd) *nFwd = nf;
    if (nBwd) *nBwd = nbw;
    return ICSF_OK;
}

int icsf_mesh_geometry(int nPoints, const double* points, int nFaces, const int* faceOffsets, const int* faceVerts,
                       int nInternalFaces, const int* owner, const int* neighbour, double* C, double* V, double* Sf,
                       double* Cf, double* weights, double* deltaCoeffs, double* nonOrthDeltaCoeffs,
                       double* corrVecs) {
    HostMesh h;
    const int zero = 0, nb = nFaces - nInternalFaces;
    std::string err = buildHostMesh(h, nPoints, points, nFaces, faceOffsets, faceVerts, nInternalFaces, owner,
                                    neighbour, 1, &nInternalFaces, &nb, &zero);
    if (!err.empty()) return fail(ICSF_ERR_ARG, err);
    if (C) std::copy(h.C.begin(), h.C.end(), C);
    if (V) std::copy(h.V.begin(), h.V.end(), V);
    if (Sf) std::copy(h.Sf.begin(), h.Sf.end(), Sf);
    if (Cf) std::copy(h.Cf.begin(), h.Cf.end(), Cf);
    if (weights) std::copy(h.weights.begin(), h.weights.end(), weights);
    if (deltaCoeffs) std::copy(h.deltaCoeffs.begin(), h.deltaCoeffs.end(), deltaCoeffs);
    if (nonOrthDeltaCoeffs) std::copy(h.nonOrthDeltaCoeffs.begin(), h.nonOrthDeltaCoeffs.end(), nonOrthDeltaCoeffs);
    if (corrVecs) std::copy(h.corrVecs.begin(), h.corrVecs.end(), corrVecs);
    return ICSF_OK;
}

int icsf_mesh_geometry_patches(int nPoints, const double* points, int nFaces, const int* faceOffsets,
                               const int* faceVerts, int nInternalFaces, const int* owner, const int* neighbour,
                               int nPatches, const int* patchStart, const int* patchSize, const int* patchKind,
                               const int* patchNeighbour, const double* patchTransform, double* weights,
                               double* deltaCoeffs, double* nonOrthDeltaCoeffs, double* corrVecs,
                               int* cyclicCell) {
    HostMesh h;
    std::string err = buildHostMesh(h, nPoints, points, nFaces, faceOffsets, faceVerts, nInternalFaces, owner,
                                    neighbour, nPatches, patchStart, patchSize, patchKind, patchNeighbour,
                                    patchTransform);
    if (!err.empty()) return fail(ICSF_ERR_ARG, err);
    if (weights) std::copy(h.weights.begin(), h.weights.end(), weights);
    if (deltaCoeffs) std::copy(h.deltaCoeffs.begin(), h.deltaCoeffs.end(), deltaCoeffs);
    if (nonOrthDeltaCoeffs) std::copy(h.nonOrthDeltaCoeffs.begin(), h.nonOrthDeltaCoeffs.end(), nonOrthDeltaCoeffs);
    if (corrVecs) std::copy(h.corrVecs.begin(), h.corrVecs.end(), corrVecs);
    if (cyclicCell)
        for (int b = 0; b < nFaces - nInternalFaces; ++b)
            cyclicCell[b] = h.cycOf[b] < 0 ? -1 : h.cycCell[h.cycOf[b]];
    return ICSF_OK;
}

int icsf_mesh_create(int nPoints, const double* points, int nFaces, const int* faceOffsets, const int* faceVerts,
                     int nInternalFaces, const int* owner, const int* neighbour, int nPatches,
                     const int* patchStart, const int* patchSize, const int* patchKind, const int* patchNeighbour,
                     const double* patchTransform, icsf_mesh** out) {
    if (!out) return fail(ICSF_ERR_ARG, "icsf_mesh_create: out is NULL");
    *out = nullptr;
    RC(checkDevice());
    std::unique_ptr<icsf_mesh> M(new icsf_mesh);
    std::string err = buildHostMesh(M->h, nPoints, points, nFaces, faceOffsets, faceVerts, nInternalFaces, owner,
                                    neighbour, nPatches, patchStart, patchSize, patchKind, patchNeighbour,
                                    patchTransform);
    if (!err.empty()) return fail(ICSF_ERR_ARG, err);
    HostMesh& h = M->h;
    buildSell(h.nCells, h.nInternal, h.owner.data(), h.neighbour.data(), M->sell);
    // sweep levels -> cell lists
    std::vector<int> fl, bl;
    int nf = 0, nbw = 0;
    sweepLevels(h.nCells, h.nInternal, h.owner.data(), h.neighbour.data(), fl, bl, nf, nbw);
    Levels& L = M->lev;
    L.nFwd = nf;
    L.nBwd = nbw;
    auto bucket = [&](const std::vector<int>& lev, int n, std::vector<int>& ptr, std::vector<int>& cells) {
        ptr.assign(n + 1, 0);
        for (int c = 0; c < h.nCells; ++c) ptr[lev[c] + 1]++;
        for (int l = 0; l < n; ++l) ptr[l + 1] += ptr[l];
        cells.assign(h.nCells, 0);
        std::vector<int> pos(ptr.begin(), ptr.end() - 1);
        for (int c = 0; c < h.nCells; ++c) cells[pos[lev[c]]++] = c;
    };
    std::vector<int> fc, bc;
    bucket(fl, nf, L.fwdPtrHost, fc);
    bucket(bl, nbw, L.bwdPtrHost, bc);
    L.twoColour = nf == 2 && nbw == 2 && L.fwdPtrHost[1] == L.bwdPtrHost[2] - L.bwdPtrHost[1] &&
                  std::equal(fc.begin() + L.fwdPtrHost[1], fc.end(), bc.begin());
    DevBuf& B = M->buf;
    MeshView& v = M->view;
    v.nC = h.nCells;
    v.nF = h.nFaces;
    v.nI = h.nInternal;
    v.nB = h.nFaces - h.nInternal;
    int *own, *nei, *cfOff, *cf, *bKind, *slicePtr, *col, *rowLen, *nLower;
    double *Sf, *magSf, *Cf, *w, *dc, *nodc, *corr, *C, *V;
    RC(B.upload(own, h.owner.data(), h.owner.size()));
    RC(B.upload(nei, h.neighbour.data(), h.neighbour.size()));
    RC(B.upload(Sf, h.Sf.data(), h.Sf.size()));
    RC(B.upload(magSf, h.magSf.data(), h.magSf.size()));
    RC(B.upload(Cf, h.Cf.data(), h.Cf.size()));
    RC(B.upload(w, h.weights.data(), h.weights.size()));
    RC(B.upload(dc, h.deltaCoeffs.data(), h.deltaCoeffs.size()));
    RC(B.upload(nodc, h.nonOrthDeltaCoeffs.data(), h.nonOrthDeltaCoeffs.size()));
    RC(B.upload(corr, h.corrVecs.data(), h.corrVecs.size()));
    RC(B.upload(C, h.C.data(), h.C.size()));
    RC(B.upload(V, h.V.data(), h.V.size()));
    RC(B.upload(cfOff, h.cellFaceOff.data(), h.cellFaceOff.size()));
    RC(B.upload(cf, h.cellFaces.data(), h.cellFaces.size()));
    std::vector<int> kinds(std::max(v.nB, 1), 0);
    for (size_t p = 0; p < h.patchStart.size(); ++p)
        for (int f = h.patchStart[p]; f < h.patchStart[p] + h.patchSize[p]; ++f) kinds[f - h.nInternal] = h.patchKind[p];
    RC(B.upload(bKind, kinds.data(), kinds.size()));
    RC(B.upload(slicePtr, M->sell.slicePtr.data(), M->sell.slicePtr.size()));
    RC(B.upload(col, M->sell.col.data(), M->sell.col.size()));
    RC(B.upload(rowLen, M->sell.rowLen.data(), M->sell.rowLen.size()));
    RC(B.upload(nLower, M->sell.nLower.data(), M->sell.nLower.size()));
    double* sgeo;
    {
        const size_t nSlot = (size_t)M->sell.slicePtr.back() * SLICE;
        std::vector<double> sg(std::max<size_t>(nSlot, 1) * NSG, 0.0);
        for (int f = 0; f < h.nInternal; ++f)
            for (const int slot : {M->sell.slotUpper[f], M->sell.slotLower[f]}) {
                for (int d = 0; d < 3; ++d) {
                    sg[sgOff(slot, SG_SX + d)] = h.Sf[3 * (size_t)f + d];
                    sg[sgOff(slot, SG_CX + d)] = h.Cf[3 * (size_t)f + d];
                }
                sg[sgOff(slot, SG_W)] = h.weights[f];
                sg[sgOff(slot, SG_MAG)] = h.magSf[f];
            }
        RC(B.upload(sgeo, sg.data(), sg.size()));
    }
    RC(B.upload(L.fwdCells, fc.data(), fc.size()));
    RC(B.upload(L.bwdCells, bc.data(), bc.size()));
    v.own = own; v.nei = nei; v.Sf = Sf; v.magSf = magSf; v.Cf = Cf; v.w = w; v.dc = dc; v.nodc = nodc;
    v.corr = corr; v.C = C; v.V = V; v.cfOff = cfOff; v.cf = cf; v.bKind = bKind;
    v.nSlices = M->sell.nSlices; v.slicePtr = slicePtr; v.col = col; v.rowLen = rowLen; v.nLower = nLower;
    v.sgeo = sgeo;
    v.emptyDir = h.emptyDir;
    v.nCyc = (int)h.cycFace.size();
    v.cycBase = (size_t)M->sell.slicePtr.back() * SLICE * NB;
    {
        int *cycOf, *cycCell, *cycRowOff, *cycRowK;
        double *cycR, *cycDelta;
        auto up = [&](auto*& d, const auto& hv) {
            return hv.empty() ? B.alloc(d, 1) : B.upload(d, hv.data(), hv.size());
        };
        RC(up(cycOf, h.cycOf)); RC(up(cycCell, h.cycCell)); RC(up(cycR, h.cycR));
        RC(up(cycDelta, h.cycDelta)); RC(up(cycRowOff, h.cycRowOff)); RC(up(cycRowK, h.cycRowK));
        v.cycOf = cycOf; v.cycCell = cycCell; v.cycR = cycR; v.cycDelta = cycDelta;
        v.cycRowOff = cycRowOff; v.cycRowK = cycRowK;
    }
    *out = M.release();
    return ICSF_OK;
}

int icsf_mesh_destroy(icsf_mesh* mesh) {
    delete mesh;
    return ICSF_OK;
}

int icsf_mesh_info(const icsf_mesh* m, int* info) {
    if (!m || !info) return fail(ICSF_ERR_ARG, "icsf_mesh_info: NULL argument");
    info[0] = m->h.nCells;
    info[1] = m->h.nFaces;
    info[2] = m->h.nInternal;
    info[3] = m->lev.nFwd;
    info[4] = m->lev.nBwd;
    info[5] = m->sell.nSlices;
    info[6] = m->sell.slicePtr.empty() ? 0 : m->sell.slicePtr.back() * SLICE;
    info[7] = m->h.emptyDir;
    return ICSF_OK;
}

namespace {
// A case; harmonic-balance instances (shared != nullptr) run on the shared
// stream and keep no Krylov storage of their own.
int createCase(icsf_mesh* mesh, const char* thermo, const char* schemes, const char* solution, hipStream_t shared,
               icsf_case** out) {
    *out = nullptr;
    RC(checkDevice());
    std::unique_ptr<icsf_case> C(new icsf_case);
    C->mesh = mesh;
    HIPCHK(hipGetDevice(&C->device));
    RC(readSettings(C.get(), thermo, schemes, solution));
    if (shared) {
        C->st = shared;
        C->ownStream = false;
    } else {
        HIPCHK(hipStreamCreate(&C->st));
    }
    const HostMesh& h = mesh->h;
    const int nC = h.nCells, nB = h.nFaces - h.nInternal;
    const size_t nBs = (size_t)std::max(nB, 1);
    C->bPT.assign(nBs, BC_ZG);
    C->bUT.assign(nBs, BC_ZG);
    C->bTT.assign(nBs, BC_ZG);
    C->bPv.assign(nBs, 0.0);
    C->bUv.assign(3 * nBs, 0.0);
    C->bTv.assign(nBs, 0.0);
    C->bInit.assign(5 * nBs, std::nan(""));
    DevBuf& B = C->buf;
    RC(B.alloc(C->d_bPT, nBs)); RC(B.alloc(C->d_bUT, nBs)); RC(B.alloc(C->d_bTT, nBs));
    RC(B.alloc(C->d_bPv, nBs)); RC(B.alloc(C->d_bUv, 3 * nBs)); RC(B.alloc(C->d_bTv, nBs));
    RC(B.alloc(C->d_bInit, 5 * nBs));
    C->mv = mesh->view;
    C->mv.bPT = C->d_bPT; C->mv.bUT = C->d_bUT; C->mv.bTT = C->d_bTT;
    C->mv.bPv = C->d_bPv; C->mv.bUv = C->d_bUv; C->mv.bTv = C->d_bTv;
    StateView& s = C->sv;
    RC(B.alloc(s.W, (size_t)5 * nC));
    RC(B.alloc(s.fld, (size_t)NFLD * nC));
    RC(B.alloc(s.T, (size_t)nC));
    RC(B.alloc(s.ecalc, (size_t)nC));
    RC(B.alloc(s.bfld, (size_t)NFLD * nBs));
    RC(B.alloc(s.bT, nBs));
    RC(B.alloc(s.grd, (size_t)NFLD * 3 * nC));
    RC(B.alloc(s.gv, (size_t)NGV * nC));
    RC(B.alloc(s.pseudoCo, (size_t)nC));
    RC(B.alloc(s.rDT, (size_t)nC));
    RC(B.alloc(s.b, (size_t)5 * nC));
    RC(B.alloc(s.ft, (size_t)NFT * h.nFaces));
    RC(B.alloc(s.bcV, 5 * nBs));
    RC(B.alloc(s.bcPhi, nBs));
    RC(B.alloc(s.bcIn, nBs));
    HIPCHK(hipMemset(s.bcV, 0, 5 * nBs * sizeof(double)));
    HIPCHK(hipMemset(s.bcPhi, 0, nBs * sizeof(double)));
    HIPCHK(hipMemset(s.bcIn, 0, nBs * sizeof(int)));
    s.bcInit = C->d_bInit;
    HIPCHK(hipMemset(s.ft, 0, (size_t)NFT * h.nFaces * sizeof(double)));
    HIPCHK(hipMemset(s.bfld, 0, NFLD * nBs * sizeof(double)));
    HIPCHK(hipMemset(s.bT, 0, nBs * sizeof(double)));
    k_fill<<<nblk(nC, 256), 256, 0, C->st>>>(nC, C->co, s.pseudoCo);
    // Jacobian storage
    const size_t slots = (size_t)mesh->sell.slicePtr.back() * SLICE;
    const size_t dgN = (size_t)mesh->sell.nSlices * SLICE * NB;
    const size_t eb = C->blkFp32 ? 4 : 8;
    {
        char* p;
        // sliced-ELL blocks (or their generators), then the interface blocks
        // of the cyclic faces
        C->mv.blkGen = C->blkGen ? 1 : 0;
        C->mv.gamma = C->ph.gamma;
        C->mv.cycBase = slots * (C->blkGen ? NGEN : NB);
        const size_t nBlk = C->mv.cycBase + (size_t)mesh->view.nCyc * NB;
        RC(B.alloc(p, nBlk * eb)); C->blk = p;
        HIPCHK(hipMemset(C->blk, 0, nBlk * eb));
        RC(B.alloc(p, dgN * eb)); C->dg = p;
        HIPCHK(hipMemset(C->dg, 0, dgN * eb));
        if (C->precond == 1) { RC(B.alloc(p, dgN * eb)); C->dinv = p; }
    }
    RC(B.alloc(C->rD, (size_t)nC));
    RC(B.alloc(C->derr, 1));
    HIPCHK(hipMemset(C->derr, 0, sizeof(int)));
    C->sv.err = C->derr;
    RC(B.alloc(C->dctl, NCTL));
    RC(B.alloc(C->dctlInit, 5));
    {
        int z[NCTL] = {0};
        z[CTL_FIRST] = 1;
        HIPCHK(hipMemcpy(C->dctl, z, sizeof(z), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(C->dctlInit, 0, 5 * sizeof(double)));
    }
    C->sv.halt = C->dctl + CTL_FIN;
    RC(B.alloc(C->fac, (size_t)nC));
    if (shared) {
        HIPCHK(hipStreamSynchronize(C->st));
        *out = C.release();
        return ICSF_OK;
    }
    // Krylov storage
    const size_t vs = (size_t)5 * nC;
    const int nOuter = C->nDirs;
    RC(B.alloc(C->V, vs * (size_t)(nOuter + 1)));
    RC(B.alloc(C->r, vs));
    RC(B.alloc(C->t1, vs));
    RC(B.alloc(C->t2, vs));
    RC(B.alloc(C->dW, vs));
    RC(B.alloc(C->dres, 10));
    if (C->solver == 1) {
        RC(B.alloc(C->Z, vs * (size_t)nOuter));
        if (C->mixed) {
            RC(B.alloc(C->V32, vs * (size_t)(C->innerDirs + 1)));
            RC(B.alloc(C->r32, vs));
            RC(B.alloc(C->r32b, vs));
        }
    }
    {
        char* p;
        RC(B.alloc(p, gmresScalarsBytes())); C->g = (GmresScalars*)p;
        RC(B.alloc(p, gmresScalarsBytes())); C->g2 = (GmresScalars*)p;
        HIPCHK(hipMemset(C->g, 0, gmresScalarsBytes()));
        HIPCHK(hipMemset(C->g2, 0, gmresScalarsBytes()));
    }
    RC(B.alloc(C->wk.part, (size_t)1024 * 64));
    HIPCHK(hipStreamSynchronize(C->st));
    *out = C.release();
    return ICSF_OK;
}
}  // namespace

int icsf_case_create(icsf_mesh* mesh, const char* thermo, const char* schemes, const char* solution,
                     icsf_case** out) {
    if (!mesh || !out) return fail(ICSF_ERR_ARG, "icsf_case_create: NULL argument");
    return createCase(mesh, thermo, schemes, solution, nullptr, out);
}

int icsf_case_destroy(icsf_case* c) {
    if (c && c->st) (void)hipStreamSynchronize(c->st);
    delete c;
    return ICSF_OK;
}

int icsf_set_boundary(icsf_case* c, int patch, int pType, int UType, int TType, const double* pValue,
                      const double* UValue, const double* TValue) {
    if (!c) return fail(ICSF_ERR_ARG, "NULL case");
    const HostMesh& h = c->mesh->h;
    if (patch < 0 || patch >= (int)h.patchStart.size()) return fail(ICSF_ERR_ARG, "patch index out of range");
    if (h.patchKind[patch] == PATCH_CYCLIC)
        return fail(ICSF_ERR_ARG, "patch " + std::to_string(patch) + " is cyclic: its values come from the "
                                  "coupled cells");
    const bool pOk = pType == BC_ZG || pType == BC_FIXED || pType == BC_TOTAL_P;
    const bool uOk = UType == BC_ZG || UType == BC_FIXED || UType == BC_SLIP || UType == BC_PIOV;
    const bool tOk = TType == BC_ZG || TType == BC_FIXED || TType == BC_TOTAL_T;
    if (!pOk || !uOk || !tOk) return fail(ICSF_ERR_KEYWORD, "unsupported boundary-condition type");
    const bool pv = pType == BC_FIXED || pType == BC_TOTAL_P, uv = UType == BC_FIXED,
               tv = TType == BC_FIXED || TType == BC_TOTAL_T;
    if ((pv && !pValue) || (uv && !UValue) || (tv && !TValue))
        return fail(ICSF_ERR_ARG, "fixedValue / totalPressure / totalTemperature condition without values");
    const bool uIn = UValue && (uv || UType == BC_PIOV);  // fixed U or tangentialVelocity
    const int s = h.patchStart[patch] - h.nInternal, n = h.patchSize[patch];
    for (int i = 0; i < n; ++i) {
        if ((pv && !std::isfinite(pValue[i])) ||
            (uIn && !(std::isfinite(UValue[3 * i]) && std::isfinite(UValue[3 * i + 1]) &&
                      std::isfinite(UValue[3 * i + 2]))) ||
            (tv && !std::isfinite(TValue[i])))
            return fail(ICSF_ERR_NONFINITE, "boundary values must be finite");
        if ((pType == BC_TOTAL_P && !(pValue[i] > 0)) || (tv && !(TValue[i] > 0)))
            return fail(ICSF_ERR_NONFINITE, "fixed / total temperature and total pressure must be positive");
    }
    if ((pType == BC_TOTAL_P || TType == BC_TOTAL_T) && !(c->ph.gamma > 1))
        return fail(ICSF_ERR_ARG, "totalPressure / totalTemperature need gamma > 1");
    if (c->mrf)
        for (int i = 0; i < n; ++i)
            if (c->mrfIncl[s + i] && UType != BC_FIXED)
                return fail(ICSF_ERR_ARG, "patch " + std::to_string(patch) + " rotates with the MRF zone: its U "
                                          "condition must be fixedValue");
    for (int i = 0; i < n; ++i) {
        c->bPT[s + i] = pType;
        c->bUT[s + i] = UType;
        c->bTT[s + i] = TType;
        if (pv) c->bPv[s + i] = pValue[i];
        for (int d = 0; d < 3; ++d) c->bUv[3 * (s + i) + d] = uIn ? UValue[3 * i + d] : 0.0;
        if (tv) c->bTv[s + i] = TValue[i];
        // the zone's rotating patches take the frame velocity
        // (MRFCoupledZone::correctBoundaryVelocity, every outer iteration)
        if (c->mrf && c->mrfIncl[s + i])
            for (int d = 0; d < 3; ++d) c->bUv[3 * (s + i) + d] = c->mrfUb[3 * (s + i) + d];
    }
    c->bcDirty = true;
    return ICSF_OK;
}

}  // extern "C"

namespace {

// One frame zone's contribution (rotating or translating): faceU & Sf/|Sf|
// per face, the moving boundary faces and their U.  Face types of
// setMRFFaces (MRFCoupledZone.C:42-189, MRFTranslatingZone.C the same):
// 1 internal faces with a zone cell on either side and the zone cells'
// faces on the other non-empty, non-coupled patches, 2 the zone cells'
// faces on coupled or non-moving patches; faceU on both (faceU,
// MRFCoupledZone.C:549-577), its normal component um = faceU & Sf/|Sf|
// (outerLoop.H:18-21); the type-1 boundary faces move with the frame
// (correctBoundaryVelocity).
int framePart(icsf_case* c, const char* who, int nZone, const int* zone, int nNonMov, const int* nonMov,
              const std::function<V3(int)>& faceU, FramePart& out) {
    const HostMesh& h = c->mesh->h;
    const int nC = h.nCells, nF = h.nFaces, nI = h.nInternal, nB = nF - nI, nP = (int)h.patchStart.size();
    std::vector<char> zc(nC, 0), excl(nP, 0);
    for (int i = 0; i < nZone; ++i) {
        if (zone[i] < 0 || zone[i] >= nC) return fail(ICSF_ERR_ARG, std::string(who) + ": zone cell out of range");
        zc[zone[i]] = 1;
    }
    for (int i = 0; i < nNonMov; ++i) {
        if (nonMov[i] < 0 || nonMov[i] >= nP) return fail(ICSF_ERR_ARG, std::string(who) + ": patch index out of range");
        excl[nonMov[i]] = 1;
    }
    std::vector<int> type(nF, 0);
    for (int f = 0; f < nI; ++f)
        if (zc[h.owner[f]] || zc[h.neighbour[f]]) type[f] = 1;
    for (int p = 0; p < nP; ++p)
        for (int f = h.patchStart[p]; f < h.patchStart[p] + h.patchSize[p]; ++f) {
            if (!zc[h.owner[f]]) continue;
            if (h.patchKind[p] == PATCH_CYCLIC || excl[p]) type[f] = 2;
            else if (h.patchKind[p] != PATCH_EMPTY) type[f] = 1;
        }
    for (int f = nI; f < nF; ++f)
        if (type[f] == 1 && c->bUT[f - nI] != BC_FIXED)
            return fail(ICSF_ERR_ARG, std::string(who) + ": a patch that moves with the zone needs a fixedValue U "
                                      "condition (list it among the non-moving patches otherwise)");
    out.on = true;
    out.un.assign(nF, 0.0);
    out.incl.assign(std::max(nB, 1), 0);
    out.Ub.assign((size_t)3 * std::max(nB, 1), 0.0);
    out.set.assign(nF, 0);
    for (int f = 0; f < nF; ++f) {
        V3 fu{0, 0, 0};
        if (type[f]) fu = faceU(f);
        out.set[f] = type[f] != 0;
        out.un[f] = dot(fu, V3{h.Sf[3 * f], h.Sf[3 * f + 1], h.Sf[3 * f + 2]} / h.magSf[f]);
        if (f >= nI && type[f] == 1) {
            const int b = f - nI;
            out.incl[b] = 1;
            out.Ub[3 * b] = fu.x; out.Ub[3 * b + 1] = fu.y; out.Ub[3 * b + 2] = fu.z;
        }
    }
    return ICSF_OK;
}

// The rotating zones combined as MRFCoupledZoneList combines them
// (MRFCoupledZoneList.C:55-142, MRFCoupledZone.C:307-318,563-627): each
// zone in list order assigns faceU on its faces, Omega in its cells and U
// on its moving patch faces, so a later zone overrides an earlier one
// where they meet.
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

// The case's frame arrays from its zones: um of the rotating zones plus um
// of the translating one, the moving faces' U (the translating zone's
// after the rotating ones', outerLoop.H:6-7) written into their fixed
// values; Omega from the rotating zones.  The frame divergence adds a
// diagonal term to every off-diagonal block: with generator storage it is
// kept as one more value per slot (MeshView::genMo).
int frameCompose(icsf_case* c) {
    const HostMesh& h = c->mesh->h;
    const int nF = h.nFaces, nI = h.nInternal, nB = nF - nI;
    const FramePart &R = c->frot, &T = c->ftr;
    std::vector<double> um(nF, 0.0);
    for (int f = 0; f < nF; ++f) um[f] = R.on && T.on ? R.un[f] + T.un[f] : (R.on ? R.un[f] : T.un[f]);
    c->mrfIncl.assign(std::max(nB, 1), 0);
    c->mrfUb.assign((size_t)3 * std::max(nB, 1), 0.0);
    for (int b = 0; b < nB; ++b)
        for (const FramePart* P : {&R, &T})
            if (P->on && P->incl[b]) {
                c->mrfIncl[b] = 1;
                for (int d = 0; d < 3; ++d) c->mrfUb[3 * b + d] = c->bUv[3 * b + d] = P->Ub[3 * b + d];
            }
    HIPCHK(hipStreamSynchronize(c->st));
    if (!c->d_umrf) RC(c->buf.alloc(c->d_umrf, (size_t)nF));
    HIPCHK(hipMemcpy(c->d_umrf, um.data(), (size_t)nF * sizeof(double), hipMemcpyHostToDevice));
    if (c->blkGen && !c->d_genMo) {
        // generator storage keeps its ten generators per slot; the frame
        // divergence's diagonal term goes beside them, one value per slot
        const size_t slots = (size_t)c->mesh->sell.slicePtr.back() * SLICE;
        const size_t eb = c->blkFp32 ? 4 : 8;
        char* p;
        RC(c->buf.alloc(p, slots * eb));
        HIPCHK(hipMemset(p, 0, slots * eb));
        c->d_genMo = p;
        c->mv.genMo = p;
    }
    c->mv.umrf = c->d_umrf;
    c->mv.omC = R.on ? c->d_omC : nullptr;
    c->mrf = true;
    c->bcDirty = true;
    c->haveResidual = c->assembled = false;
    // a captured step holds the mesh view without the zone
    if (c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
    }
    return ICSF_OK;
}

}  // namespace

extern "C" {

int icsf_set_mrf(icsf_case* c, int nZoneCells, const int* zoneCells, const double* origin, const double* axis,
                 double omega, int nNonRotating, const int* nonRotatingPatches) {
    if (!c || (nZoneCells > 0 && !zoneCells) || !origin || !axis || (nNonRotating > 0 && !nonRotatingPatches))
        return fail(ICSF_ERR_ARG, "icsf_set_mrf: NULL argument");
    if (!std::isfinite(omega) || !std::isfinite(origin[0]) || !std::isfinite(origin[1]) || !std::isfinite(origin[2]) ||
        !std::isfinite(axis[0]) || !std::isfinite(axis[1]) || !std::isfinite(axis[2]))
        return fail(ICSF_ERR_NONFINITE, "icsf_set_mrf: origin, axis and omega must be finite");
    const HostMesh& h = c->mesh->h;
    const int nC = h.nCells;
    // MRFCoupledZone: axis normalised, Omega = omega axis, faceU = Omega ^
    // (Cf - origin)
    V3 ax{axis[0], axis[1], axis[2]};
    const double am = mag(ax);
    ax = am < kROOTVSMALL ? V3{0, 0, 0} : ax / am;
    const V3 Om = omega * ax, o{origin[0], origin[1], origin[2]};
    FramePart P;
    RC(framePart(c, "icsf_set_mrf", nZoneCells, zoneCells, nNonRotating, nonRotatingPatches,
                 [&](int f) { return cross(Om, V3{h.Cf[3 * f], h.Cf[3 * f + 1], h.Cf[3 * f + 2]} - o); }, P));
    std::vector<double>& omc = c->omcHost;
    if (omc.empty()) omc.assign((size_t)3 * nC, 0.0);
    for (int i = 0; i < nZoneCells; ++i) {
        const int z = zoneCells[i];
        omc[3 * (size_t)z] = Om.x; omc[3 * (size_t)z + 1] = Om.y; omc[3 * (size_t)z + 2] = Om.z;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    if (!c->d_omC) RC(c->buf.alloc(c->d_omC, omc.size()));
    HIPCHK(hipMemcpy(c->d_omC, omc.data(), omc.size() * sizeof(double), hipMemcpyHostToDevice));
    c->rotZones.push_back(std::move(P));
    c->frot = rotCombined(c->rotZones);
    return frameCompose(c);
}

int icsf_set_translating_frame(icsf_case* c, int nZoneCells, const int* zoneCells, const double* velocity,
                               int nNonMoving, const int* nonMovingPatches) {
    if (!c || (nZoneCells > 0 && !zoneCells) || !velocity || (nNonMoving > 0 && !nonMovingPatches))
        return fail(ICSF_ERR_ARG, "icsf_set_translating_frame: NULL argument");
    if (c->ftr.on) return fail(ICSF_ERR_STATE, "icsf_set_translating_frame: the case already has a translating zone");
    if (!std::isfinite(velocity[0]) || !std::isfinite(velocity[1]) || !std::isfinite(velocity[2]))
        return fail(ICSF_ERR_NONFINITE, "icsf_set_translating_frame: the frame velocity must be finite");
    const V3 V{velocity[0], velocity[1], velocity[2]};
    FramePart P;
    RC(framePart(c, "icsf_set_translating_frame", nZoneCells, zoneCells, nNonMoving, nonMovingPatches,
                 [&](int) { return V; }, P));
    c->ftr = std::move(P);
    return frameCompose(c);
}

int icsf_set_boundary_values(icsf_case* c, int patch, const double* pValue, const double* UValue,
                             const double* TValue) {
    if (!c) return fail(ICSF_ERR_ARG, "NULL case");
    const HostMesh& h = c->mesh->h;
    if (patch < 0 || patch >= (int)h.patchStart.size()) return fail(ICSF_ERR_ARG, "patch index out of range");
    const int s = h.patchStart[patch] - h.nInternal, n = h.patchSize[patch];
    for (int i = 0; i < n; ++i) {
        if ((pValue && !(pValue[i] > 0 && std::isfinite(pValue[i]))) ||
            (TValue && !(TValue[i] > 0 && std::isfinite(TValue[i]))) ||
            (UValue && !(std::isfinite(UValue[3 * i]) && std::isfinite(UValue[3 * i + 1]) &&
                         std::isfinite(UValue[3 * i + 2]))))
            return fail(ICSF_ERR_NONFINITE, "initial boundary values must be finite, p and T positive");
    }
    for (int i = 0; i < n; ++i) {
        double* q = &c->bInit[5 * (size_t)(s + i)];
        if (pValue) q[0] = pValue[i];
        if (UValue) for (int d = 0; d < 3; ++d) q[1 + d] = UValue[3 * i + d];
        if (TValue) q[4] = TValue[i];
    }
    return ICSF_OK;
}

int icsf_set_fields(icsf_case* c, const double* rho, const double* U, const double* T) {
    if (!c || !rho || !U || !T) return fail(ICSF_ERR_ARG, "icsf_set_fields: NULL argument");
    const int nC = c->mv.nC;
    for (int i = 0; i < nC; ++i)
        if (!(rho[i] > 0) || !(T[i] > 0) || !std::isfinite(U[3 * i]) || !std::isfinite(U[3 * i + 1]) ||
            !std::isfinite(U[3 * i + 2]))
            return fail(ICSF_ERR_NONFINITE, "initial state must have finite U and positive rho, T");
    double *dr, *du, *dt;
    DevBuf tmp;
    RC(tmp.upload(dr, rho, nC));
    RC(tmp.upload(du, U, (size_t)3 * nC));
    RC(tmp.upload(dt, T, nC));
    // initial values of the derived conditions (createFields.H: the
    // fields' `value` entries; defaults p0, the owner cell's U, T0)
    {
        const HostMesh& h = c->mesh->h;
        const int nB = h.nFaces - h.nInternal;
        std::vector<double> q((size_t)5 * std::max(nB, 1), 0.0);
        for (int b = 0; b < nB; ++b) {
            const double* g = &c->bInit[5 * (size_t)b];
            const int o = h.owner[h.nInternal + b];
            q[5 * (size_t)b] = std::isnan(g[0]) ? c->bPv[b] : g[0];
            for (int d = 0; d < 3; ++d) q[5 * (size_t)b + 1 + d] = std::isnan(g[1 + d]) ? U[3 * (size_t)o + d] : g[1 + d];
            q[5 * (size_t)b + 4] = std::isnan(g[4]) ? c->bTv[b] : g[4];
        }
        HIPCHK(hipMemcpy(c->d_bInit, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    RC(uploadBC(c));
    launchUpdate(c->mv, c->ph, c->sv, nullptr, 0, dr, du, dt, c->st);
    HIPCHK(hipStreamSynchronize(c->st));
    c->haveFields = true;
    c->haveResidual = c->assembled = false;
    return ICSF_OK;
}

int icsf_get_fields(icsf_case* c, double* rho, double* U, double* T, double* p, double* W) {
    if (!c) return fail(ICSF_ERR_ARG, "NULL case");
    const int nC = c->mv.nC;
    HIPCHK(hipStreamSynchronize(c->st));
    std::vector<double> fld((size_t)NFLD * nC);
    HIPCHK(hipMemcpy(fld.data(), c->sv.fld, fld.size() * sizeof(double), hipMemcpyDeviceToHost));
    if (rho) for (int i = 0; i < nC; ++i) rho[i] = fld[(size_t)i * NFLD + F_RHO];
    if (p) for (int i = 0; i < nC; ++i) p[i] = fld[(size_t)i * NFLD + F_P];
    if (U)
        for (int i = 0; i < nC; ++i)
            for (int d = 0; d < 3; ++d) U[3 * i + d] = fld[(size_t)i * NFLD + F_UX + d];
    if (T) HIPCHK(hipMemcpy(T, c->sv.T, (size_t)nC * sizeof(double), hipMemcpyDeviceToHost));
    if (W) HIPCHK(hipMemcpy(W, c->sv.W, (size_t)5 * nC * sizeof(double), hipMemcpyDeviceToHost));
    return ICSF_OK;
}

int icsf_calc_flux(icsf_case* c, double* phi, double* phiUp, double* phiEp) {
    if (!c || !phi || !phiUp || !phiEp) return fail(ICSF_ERR_ARG, "icsf_calc_flux: NULL argument");
    if (!c->haveFields) return fail(ICSF_ERR_STATE, "fields not set");
    RC(residualStage(c));
    const int nF = c->mv.nF;
    double *a, *b, *d;
    DevBuf tmp;
    RC(tmp.alloc(a, nF)); RC(tmp.alloc(b, (size_t)3 * nF)); RC(tmp.alloc(d, nF));
    launchFaceFlux(c->mv, c->ph, c->sv, a, b, d, c->st);
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(phi, a, nF * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(phiUp, b, 3 * (size_t)nF * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(phiEp, d, nF * sizeof(double), hipMemcpyDeviceToHost));
    return ICSF_OK;
}

int icsf_residual(icsf_case* c, double* b, double* rDT) {
    if (!c) return fail(ICSF_ERR_ARG, "NULL case");
    if (!c->haveFields) return fail(ICSF_ERR_STATE, "fields not set");
    RC(residualStage(c));
    HIPCHK(hipStreamSynchronize(c->st));
    const int nC = c->mv.nC;
    if (b) HIPCHK(hipMemcpy(b, c->sv.b, (size_t)5 * nC * sizeof(double), hipMemcpyDeviceToHost));
    if (rDT) HIPCHK(hipMemcpy(rDT, c->sv.rDT, (size_t)nC * sizeof(double), hipMemcpyDeviceToHost));
    if (b)
        for (size_t i = 0; i < (size_t)5 * nC; ++i)
            if (!std::isfinite(b[i])) return fail(ICSF_ERR_NONFINITE, "non-finite residual");
    return ICSF_OK;
}

// The geometry adjoint covers the inviscid residual only.
static int geometryAdjointSupported(const icsf_case* c, const double* gSf, const double* gCf) {
    if (c->ph.viscous && (gSf || gCf))
        return fail(ICSF_ERR_ARG, "residual VJP: the Sf/Cf adjoint is built for inviscid residuals (pass NULL)");
    return ICSF_OK;
}

int icsf_residual_vjp_device(icsf_case* c, const double* W, const double* v, double* gW, double* gSf,
                             double* gCf) {
    if (!c || !v || !gW) return fail(ICSF_ERR_ARG, "icsf_residual_vjp_device: NULL argument");
    RC(geometryAdjointSupported(c, gSf, gCf));
    return vjpDevice(c, W, v, gW, gSf, gCf, gSf || gCf);
}

int icsf_residual_vjp(icsf_case* c, const double* W, const double* v, double* gW, double* gSf, double* gCf) {
    if (!c || !v || !gW) return fail(ICSF_ERR_ARG, "icsf_residual_vjp: NULL argument");
    RC(geometryAdjointSupported(c, gSf, gCf));
    const size_t n5 = 5 * (size_t)c->mv.nC, n3 = 3 * (size_t)c->mv.nF;
    DevBuf tmp;
    double *dW = nullptr, *dv, *dg;
    if (W) RC(tmp.upload(dW, W, n5));
    RC(tmp.upload(dv, v, n5));
    RC(tmp.alloc(dg, n5));
    RC(vjpDevice(c, dW, dv, dg, nullptr, nullptr, gSf || gCf));
    HIPCHK(hipMemcpy(gW, dg, n5 * sizeof(double), hipMemcpyDeviceToHost));
    if (gSf) HIPCHK(hipMemcpy(gSf, c->adjGS, n3 * sizeof(double), hipMemcpyDeviceToHost));
    if (gCf) HIPCHK(hipMemcpy(gCf, c->adjGC, n3 * sizeof(double), hipMemcpyDeviceToHost));
    return ICSF_OK;
}

int icsf_set_state(icsf_case* c, const double* W) {
    if (!c || !W) return fail(ICSF_ERR_ARG, "icsf_set_state: NULL argument");
    if (!c->haveFields) return fail(ICSF_ERR_STATE, "fields not set (icsf_set_fields initialises the boundary state)");
    const size_t n5 = 5 * (size_t)c->mv.nC;
    for (size_t i = 0; i < n5; ++i)
        if (!std::isfinite(W[i])) return fail(ICSF_ERR_NONFINITE, "icsf_set_state: non-finite conserved state");
    DevBuf tmp;
    double* dW;
    RC(tmp.upload(dW, W, n5));
    RC(uploadBC(c));
    RC(stateFromW(c, dW));
    HIPCHK(hipStreamSynchronize(c->st));
    return takeDeviceErrors(c->st, {c->derr});
}

int icsf_get_boundary_values(icsf_case* c, double* p, double* U, double* T) {
    if (!c) return fail(ICSF_ERR_ARG, "NULL case");
    if (!c->haveFields) return fail(ICSF_ERR_STATE, "fields not set");
    const int nB = c->mv.nB;
    if (nB == 0) return ICSF_OK;
    std::vector<double> v((size_t)5 * nB);
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(v.data(), c->sv.bcV, v.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int b = 0; b < nB; ++b) {
        if (p) p[b] = v[5 * (size_t)b];
        if (U) for (int d = 0; d < 3; ++d) U[3 * (size_t)b + d] = v[5 * (size_t)b + 1 + d];
        if (T) T[b] = v[5 * (size_t)b + 4];
    }
    return ICSF_OK;
}

int icsf_residual_records_device(icsf_case* c, const double* W, int* faceBranches, int* limiterElement,
                                 int* limiterFace) {
    if (!c || !faceBranches || !limiterElement || !limiterFace)
        return fail(ICSF_ERR_ARG, "icsf_residual_records_device: NULL argument");
    if (!W && !c->haveFields) return fail(ICSF_ERR_STATE, "fields not set");
    RC(adjointPrepare(c));
    RC(uploadBC(c));
    RC(stateFromW(c, W));
    const size_t nC = c->mv.nC, nF = c->mv.nF;
    DevBuf tmp;
    double *v, *gW;
    RC(tmp.alloc(v, 5 * nC));
    RC(tmp.alloc(gW, 5 * nC));
    HIPCHK(hipMemsetAsync(v, 0, 5 * nC * sizeof(double), c->st));
    AdjView a = c->adj;
    a.fbr = faceBranches;
    launchAdjointForward(c->mv, c->ph, c->sv, a.gun, a.lim, c->st);
    launchResidualVjp(c->mv, c->ph, c->sv, a, v, gW, c->adjGS, c->adjGC, false, c->st);
    HIPCHK(hipMemcpyAsync(limiterElement, a.tgt, nC * NFLD * sizeof(int), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(limiterFace, a.fst, nC * NFLD * sizeof(int), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    (void)nF;
    return takeDeviceErrors(c->st, {c->derr});
}

int icsf_residual_vjp_device_noalloc(icsf_case* c, const double* W, const double* v, double* gW, double* gSf,
                                     double* gCf) {
    if (!c || !v || !gW) return fail(ICSF_ERR_ARG, "icsf_residual_vjp_device_noalloc: NULL argument");
    RC(geometryAdjointSupported(c, gSf, gCf));
    if (!W && !c->haveFields) return fail(ICSF_ERR_STATE, "fields not set");
    RC(adjointSupported(c));
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != c->device) return fail(ICSF_ERR_ARG, "icsf_residual_vjp_device_noalloc: the case's device is not current");
    const size_t nC = c->mv.nC, nF = c->mv.nF, nB = std::max(c->mv.nB, 1);
    // the adjoint's work buffers carved from the fixed scratch array
    ScratchCarve sc{nullptr, 0};
    auto carve = [&](AdjView& a, double*& gS, double*& gC) {
        a.W0 = sc.take<double>(5 * nC);
        a.gun = sc.take<double>(nC * NFLD * 3);
        a.lim = sc.take<double>(nC * NFLD);
        a.fLR = sc.take<double>(nF * 2 * NFLD);
        a.G = sc.take<double>(nC * NFLD * 3);
        a.qc = sc.take<double>(nC * NFLD);
        a.tgt = sc.take<int>(nC * NFLD);
        a.tval = sc.take<double>(nC * NFLD);
        a.fst = sc.take<int>(nC * NFLD);
        a.fext = sc.take<double>(nC * NFLD);
        a.qf = sc.take<double>(nF * NFLD);
        a.bprim = sc.take<double>(nB * 5);
        gS = sc.take<double>(3 * nF);
        gC = sc.take<double>(3 * nF);
        if (c->ph.viscous) {
            a.vis = sc.take<double>(nF * NVIS);
            a.qe = sc.take<double>(nF);
            a.Ge = sc.take<double>(nC * 3);
        }
    };
    AdjView a{};
    double *gS = nullptr, *gC = nullptr;
    carve(a, gS, gC);  // sizes only
    if (sc.off > kScratchBytes)
        return fail(ICSF_ERR_ARG, "icsf_residual_vjp_device_noalloc: the case needs " + std::to_string(sc.off) +
                                      " bytes of work space, more than the fixed " + std::to_string(kScratchBytes) +
                                      "-byte scratch array (use icsf_residual_vjp_device)");
    char* base = nullptr;
    RC(scratchAddress(dev, &base));
    sc = ScratchCarve{base, 0};
    carve(a, gS, gC);
    RC(uploadBC(c));
    const size_t n5 = 5 * nC;
    if (W) HIPCHK(hipMemcpyAsync(c->sv.W, W, n5 * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(a.W0, c->sv.W, n5 * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    launchUpdate(c->mv, c->ph, c->sv, nullptr, 1, nullptr, nullptr, nullptr, c->st);
    c->haveFields = true;
    c->haveResidual = c->assembled = false;
    launchAdjointForward(c->mv, c->ph, c->sv, a.gun, a.lim, c->st);
    launchResidualVjp(c->mv, c->ph, c->sv, a, v, gW, gSf ? gSf : gS, gCf ? gCf : gC, gSf || gCf, c->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->st));
    return takeDeviceErrors(c->st, {c->derr});
}

int icsf_residual_device(icsf_case* c, const double* W, double* b) {
    if (!c || !b) return fail(ICSF_ERR_ARG, "icsf_residual_device: NULL argument");
    if (!W && !c->haveFields) return fail(ICSF_ERR_STATE, "fields not set");
    RC(uploadBC(c));
    RC(stateFromW(c, W));
    RC(residualStage(c));
    HIPCHK(hipMemcpyAsync(b, c->sv.b, 5 * (size_t)c->mv.nC * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return takeDeviceErrors(c->st, {c->derr});
}

int icsf_assemble_jacobian(icsf_case* c) {
    if (!c) return fail(ICSF_ERR_ARG, "NULL case");
    if (!c->haveResidual) return fail(ICSF_ERR_STATE, "icsf_residual must run before the Jacobian assembly");
    RC(jacobianStage(c));
    return takeDeviceErrors(c->st, {c->derr});
}

int icsf_get_jacobian(icsf_case* c, double* diag, double* upper, double* lower) {
    if (!c) return fail(ICSF_ERR_ARG, "NULL case");
    if (!c->assembled) return fail(ICSF_ERR_STATE, "Jacobian not assembled");
    const icsf_mesh* M = c->mesh;
    const size_t slots = (size_t)M->sell.slicePtr.back() * SLICE;
    const size_t dgN = (size_t)M->sell.nSlices * SLICE * NB;
    const size_t eb = c->blkFp32 ? 4 : 8;
    std::vector<char> hb(slots * NB * eb), hd(dgN * eb);
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->mv.blkGen) {
        // generator storage: expand the blocks on the device
        void* tmp = nullptr;
        HIPCHK(hipMalloc(&tmp, hb.size()));
        launchExpandBlocks(c->mv, c->blkFp32, c->blk, tmp, c->st);
        const hipError_t e1 = hipStreamSynchronize(c->st);
        const hipError_t e2 = e1 == hipSuccess ? hipMemcpy(hb.data(), tmp, hb.size(), hipMemcpyDeviceToHost) : e1;
        (void)hipFree(tmp);
        HIPCHK(e2);
    } else {
        HIPCHK(hipMemcpy(hb.data(), c->blk, hb.size(), hipMemcpyDeviceToHost));
    }
    HIPCHK(hipMemcpy(hd.data(), c->dg, hd.size(), hipMemcpyDeviceToHost));
    auto get = [&](const std::vector<char>& v, size_t i) -> double {
        return c->blkFp32 ? (double)((const float*)v.data())[i] : ((const double*)v.data())[i];
    };
    const int nC = M->h.nCells, nI = M->h.nInternal;
    if (diag)
        for (int r = 0; r < nC; ++r)
            for (int e = 0; e < NB; ++e) diag[(size_t)r * NB + e] = get(hd, diagOff(r, e));
    for (int f = 0; f < nI; ++f)
        for (int e = 0; e < NB; ++e) {
            if (upper) upper[(size_t)f * NB + e] = get(hb, sellOff(M->sell.slotUpper[f], e));
            if (lower) lower[(size_t)f * NB + e] = get(hb, sellOff(M->sell.slotLower[f], e));
        }
    return ICSF_OK;
}

int icsf_get_tuning(icsf_case* c, int* params) {
    if (!c || !params) return fail(ICSF_ERR_ARG, "icsf_get_tuning: NULL argument");
    const MeshView& m = c->mv;
    params[0] = m.bsOp; params[1] = m.bsFace; params[2] = m.bsCell; params[3] = m.bsJac; params[4] = m.mdCh;
    params[5] = c->tuned ? 1 : 0;
    params[6] = m.occOp; params[7] = m.occFace; params[8] = m.occCell; params[9] = m.occJac;
    return ICSF_OK;
}

int icsf_set_tuning(icsf_case* c, const int* params) {
    if (!c || !params) return fail(ICSF_ERR_ARG, "icsf_set_tuning: NULL argument");
    const auto wg = [](int b) { return b == 64 || b == 128 || b == 256; };
    for (int i : {0, 1, 2, 3})
        if (!wg(params[i])) return fail(ICSF_ERR_ARG, "icsf_set_tuning: workgroup sizes are 64, 128 or 256");
    if (params[4] != 8 && params[4] != 16 && params[4] != 32)
        return fail(ICSF_ERR_ARG, "icsf_set_tuning: projection passes take 8, 16 or 32 vectors");
    if (params[6] < 0 || params[6] > 1 || params[7] < 0 || params[7] > 2 || params[8] < 0 || params[8] > 2 ||
        params[9] < 0 || params[9] > 2)
        return fail(ICSF_ERR_ARG, "icsf_set_tuning: register budget out of range");
    MeshView& m = c->mv;
    m.bsOp = params[0]; m.bsFace = params[1]; m.bsCell = params[2]; m.bsJac = params[3]; m.mdCh = params[4];
    m.occOp = params[6]; m.occFace = params[7]; m.occCell = params[8]; m.occJac = params[9];
    c->tuned = true;
    // a captured step holds the old launch parameters
    if (c->gexec) {
        HIPCHK(hipStreamSynchronize(c->st));
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
    }
    return ICSF_OK;
}

int icsf_get_interface_blocks(icsf_case* c, double* blocks) {
    if (!c || !blocks) return fail(ICSF_ERR_ARG, "icsf_get_interface_blocks: NULL argument");
    if (!c->assembled) return fail(ICSF_ERR_STATE, "Jacobian not assembled");
    const HostMesh& h = c->mesh->h;
    const int nB = h.nFaces - h.nInternal, nCyc = c->mv.nCyc;
    std::fill(blocks, blocks + (size_t)nB * NB, 0.0);
    if (nCyc == 0) return ICSF_OK;
    const size_t eb = c->blkFp32 ? 4 : 8;
    std::vector<char> hv((size_t)nCyc * NB * eb);
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(hv.data(), (const char*)c->blk + c->mv.cycBase * eb, hv.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < nCyc; ++k) {
        const int b = h.cycFace[k] - h.nInternal;
        for (int e = 0; e < NB; ++e) {
            const size_t i = (size_t)k * NB + e;
            blocks[(size_t)b * NB + e] = c->blkFp32 ? (double)((const float*)hv.data())[i]
                                                    : ((const double*)hv.data())[i];
        }
    }
    return ICSF_OK;
}

int icsf_set_jacobian(icsf_case* c, const double* diag, const double* upper, const double* lower) {
    if (!c || !diag || !upper || !lower) return fail(ICSF_ERR_ARG, "icsf_set_jacobian: NULL argument");
    if (c->mv.blkGen)
        return fail(ICSF_ERR_STATE, "icsf_set_jacobian: the case stores its off-diagonal blocks as generators; "
                                    "create it with jacobianStorage blocks to supply blocks");
    const icsf_mesh* M = c->mesh;
    const size_t slots = (size_t)M->sell.slicePtr.back() * SLICE;
    const size_t dgN = (size_t)M->sell.nSlices * SLICE * NB;
    const int nC = M->h.nCells, nI = M->h.nInternal;
    std::vector<double> hb(slots * NB, 0.0), hd(dgN, 0.0), rD(nC);
    for (int r = 0; r < nC; ++r)
        for (int e = 0; e < NB; ++e) hd[diagOff(r, e)] = diag[(size_t)r * NB + e];
    for (int f = 0; f < nI; ++f)
        for (int e = 0; e < NB; ++e) {
            hb[sellOff(M->sell.slotUpper[f], e)] = upper[(size_t)f * NB + e];
            hb[sellOff(M->sell.slotLower[f], e)] = lower[(size_t)f * NB + e];
        }
    for (int r = 0; r < nC; ++r) {
        const double* D = diag + (size_t)r * NB;
        double x = kGREAT;
        for (int e : {0, 6, 12, 18, 24}) x = 1.0 / std::max(1.0 / x, std::fabs(D[e]));
        if (!(x >= kVSMALL)) return fail(ICSF_ERR_ZERO_DIAG, "All diagonals of coupledMatrix are zero.");
        rD[r] = x;
    }
    HIPCHK(hipStreamSynchronize(c->st));
    if (c->blkFp32) {
        std::vector<float> fb(hb.begin(), hb.end()), fd(hd.begin(), hd.end());
        HIPCHK(hipMemcpy(c->blk, fb.data(), fb.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->dg, fd.data(), fd.size() * 4, hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMemcpy(c->blk, hb.data(), hb.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->dg, hd.data(), hd.size() * 8, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c->rD, rD.data(), (size_t)nC * 8, hipMemcpyHostToDevice));
    if (c->precond == 1) launchJacobiSetup(nC, c->blkFp32, c->dg, c->dinv, c->st);
    HIPCHK(hipStreamSynchronize(c->st));
    c->assembled = true;
    return ICSF_OK;
}

int icsf_matrix_mul(icsf_case* c, const double* x, double* y) {
    if (!c || !x || !y) return fail(ICSF_ERR_ARG, "icsf_matrix_mul: NULL argument");
    if (!c->assembled) return fail(ICSF_ERR_STATE, "Jacobian not assembled");
    const size_t vs = (size_t)5 * c->mv.nC;
    HIPCHK(hipMemcpyAsync(c->t1, x, vs * 8, hipMemcpyHostToDevice, c->st));
    launchSpmv(c->mv, c->blkFp32, 0, c->blk, c->dg, c->t1, nullptr, nullptr, c->t2, c->st);
    HIPCHK(hipMemcpyAsync(y, c->t2, vs * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return ICSF_OK;
}

int icsf_precondition(icsf_case* c, double* r) {
    if (!c || !r) return fail(ICSF_ERR_ARG, "icsf_precondition: NULL argument");
    if (!c->assembled) return fail(ICSF_ERR_STATE, "Jacobian not assembled");
    const size_t vs = (size_t)5 * c->mv.nC;
    HIPCHK(hipMemcpyAsync(c->t1, r, vs * 8, hipMemcpyHostToDevice, c->st));
    applyPrecond(c, 0, c->t1);
    HIPCHK(hipMemcpyAsync(r, c->t1, vs * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return ICSF_OK;
}

int icsf_solve_for_incr(icsf_case* c, const double* W, const double* b, double* dW, double* residuals,
                        int* nIter) {
    if (!c) return fail(ICSF_ERR_ARG, "NULL case");
    if (!c->assembled) return fail(ICSF_ERR_STATE, "Jacobian not assembled");
    const size_t vs = (size_t)5 * c->mv.nC;
    DevBuf tmp;
    const double* dWsrc = c->sv.W;
    const double* db = c->sv.b;
    if (W) {
        double* p;
        RC(tmp.upload(p, W, vs));
        dWsrc = p;
    }
    if (b) {
        double* p;
        RC(tmp.upload(p, b, vs));
        db = p;
    }
    int it = 0;
    RC(solveStage(c, dWsrc, db, it));
    HIPCHK(hipStreamSynchronize(c->st));
    if (dW) HIPCHK(hipMemcpy(dW, c->dW, vs * 8, hipMemcpyDeviceToHost));
    if (residuals) gmresReadResiduals(c->g, residuals, c->st);
    if (nIter) *nIter = it;
    return ICSF_OK;
}

int icsf_pseudo_steps(icsf_case* c, int nSteps, double* residuals, int* nIter) {
    if (!c || nSteps < 0) return fail(ICSF_ERR_ARG, "icsf_pseudo_steps: bad arguments");
    if (!c->haveFields) return fail(ICSF_ERR_STATE, "fields not set");
    // flags left by earlier calls (e.g. a state handed to the VJP) do not
    // belong to these steps
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemset(c->derr, 0, sizeof(int)));
    for (int i = 0; i < nSteps; ++i) {
        int it = 1;
        if (graphable(c)) RC(graphStep(c));
        else RC(pseudoStep(c, it));
        if (residuals) gmresReadResiduals(c->g, residuals + 10 * (size_t)i, c->st);
        if (nIter) nIter[i] = it;
    }
    HIPCHK(hipGetLastError());
    const int rc = takeDeviceErrors(c->st, {c->derr});
    resolveEvents(c);
    return rc;
}

int icsf_pseudo_time_loop(icsf_case* c, int maxSteps, int batch, int* stepsRun, int* converged) {
    if (!c || maxSteps < 0 || batch < 1) return fail(ICSF_ERR_ARG, "icsf_pseudo_time_loop: bad arguments");
    if (!c->haveFields) return fail(ICSF_ERR_STATE, "fields not set");
    if (std::isnan(c->pseudoTol) || std::isnan(c->pseudoTolRel))
        return fail(ICSF_ERR_KEYWORD, "pseudoTime: keywords pseudoTol and pseudoTolRel are required "
                                      "(pseudotimeControl.C:56-58)");
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemset(c->derr, 0, sizeof(int)));
    // corr_ continues from the steps run so far (solnControl.setCorr of a
    // restart, initialise.H); the loop starts enabled and not converged
    {
        int h[NCTL];
        HIPCHK(hipMemcpy(h, c->dctl, sizeof(h), hipMemcpyDeviceToHost));
        h[CTL_CORR] = c->corr;
        h[CTL_CONV] = h[CTL_FIN] = h[CTL_RUN] = 0;
        h[CTL_EN] = 1;
        HIPCHK(hipMemcpy(c->dctl, h, sizeof(h), hipMemcpyHostToDevice));
    }
    const int corr0 = c->corr;
    int issued = 0, h[NCTL] = {0};
    int rc = ICSF_OK;
    while (issued < maxSteps) {
        const int n = std::min(batch, maxSteps - issued);
        for (int i = 0; i < n && rc == ICSF_OK; ++i) {
            int it = 1;
            rc = graphable(c) ? graphStep(c) : pseudoStep(c, it);
        }
        if (rc) break;
        issued += n;
        HIPCHK(hipMemcpyAsync(h, c->dctl, sizeof(h), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if (h[CTL_FIN]) break;
    }
    // pseudotimeControl resets converged_ when the loop ends; the halt flag
    // is cleared so that later calls change the state again
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(h, c->dctl, sizeof(h), hipMemcpyDeviceToHost));
    const int run = h[CTL_RUN];
    const bool conv = h[CTL_CONV] || h[CTL_FIN];
    h[CTL_CONV] = h[CTL_FIN] = h[CTL_EN] = 0;
    HIPCHK(hipMemcpy(c->dctl, h, sizeof(h), hipMemcpyHostToDevice));
    c->corr = corr0 + run;
    if (stepsRun) *stepsRun = run;
    if (converged) *converged = conv ? 1 : 0;
    if (rc) return rc;
    HIPCHK(hipGetLastError());
    rc = takeDeviceErrors(c->st, {c->derr});
    resolveEvents(c);
    return rc;
}

int icsf_get_pseudo_co(icsf_case* c, double* co) {
    if (!c || !co) return fail(ICSF_ERR_ARG, "NULL argument");
    HIPCHK(hipStreamSynchronize(c->st));
    HIPCHK(hipMemcpy(co, c->sv.pseudoCo, (size_t)c->mv.nC * 8, hipMemcpyDeviceToHost));
    return ICSF_OK;
}

int icsf_synchronize(icsf_case* c) {
    if (!c) return fail(ICSF_ERR_ARG, "NULL case");
    HIPCHK(hipStreamSynchronize(c->st));
    return ICSF_OK;
}

int icsf_timing_enable(icsf_case* c, int mask) {
    if (!c) return fail(ICSF_ERR_ARG, "NULL case");
    resolveEvents(c);
    c->timing = mask != 0;
    c->tmask = mask;
    for (int k = 0; k < NTIMER; ++k) { c->tms[k] = 0; c->tcnt[k] = 0; }
    return ICSF_OK;
}

int icsf_timing_read(icsf_case* c, int n, double* ms, long long* count) {
    if (!c) return fail(ICSF_ERR_ARG, "NULL case");
    resolveEvents(c);
    for (int k = 0; k < n && k < NTIMER; ++k) {
        if (ms) ms[k] = c->tms[k];
        if (count) count[k] = c->tcnt[k];
    }
    return ICSF_OK;
}


// ------------------------------------------------------ harmonic balance --
int icsf_hb_operator(int nFreq, const double* frequencies, const int* harmonics, int nInstants,
                     double selectedPeriod, double* instants, double* D, double* condition) {
    if (nFreq < 0 || (nFreq > 0 && (!frequencies || !harmonics)) || nInstants < 1 || !D)
        return fail(ICSF_ERR_ARG, "icsf_hb_operator: bad argument");
    std::vector<double> f(frequencies, frequencies + nFreq);
    std::vector<int> hm(harmonics, harmonics + nFreq);
    for (int k : hm)
        if (k < 0) return fail(ICSF_ERR_ARG, "icsf_hb_operator: negative harmonic count");
    const std::vector<double> om = hbOmegaList(f, hm);
    if ((int)om.size() > nInstants)
        return fail(ICSF_ERR_ARG, "specified number of instants is lower than frequency vector dimension");
    double cond = 0;
    const std::vector<double> t = hbInstants(nInstants, selectedPeriod, om, &cond);
    std::vector<double> Dv;
    std::string err;
    if (!hbOperator(om, t, Dv, err)) return fail(ICSF_ERR_ARG, err);
    if (instants) std::copy(t.begin(), t.end(), instants);
    std::copy(Dv.begin(), Dv.end(), D);
    if (condition) *condition = cond;
    return ICSF_OK;
}

int icsf_hb_reconstruct_weights(int nFreq, const double* frequencies, const int* harmonics, int nInstants,
                                const double* instants, double time, double* weights) {
    if (nFreq < 0 || (nFreq > 0 && (!frequencies || !harmonics)) || nInstants < 1 || !instants || !weights)
        return fail(ICSF_ERR_ARG, "icsf_hb_reconstruct_weights: bad argument");
    std::vector<double> f(frequencies, frequencies + nFreq);
    std::vector<int> hm(harmonics, harmonics + nFreq);
    for (int k : hm)
        if (k < 0) return fail(ICSF_ERR_ARG, "icsf_hb_reconstruct_weights: negative harmonic count");
    const std::vector<double> om = hbOmegaList(f, hm);
    std::vector<double> w;
    std::string err;
    if (!hbReconstructWeights(om, std::vector<double>(instants, instants + nInstants), time, w, err))
        return fail(ICSF_ERR_ARG, err);
    std::copy(w.begin(), w.end(), weights);
    return ICSF_OK;
}

int icsf_hb_reconstruct(icsf_hb* h, const double* weights, double* rho, double* U, double* T, double* p) {
    if (!h || !weights) return fail(ICSF_ERR_ARG, "icsf_hb_reconstruct: NULL argument");
    for (int J = 0; J < h->nO; ++J)
        if (!std::isfinite(weights[J])) return fail(ICSF_ERR_ARG, "icsf_hb_reconstruct: non-finite weight");
    const int nC = h->inst[0]->mv.nC;
    DevBuf tmp;
    double *dw, *out;
    RC(tmp.upload(dw, weights, (size_t)h->nO));
    RC(tmp.alloc(out, (size_t)6 * nC));
    launchHBReconstruct(nC, h->nO, h->dInst, dw, out, h->st);
    HIPCHK(hipGetLastError());
    std::vector<double> o((size_t)6 * nC);
    HIPCHK(hipMemcpyAsync(o.data(), out, o.size() * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (int c = 0; c < nC; ++c) {
        const double* q = &o[6 * (size_t)c];
        if (rho) rho[c] = q[0];
        if (U) for (int d = 0; d < 3; ++d) U[3 * (size_t)c + d] = q[1 + d];
        if (T) T[c] = q[4];
        if (p) p[c] = q[5];
    }
    return ICSF_OK;
}

int icsf_hb_create(icsf_mesh* mesh, const char* thermo, const char* schemes, const char* solution, int nInstants,
                   const double* D, icsf_hb** out) {
    if (!mesh || !D || !out) return fail(ICSF_ERR_ARG, "icsf_hb_create: NULL argument");
    if (nInstants < 1 || nInstants > HB_MAXO) return fail(ICSF_ERR_ARG, "icsf_hb_create: 1 <= nInstants <= 16");
    *out = nullptr;
    RC(checkDevice());
    std::unique_ptr<icsf_hb> H(new icsf_hb);
    H->mesh = mesh;
    H->nO = nInstants;
    H->D.assign(D, D + (size_t)nInstants * nInstants);
    HIPCHK(hipStreamCreate(&H->st));
    for (int J = 0; J < nInstants; ++J) {
        icsf_case* c = nullptr;
        RC(createCase(mesh, thermo, schemes, solution, H->st, &c));
        H->inst.push_back(c);
    }
    const icsf_case* c0 = H->inst[0];
    const bool fp64Gmres = c0->solver == 0 && !c0->mixed && !c0->blkFp32;
    const bool mixedFgmres = c0->solver == 1 && c0->mixed && c0->blkFp32;
    if (c0->precond != 0 || !(fp64Gmres || mixedFgmres))
        return fail(ICSF_ERR_KEYWORD, "harmonic balance: flowSolver must be GMRES in double precision or FGMRES "
                                      "with precision mixed, preconditioner LUSGS");
    H->solver = c0->solver; H->mixed = c0->mixed; H->innerDirs = c0->innerDirs;
    H->nDirs = c0->nDirs; H->maxIter = c0->maxIter; H->minIter = c0->minIter; H->orth = c0->orth;
    H->fuse = c0->fuse;
    H->useGraph = c0->useGraph;
    H->tol = c0->tol; H->relTol = c0->relTol;
    const size_t nC = mesh->h.nCells, vs = 5 * nC, vA = vs * nInstants;
    DevBuf& B = H->buf;
    RC(B.upload(H->dD, H->D.data(), H->D.size()));
    RC(B.alloc(H->Wall, vA)); RC(B.alloc(H->ball, vA)); RC(B.alloc(H->r, vA)); RC(B.alloc(H->r2, vA));
    RC(B.alloc(H->dW, vA)); RC(B.alloc(H->t1, vA)); RC(B.alloc(H->t2, vA));
    RC(B.alloc(H->V, vA * (size_t)(H->nDirs + 1)));
    if (H->mixed) {
        RC(B.alloc(H->Z, vA * (size_t)H->nDirs));
        RC(B.alloc(H->r32, vA));
        RC(B.alloc(H->r32b, vA));
        RC(B.alloc(H->V32, vA * (size_t)(H->innerDirs + 1)));
    }
    RC(B.alloc(H->rD, nC));
    RC(B.alloc(H->dres, 10 * (size_t)HB_MAXO));
    RC(B.alloc(H->derr, 1));
    HIPCHK(hipMemset(H->derr, 0, sizeof(int)));
    RC(B.alloc(H->wk.part, (size_t)1024 * 64));
    {
        char* p;
        RC(B.alloc(p, gmresScalarsBytes())); H->g = (GmresScalars*)p;
        HIPCHK(hipMemset(H->g, 0, gmresScalarsBytes()));
        RC(B.alloc(p, gmresScalarsBytes())); H->g2 = (GmresScalars*)p;
        HIPCHK(hipMemset(H->g2, 0, gmresScalarsBytes()));
        for (int J = 0; J < nInstants; ++J) {
            RC(B.alloc(p, gmresScalarsBytes()));
            HIPCHK(hipMemset(p, 0, gmresScalarsBytes()));
            H->gJ.push_back((GmresScalars*)p);
        }
    }
    // the registered e of every instance, which phase-lagged patches read
    for (icsf_case* c : H->inst) {
        double* e;
        RC(B.alloc(e, nC));
        HIPCHK(hipMemset(e, 0, nC * sizeof(double)));
        c->sv.e = e;
    }
    H->hbb.nO = nInstants;
    H->hbb.fp32 = c0->blkFp32;
    H->hbb.D = H->dD;
    std::vector<HBInst> hi(nInstants);
    for (int J = 0; J < nInstants; ++J) {
        icsf_case* c = H->inst[J];
        H->hbb.blk[J] = c->blk;
        H->hbb.dg[J] = c->dg;
        hi[J] = HBInst{c->mv, c->sv, c->blk, c->dg, c->rD, c->derr};
    }
    RC(B.alloc(H->dInst, (size_t)nInstants));
    HIPCHK(hipMemcpy(H->dInst, hi.data(), sizeof(HBInst) * nInstants, hipMemcpyHostToDevice));
    *out = H.release();
    return ICSF_OK;
}

int icsf_hb_destroy(icsf_hb* h) {
    delete h;
    return ICSF_OK;
}

int icsf_hb_set_boundary(icsf_hb* h, int instant, int patch, int pType, int UType, int TType, const double* pValue,
                         const double* UValue, const double* TValue) {
    RC(hbCheckInstance(h, instant));
    return icsf_set_boundary(h->inst[instant], patch, pType, UType, TType, pValue, UValue, TValue);
}

static int hbRefreshInstances(icsf_hb* h);

int icsf_hb_set_mrf(icsf_hb* h, int nZoneCells, const int* zoneCells, const double* origin, const double* axis,
                    double omega, int nNonRotating, const int* nonRotatingPatches) {
    if (!h) return fail(ICSF_ERR_ARG, "NULL harmonic-balance case");
    for (icsf_case* c : h->inst)
        RC(icsf_set_mrf(c, nZoneCells, zoneCells, origin, axis, omega, nNonRotating, nonRotatingPatches));
    return hbRefreshInstances(h);
}

int icsf_hb_set_translating_frame(icsf_hb* h, int nZoneCells, const int* zoneCells, const double* velocity,
                                  int nNonMoving, const int* nonMovingPatches) {
    if (!h) return fail(ICSF_ERR_ARG, "NULL harmonic-balance case");
    for (icsf_case* c : h->inst)
        RC(icsf_set_translating_frame(c, nZoneCells, zoneCells, velocity, nNonMoving, nonMovingPatches));
    return hbRefreshInstances(h);
}

int icsf_hb_phase_lag_operator(int nFreq, const double* frequencies, const int* harmonics, int nInstants,
                               const double* instants, double ibpa, double* Dl) {
    if (nFreq < 0 || (nFreq > 0 && (!frequencies || !harmonics)) || nInstants < 1 || !instants || !Dl ||
        !std::isfinite(ibpa))
        return fail(ICSF_ERR_ARG, "icsf_hb_phase_lag_operator: bad argument");
    std::vector<double> f(frequencies, frequencies + nFreq);
    std::vector<int> hm(harmonics, harmonics + nFreq);
    for (int k : hm)
        if (k < 0) return fail(ICSF_ERR_ARG, "icsf_hb_phase_lag_operator: negative harmonic count");
    const std::vector<double> om = hbOmegaList(f, hm);
    std::vector<double> d;
    std::string err;
    if (!hbLagOperator(om, std::vector<double>(instants, instants + nInstants), ibpa, d, err))
        return fail(ICSF_ERR_ARG, err);
    std::copy(d.begin(), d.end(), Dl);
    return ICSF_OK;
}

int icsf_hb_set_phase_lag(icsf_hb* h, int patch, const double* Dl, int cylCoords, const double* rotationAxis,
                          const double* rotationCentre) {
    if (!h || !Dl) return fail(ICSF_ERR_ARG, "icsf_hb_set_phase_lag: NULL argument");
    const HostMesh& hm = h->mesh->h;
    if (patch < 0 || patch >= (int)hm.patchStart.size() || hm.patchKind[patch] != ICSF_CYCLIC)
        return fail(ICSF_ERR_ARG, "icsf_hb_set_phase_lag: not a cyclic patch");
    const int nO = h->nO;
    for (int i = 0; i < nO * nO; ++i)
        if (!std::isfinite(Dl[i])) return fail(ICSF_ERR_ARG, "icsf_hb_set_phase_lag: non-finite time shift");
    // the viscous terms take grad(U) whole: the reference's unrotated lagged
    // velocity components would need component gradients of their own
    if (h->inst[0]->ph.viscous && !cylCoords)
        return fail(ICSF_ERR_ARG, "icsf_hb_set_phase_lag: viscous cases need cylCoords");
    double geo[8] = {0, 0, 1, 0, 0, 0, 0, 0};
    if (cylCoords) {
        if (!rotationAxis || !rotationCentre)
            return fail(ICSF_ERR_ARG, "icsf_hb_set_phase_lag: cylCoords needs the rotation axis and centre");
        const double* a = rotationAxis;
        const double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        if (!(la > 0) || !std::isfinite(la)) return fail(ICSF_ERR_ARG, "icsf_hb_set_phase_lag: zero rotation axis");
        for (int d = 0; d < 3; ++d) { geo[d] = a[d] / la; geo[3 + d] = rotationCentre[d]; }
        geo[6] = 1.0;
    }
    const int nCyc = (int)hm.cycFace.size();
    const int sideIx = (int)(h->plGeo.size() / 8);
    if (h->plSide.empty()) h->plSide.assign(nCyc, -1);
    for (int f = hm.patchStart[patch]; f < hm.patchStart[patch] + hm.patchSize[patch]; ++f)
        h->plSide[hm.cycOf[f - hm.nInternal]] = sideIx;
    h->plD.insert(h->plD.end(), Dl, Dl + (size_t)nO * nO);
    h->plGeo.insert(h->plGeo.end(), geo, geo + 8);
    HIPCHK(hipStreamSynchronize(h->st));
    for (void* p : {(void*)h->dplSide, (void*)h->dplD, (void*)h->dplGeo})
        if (p) h->buf.release(p);
    RC(h->buf.upload(h->dplSide, h->plSide.data(), h->plSide.size()));
    RC(h->buf.upload(h->dplD, h->plD.data(), h->plD.size()));
    RC(h->buf.upload(h->dplGeo, h->plGeo.data(), h->plGeo.size()));
    if (!h->dplFar) {
        std::vector<double> cf(3 * (size_t)nCyc);
        for (int kc = 0; kc < nCyc; ++kc)
            for (int d = 0; d < 3; ++d) cf[3 * (size_t)kc + d] = hm.Cf[3 * (size_t)hm.cycPair[kc] + d];
        RC(h->buf.upload(h->dplCf, cf.data(), cf.size()));
        RC(h->buf.alloc(h->dplFar, (size_t)nO * nCyc * PLW));
        HIPCHK(hipMemset(h->dplFar, 0, (size_t)nO * nCyc * PLW * sizeof(double)));
    }
    for (int J = 0; J < nO; ++J) {
        icsf_case* c = h->inst[J];
        c->mv.plSide = h->dplSide;
        c->sv.plFar = h->dplFar + (size_t)J * nCyc * PLW;
        c->haveResidual = c->assembled = false;
    }
    return hbRefreshInstances(h);
}

static int hbRefreshInstances(icsf_hb* h) {
    // the batched launches read the instances' views from device memory
    // (with the frame arrays, and full-block storage)
    std::vector<HBInst> hi(h->nO);
    for (int J = 0; J < h->nO; ++J) {
        icsf_case* c = h->inst[J];
        hi[J] = HBInst{c->mv, c->sv, c->blk, c->dg, c->rD, c->derr};
        h->hbb.blk[J] = c->blk;
    }
    HIPCHK(hipMemcpy(h->dInst, hi.data(), sizeof(HBInst) * h->nO, hipMemcpyHostToDevice));
    h->assembled = false;
    if (h->gexec) {
        HIPCHK(hipStreamSynchronize(h->st));
        (void)hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
    }
    return ICSF_OK;
}

int icsf_hb_set_boundary_values(icsf_hb* h, int instant, int patch, const double* pValue, const double* UValue,
                                const double* TValue) {
    RC(hbCheckInstance(h, instant));
    return icsf_set_boundary_values(h->inst[instant], patch, pValue, UValue, TValue);
}

int icsf_hb_set_fields(icsf_hb* h, int instant, const double* rho, const double* U, const double* T) {
    RC(hbCheckInstance(h, instant));
    h->assembled = false;
    return icsf_set_fields(h->inst[instant], rho, U, T);
}

int icsf_hb_get_fields(icsf_hb* h, int instant, double* rho, double* U, double* T, double* p, double* W) {
    RC(hbCheckInstance(h, instant));
    return icsf_get_fields(h->inst[instant], rho, U, T, p, W);
}

// error flags of the coupled diagonal and of every instance
static std::vector<int*> hbErrorFlags(const icsf_hb* h) {
    std::vector<int*> f{h->derr};
    for (icsf_case* c : h->inst) f.push_back(c->derr);
    return f;
}

static int hbReady(icsf_hb* h) {
    if (!h) return fail(ICSF_ERR_ARG, "NULL HB case");
    for (icsf_case* c : h->inst)
        if (!c->haveFields) return fail(ICSF_ERR_STATE, "fields not set for every instant");
    return ICSF_OK;
}

int icsf_hb_residual(icsf_hb* h, double* b) {
    RC(hbReady(h));
    RC(hbResidualStage(h));
    HIPCHK(hipStreamSynchronize(h->st));
    if (b)
        HIPCHK(hipMemcpy(b, h->ball, 5 * (size_t)h->inst[0]->mv.nC * h->nO * sizeof(double),
                         hipMemcpyDeviceToHost));
    return ICSF_OK;
}

int icsf_hb_assemble_jacobian(icsf_hb* h) {
    RC(hbReady(h));
    for (icsf_case* c : h->inst)
        if (!c->haveResidual) return fail(ICSF_ERR_STATE, "icsf_hb_residual must run before the Jacobian assembly");
    RC(hbJacobianStage(h));
    return takeDeviceErrors(h->st, hbErrorFlags(h));
}

int icsf_hb_matrix_mul(icsf_hb* h, const double* x, double* y) {
    if (!h || !x || !y) return fail(ICSF_ERR_ARG, "icsf_hb_matrix_mul: NULL argument");
    if (!h->assembled) return fail(ICSF_ERR_STATE, "HB Jacobian not assembled");
    const size_t vA = 5 * (size_t)h->inst[0]->mv.nC * h->nO;
    HIPCHK(hipMemcpy(h->t1, x, vA * sizeof(double), hipMemcpyHostToDevice));
    launchSpmvHB(h->inst[0]->mv, h->hbb, 0, h->t1, nullptr, nullptr, h->t2, h->st);
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemcpy(y, h->t2, vA * sizeof(double), hipMemcpyDeviceToHost));
    return ICSF_OK;
}

int icsf_hb_precondition(icsf_hb* h, double* r) {
    if (!h || !r) return fail(ICSF_ERR_ARG, "icsf_hb_precondition: NULL argument");
    if (!h->assembled) return fail(ICSF_ERR_STATE, "HB Jacobian not assembled");
    const size_t vA = 5 * (size_t)h->inst[0]->mv.nC * h->nO;
    HIPCHK(hipMemcpy(h->t1, r, vA * sizeof(double), hipMemcpyHostToDevice));
    hbLusgs(h, h->t1);
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemcpy(r, h->t1, vA * sizeof(double), hipMemcpyDeviceToHost));
    return ICSF_OK;
}

int icsf_hb_solve_for_incr(icsf_hb* h, const double* b, double* dW, double* residuals, int* nIter) {
    if (!h || !b || !dW) return fail(ICSF_ERR_ARG, "icsf_hb_solve_for_incr: NULL argument");
    if (!h->assembled) return fail(ICSF_ERR_STATE, "HB Jacobian not assembled");
    const size_t vs = 5 * (size_t)h->inst[0]->mv.nC, vA = vs * h->nO;
    for (int J = 0; J < h->nO; ++J) launchCopy(vs, 0, h->inst[J]->sv.W, h->Wall + J * vs, h->st);
    HIPCHK(hipMemcpy(h->ball, b, vA * sizeof(double), hipMemcpyHostToDevice));
    int it = 0;
    if (h->solver == 1) RC(hbSolveMixed(h, h->ball, it));
    else RC(hbSolve(h, h->ball, it));
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemcpy(dW, h->dW, vA * sizeof(double), hipMemcpyDeviceToHost));
    if (residuals) {
        std::vector<double> res;
        hbReadResiduals(h, res);
        std::copy(res.begin(), res.end(), residuals);
    }
    if (nIter) *nIter = it;
    return ICSF_OK;
}

int icsf_hb_pseudo_steps(icsf_hb* h, int nSteps, double* residuals) {
    RC(hbReady(h));
    if (nSteps < 0) return fail(ICSF_ERR_ARG, "icsf_hb_pseudo_steps: nSteps < 0");
    HIPCHK(hipStreamSynchronize(h->st));
    HIPCHK(hipMemset(h->derr, 0, sizeof(int)));
    for (icsf_case* c : h->inst) HIPCHK(hipMemset(c->derr, 0, sizeof(int)));
    for (int i = 0; i < nSteps; ++i) {
        if (hbGraphable(h)) RC(hbGraphStep(h));
        else RC(hbPseudoStep(h));
        if (residuals) {
            std::vector<double> res;
            hbReadResiduals(h, res);
            std::copy(res.begin(), res.end(), residuals + (size_t)i * 10 * h->nO);
        }
    }
    HIPCHK(hipGetLastError());
    return takeDeviceErrors(h->st, hbErrorFlags(h));
}

int icsf_hb_synchronize(icsf_hb* h) {
    if (!h) return fail(ICSF_ERR_ARG, "NULL HB case");
    HIPCHK(hipStreamSynchronize(h->st));
    return ICSF_OK;
}

}  // extern "C"
