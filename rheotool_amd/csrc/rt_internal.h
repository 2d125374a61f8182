// rt_internal.h -- internal types shared by the host runtime and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "rheo_gpu.h"

namespace rt {

// ----------------------------------------------------------------- errors
struct Error {
    int code;
    std::string msg;
};

#define RT_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::rt::Error{RT_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

// ----------------------------------------------------------------- context
struct Ctx;

// Device buffer with budget accounting; move-only.  A view (view = true)
// points into memory owned elsewhere (the no-allocation scratch array) and
// releases nothing.
struct DBuf {
    Ctx* ctx = nullptr;
    void* p = nullptr;
    size_t bytes = 0;
    bool view = false;
    DBuf() = default;
    static DBuf viewOf(void* p, size_t bytes) {
        DBuf b;
        b.p = p;
        b.bytes = bytes;
        b.view = true;
        return b;
    }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept { *this = std::move(o); }
    DBuf& operator=(DBuf&& o) noexcept;
    ~DBuf() { release(); }
    void release();
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct KernelStat {
    int64_t launches = 0;
    double ms = 0;
    double bytes = 0;
};

struct Ctx {
    // The handle and every mesh, model and server made on this context hold a
    // reference; the context is torn down when the last one goes, so objects
    // destroyed after rt_ctx_destroy still release their memory safely.
    int refs = 1;
    rt_ctx* handle = nullptr;
    int device = 0;
    int numCUs = 256;
    hipStream_t stream = nullptr;
    size_t budget = 0, inUse = 0;
    std::string lastError;
    // profiling
    bool profiling = false;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> eventPool;
    std::map<std::string, KernelStat> stats;
    // pinned host scratch for status read-back
    void* pinned = nullptr;
    size_t pinnedBytes = 0;
    // device staging buffer for host AoS uploads/downloads
    DBuf stage;
    // Released device buffers are cached by size class and reused by later
    // allocations on this context: hipFree synchronises the whole device, which
    // would stall every other stream (the serving loop destroys a model per
    // request).  Buffers are reused only on this context's stream, in order.
    std::multimap<size_t, void*> cache;
    size_t cachedBytes = 0;
    size_t cacheCap = size_t(8) << 30;
    // pinned 4 KiB blocks for the per-solve status read-back, reused likewise
    std::vector<void*> pinnedBlocks;
    static constexpr size_t kPinnedBlock = 4096;

    DBuf alloc(size_t bytes);
    void recycle(void* p, size_t bytes);
    void trimCache(size_t keep);  // frees cached buffers until at most keep bytes remain
    void* takePinnedBlock();
    void givePinnedBlock(void* p);
    void* pinnedHost(size_t bytes);
    hipEvent_t getEvent();
    void flushProfile();  // synchronises and folds pending events into stats
};

// Launch helper that records HIP events around the launch when profiling.
struct ProfScope {
    Ctx* ctx;
    const char* name;
    double bytes;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(Ctx* c, const char* n, double by) : ctx(c), name(n), bytes(by) {
        if (ctx->profiling) {
            a = ctx->getEvent();
            b = ctx->getEvent();
            RT_HIP(hipEventRecord(a, ctx->stream));
        }
    }
    ~ProfScope() noexcept(false) {
        if (a) {
            RT_HIP(hipEventRecord(b, ctx->stream));
            ctx->pending.push_back({name, a, b, bytes});
        }
    }
};

// ------------------------------------------------------------------- mesh
// Device layout.  Rows are cells in the device ("internal") numbering.
// Sliced ELL with 64-row slices: entry slot k of row r = 64*s + lane lives at
// ellOff[s] + 64*k + lane.  Two orders of the same entries share the slots:
//  * column order (ellCol, matrix values): slot k holds the k-th smallest
//    column, so slot k is the same kind of neighbour across the lanes of a
//    wavefront (coalesced gathers in the Krylov kernels);
//  * face order (ellFid, ellSlot, eSf, eW, ePhi): entry k is the row's k-th
//    face by increasing face id -- the summation order of the reference's
//    face loops -- and ellSlot maps it to its column-order slot.
// fid >= 0: the row is the owner of face fid; fid < 0: the row is the
// neighbour of face (-fid-1); RT_FID_PAD marks padding (column = row).
constexpr int32_t RT_FID_PAD = INT32_MIN;
constexpr int SLICE = 64;
constexpr int NCMP = 6;  // symmTensor components, interleaved per row

struct Mesh {
    Ctx* ctx = nullptr;
    int32_t N = 0, F = 0, NB = 0, nPatches = 0;
    int32_t nSlices = 0;
    int64_t nEll = 0;  // total ELL entries incl. padding
    int maxWidth = 0;
    int emptyDir = -1;
    int32_t valid[6];
    bool renumbered = false;
    std::vector<int32_t> perm;   // external cell -> internal row
    std::vector<int32_t> iperm;  // internal row -> external cell
    std::vector<rt_patch> patches;
    std::vector<int32_t> hOwner, hNeighbour;  // external numbering
    // device arrays
    DBuf ellOff;    // int64 [nSlices+1]
    DBuf ellWidth;  // int32 [nSlices]
    DBuf ellCol;    // int32 [nEll]  column order
    DBuf ellFid;    // int32 [nEll]  face order
    DBuf ellSlot;   // uint8 [nEll]  face-order entry -> column-order slot
    DBuf eSf;       // double [3][nEll] face order, Sf signed for the row's side
    DBuf eW;        // double [nEll] face order, linear weight of the face
    DBuf bPtr;      // int32 [N+1]   boundary faces of each row (non-empty patches)
    DBuf bFace;     // int32 [...]   boundary face index b (0..NB-1), ascending
    DBuf bPatch;    // int32 [NB]    patch of every boundary face
    DBuf bSf;       // double [3][NB] boundary face area vectors
    DBuf V;         // double [N] internal order
    DBuf C4;        // double [N][4] cell centres (x y z pad); empty when not given
    DBuf bDc;       // double [NB] patch delta coefficients 1/|n (n & (Cf - C_P))|; empty
                    // unless cell and boundary face centres are given
    DBuf permD;     // int32 [N] internal row -> external cell
    double meanAbsBandwidth = 0;
    // two-colourable cell graphs (hexahedral and quadrilateral meshes) are
    // stored colour by colour: rows [0, nRed) have only neighbours in
    // [nRed, N) and the other way round (0: not coloured)
    int32_t nRed = 0;
};

// --------------------------------------------------------- solver controls
enum Precond { PC_NONE = 0, PC_JACOBI = 1, PC_MCILU0 = 2 };
enum Solver { SV_PBICGSTAB = 0, SV_PBICG = 1 };
struct Controls {
    std::string solverName = "PBiCGStab";
    int solver = SV_PBICGSTAB;
    std::string precondName = "DILU";
    int precond = PC_JACOBI;
    bool precondStrict = false;  // multicolourILU0 named: refuse a mesh or solver it cannot run on
    double tolerance = 1e-6, relTol = 0.0;
    int maxIter = 1000, minIter = 0;
    bool mixed = false;
    double innerRelTol = 1e-5;  // FP32 inner solve reduction per refinement sweep
    int maxRefine = 12;
};
// Parses an fvSolution-style entry; throws Error.
Controls parseControls(const char* text);

// Per-component solver state living on the device, written by the last
// workgroup of each reduction kernel and read by the next kernels.
struct alignas(64) CompState {
    double rho, rhoOld, alpha, omega, beta;
    double normFactor, initRes, finalRes, xbar;
    double tmp0, tmp1, tmp2;
    int32_t iter, active, converged, solved;
    int32_t exitAtS, needRefine, nRefine, innerIter;
    double innerTarget;  // inner (FP32) residual target per sweep
    int32_t foldE;       // mixed: e holds an inner solve's correction not yet added to x
    int32_t pad0;
    double pad[2];
};
static_assert(sizeof(CompState) % 8 == 0, "CompState layout");

// controls the device side of the solve needs
struct KryCtl {
    double tol, relTol, innerRelTol;
    int32_t maxIter, minIter, maxInner, maxRefine;
    int32_t mixed;
    int32_t bicg;   // PBiCG (1) instead of PBiCGStab (0)
    int32_t precond;  // PC_NONE, PC_JACOBI or PC_MCILU0
    // red-black ILU0 (PC_MCILU0 on a colour-blocked mesh): the Krylov solve runs
    // on the Schur complement of the second colour; every solve starts and
    // ends with a sweep that eliminates the first colour exactly
    int32_t rb;
};

struct SolveState {
    CompState c[6];
    int32_t anyActive;
    double misc[8];  // misc[0]: ELL entries in live slots of the solve's matrix (set by the assembly)
};

struct SolveRun;  // host side of a resumable solve (kernels_krylov.hip)
struct SolveRunDeleter {
    void operator()(SolveRun* r) const;
};

// Workspace of the lockstep Krylov solver (all 6 components, SoA).
struct Krylov {
    Ctx* ctx = nullptr;
    int32_t N = 0;
    int grid = 0;  // workgroups of the row kernels
    DBuf r, r0, p, v, t;   // FP64 [N][6]
    DBuf rs, r0s, ps, vs, ts;  // FP32 [N][6] (mixed precision inner solve)
    DBuf es;               // FP32 [N][6] the inner solve's correction to x (mixed)
    // PBiCG: rA = r, rT = r0, pA = p, pT = v, wA = t, wT = s
    DBuf s64, s32;         // [N][6] wT in FP64 / FP32
    DBuf atval;            // double [nEll] A^T off-diagonals, column order (written by the assembly)
    DBuf a32, at32;        // float [nEll] A and A^T off-diagonals of the mixed PBiCG
    DBuf ahat32;           // float [nEll]  column-scaled off-diagonals A D^-1
    DBuf ahat64;           // double [nEll]
    DBuf rD;               // double [N]    1/diag
    DBuf rDs;              // double [N]    red-black: 1/diag of the Schur complement
                           // (the ILU0 diagonal) where it differs from 1/diag
    DBuf partials;         // double [grid][24]
    DBuf state;            // SolveState
    std::unique_ptr<SolveRun, SolveRunDeleter> run;
    // set when the buffers above are views into the no-allocation scratch
    // array: ensure() then refuses to allocate instead of calling hipMalloc
    bool noAlloc = false;
    void ensure(Mesh& m, bool mixed, bool bicg = false);
};

// ------------------------------------------------------------------- model
// GaussDefCmpw limiters (limiters.H:38-100); DS_NONE drops the convection term
enum DivScheme { DS_UPWIND = 0, DS_CUBISTA = 1, DS_MINMOD = 2, DS_SMART = 3, DS_WACEB = 4, DS_SUPERBEE = 5,
                 DS_NONE = 6 };
enum ModelKind { M_OLDROYDB = 0, M_GIESEKUS = 1, M_FENEP = 2, M_PTT_LINEAR = 3, M_PTT_EXP = 4,
                 M_FENECR = 5, M_WMCY = 6, M_ROLIEPOLY = 7, M_XPP = 8, M_SARAMITO = 9,
                 // BMPLog: the theta equation, and its scalar fluidity equation (an
                 // owned sub-model holding Phi in component xx of its theta)
                 M_BMP = 10, M_FLUIDITY = 11,
                 // Oldroyd-BRootk (theta = C^(1/k)); PTTLog with the generalized
                 // (Mittag-Leffler) destruction function
                 M_ROOTK = 12, M_PTT_GEN = 13 };
// models whose tau needs a per-cell factor of the old state (Model::fOld)
inline bool modelKeepsFactor(int kind) { return kind == M_FENEP || kind == M_FENECR || kind == M_WMCY; }

struct ModelParams {
    int kind = M_OLDROYDB;
    double rho = 1, etaS = 0, etaP = 0, lambda = 1;
    double alpha = 0, L2 = 0, epsilon = 0, zeta = 0;
    int ofInvQuirk = 0;
    double lambdaR = 1, lambdaD = 1, beta = 0, delta = 0, chiMax = 0;  // Rolie-Poly
    double lambdaS = 1, lambdaB = 1, q = 1, nXPP = 0;                 // XPP (+ alpha)
    double wmM = 1, wmN = 1, wmK = 0, wmL = 0, wmA = 1, wmB = 1;      // White-Metzner CY
    // Saramito: yield stress, power, consistency, valid dimensions, PTT function
    double tau0 = 0, sarN = 1, sarK = 0, dimX = 1, dimY = 1, dimZ = 1;
    int sarPTT = 0;  // 0 none, 1 linear, 2 exponential
    // BMP: modulus, rate constant, fluidity at rest and at infinite shear
    double G0 = 1, kBMP = 0, Phi0 = 0, PhiInf = 0;
    // Oldroyd-BRootk: the root k (Oldroyd_BRootk.C:121)
    double rootK = 1;
    // PTT generalized (PTTLog.C:135-152): gamma(beta), the number of series
    // terms, the relative tolerance, and the device table
    // mlGamma[j] = gamma(alpha j + beta) of the series denominators
    double mlGammaBeta = 1, mlRtol = 1e-12;
    int32_t mlMaxIter = 0;
    const double* mlGamma = nullptr;
};

struct Model {
    Ctx* ctx = nullptr;
    Mesh* mesh = nullptr;
    std::string name, typeName;
    ModelParams prm;
    Controls ctl;
    double relax = 0;
    // the momentum coupling of divTau (constitutiveEq.C:72-132): stabilization
    // option (0 none, 1 BSD, 2 coupling) and the snGrad scheme of the
    // laplacian(eta,U) entry (0 corrected, 1 orthogonal, 2 uncorrected)
    int stab = 0;
    int lapScheme = 0;
    // device state in internal numbering
    DBuf theta, tau;        // [N][6] interleaved
    DBuf eigVal;            // [3][N] conformation eigenvalues (diagonal of eigVals)
    DBuf eigVec;            // [9][N]
    DBuf U;                 // [N][4] interleaved (4th slot unused)
    DBuf Ub;                // [3][NB]
    DBuf phi;               // [F+NB] caller's face order
    DBuf ePhi;              // [nEll] face flux per ELL entry (face order)
    DBuf gradU;             // [9][N] when supplied
    DBuf diag, aval, b, fOld;  // matrix (column order) / sources [N][6] of the step
    DBuf slotMask;             // uint32 [nSlices] live slots of aval
    DBuf thetaOld;             // [N][6] theta at the start of the step (inner iterations)
    bool iterating = false;    // thetaOld holds the current step's oldTime
    int scheme = 0;            // DivScheme of div(phi,theta)
    bool bounded = false;      // the `bounded` wrapper: diagonal minus the cell's net outflux
    DBuf mlGamma;              // PTT generalized: the gamma table of the series (prm.mlGamma)
    std::vector<double> mlGammaHost;
    DBuf gradTh;               // double [N][18] component gradients (limited schemes)
    DBuf souT;                 // double [N][6] explicit correction source
    DBuf bcType;            // int32 [nPatches]
    DBuf bcVal;             // double [6][NB]
    DBuf patchT;            // double [nPatches][9] faceT of the wedge patches
    std::vector<int32_t> hBcType;
    bool hasFlow = false, hasState = false;
    Krylov kry;
    rt_solver_perf perf[6];
    struct SolveArgs* pendingSolve = nullptr;  // set between correct begin/advance
    // multiMode (multiMode.C): the modes own their state; this model holds the
    // sum of their tau and forwards the setters
    std::vector<rt_model*> modes;
    bool isMulti() const { return !modes.empty(); }
    // BMPLog (BMPLog.C:142-163): the fluidity equation, solved before theta each
    // step; in the fluidity sub-model, the BMP model's tau its source reads
    rt_model* fluid = nullptr;
    const DBuf* parentTau = nullptr;
    bool scalar() const { return prm.kind == M_FLUIDITY; }
    int32_t valid6[6] = {0, 0, 0, 0, 0, 0};  // components solved in the last step
};

// ---------------------------------------------------------------- kernels
struct AssembleArgs {
    int32_t N;
    const int64_t* ellOff;
    const int32_t* ellWidth;
    const int32_t* ellCol;
    const int32_t* ellFid;
    const uint8_t* ellSlot;
    const double* eSf;    // [3][nEll]
    const double* eW;
    const double* ePhi;
    int64_t nEll;
    const int32_t* bPtr;
    const int32_t* bFace;
    const int32_t* bPatch;
    const int32_t* bcType;
    const double* bcVal;  // [6][NB]
    const double* patchT; // [nPatches][9] wedge faceT (row-major)
    int32_t NB;
    const double* bSf;    // [3][NB]
    const double* phiB;   // [NB] boundary face fluxes
    const double* V;
    const double* U;      // [N][4]
    const double* Ub;     // [3][NB]
    const double* gradU;  // [9][N], nullable
    const double* theta;  // [N][6]
    const double* thetaOld;  // [N][6] theta.oldTime() for inner iterations, NULL: theta
    const double* eigVal;
    const double* eigVec;
    const double* tauOld;  // [N][6] the model's tau before the step (Saramito, BMP fluidity), else NULL
    const double* aux;     // [N][6] BMP: the step's fluidity in component xx, else NULL
    double rDeltaT;
    double relax;         // <= 0: none
    ModelParams prm;
    int32_t invMask;      // [OF9] inv() first-cell shift bits (ofInvQuirk)
    double* aval;         // the solve's matrix, column order
    uint32_t* slotMask;   // [nSlices] live column-order slots of aval
    double* diag;
    double* b;            // [N][6]
    double* fOld;         // FENE-P f, nullable
    double* atval;        // A^T off-diagonals in column order (PBiCG), nullable
    const double* souT;   // [N][6] deferred-correction source, NULL for none
    int32_t noConv;       // the "none" scheme: no convection term
    int32_t bounded;      // subtract V*(net outflux/V) from the diagonal
    double* partials;     // [grid][19] residual sums, entries in live slots
    CompState* cs;
    double* nnzOut;       // SolveState::misc[0]
    int32_t valid[6];
    // initial residual of the solve, fused into the assembly
    void* r;              // [N][6] Krylov precision
    void* r0;
    double* rD;           // [N] 1/diag
    KryCtl ctl;
    int32_t mixed;        // r/r0 in FP32
    int32_t nRed;         // colour-blocked rows: the first colour's count (tile order), 0: none
};

void launch_assemble(Ctx* ctx, const AssembleArgs& a, int grid);
// Gauss-linear gradients of the 6 theta components ([N][18]) and the
// deferred correction source souT ([N][6]) of a limited GaussDefCmpw scheme
void launch_defc(Ctx* ctx, const AssembleArgs& a, int scheme, const double* C4, double* gradTh, double* souT,
                 int grid);
// fvc::div(tau/rho) -> out [N][3]; tauB [6][NB] patch values or NULL (zeroGradient)
// with stab (0 none, 1 BSD, 2 coupling) the stabilization term of divTau, lap
// the snGrad scheme of its laplacian (0 corrected, 1 orthogonal, 2
// uncorrected), gam = etaP/rho; G [N][9] receives grad(U), C4 cell centres,
// bDc [NB] the patch delta coefficients
void launch_div_tau(Ctx* ctx, const AssembleArgs& a, const double* tau, const double* tauB, double rho,
                    int stab, int lap, double gam, double* G, const double* C4, const double* bDc, double* out,
                    int grid);
// tauTotal() -> out [N][6]
void launch_tau_total(Ctx* ctx, const AssembleArgs& a, const double* tau, double etaS, double* out, int grid);
void launch_eig_tau(Ctx* ctx, int32_t N, const ModelParams& prm, const double* theta,
                    double* eigVal, double* eigVec, double* tau, const double* fOld);
void launch_aos_to_soa(Ctx* ctx, int32_t N, int ncomp, const double* aos, const int32_t* perm,
                       double* soa);
void launch_soa_to_aos(Ctx* ctx, int32_t N, int ncomp, const double* soa, const int32_t* perm,
                       double* aos);
// gather rows of an interleaved [N][nc] array: out[r] = in[perm[r]] (perm null: copy),
// padding the destination stride ncOut >= nc with zeros
void launch_permute_rows(Ctx* ctx, int32_t N, int nc, int ncOut, const double* in, const int32_t* perm,
                         double* out);
// scatter rows back: out[perm[r]] = in[r] (first nc of ncIn components)
void launch_unpermute_rows(Ctx* ctx, int32_t N, int ncIn, int nc, const double* in, const int32_t* perm,
                           double* out);
// per-entry face flux: ePhi[e] = phi[face of entry e]
void launch_face_to_entry(Ctx* ctx, const Mesh& m, const double* phi, double* ePhi);
void launch_eigvals_from_tensor(Ctx* ctx, int32_t N, const double* aos9, const int32_t* perm,
                                double* eigVal);
void launch_eigvals_to_tensor(Ctx* ctx, int32_t N, const double* eigVal, const int32_t* perm,
                              double* aos9);
// out[i] += in[i], i < n
void launch_add(Ctx* ctx, int64_t n, const double* in, double* out);
void launch_fill(Ctx* ctx, double* p, int64_t n, double v);
void launch_copy(Ctx* ctx, const void* src, void* dst, size_t bytes);

// user-supplied LDU -> ELL values and diagonal (the sparseSolver surface)
void launch_ldu_to_ell(Ctx* ctx, const Mesh& m, const double* lower, const double* upper,
                       double* aval);

struct SolveArgs {
    Mesh* mesh;
    Krylov* kry;
    const double* diag;  // [N]
    const double* aval;  // ELL values
    const uint32_t* slotMask;  // [nSlices] live slots of aval; NULL: all live
    const double* atval;       // A^T off-diagonals (PBiCG), column order
    const double* b;     // [N][6]
    double* x;           // [N][6] (theta in place)
    const int32_t* valid;
    const Controls* ctl;
    bool xbarKnown;      // CompState.xbar already set
    bool initDone;       // r, r0, rD and the initial state already written (by the assembly)
    // red-black: 1/diag of the Schur complement, NULL when it is 1/diag (every
    // product A_ij A_ji vanishes, as for the upwind theta matrix)
    const double* rDs = nullptr;
    // pageable host block for the status read-back; NULL: a page-locked block
    // of the context and an event (the blocking path then waits on the stream)
    SolveState* hostStatus = nullptr;
};
KryCtl kryCtl(const Controls& C, bool mixed, const Mesh& m);
// red-black ILU0 of a general LDU matrix: out[0, N) = 1/Ds (the Schur
// complement's diagonal on the second colour, 1/D on the first), out[N, 2N) =
// D/Ds; lower/upper per internal face
void launch_rb_schur_diag(Ctx* ctx, const Mesh& m, const double* lower, const double* upper, const double* diag,
                          double* out);
// mean of theta per component -> CompState.xbar (gAverage(psi) of the normalisation)
struct Dict;
// a model (or a multiMode and its modes) from its parameters dictionary
void createModel(Model& M, Ctx* c, Mesh* mesh, const std::string& name, const Dict& d, const std::string& dn);
// state and flow uploads (rt_model_set_state / rt_model_set_flow); wait =
// false returns before the copies have read the host arrays
void modelUploadState(Model& M, const double* theta, const double* eigVals, const double* eigVecs,
                      const double* tau, bool wait);
void modelUploadFlow(Model& M, const double* U, const double* Ub, const double* phi, bool wait);
void destroyModes(Model& M);
void allocModelState(Model& M);
void launch_mean6(Ctx* ctx, int32_t N, const double* x, double* partials,
                  CompState* cs, int grid);
// Lockstep PBiCGStab over the valid components; fills perf.
void solve_lockstep(Ctx* ctx, const SolveArgs& a, rt_solver_perf perf[6]);
// The same as a resumable job: begin launches the initial residual;
// advance launches the next round once the previous round's state is on the
// host (block = false returns false immediately when it is not yet) and
// returns true when every component is finished; finish fills perf.
void solve_begin(Ctx* ctx, const SolveArgs& a);
bool solve_advance(Ctx* ctx, Krylov& K, bool block);
void solve_finish(Krylov& K, rt_solver_perf perf[6]);

int grid_for(const Ctx* ctx, int32_t N);

// Grouped Krylov rounds (the serving loop): one launch per kernel runs the
// lockstep PBiCGStab of every member solve, each member on the workgroups its
// own launch would use and with its own fixed-order reductions, so every member
// computes the bits it computes alone.  Members share the precision and the
// preconditioner form; each is begun (solve_begin) on its own stream, and the
// rounds run on the group's stream.
struct SolveGroup {
    Ctx* ctx = nullptr;  // the group's stream and table memory
    bool mixed = false, rb = false;
    struct Member {
        Krylov* K;
        hipStream_t home;  // the stream the solve was begun on (and continues on)
    };
    std::vector<Member> members;
    bool dirty = true;   // tables to upload before the next launch
    DBuf tables;
    void* hostTables = nullptr;  // page-locked staging of the tables
    size_t hostBytes = 0;
    bool launched = false;       // rounds queued whose member states are not all back
    int64_t launches = 0;        // grouped rounds launched (for tests and measurements)
    ~SolveGroup();
};
// a begun solve that grouped rounds can run (PBiCGStab), and its group class
bool solve_groupable(const Krylov& K, bool* mixed, bool* rb);
// adds a begun solve: the group's stream waits for the work queued on `home`
void group_join(SolveGroup& g, Krylov& K, hipStream_t home);
// Polls the members' states without blocking.  Members whose solve has
// finished leave the group (their `home` stream first waits for the group's
// queued work, which may still read them) and are appended to `finished`.
// When every member's state is back, launches the next grouped rounds.
void group_advance(SolveGroup& g, std::vector<Krylov*>& finished);
// takes a member out (a failed request): waits for the group's queued work
void group_remove(SolveGroup& g, Krylov& K);
// looks up the device address of the no-allocation scratch array on the
// current device, once per device (rt_noalloc.cpp)
void noallocInit(int device);

// correct() as a resumable job (the serving loop interleaves several):
// begin launches the assembly and the initial residual; advance returns true
// once the solve has finished and the eigen/tau kernel has been launched.
// iteration < 0: a plain step (theta.oldTime() is theta); 0 starts a step with
// inner iterations (saves theta.oldTime()); > 0 repeats the step from the
// current iterate with the saved oldTime
void modelCorrectBegin(Model& M, double deltaT, const double* gradU, int iteration = -1);
bool modelCorrectAdvance(Model& M, bool block);
// after the solve: its performance record, then calcEig and tau (launched)
void modelCorrectFinish(Model& M);
void modelCorrect(Model& M, double deltaT, const double* gradU, int iteration = -1);
// device bytes a model of this size allocates (for admission control)
size_t modelBytesEstimate(int32_t N, int32_t F, int32_t NB, int64_t nEll, bool mixed, bool gradU = false,
                          bool download = false);

}  // namespace rt

struct rt_ctx {
    rt::Ctx impl;
};
namespace rt {
inline void ctxRef(Ctx* c) { ++c->refs; }
void ctxUnref(Ctx* c);  // tears the context down at the last reference
}  // namespace rt
struct rt_mesh {
    rt::Mesh impl;
};
struct rt_model {
    rt::Model impl;
};
