/*
 * rheo_gpu.h -- C-ABI of the MI355X (gfx950) log-conformation stress step.
 *
 * This is the drop-in boundary for rheoTool's viscoelastic stress step:
 * constitutiveEq::correct() of the *Log models and the segregated lduMatrix
 * solve of the six theta components behind fvMatrix::solve().  Plain C types
 * only: pointers, sizes, status codes.  Host arrays passed in are borrowed for
 * the duration of the call; the library never frees caller memory.
 *
 * Layouts of host arrays are OpenFOAM's (AoS per cell/face):
 *   vector      3 doubles  x y z
 *   symmTensor  6 doubles  xx xy xz yy yz zz
 *   tensor      9 doubles  xx xy xz yx yy yz zx zy zz (row-major)
 * Reference paths are relative to rheoTool's of90/src/libs/.
 *
 * Threading: a context binds one device and one HIP stream; calls on one
 * context are not thread-safe, separate contexts are independent (as the
 * reference's single-threaded-per-rank execution).
 */
#ifndef RHEO_GPU_H
#define RHEO_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes
 * Reference: FatalError aborts the process (unknown model type
 * constitutiveEquations/constitutiveEqs/constitutiveEq/newConstitutiveEq.C:47-58,
 * unknown solver sparseMatrixSolvers/segregated/sparseSolver.C:48-56, missing
 * dictionary keys via dict.lookup).  Here they become status codes and the
 * same message text is kept for rt_last_error().  Non-convergence is not an
 * error in the reference (maxIter just stops); it is reported as the warning
 * bit RT_W_NOT_CONVERGED on an otherwise successful call. */
enum rt_status {
    RT_OK = 0,
    RT_E_UNKNOWN_TYPE = 1, /* unknown constitutiveEq / solver / preconditioner keyword */
    RT_E_BAD_ARG = 2,      /* NULL or inconsistent argument, missing dictionary key   */
    RT_E_OOM = 3,          /* device allocation failed or HBM budget exceeded          */
    RT_E_HIP = 4,          /* HIP runtime error                                        */
    RT_E_UNSUPPORTED = 5,  /* valid in the reference but outside this path (e.g.
                              transform patches, coupled interfaces)                    */
    RT_E_NO_DEVICE = 6,    /* no gfx950 device visible                                 */
    RT_W_NOT_CONVERGED = 0x100
};

typedef struct rt_ctx rt_ctx;
typedef struct rt_mesh rt_mesh;
typedef struct rt_model rt_model;

/* ---------------------------------------------------------------- context */

/* Number of visible HIP devices (0 when none). */
int rt_device_count(void);

/* Creates a context on `device` with one HIP stream.  hbm_budget bytes caps
 * the library's own device allocations (0 = no cap).  Replaces nothing in the
 * reference (CPU-only); it is the device/stream owner every call needs. */
int rt_ctx_create(int device, size_t hbm_budget, rt_ctx** out);
void rt_ctx_destroy(rt_ctx* ctx);

/* Last error message of the context (empty string when none). */
const char* rt_last_error(const rt_ctx* ctx);

/* Device bytes currently allocated by the context. */
size_t rt_ctx_bytes_in_use(const rt_ctx* ctx);

/* Kernel timing: when enabled, every launch is bracketed by HIP events on the
 * context's stream.  rt_ctx_kernel_stats fills up to `cap` records (name,
 * launches, total ms) and returns the number of distinct kernels. */
int rt_ctx_set_profiling(rt_ctx* ctx, int enable);
typedef struct {
    char name[48];
    int64_t launches;
    double total_ms;
    double bytes; /* algorithmic bytes summed over the launches */
} rt_kernel_stat;
int rt_ctx_kernel_stats(rt_ctx* ctx, rt_kernel_stat* out, int cap);
void rt_ctx_reset_stats(rt_ctx* ctx);

/* HBM calibration: copies `bytes` between two device buffers with a plain
 * streaming kernel `reps` times and returns the mean kernel time in *ms.
 * No reference counterpart; it gives the copy bandwidth of the device and the
 * known-size kernel that hardware byte counters are scaled against. */
int rt_ctx_copy_bench(rt_ctx* ctx, size_t bytes, int reps, double* ms);

/* ------------------------------------------------------------------- mesh
 * The primitiveMesh/lduAddressing the reference's path runs on:
 * lowerAddr(=owner)/upperAddr(=neighbour) per internal face with
 * owner < neighbour in upper-triangular order, boundary faces grouped by
 * patch after the internal ones, patchAddr = owner of each boundary face
 * (sparseMatrixSolvers/segregated/eigenSolver/eigenSolver.C:600-641).
 * Geometry: face area vectors Sf (out of the owner), linear weights w,
 * cell volumes V -- the inputs of fvc::grad(U) Gauss linear
 * (constitutiveEqs/utils/boilerLog.H:1) and of fvm::ddt -- and the cell
 * centres C, which only the limited GaussDefCmpw schemes read
 * (gaussDefCmpwConvectionScheme.C:251-266, mesh.C()). */
enum rt_patch_kind { RT_PATCH_PATCH = 0, RT_PATCH_WALL = 1, RT_PATCH_EMPTY = 2 };

typedef struct {
    int32_t start; /* first face of the patch, counted from the first boundary face */
    int32_t size;
    int32_t kind;  /* rt_patch_kind; empty patches contribute nothing (2D cases) */
} rt_patch;

typedef struct {
    int32_t nCells;
    int32_t nInternalFaces;
    int32_t nBoundaryFaces;
    int32_t nPatches;
    const int32_t* owner;     /* [nInternalFaces + nBoundaryFaces] */
    const int32_t* neighbour; /* [nInternalFaces] */
    const double* Sf;         /* [3*(nInternalFaces + nBoundaryFaces)] */
    const double* weights;    /* [nInternalFaces] */
    const double* V;          /* [nCells] */
    const rt_patch* patches;  /* [nPatches] */
    const double* C;          /* [3*nCells] cell centres; may be NULL (then only the
                                 upwind and none schemes are available) */
    const double* Cb;         /* [3*nBoundaryFaces] boundary face centres Cf; may be
                                 NULL (then divTau's BSD and coupling terms, which
                                 need the patch delta coefficients, are unavailable) */
} rt_mesh_desc;

enum rt_renumber { RT_RENUMBER_AUTO = 0, RT_RENUMBER_NONE = 1, RT_RENUMBER_RCM = 2, RT_RENUMBER_SFC = 3,
                   /* flag: keep a two-colourable mesh in one block (no red-black ILU0) */
                   RT_RENUMBER_NO_COLOUR = 16 };

/* Uploads a mesh.  The device copy is renumbered for locality when `renumber`
 * asks for it: RCM = reverse Cuthill-McKee on the face graph, SFC = Morton
 * (Z-order) curve of the cell centres (needs desc->C), which keeps a cell's
 * neighbours in every direction close in memory.  RT_RENUMBER_AUTO picks RCM
 * when the input ordering has poor locality (on a 256^3 box RCM and SFC give
 * the same mixed-precision step time; SFC's FP64 step is 4% faster).  A mesh
 * whose cell graph is two-colourable (any hexahedral or quadrilateral mesh) is
 * then stored colour by colour, each colour in that order, for the red-black
 * ILU0 preconditioner (multicolourILU0); RT_RENUMBER_NO_COLOUR keeps one
 * block.  Every field crossing the C-ABI stays in the caller's numbering. */
int rt_mesh_create(rt_ctx* ctx, const rt_mesh_desc* desc, int renumber, rt_mesh** out);
void rt_mesh_destroy(rt_mesh* mesh);

/* validComponents of a symmTensor on this mesh ([OF9]
 * fvMesh::validComponents<symmTensor>(), restated in
 * sparseMatrixSolvers/segregated/eigenSolver/eigenSolver.C:69-78): +1 solved,
 * -1 frozen (xz, yz of a 2D case). */
int rt_mesh_valid_components(const rt_mesh* mesh, int32_t out[6]);

/* ------------------------------------------------------- solver controls
 * The fvSolution `solvers { theta { ... } }` entry
 * (e.g. tutorials/rheoFoam/Channel/Oldroyd-BLog/system/fvSolution:29-42),
 * as OpenFOAM dictionary text:
 *   solver PBiCGStab;  (PBiCG accepted: solved with PBiCGStab)
 *   preconditioner DILU;  (DILU, DIC, diagonal, Jacobi, multicolourILU0 ->
 *                          right Jacobi on the device; none)
 *   tolerance 1e-10; relTol 0; maxIter 1000; minIter 0;
 *   precision double;  (or "mixed": FP32 Krylov + FP64 iterative refinement)
 * Unknown solver / preconditioner keywords -> RT_E_UNKNOWN_TYPE listing the
 * valid ones (sparseSolver.C:48-56 style). */
typedef struct {
    double initialResidual;
    double finalResidual;
    int32_t nIterations;  /* Krylov iterations (inner iterations for "mixed") */
    int32_t nRefinements; /* FP64 true-residual sweeps: refinements ("mixed")
                             and restarts after a PBiCGStab breakdown */
    int32_t converged;
    int32_t solved;       /* 0 for a frozen (invalid) component */
} rt_solver_perf;

/* ----------------------------------------------------------------- model
 * constitutiveEq run-time selection:
 *   constitutiveEq::New(name, U, phi, dict)   newConstitutiveEq.C:32-61
 * keyed by dict.lookup("type"): "Oldroyd-BLog" (Oldroyd_BLog.H:132),
 * "GiesekusLog" (GiesekusLog.H:133), "FENE-PLog" (FENE_PLog.H:133),
 * "PTTLog" (PTTLog.H:167, destructionFunctionType linear|exponential),
 * "FENE-CRLog" (FENE_CRLog.H:134), "WhiteMetznerCYLog" (WhiteMetznerCYLog.H:148),
 * "Rolie-PolyLog" (RoliePolyLog.H:146), "XPomPomLog" (XPomPomLog.H:146),
 * "SaramitoLog" (SaramitoLog.H:137; its plasticity factor reads the model's
 * tau before each step, so give tau to rt_model_set_state, as the reference
 * reads it MUST_READ), "multiMode" (multiMode.C:37).
 * `dict` is the text of the `parameters { ... }` sub-dictionary of
 * constant/constitutiveProperties in OpenFOAM syntax, dimension sets
 * allowed (e.g. "type Oldroyd-BLog; rho rho [1 -3 0 0 0 0 0] 1; etaS ...;
 * etaP ...; lambda ...; stabilization coupling;").  Keys read as the
 * reference constructors do (Oldroyd_BLog.C:114-121 etc.); a missing key is
 * RT_E_BAD_ARG with the key named.  The model owns theta, tau, eigVals and
 * eigVecs on the device (Oldroyd_BLog.C:52-113); U and phi are referenced. */
int rt_model_create(rt_ctx* ctx, const char* name, const rt_mesh* mesh, const char* dict,
                    rt_model** out);
void rt_model_destroy(rt_model* model);

/* Type name of the selected model (TypeName of the reference class). */
const char* rt_model_type(const rt_model* model);

/* State: theta [6N] (MUST_READ); eigVals/eigVecs [9N] each (READ_IF_PRESENT:
 * NULL leaves them as they are, which is the identity tensor until they are
 * first set, as the reference's default); tau [6N] may be NULL. */
int rt_model_set_state(rt_model* model, const double* theta, const double* eigVals,
                       const double* eigVecs, const double* tau);

/* Boundary condition of theta on one patch: 0 zeroGradient, 1 fixedValue
 * (values [6*patchSize]), 2 empty, 3 symmetryPlane / symmetry (the patch value
 * (theta_P + R theta_P R^T)/2 with R = I - 2 n n, as OpenFOAM's
 * basicSymmetryFvPatchField evaluates it), 4 wedge (the patch value
 * T theta_P T^T, T the rotation that takes the wedge's centre-plane normal
 * onto the patch's mean face normal, as OpenFOAM's wedgePolyPatch /
 * wedgeFvPatchField compute it; the patch must be planar and its centre
 * plane normal to a coordinate axis, else RT_E_BAD_ARG).  On the transform
 * patches (3 and 4) the solve's coefficients differ from the reference's
 * per-component ones only through the patch flux, which these conditions
 * make zero.  values is read only for fixedValue.  Default: zeroGradient
 * (empty on empty patches). */
enum rt_bc { RT_BC_ZEROGRADIENT = 0, RT_BC_FIXEDVALUE = 1, RT_BC_EMPTY = 2, RT_BC_SYMMETRY = 3,
             RT_BC_WEDGE = 4 };
int rt_model_set_bc(rt_model* model, int patch, int bcType, const double* values);

/* The referenced flow fields (constitutiveEq.H:72-75 hold U and phi by
 * reference): U [3N], Ub [3*nBoundaryFaces] (boundary values of U, used by
 * the Gauss gradient), phi [nInternalFaces + nBoundaryFaces].  Copied to the
 * device; call again whenever the caller's U/phi change. */
int rt_model_set_flow(rt_model* model, const double* U, const double* Ub, const double* phi);

/* Solver controls for the theta equation (text as above). */
int rt_model_set_solver(rt_model* model, const char* controls);

/* Relaxation factor of the theta equation ([OF9] fvMatrix::relax(),
 * Oldroyd_BLog.C:166): <= 0 (default) means no relaxationFactors entry. */
int rt_model_set_relax(rt_model* model, double factor);

/* The fvSchemes entry div(phi,theta) (e.g. tutorials/rheoFoam/Channel/
 * Oldroyd-BLog/system/fvSchemes:36 "GaussDefCmpw cubista"): the text after
 * the key, "GaussDefCmpw <limiter>" or the limiter alone, one of upwind
 * (default), cubista, minmod, smart, waceb, superbee, none
 * (gaussDefCmpwConvectionScheme.H:68 scheme_, limiters.H:38-100).  The
 * limited schemes add the explicit deferred correction of
 * gaussDefCmpwConvectionScheme.C:120-169 and need the mesh's cell centres;
 * none drops the convection term (:90).  Unknown names -> RT_E_UNKNOWN_TYPE
 * with the reference's list (limiters.H:88-97).  A leading "bounded"
 * (e.g. tutorials/rheoFilmFoam/UCM/system/fvSchemes:35
 * "bounded GaussDefCmpw cubista") applies OpenFOAM's bounded
 * convection wrapper: the matrix diagonal loses the cell's net outflux
 * (fvm::Sp(fvc::surfaceIntegrate(phi), theta)).  "Gauss upwind" is accepted
 * as the upwind scheme (the same matrix). */
int rt_model_set_scheme(rt_model* model, const char* divScheme);

/* multiMode (constitutiveEqs/multiMode/multiMode.C): a model created from
 * "type multiMode; models ( M1 { type ...; ... } M2 { ... } );" owns one
 * model per entry, named name + keyword (:43-83).  Its setters (bc, flow,
 * solver, relax, scheme) apply to every mode, correct() runs the modes in
 * turn (:250-260), RT_TAU and rt_model_div_tau / rt_model_tau_total are the
 * sums over the modes (:141-151, :219-227; etaS summed :80), and
 * rt_model_solver_perf reports the worst mode per component.  The state of
 * each mode (theta MUST_READ etc.) is set and read through its handle.
 * rt_model_mode_count: the number of modes (0 for a single model).
 * rt_model_mode: mode i, owned by the multiMode model (valid until it is
 * destroyed; never pass it to rt_model_destroy). */
int rt_model_mode_count(const rt_model* model);
rt_model* rt_model_mode(rt_model* model, int i);

/* BMPLog's fluidity equation (BMPLog.C:98-111 Phi<name> MUST_READ, 151-163:
 * ddt + div(phi,Phi) == -Sp(1/lambda) + Phi0/lambda + k (PhiInf - Phi)(tau &&
 * symm(L))), solved before theta in every rt_model_correct, as a handle of its
 * own for its fvSolution/fvSchemes entries (`Phi`, `div(phi,Phi)`): it takes
 * rt_model_set_state (theta = Phi [N]; eigen state and tau ignored),
 * rt_model_set_bc (fixedValue values [n], one per face), rt_model_set_solver,
 * rt_model_set_scheme, rt_model_set_relax, rt_model_get(RT_THETA) -> Phi [N]
 * and rt_model_solver_perf (component 0).  The flow comes from the BMP model.
 * NULL for other models; owned by the BMP model. */
rt_model* rt_model_fluidity(rt_model* model);

/* ------------------------------------------- consumers of tau (momentum eq.)
 * constitutiveEq::divTau(U) (constitutiveEquations/constitutiveEqs/
 * constitutiveEq/constitutiveEq.C:72-132, the segregated branch) returns
 *   explicit(U) + fvm::laplacian(nu, U)
 * with the model's `stabilization` option:
 *   none      explicit = fvc::div(tau/rho, "div(tau)")                      nu = etaS/rho
 *   BSD       explicit = fvc::div(tau/rho) - fvc::laplacian(etaP/rho, U)    nu = (etaP+etaS)/rho
 *   coupling  explicit = fvc::div(tau/rho) - fvc::div((etaP/rho) grad(U))   nu = (etaP+etaS)/rho
 * (div(tau) and div(grad(U)) Gauss linear, grad(U) Gauss linear from the flow
 * last set).  rt_model_div_tau writes the explicit field from the model's
 * current tau on the device, out [3N] (to be added to the momentum source as
 * the caller's fvc terms are); rt_model_div_tau_coeff gives nu, the
 * coefficient of the implicit fvm::laplacian the caller's momentum matrix
 * holds.  tauB [6*nBoundaryFaces]: tau's patch values; NULL takes the
 * cell value (zeroGradient), and on the symmetryPlane/symmetry and wedge
 * theta patches the transform of the cell value that theta takes there (an
 * OpenFOAM constraint patch gives every field its type).  U's patch values (rt_model_set_flow's Ub) are read as a
 * fixedValue patch's (snGrad = (Ub - U_P) |n (n & (Cf - C_P))|^-1); a
 * zeroGradient velocity patch gives Ub = U_P, hence a zero snGrad, as the
 * reference's.  BSD and coupling need the mesh's cell and boundary face
 * centres.  multiMode sums the modes' terms and coefficients
 * (multiMode.C:141-151). */
int rt_model_div_tau(rt_model* model, const double* tauB, double* out);
int rt_model_div_tau_coeff(const rt_model* model, double* nu);
/* The fvSchemes entry laplacian(eta,U) that BSD's explicit laplacian uses
 * (e.g. tutorials/rheoFoam/Channel/Oldroyd-BLog/system/fvSchemes:44
 * "Gauss linear corrected"): "Gauss linear corrected" (default),
 * "Gauss linear orthogonal" or "Gauss linear uncorrected". */
int rt_model_set_laplacian_scheme(rt_model* model, const char* laplacianScheme);
/* constitutiveEq::tauTotal() (constitutiveEq.C:534-547):
 * tau + etaS (symm(L + L^T) - 2/3 tr(L) I), L = fvc::grad(U) Gauss linear
 * from the flow last set.  out [6N]. */
int rt_model_tau_total(rt_model* model, double* out);

/* constitutiveEq::correct(alpha, gradU)  (constitutiveEq.H:346-350,
 * Oldroyd_BLog.C:127-179): decompose grad(U)^T, assemble theta's
 * ddt + upwind div == source, solve the valid components, eigen-decompose
 * theta, rebuild tau.  gradU [9N] may be NULL (Gauss linear from U); alpha
 * is ignored, as by every single-phase *Log model.  deltaT is
 * runTime.deltaT(). */
int rt_model_correct(rt_model* model, double deltaT, const double* gradU);
/* correct() as an inner iteration of a time step: rheoTestFoam calls
 * correct() nInIter times per step (solvers/rheoTestFoam/rheoTestFoam.C:64-68),
 * each solve's ddt using theta.oldTime() of the step while the explicit terms
 * use the latest iterate.  iteration 0 starts the step, 1, 2, ... repeat it;
 * a negative iteration is rt_model_correct. */
int rt_model_correct_iter(rt_model* model, double deltaT, const double* gradU, int iteration);

enum rt_field { RT_THETA = 0, RT_TAU = 1, RT_EIGVALS = 2, RT_EIGVECS = 3 };
/* Copies a model field to the host: theta/tau [6N], eigVals/eigVecs [9N]. */
int rt_model_get(const rt_model* model, int field, double* out);

/* Solver performance of the last correct(), one record per symmTensor
 * component (fvMatrix::solveSegregated's SolverPerformance vector). */
int rt_model_solver_perf(const rt_model* model, rt_solver_perf perf[6]);

/* ------------------------------------------------ segregated solve surface
 * sparseSolver<symmTensor>::solve(fvMatrix<symmTensor>&)
 * (sparseMatrixSolvers/segregated/sparseSolver.H:190-193) and
 * fvMatrix<symmTensor>::solve() behind it: one call solves every valid
 * component of A psi = source with A's LDU coefficients shared by the
 * components.  Boundary coefficients as the fvMatrix holds them:
 * internalCoeffs are added to the diagonal and boundaryCoeffs to the source
 * (sparseSolver.C:72-127) on non-empty patches.  psi [6N] is updated in
 * place. */
typedef struct {
    const double* diag;           /* [nCells] */
    const double* lower;          /* [nInternalFaces] */
    const double* upper;          /* [nInternalFaces] */
    const double* source;         /* [6*nCells] */
    const double* internalCoeffs; /* [6*nBoundaryFaces], may be NULL */
    const double* boundaryCoeffs; /* [6*nBoundaryFaces], may be NULL */
} rt_ldu_symm;

int rt_symm_solve(rt_ctx* ctx, const rt_mesh* mesh, const rt_ldu_symm* sys, double* psi,
                  const char* controls, rt_solver_perf perf[6]);

/* rt_symm_solve without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and kernels as rt_symm_solve, but it calls no hipMalloc,
 * hipHostMalloc or hipFree: the device copies of the system, the ELL values,
 * the Krylov vectors, the reduction partials, the solver state and the
 * staging live in one device array of RT_NOALLOC_SCRATCH_BYTES that the
 * library holds on each device from the first rt_ctx_create there until the
 * process ends, and the solver status comes back through pageable memory
 * (the call waits on the context's stream).  A mesh whose solve needs more
 * (about 180k cells in FP64, 280k with "precision mixed") is refused with
 * RT_E_BAD_ARG.  Calls that overlap on one device share the array and race;
 * every call clears the state it needs.  Kernel profiling
 * (rt_ctx_set_profiling) creates timing events and is meant to be off. */
#define RT_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int rt_symm_solve_noalloc(rt_ctx* ctx, const rt_mesh* mesh, const rt_ldu_symm* sys, double* psi,
                          const char* controls, rt_solver_perf perf[6]);

/* ------------------------------------------------------------------ serving
 * A device-resident request loop: independent stress steps (different
 * meshes, models, Weissenberg numbers) are admitted against an HBM budget
 * and advanced together.  No reference counterpart (the reference runs one
 * case per process); it serves callers that batch many small cases. */
typedef struct rt_server rt_server;

typedef struct {
    const rt_mesh* mesh;  /* an uploaded mesh (rt_mesh_create); requests may share one */
    const char* dict;     /* constitutiveProperties parameters text */
    const char* controls; /* fvSolution text, NULL for the server default */
    const double* theta;  /* [6N] */
    const double* eigVals; /* [9N] each; both NULL: computed from theta on the device */
    const double* eigVecs; /* (calcEig, constitutiveEq.C:360-428, as the previous step left them) */
    const double* U;      /* [3N] */
    const double* Ub;     /* [3*nBoundaryFaces] */
    const double* phi;    /* [nInternalFaces + nBoundaryFaces] */
    double deltaT;
    double* thetaOut;     /* [6N] result, written when the request finishes; may be NULL */
    double* tauOut;       /* [6N] result, may be NULL */
    const double* gradU;  /* [9N] prescribed grad(U) (the gradU argument of correct(),
                             constitutiveEq.H:346-350); NULL: Gauss linear from U */
} rt_request;

int rt_server_create(rt_ctx* ctx, size_t hbm_budget, int max_inflight, rt_server** out);
/* Device bytes the server reserves against its HBM budget for a request on
 * `mesh` (model state, flow, matrix, Krylov workspace, staging, the prescribed
 * gradient when gradU != 0, the result buffers): an upper bound of what the
 * request allocates. */
size_t rt_request_bytes_estimate(const rt_mesh* mesh, int mixed, int gradU);
void rt_server_destroy(rt_server* server);
/* Queues the request; it is admitted when a slot and the HBM budget allow.
 * The strings are copied; the host arrays are NOT: they are read when the
 * request is admitted and written (thetaOut, tauOut) when it finishes, so they
 * must stay valid, and the inputs unchanged, until rt_server_poll reports
 * the ticket finished or failed.  *ticket identifies the request. */
int rt_server_submit(rt_server* server, const rt_request* req, int64_t* ticket);
/* Advances every in-flight request by one batched round of device work and
 * admits queued ones as memory frees up.  Returns the number of requests
 * still in flight or queued. */
int rt_server_progress(rt_server* server);
/* 1 when the request is finished (then the request's thetaOut/tauOut hold the
 * results, theta/tau [6N] get a copy of them when non-NULL, and *latency_ms is
 * submit-to-finish); 0 when still pending;
 * -(status code) when the request failed (message in rt_last_error of the
 * server's context). */
int rt_server_poll(rt_server* server, int64_t ticket, double* theta, double* tau,
                   rt_solver_perf perf[6], double* latency_ms);
/* Writes up to max tickets that finished or failed since the last call,
 * oldest first, and returns how many; a caller then polls only those instead
 * of every pending ticket. */
int rt_server_finished(rt_server* server, int64_t* tickets, int max);
/* Batching of the in-flight requests' Krylov solves (on by default): one
 * launch per kernel runs the rounds of every request of the same precision and
 * preconditioner form; off, each request launches its own rounds.  Results are
 * the same bits either way.  Changed only while no request is in flight. */
int rt_server_set_batching(rt_server* server, int enable);
/* Batched Krylov rounds launched so far (each covers every request then in its batch). */
int64_t rt_server_batched_rounds(const rt_server* server);

/* Page-locked host memory (hipHostMalloc on the context's device) for a
 * serving caller's request and result arrays.  Transfers from and to it run
 * asynchronously on the device's DMA engines, so the server's uploads at
 * admission and its result downloads overlap the kernels of other requests;
 * from pageable arrays each transfer holds the calling thread until it is
 * done.  No counterpart in the reference (an OpenFOAM rank's fields live in
 * pageable memory and never cross a bus).  rt_host_free(NULL) does nothing. */
int rt_host_alloc(rt_ctx* ctx, size_t bytes, void** out);
void rt_host_free(void* p);

#ifdef __cplusplus
}
#endif
#endif /* RHEO_GPU_H */
