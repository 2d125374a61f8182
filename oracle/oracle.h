/*
 * oracle.h -- CPU FP64 restatement of rheoTool's log-conformation stress step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product path (librheo_gpu.so) never calls into it.
 *
 * Each function follows the reference text cited beside it (paths relative to
 * the rheoTool tree, of90/src/libs/...).  Where the reference calls
 * OpenFOAM-9 or Eigen (third-party, not in the reference tree) the behaviour
 * restated is marked [OF9] / [Eigen 3.2.9].
 *
 * Layouts (host, AoS, as OpenFOAM stores them):
 *   symmTensor: 6 doubles xx xy xz yy yz zz
 *   tensor:     9 doubles row-major xx xy xz yx yy yz zx zy zz
 *   vector:     3 doubles
 */
#ifndef RHEO_ORACLE_H
#define RHEO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LDU-addressed mesh (primitiveMesh conventions, see rtcase.h) */
typedef struct {
    int32_t nCells, nInternal, nBoundary, nPatches;
    const int32_t* owner;      /* [nInternal + nBoundary] */
    const int32_t* neighbour;  /* [nInternal] */
    const double* Sf;          /* [3*(nInternal + nBoundary)] */
    const double* w;           /* [nInternal] */
    const double* V;           /* [nCells] */
    const int32_t* patchStart; /* [nPatches], boundary-face index */
    const int32_t* patchSize;
    const int32_t* patchEmpty; /* 1: empty patch (no faces in the fvPatch sense) */
} orc_mesh;

enum { ORC_OLDROYDB = 0, ORC_GIESEKUS = 1, ORC_FENEP = 2, ORC_PTT_LINEAR = 3, ORC_PTT_EXP = 4,
       ORC_FENECR = 5, ORC_WMCY = 6, ORC_ROLIEPOLY = 7, ORC_XPP = 8, ORC_SARAMITO = 9,
       /* BMPLog (BMPLog.C:142-202): the theta equation, and its scalar fluidity
        * equation, held in component xx of a [6n] field (the others unused) */
       ORC_BMP = 10, ORC_FLUIDITY = 11,
       /* Oldroyd-BRootk (Oldroyd_BRootk.C:129-206): theta = C^(1/k) */
       ORC_ROOTK = 12,
       /* PTTLog with destructionFunctionType generalized (PTTLog.C:135-152,202-230):
        * the Mittag-Leffler function E_{alpha,beta} of the trace */
       ORC_PTT_GEN = 13 };

typedef struct {
    int32_t model;
    double rho, etaS, etaP, lambda;
    double alpha;   /* Giesekus mobility */
    double L2;      /* FENE-P extensibility */
    double epsilon; /* PTT */
    double zeta;    /* PTT */
    int32_t ofInvQuirk; /* 1: apply [OF9] inv(tensorField) first-cell component shift */
    /* Rolie-Poly (RoliePolyLog.C:114-121) */
    double lambdaR, lambdaD, beta, delta, chiMax;
    /* eXtended Pom-Pom (XPomPomLog.C:115-122); alpha above */
    double lambdaS, lambdaB, q, nXPP;
    /* White-Metzner Carreau-Yasuda (WhiteMetznerCYLog.C:115-124): m n K L a b */
    double wmM, wmN, wmK, wmL, wmA, wmB;
    /* Saramito (SaramitoLog.C:117-170): yield stress tau0, power n, consistency
     * k (etaP when n = 1), the valid dimensions, PTT function 0 none / 1
     * linear / 2 exponential (n = 1 only); epsilon, zeta above */
    double tau0, sarN, sarK, dimX, dimY, dimZ;
    int32_t sarPTT;
    /* BMP (BMPLog.C:113-119): modulus, rate constant, fluidity at rest and at
     * infinite shear; lambda above */
    double G0, kBMP, Phi0, PhiInf;
    /* Oldroyd-BRootk: the root k */
    double rootK;
    /* PTT generalized: alpha and beta of E_{alpha,beta}, the series' relative
     * tolerance rTolMittagLeffler and maxIterMittagLeffler (PTTLog.C:135-136) */
    double mlAlpha, mlBeta, mlRtol;
    int32_t mlMaxIter;
} orc_params;

/* ORC_BC_COUPLED: a processor patch of a decomposed run.  The upwind
 * convection puts phi_b*pos0(phi_b) on the diagonal (internalCoeffs) and
 * nothing in the source; -phi_b*(1 - pos0(phi_b)) multiplies the
 * neighbouring domain's value inside Amul (interfaceBouCoeffs). */
/* ORC_BC_SYMMETRY: symmetryPlane / symmetry ([OF9] basicSymmetryFvPatchField,
 * a transformFvPatchField): the patch value is (theta_P + R theta_P R^T)/2,
 * R = I - 2 n n; the upwind convection's internalCoeffs are phi_b times
 * 1 - |n_i||n_j| per component (i, j) and its boundaryCoeffs -phi_b times
 * (theta_b - (1 - |n_i||n_j|) theta_P) (gaussDefCmpwConvectionScheme.C:103-117
 * through the patch field's valueInternalCoeffs/valueBoundaryCoeffs). */
/* ORC_BC_WEDGE: wedge ([OF9] wedgeFvPatchField, a transformFvPatchField on a
 * wedgePolyPatch): the patch value is faceT theta_P faceT^T
 * (orc_wedge_face_t); internalCoeffs phi_b (1 - d_i d_j) per component (i, j)
 * with d = diag(I - cellT)/2, cellT = faceT & faceT, and boundaryCoeffs
 * -phi_b (theta_b - (1 - d_i d_j) theta_P). */
enum { ORC_BC_ZEROGRADIENT = 0, ORC_BC_FIXEDVALUE = 1, ORC_BC_EMPTY = 2, ORC_BC_COUPLED = 3,
       ORC_BC_SYMMETRY = 4, ORC_BC_WEDGE = 5 };

/* [OF9] wedgePolyPatch::calcGeometry: faceT (row-major) of patch p, the
 * rotation taking the centre-plane normal (the coordinate axis nearest the
 * mean unit face normal n) onto n.  Returns 0, or 1 when the centre plane is
 * not normal to a coordinate axis, 2 when n is a coordinate axis (no angle),
 * 3 when the patch has no faces. */
int orc_wedge_face_t(const orc_mesh* m, int32_t patch, double faceT[9]);

enum { ORC_PBICGSTAB = 0, ORC_PBICG = 1 };
enum { ORC_PC_NONE = 0, ORC_PC_DIAGONAL = 1, ORC_PC_DILU = 2 };

typedef struct {
    int32_t solver, precond;
    double tolerance, relTol;
    int32_t maxIter, minIter;
} orc_controls;

typedef struct {
    double initialResidual, finalResidual;
    int32_t nIterations, converged, solved;
} orc_perf;

/* [OF9] fvc::grad(U), Gauss linear (boilerLog.H:1).  Ub per boundary face. */
void orc_grad_gauss(const orc_mesh* m, const double* U, const double* Ub, double* gradU);

/* Per-cell right-hand side S = symm(Omega.theta - theta.Omega + 2B + X)
 * of the model, from the old eigen state (boilerLog.H, constitutiveEq.C:323-358,
 * Oldroyd_BLog.C:141-164, GiesekusLog.C:142-160, FENE_PLog.C:142-166,
 * PTTLog.C:184-196,241-253, FENE_CRLog.C:138-166, WhiteMetznerCYLog.C:150-190,
 * RoliePolyLog.C:140-185, XPomPomLog.C:140-185).  fOut (may be NULL) receives
 * the per-cell factor tau needs from the old state: f of FENE-P and FENE-CR,
 * etaP/lambda of White-Metzner. */
void orc_rhs(const orc_params* p, int32_t n, const double* gradU, const double* theta,
             const double* eigVals, const double* eigVecs, double* S, double* fOut);

/* orc_rhs with the model's current tau [6n] (SaramitoLog.C:181-212 reads it
 * for the plasticity factor, the BMP fluidity equation for its source; NULL
 * reads as zero) and aux [6n] (BMP: the step's fluidity in component xx).
 * fOut receives Saramito's factor fac. */
void orc_rhs_tau(const orc_params* p, int32_t n, const double* gradU, const double* theta,
                 const double* eigVals, const double* eigVecs, const double* tauOld, const double* aux,
                 double* S, double* fOut);

/* Euler ddt + upwind GaussDefCmpw div + (== S) assembly
 * (GDC/gaussDefCmpwConvectionScheme.C:70-118, [OF9] EulerDdtScheme, fvMatrix
 * operator==).  bcType/bcValue per patch/boundary face for theta.
 * Outputs: diag[N], lower/upper[nInternal], source[6N],
 * internalCoeffs/boundaryCoeffs[6*nBoundary]. */
void orc_assemble(const orc_mesh* m, const double* phi, const double* theta, const double* S,
                  double deltaT, const int32_t* bcType, const double* bcValue, double* diag,
                  double* lower, double* upper, double* source, double* internalCoeffs,
                  double* boundaryCoeffs);

/* GaussDefCmpw limited schemes (gaussDefCmpwConvectionScheme.H:68 scheme_,
 * limiters.H:38-100). */
enum { ORC_DC_UPWIND = 0, ORC_DC_CUBISTA = 1, ORC_DC_MINMOD = 2, ORC_DC_SMART = 3, ORC_DC_WACEB = 4,
       ORC_DC_SUPERBEE = 5, ORC_DC_NONE = 6 };
/* Or-ed into a scheme: OpenFOAM's `bounded` convection wrapper, which
 * subtracts fvm::Sp(fvc::surfaceIntegrate(phi), theta) from the scheme's
 * matrix ([OF9] boundedConvectionScheme.C fvmDiv). */
enum { ORC_DC_BOUNDED = 16 };

/* [OF9] fvc::grad of each symmTensor component, Gauss linear, as phifDefC
 * takes it (gaussDefCmpwConvectionScheme.C:243-250): grad[18*N], per cell the
 * 6 components (xx xy xz yy yz zz) x 3 derivatives (d/dx d/dy d/dz).
 * thetaB[6*nBoundary]: patch values of theta on boundary faces (empty
 * patches are skipped). */
void orc_grad_gauss_symm(const orc_mesh* m, const double* theta, const double* thetaB, double* grad);

/* Deferred-correction face values of phifDefC with onlyDCphi = true
 * (gaussDefCmpwConvectionScheme.C:238-283): dc[6*nInternal] = limited HRS
 * face value minus the upwind value, per component.  C[3N] cell centres. */
void orc_defc_faces(const orc_mesh* m, int32_t scheme, const double* phi, const double* theta,
                    const double* grad, const double* C, double* dc);

/* The explicit correction source of fvmDiv (gaussDefCmpwConvectionScheme.C:120-169):
 * souT[6N] = sum over faces of +-dc*phi (owner +, neighbour -); the matrix
 * source is -souT.  No coupled patches (they are out of scope). */
void orc_defc_sout(const orc_mesh* m, const double* phi, const double* dc, double* souT);

/* fvc::div(tau/rho), Gauss linear: the explicit stress term that
 * constitutiveEq::divTau hands to the momentum equation
 * (constitutiveEq.C:88-92 "fvc::div(tau()/rho(), "div(tau)")").  tauB
 * [6*nBoundary] patch values of tau (NULL: zeroGradient, the cell value).
 * out [3N]. */
void orc_div_tau(const orc_mesh* m, const double* tau, const double* tauB, double rho, double* out);

/* The explicit part of constitutiveEq::divTau(U) for a stabilization option
 * (constitutiveEq.C:72-132, segregated branch):
 *   none      fvc::div(tau/rho)
 *   BSD       fvc::div(tau/rho) - fvc::laplacian(etaP/rho, U)   "laplacian(eta,U)"
 *   coupling  fvc::div(tau/rho) - fvc::div((etaP/rho)*fvc::grad(U))  "div(grad(U))"
 * with Gauss linear div/grad and the laplacian's snGrad scheme lapScheme
 * (ORC_SNGRAD_*).  [OF9] semantics restated: linear interpolation
 * w*(a_o - a_n) + a_n; deltaCoeffs 1/|C_n - C_o| (orthogonal), nonOrth
 * 1/max(n.d, 0.05|d|) with the correction (n - d*dc) & interpolate(grad U)
 * (corrected); boundary faces: delta = n (n & (Cf - C_P)), patch snGrad
 * (Ub - U_P)/|delta| (fixedValue; zeroGradient patches give Ub = U_P, hence 0);
 * grad(U) on a patch face: G_P + n*(snGrad - n & G_P)
 * (gaussGrad::correctBoundaryConditions).  C [3N] cell centres, Cb [3NB]
 * boundary face centres, U [3N], Ub [3NB], tau [6N], tauB [6NB] or NULL
 * (zeroGradient).  out [3N].  The implicit fvm::laplacian(nu, U) part is the
 * momentum matrix's own: nu = etaS/rho (none), (etaP + etaS)/rho (BSD,
 * coupling). */
enum { ORC_STAB_NONE = 0, ORC_STAB_BSD = 1, ORC_STAB_COUPLING = 2 };
enum { ORC_SNGRAD_CORRECTED = 0, ORC_SNGRAD_ORTHOGONAL = 1, ORC_SNGRAD_UNCORRECTED = 2 };
void orc_div_tau_stab(const orc_mesh* m, const double* C, const double* Cb, const double* U, const double* Ub,
                      const double* tau, const double* tauB, double rho, double etaP, int32_t stab,
                      int32_t lapScheme, double* out);

/* constitutiveEq::tauTotal() (constitutiveEq.C:534-547):
 * tau + etaS*(symm(L + L^T) - (2/3) tr(L) I), L = fvc::grad(U) Gauss linear.
 * out [6N]. */
void orc_tau_total(const orc_mesh* m, const double* U, const double* Ub, const double* tau, double etaS,
                   double* out);

/* [OF9] fvMatrix::relax(1) on the assembled system (Oldroyd_BLog.C:166). */
void orc_relax(const orc_mesh* m, const int32_t* bcType, double alpha, const double* psi,
               double* diag, const double* lower, const double* upper, double* source,
               const double* internalCoeffs);

/* [OF9] fvMatrix<symmTensor>::solveSegregated (restated in
 * SMS/segregated/eigenSolver/eigenSolver.C:409-563): per valid component,
 * add boundary diag/source and run the selected lduMatrix solver.
 * validComponents[6] is +1 / -1.  psi[6N] updated in place. */
void orc_solve_segregated(const orc_mesh* m, const double* diag, const double* lower,
                          const double* upper, const double* source,
                          const double* internalCoeffs, const double* boundaryCoeffs,
                          const int32_t* bcType, const int32_t validComponents[6],
                          const orc_controls* c, double* psi, orc_perf perf[6]);

/* One scalar lduMatrix solve (diag already includes boundary coefficients). */
orc_perf orc_ldu_solve(int32_t nCells, int32_t nFaces, const int32_t* l, const int32_t* u,
                       const double* diag, const double* lower, const double* upper,
                       const double* b, const orc_controls* c, double* x);

/* [SMS/segregated/sparseSolver.C:129-180] normalised residual of one component */
double orc_foam_residual(int32_t nCells, int32_t nFaces, const int32_t* l, const int32_t* u,
                         const double* diag, const double* lower, const double* upper,
                         const double* b, const double* x);

/* calcEig (constitutiveEq.C:360-428).  method 0: tridiagonal + implicit QR
 * with Eigen's scaling/deflation/ascending sort [Eigen 3.2.9
 * SelfAdjointEigenSolver]; method 1: the cyclic Jacobi of utils/jacobi.H
 * (result re-sorted ascending). */
void orc_calc_eig(int32_t n, const double* theta, double* eigVals, double* eigVecs, int32_t method);

/* tau from the new eigen state (Oldroyd_BLog.C:175, GiesekusLog.C:172,
 * FENE_PLog.C:178 with f from the old state, PTTLog.C:264). */
void orc_tau(const orc_params* p, int32_t n, const double* eigVals, const double* eigVecs,
             const double* fOld, double* tau);

/* One full constitutiveEq::correct() of a *Log model (Oldroyd_BLog.C:127-179).
 * gradU may be NULL (then Gauss linear from U/Ub).  theta/eigVals/eigVecs
 * are updated in place; tau holds the model's current tau on entry (read by
 * Saramito's plasticity factor) and the new tau on return.  relaxFactor <= 0
 * disables relax(). */
void orc_correct(const orc_mesh* m, const orc_params* p, const double* U, const double* Ub,
                 const double* phi, const double* gradU, double deltaT, const int32_t* bcType,
                 const double* bcValue, const int32_t validComponents[6],
                 const orc_controls* c, double relaxFactor, double* theta, double* eigVals,
                 double* eigVecs, double* tau, orc_perf perf[6]);

/* orc_correct with the div(phi,theta) scheme of fvSchemes: scheme is an
 * ORC_DC_* value; C[3N] cell centres (needed unless the scheme is upwind or
 * none).  ORC_DC_NONE drops the convection term (fvmDiv :90). */
void orc_correct_scheme(const orc_mesh* m, const orc_params* p, const double* U, const double* Ub,
                        const double* phi, const double* gradU, double deltaT, const int32_t* bcType,
                        const double* bcValue, const int32_t validComponents[6],
                        const orc_controls* c, double relaxFactor, int32_t scheme, const double* C,
                        double* theta, double* eigVals, double* eigVecs, double* tau, orc_perf perf[6]);

/* An inner iteration of a time step (rheoTestFoam.C:64-68 calls correct()
 * nInIter times per step): the ddt source uses thetaOld, theta.oldTime() of
 * the step, while the explicit terms use the current theta and eigen state.
 * thetaOld NULL: theta itself (a new step). */
void orc_correct_iter(const orc_mesh* m, const orc_params* p, const double* U, const double* Ub,
                      const double* phi, const double* gradU, double deltaT, const int32_t* bcType,
                      const double* bcValue, const int32_t validComponents[6],
                      const orc_controls* c, double relaxFactor, int32_t scheme, const double* C,
                      const double* thetaOld, double* theta, double* eigVals, double* eigVecs, double* tau,
                      orc_perf perf[6]);

/* orc_correct_iter with aux [6N] (BMP: the step's fluidity in component xx).
 * ORC_FLUIDITY solves BMP's fluidity equation held in component xx of theta
 * (pass validComponents {1,-1,-1,-1,-1,-1}): the implicit 1/lambda joins the
 * diagonal, tau is the BMP model's tau (read only), and eigVals/eigVecs are
 * neither read nor written. */
void orc_correct_aux(const orc_mesh* m, const orc_params* p, const double* U, const double* Ub,
                     const double* phi, const double* gradU, double deltaT, const int32_t* bcType,
                     const double* bcValue, const int32_t validComponents[6],
                     const orc_controls* c, double relaxFactor, int32_t scheme, const double* C,
                     const double* thetaOld, const double* aux, double* theta, double* eigVals,
                     double* eigVecs, double* tau, orc_perf perf[6]);

/* ---- The stress step as a decomposed (parallel) OpenFOAM run executes it ----
 * decomposePar splits the cells into nDomains domains; each domain holds its
 * cells in their original relative order, its internal faces, its share of
 * every boundary patch and one processor patch with the faces it shares with
 * other domains.  A step runs one thread per domain (one MPI rank each in
 * the reference) through the same phases as orc_correct, with:
 *   - the processor faces' U interpolated linearly from both sides
 *     (coupled patch, [OF9] processorFvPatchField);
 *   - their upwind coefficients split into the diagonal and the interface
 *     coefficient (ORC_BC_COUPLED), applied inside Amul and sumA
 *     ([OF9] lduMatrix::Amul, updateMatrixInterfaces, sumA);
 *   - DILU on each domain's own matrix, the interfaces left out
 *     ([OF9] DILUPreconditioner works on the local lduAddressing only);
 *   - residuals and inner products summed over domains (gSum).
 * With one domain the step is the serial orc_correct operation for
 * operation.  Only Gauss upwind (optionally bounded), PBiCGStab, no relax()
 * and Gauss-linear gradU are supported (the configuration timed as the CPU
 * baseline); the step returns -1 for anything else and for a thread team
 * smaller than nDomains.
 *
 * domainOf[nCells]: the domain of each cell, 0 <= d < nDomains.  Fields are
 * in the original numbering in set_fields/get; the step works on the
 * domain-major copies the object holds. */
typedef struct orc_decomp orc_decomp;
orc_decomp* orc_decomp_create(const orc_mesh* m, int32_t nDomains, const int32_t* domainOf);
void orc_decomp_destroy(orc_decomp* d);
/* U[3N], Ub[3 nBoundary], phi[nInternal + nBoundary], theta[6N],
 * eigVals/eigVecs[9N], tau[6N] (read by Saramito), bcValue[6 nBoundary] */
void orc_decomp_set_fields(orc_decomp* d, const double* U, const double* Ub, const double* phi,
                           const double* theta, const double* eigVals, const double* eigVecs,
                           const double* tau, const double* bcValue);
int32_t orc_decomp_correct(orc_decomp* d, const orc_params* p, double deltaT, const int32_t* bcType,
                           const int32_t validComponents[6], const orc_controls* c, int32_t scheme,
                           orc_perf perf[6]);
void orc_decomp_get(const orc_decomp* d, double* theta, double* eigVals, double* eigVecs, double* tau);

#ifdef __cplusplus
}
#endif
#endif
