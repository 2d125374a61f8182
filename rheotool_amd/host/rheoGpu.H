/*---------------------------------------------------------------------------*\
  rheoGpu.H -- C++ host mirror of rheoTool's stress-step interfaces, running
  on the MI355X through the C-ABI of include/rheo_gpu.h.

  The classes keep the reference's names, argument meaning and error
  behaviour so that a rheoFoam-style caller reads the same:

    constitutiveEq::New(name, U, phi, dict)   newConstitutiveEq.C:32-61
    constitutiveEq::correct(alpha, gradU)     constitutiveEq.H:346-350
    constitutiveEq::tau(), rho(), isGNF()     constitutiveEq.H:314,321,343
    constitutiveModel(U, phi)                 constitutiveModel.C:41-59,99-106
    sparseSolver<Type>::New(T, mesh, fvSolution), solve(fvMatrix<Type>&)
                                              sparseSolver.H:134-212
    fvMatrix<symmTensor>::solve()             ([OF9], called at Oldroyd_BLog.C:167)

  (paths relative to rheoTool's of90/src/libs/).  The field and mesh types
  are light host containers with OpenFOAM's layouts; the heavy state lives on
  the device.  Errors that abort the reference (FatalError) throw
  rheoGpu::FatalError with the reference's message text.
\*---------------------------------------------------------------------------*/
#ifndef rheoGpu_H
#define rheoGpu_H

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rheo_gpu.h"

namespace rheoGpu {

typedef int32_t label;
typedef double scalar;
typedef std::string word;

//- FatalError / FatalIOError of the reference as an exception
class FatalError : public std::runtime_error {
public:
    int code;
    FatalError(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

//- OpenFOAM dictionary text (constitutiveProperties, fvSolution)
class dictionary {
    std::string text_;
public:
    dictionary() = default;
    explicit dictionary(std::string text) : text_(std::move(text)) {}
    static dictionary fromFile(const std::string& path);
    const std::string& text() const { return text_; }
    //- text of a named sub-dictionary ("" when absent)
    dictionary subDict(const word& name) const;
};

//- device context shared by the meshes and models of one process
class gpuContext {
    rt_ctx* ctx_ = nullptr;
    gpuContext(int device, size_t budget);
public:
    ~gpuContext();
    static gpuContext& get(int device = 0, size_t budget = 0);
    rt_ctx* ptr() const { return ctx_; }
    void check(int rc, const char* what) const;
};

//- runTime subset used by the stress step
struct Time {
    scalar value = 0, deltaT = 0;
    scalar deltaTValue() const { return deltaT; }
};

struct patchInfo {
    word name;
    label start, size;  // start counted from the first boundary face
    int kind;           // rt_patch_kind
};

//- primitiveMesh + lduAddressing + fvSolution of one region
class fvMesh {
public:
    label nCells, nInternalFaces, nBoundaryFaces;
    std::vector<label> owner, neighbour;
    std::vector<scalar> Sf, weights, V;
    std::vector<patchInfo> boundary;
    std::vector<scalar> C;  // cell centres [3*nCells]; set before first use when the
                            // div(phi,theta) scheme is a limited GaussDefCmpw one
    std::vector<scalar> Cb; // boundary face centres [3*nBoundaryFaces]; set before first
                            // use (with C) when divTau's stabilization is BSD or coupling
    Time time;
    dictionary fvSolution;
    dictionary fvSchemes;

    fvMesh(label nCells, label nInternalFaces, label nBoundaryFaces, std::vector<label> owner,
           std::vector<label> neighbour, std::vector<scalar> Sf, std::vector<scalar> weights,
           std::vector<scalar> V, std::vector<patchInfo> patches);
    ~fvMesh();
    fvMesh(const fvMesh&) = delete;
    fvMesh& operator=(const fvMesh&) = delete;
    const rt_mesh* device() const;
    label nFaces() const { return nInternalFaces + nBoundaryFaces; }
    //- [OF9] validComponents<symmTensor>()
    std::vector<label> validComponentsSymmTensor() const;
    //- fvSolution `solvers { <field> {...} }` entry
    dictionary solverDict(const word& field) const;
private:
    mutable rt_mesh* dev_ = nullptr;
};

// the theta patch field types the device takes (symmetryPlane stands for
// OpenFOAM's symmetryPlane and symmetry, both basicSymmetry patch fields;
// wedge takes its rotation from the patch's geometry)
enum patchFieldType { zeroGradient = RT_BC_ZEROGRADIENT, fixedValue = RT_BC_FIXEDVALUE, empty = RT_BC_EMPTY,
                      symmetryPlane = RT_BC_SYMMETRY, wedge = RT_BC_WEDGE };

//- GeometricField<Type, fvPatchField, volMesh> with nCmpt components
template <int nCmpt>
struct volField {
    word name;
    const fvMesh* mesh = nullptr;
    std::vector<scalar> internal;  // [nCmpt * nCells]
    std::vector<scalar> boundary;  // [nCmpt * nBoundaryFaces]
    std::vector<int> bcType;       // per patch
    volField() = default;
    volField(const word& n, const fvMesh& m)
        : name(n), mesh(&m), internal((size_t)nCmpt * m.nCells, 0.0),
          boundary((size_t)nCmpt * m.nBoundaryFaces, 0.0), bcType(m.boundary.size(), zeroGradient) {
        for (size_t p = 0; p < m.boundary.size(); ++p)
            if (m.boundary[p].kind == RT_PATCH_EMPTY) bcType[p] = empty;
    }
    label size() const { return mesh ? mesh->nCells : 0; }
    scalar* operator[](label celli) { return internal.data() + (size_t)nCmpt * celli; }
    const scalar* operator[](label celli) const { return internal.data() + (size_t)nCmpt * celli; }
};
typedef volField<1> volScalarField;
typedef volField<3> volVectorField;
typedef volField<6> volSymmTensorField;
typedef volField<9> volTensorField;

//- surfaceScalarField: internal faces followed by boundary faces
struct surfaceScalarField {
    word name;
    const fvMesh* mesh = nullptr;
    std::vector<scalar> values;
    surfaceScalarField() = default;
    surfaceScalarField(const word& n, const fvMesh& m) : name(n), mesh(&m), values(m.nFaces(), 0.0) {}
};

//- solver performance of one component (SolverPerformance<scalar>)
struct solverPerformance {
    word fieldName;
    scalar initialResidual = 0, finalResidual = 0;
    label nIterations = 0, nRefinements = 0;
    bool converged = false, solved = false;
};

// ---------------------------------------------------------------- models
class constitutiveEq {
public:
    typedef std::function<std::unique_ptr<constitutiveEq>(const word&, const volVectorField&,
                                                          const surfaceScalarField&, const dictionary&)>
        constructorFn;
    //- dictionaryConstructorTable of the run-time selection
    static std::map<word, constructorFn>& dictionaryConstructorTable();
    template <class Model>
    struct adddictionaryConstructorToTable {
        explicit adddictionaryConstructorToTable(const word& typeName) {
            dictionaryConstructorTable()[typeName] =
                [](const word& n, const volVectorField& U, const surfaceScalarField& phi,
                   const dictionary& d) { return std::unique_ptr<constitutiveEq>(new Model(n, U, phi, d)); };
        }
    };
    //- Return a reference to the selected viscoelastic law
    static std::unique_ptr<constitutiveEq> New(const word& name, const volVectorField& U,
                                               const surfaceScalarField& phi, const dictionary& dict);

    constitutiveEq(const word& name, const volVectorField& U, const surfaceScalarField& phi)
        : name_(name), U_(U), phi_(phi) {}
    virtual ~constitutiveEq() = default;
    const word& name() const { return name_; }
    const volVectorField& U() const { return U_; }
    const surfaceScalarField& phi() const { return phi_; }

    virtual volSymmTensorField tau() const = 0;
    virtual scalar rho() const = 0;
    virtual bool isGNF() const = 0;
    virtual void correct(const volScalarField* alpha = nullptr, const volTensorField* gradU = nullptr) = 0;

private:
    word name_;
    const volVectorField& U_;
    const surfaceScalarField& phi_;
};

//- The log-conformation models: state on the device, correct() through the C-ABI
class logConformationEq : public constitutiveEq {
public:
    logConformationEq(const word& typeName, const word& name, const volVectorField& U,
                      const surfaceScalarField& phi, const dictionary& dict);
    ~logConformationEq() override;
    volSymmTensorField tau() const override;
    scalar rho() const override { return rho_; }
    bool isGNF() const override { return false; }
    void correct(const volScalarField* alpha = nullptr, const volTensorField* gradU = nullptr) override;
    //- theta (MUST_READ) and the eigen state (READ_IF_PRESENT) for restarts; tau
    //  for the models that read it (SaramitoLog, BMPLog: MUST_READ)
    void setState(const volSymmTensorField& theta, const volTensorField* eigVals,
                  const volTensorField* eigVecs, const volSymmTensorField* tau = nullptr);
    //- BMPLog's fluidity Phi<name> (MUST_READ, BMPLog.C:98-111), with its patch conditions
    void setFluidity(const volScalarField& Phi);
    volScalarField fluidity() const;
    volSymmTensorField theta() const;
    volTensorField eigVals() const;
    volTensorField eigVecs() const;
    //- the explicit part of divTau(U) (constitutiveEq.C:72-132) for the
    //  stabilization option: fvc::div(tau()/rho()) (none), minus
    //  fvc::laplacian(etaP/rho, U) with fvSchemes laplacian(eta,U) (BSD), or minus
    //  fvc::div((etaP/rho) grad(U)) (coupling); tau's patch values from
    //  tauBoundary ([6*nBoundaryFaces]) or zeroGradient when it is empty
    volVectorField fvcDivTau(const std::vector<scalar>& tauBoundary = std::vector<scalar>()) const;
    //- nu of the implicit fvm::laplacian(nu, U) divTau(U) holds: etaS/rho (none),
    //  (etaP + etaS)/rho (BSD, coupling)
    scalar divTauCoeff() const;
    //- tauTotal() (constitutiveEq.C:534-547)
    volSymmTensorField tauTotal() const;
    const std::vector<solverPerformance>& solverPerf() const { return perf_; }
    const word& type() const { return type_; }

private:
    word type_;
    rt_model* model_ = nullptr;
    scalar rho_ = 1;
    std::vector<solverPerformance> perf_;
    volSymmTensorField getSymm(int which, const word& n) const;
    volTensorField getTensor(int which, const word& n) const;
};

//- multiMode (multiMode.C): one log model per entry of `models`; tau() is the
//  sum over the modes, whose state is set and read per mode
class multiMode : public constitutiveEq {
public:
    static const char* typeName_() { return "multiMode"; }
    multiMode(const word& name, const volVectorField& U, const surfaceScalarField& phi, const dictionary& dict);
    ~multiMode() override;
    volSymmTensorField tau() const override;
    scalar rho() const override { return rho_; }
    bool isGNF() const override { return false; }
    void correct(const volScalarField* alpha = nullptr, const volTensorField* gradU = nullptr) override;
    label size() const { return (label)modes_.size(); }
    void setState(label mode, const volSymmTensorField& theta, const volTensorField* eigVals,
                  const volTensorField* eigVecs);
    volSymmTensorField theta(label mode) const;
    volSymmTensorField tau(label mode) const;

private:
    rt_model* model_ = nullptr;
    std::vector<rt_model*> modes_;  // owned by model_
    scalar rho_ = 1;
};

#define RHEOGPU_LOG_MODEL(Class, TypeNameStr)                                                   \
    class Class : public logConformationEq {                                                 \
    public:                                                                                  \
        static const char* typeName_() { return TypeNameStr; }                               \
        Class(const word& n, const volVectorField& U, const surfaceScalarField& phi,         \
              const dictionary& d)                                                           \
            : logConformationEq(TypeNameStr, n, U, phi, d) {}                                \
    };

RHEOGPU_LOG_MODEL(Oldroyd_BLog, "Oldroyd-BLog")
RHEOGPU_LOG_MODEL(GiesekusLog, "GiesekusLog")
RHEOGPU_LOG_MODEL(FENE_PLog, "FENE-PLog")
RHEOGPU_LOG_MODEL(PTTLog, "PTTLog")
RHEOGPU_LOG_MODEL(FENE_CRLog, "FENE-CRLog")
RHEOGPU_LOG_MODEL(WhiteMetznerCYLog, "WhiteMetznerCYLog")
RHEOGPU_LOG_MODEL(RoliePolyLog, "Rolie-PolyLog")
RHEOGPU_LOG_MODEL(XPomPomLog, "XPomPomLog")
RHEOGPU_LOG_MODEL(SaramitoLog, "SaramitoLog")
RHEOGPU_LOG_MODEL(BMPLog, "BMPLog")
// the root-k conformation form (theta = C^(1/k)) runs on the same device path
RHEOGPU_LOG_MODEL(Oldroyd_BRootk, "Oldroyd-BRootk")

//- IOdictionary wrapper reading constitutiveProperties (constitutiveModel.C:41-59)
class constitutiveModel {
public:
    constitutiveModel(const volVectorField& U, const surfaceScalarField& phi,
                      const dictionary& constitutiveProperties);
    volSymmTensorField tau() const { return eqPtr_->tau(); }
    scalar rho() const { return eqPtr_->rho(); }
    bool isGNF() const { return eqPtr_->isGNF(); }
    void correct(const volScalarField* alpha = nullptr, const volTensorField* gradU = nullptr) {
        eqPtr_->correct(alpha, gradU);
    }
    constitutiveEq& eq() { return *eqPtr_; }
private:
    std::unique_ptr<constitutiveEq> eqPtr_;
};

// ------------------------------------------------------------ solve surface
struct symmTensor {};

template <class Type>
class fvMatrix;

//- fvMatrix<symmTensor>: LDU coefficients + source + boundary coefficients
template <>
class fvMatrix<symmTensor> {
public:
    explicit fvMatrix(volSymmTensorField& psi);
    volSymmTensorField& psi() { return *psi_; }
    std::vector<scalar>& diag() { return diag_; }
    std::vector<scalar>& lower() { return lower_; }
    std::vector<scalar>& upper() { return upper_; }
    std::vector<scalar>& source() { return source_; }                  // [6 nCells]
    std::vector<scalar>& internalCoeffs() { return internalCoeffs_; }  // [6 nBoundaryFaces]
    std::vector<scalar>& boundaryCoeffs() { return boundaryCoeffs_; }  // [6 nBoundaryFaces]
    //- [OF9] fvMatrix::solve(): controls from fvSolution solvers.<psi name>
    std::vector<solverPerformance> solve();
    std::vector<solverPerformance> solve(const dictionary& solverControls);
private:
    volSymmTensorField* psi_;
    std::vector<scalar> diag_, lower_, upper_, source_, internalCoeffs_, boundaryCoeffs_;
};

template <class Type>
class sparseSolver {
public:
    typedef std::function<std::unique_ptr<sparseSolver>(const volSymmTensorField&, const fvMesh&,
                                                        const dictionary&)>
        constructorFn;
    static std::map<word, constructorFn>& dictionaryConstructorTable();
    //- selected by the solverType key of fvSolution solvers.<field> (default openFoamSolver)
    static std::unique_ptr<sparseSolver> New(const volSymmTensorField& T, const fvMesh& mesh,
                                             const dictionary& fvSolution);
    sparseSolver(const fvMesh& mesh, const dictionary& fvSolution) : mesh_(mesh), fvSolution_(fvSolution) {}
    virtual ~sparseSolver() = default;
    virtual bool isExternalSolver() const = 0;
    virtual void solve(fvMatrix<Type>& matrix) = 0;
    virtual void solve(fvMatrix<Type>& matrix, const dictionary& dict) = 0;
    const fvMesh& mesh() const { return mesh_; }
    const std::vector<solverPerformance>& performance() const { return perf_; }
protected:
    const fvMesh& mesh_;
    dictionary fvSolution_;
    std::vector<solverPerformance> perf_;
};

}  // namespace rheoGpu

#endif
