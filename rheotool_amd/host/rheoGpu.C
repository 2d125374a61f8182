/*---------------------------------------------------------------------------*\
  rheoGpu.C -- implementation of the C++ host mirror (see rheoGpu.H).
\*---------------------------------------------------------------------------*/
#include "rheoGpu.H"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <regex>
#include <sstream>

namespace rheoGpu {

// ------------------------------------------------------------ dictionary text
namespace {

std::string stripComments(const std::string& s) {
    std::string o;
    o.reserve(s.size());
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '/' && i + 1 < s.size() && s[i + 1] == '/') {
            while (i < s.size() && s[i] != '\n') ++i;
            o += '\n';
        } else if (s[i] == '/' && i + 1 < s.size() && s[i + 1] == '*') {
            size_t e = s.find("*/", i + 2);
            i = e == std::string::npos ? s.size() : e + 1;
        } else {
            o += s[i];
        }
    }
    return o;
}

// top-level `key { ... }` blocks of a dictionary body: key -> inner text
std::vector<std::pair<std::string, std::string>> topBlocks(const std::string& raw) {
    const std::string s = stripComments(raw);
    std::vector<std::pair<std::string, std::string>> out;
    int depth = 0;
    size_t i = 0;
    std::string lastTok;
    while (i < s.size()) {
        char c = s[i];
        if (c == '{') {
            if (depth == 0) {
                int d = 1;
                size_t j = i + 1;
                while (j < s.size() && d > 0) {
                    if (s[j] == '{') ++d;
                    if (s[j] == '}') --d;
                    ++j;
                }
                out.push_back({lastTok, s.substr(i + 1, j - i - 2)});
                i = j;
                lastTok.clear();
                continue;
            }
            ++depth;
        } else if (c == '}') {
            --depth;
        } else if (c == ';') {
            lastTok.clear();
        } else if (!std::isspace((unsigned char)c) && depth == 0) {
            size_t j = i;
            if (c == '"') {
                j = s.find('"', i + 1);
                lastTok = s.substr(i + 1, j - i - 1);
                i = j + 1;
                continue;
            }
            while (j < s.size() && !std::isspace((unsigned char)s[j]) && s[j] != '{' && s[j] != ';' &&
                   s[j] != '}')
                ++j;
            lastTok = s.substr(i, j - i);
            i = j;
            continue;
        }
        ++i;
    }
    return out;
}

// text of a top-level `key ...;` entry (without the key), "" when absent
bool topEntry(const std::string& raw, const std::string& key, std::string& val) {
    const std::string s = stripComments(raw);
    int depth = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '{') ++depth;
        else if (s[i] == '}') --depth;
        else if (depth == 0 && s.compare(i, key.size(), key) == 0 &&
                 (i == 0 || std::isspace((unsigned char)s[i - 1]) || s[i - 1] == ';') &&
                 i + key.size() < s.size() && std::isspace((unsigned char)s[i + key.size()])) {
            size_t e = s.find(';', i);
            if (e == std::string::npos) return false;
            val = s.substr(i + key.size(), e - i - key.size());
            return true;
        }
    }
    return false;
}

bool lookupWord(const std::string& raw, const std::string& key, std::string& w) {
    std::string v;
    if (!topEntry(raw, key, v)) return false;
    std::istringstream is(v);
    return (bool)(is >> w);
}

bool lookupScalar(const std::string& raw, const std::string& key, double& x) {
    std::string v;
    if (!topEntry(raw, key, v)) return false;
    // the last number outside [...] (skips "name [dims]")
    std::string clean;
    int d = 0;
    for (char c : v) {
        if (c == '[') ++d;
        else if (c == ']') --d;
        else if (d == 0) clean += c;
    }
    std::istringstream is(clean);
    std::string tok;
    bool ok = false;
    while (is >> tok) {
        char* end = nullptr;
        double y = std::strtod(tok.c_str(), &end);
        if (end && *end == '\0') {
            x = y;
            ok = true;
        }
    }
    return ok;
}

}  // namespace

dictionary dictionary::fromFile(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw FatalError(RT_E_BAD_ARG, "cannot open file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return dictionary(ss.str());
}

dictionary dictionary::subDict(const word& name) const {
    for (auto& kv : topBlocks(text_))
        if (kv.first == name) return dictionary(kv.second);
    return dictionary();
}

// --------------------------------------------------------------- context
gpuContext::gpuContext(int device, size_t budget) {
    int rc = rt_ctx_create(device, budget, &ctx_);
    if (rc != RT_OK) throw FatalError(rc, "rheoGpu: cannot create a context on device " + std::to_string(device));
}
gpuContext::~gpuContext() {
    if (ctx_) rt_ctx_destroy(ctx_);
}
gpuContext& gpuContext::get(int device, size_t budget) {
    static std::map<int, std::unique_ptr<gpuContext>> ctxs;
    auto& p = ctxs[device];
    if (!p) p.reset(new gpuContext(device, budget));
    return *p;
}
void gpuContext::check(int rc, const char* what) const {
    if (rc != RT_OK && rc != RT_W_NOT_CONVERGED)
        throw FatalError(rc, std::string(what) + ": " + rt_last_error(ctx_));
}

// ------------------------------------------------------------------- mesh
fvMesh::fvMesh(label nC, label nI, label nB, std::vector<label> own, std::vector<label> nei,
               std::vector<scalar> sf, std::vector<scalar> w, std::vector<scalar> vol,
               std::vector<patchInfo> patches)
    : nCells(nC), nInternalFaces(nI), nBoundaryFaces(nB), owner(std::move(own)),
      neighbour(std::move(nei)), Sf(std::move(sf)), weights(std::move(w)), V(std::move(vol)),
      boundary(std::move(patches)) {}

fvMesh::~fvMesh() {
    if (dev_) rt_mesh_destroy(dev_);
}

const rt_mesh* fvMesh::device() const {
    if (!dev_) {
        std::vector<rt_patch> p;
        for (auto& b : boundary) p.push_back(rt_patch{b.start, b.size, b.kind});
        rt_mesh_desc d{nCells, nInternalFaces, nBoundaryFaces, (int32_t)p.size(), owner.data(),
                       neighbour.data(), Sf.data(), weights.data(), V.data(), p.data(),
                       C.empty() ? nullptr : C.data(), Cb.empty() ? nullptr : Cb.data()};
        gpuContext& c = gpuContext::get();
        c.check(rt_mesh_create(c.ptr(), &d, RT_RENUMBER_AUTO, &dev_), "fvMesh");
    }
    return dev_;
}

std::vector<label> fvMesh::validComponentsSymmTensor() const {
    int32_t v[6];
    rt_mesh_valid_components(device(), v);
    return std::vector<label>(v, v + 6);
}

dictionary fvMesh::solverDict(const word& field) const {
    const dictionary solvers = fvSolution.subDict("solvers");
    for (auto& kv : topBlocks(solvers.text())) {
        if (kv.first == field) return dictionary(kv.second);
        try {
            if (std::regex_match(field, std::regex(kv.first))) return dictionary(kv.second);
        } catch (const std::regex_error&) {
        }
    }
    return dictionary();
}

// ----------------------------------------------------------------- models
std::map<word, constitutiveEq::constructorFn>& constitutiveEq::dictionaryConstructorTable() {
    static std::map<word, constructorFn> table;
    return table;
}

static constitutiveEq::adddictionaryConstructorToTable<Oldroyd_BLog> addOB("Oldroyd-BLog");
static constitutiveEq::adddictionaryConstructorToTable<GiesekusLog> addGiesekus("GiesekusLog");
static constitutiveEq::adddictionaryConstructorToTable<FENE_PLog> addFENEP("FENE-PLog");
static constitutiveEq::adddictionaryConstructorToTable<PTTLog> addPTT("PTTLog");
static constitutiveEq::adddictionaryConstructorToTable<FENE_CRLog> addFENECR("FENE-CRLog");
static constitutiveEq::adddictionaryConstructorToTable<WhiteMetznerCYLog> addWMCY("WhiteMetznerCYLog");
static constitutiveEq::adddictionaryConstructorToTable<RoliePolyLog> addRP("Rolie-PolyLog");
static constitutiveEq::adddictionaryConstructorToTable<XPomPomLog> addXPP("XPomPomLog");
static constitutiveEq::adddictionaryConstructorToTable<SaramitoLog> addSaramito("SaramitoLog");
static constitutiveEq::adddictionaryConstructorToTable<BMPLog> addBMP("BMPLog");
static constitutiveEq::adddictionaryConstructorToTable<Oldroyd_BRootk> addRootk("Oldroyd-BRootk");
static constitutiveEq::adddictionaryConstructorToTable<multiMode> addMulti("multiMode");

multiMode::multiMode(const word& name, const volVectorField& U, const surfaceScalarField& phi,
                     const dictionary& dict)
    : constitutiveEq(name, U, phi) {
    gpuContext& c = gpuContext::get();
    c.check(rt_model_create(c.ptr(), name.c_str(), U.mesh->device(), dict.text().c_str(), &model_),
            "constitutiveEq::New");
    for (int i = 0; i < rt_model_mode_count(model_); ++i) modes_.push_back(rt_model_mode(model_, i));
    // multiMode::rho() averages the modes' densities (multiMode.C:229-242)
    std::vector<double> rhos;
    std::regex re("rho\\s+(rho\\s+)?(\\[[^\\]]*\\]\\s*)?([-+0-9.eE]+)\\s*;");
    for (std::sregex_iterator it(dict.text().begin(), dict.text().end(), re), e; it != e; ++it)
        rhos.push_back(std::stod((*it)[3]));
    if (!rhos.empty()) {
        rho_ = 0;
        for (double r : rhos) rho_ += r;
        rho_ /= rhos.size();
    }
    volSymmTensorField th("theta", *U.mesh);
    for (rt_model* m : modes_) c.check(rt_model_set_state(m, th.internal.data(), nullptr, nullptr, nullptr), "theta");
}
multiMode::~multiMode() {
    if (model_) rt_model_destroy(model_);
}
void multiMode::setState(label mode, const volSymmTensorField& theta, const volTensorField* eigVals,
                         const volTensorField* eigVecs) {
    gpuContext::get().check(rt_model_set_state(modes_.at(mode), theta.internal.data(),
                                               eigVals ? eigVals->internal.data() : nullptr,
                                               eigVecs ? eigVecs->internal.data() : nullptr, nullptr),
                            "multiMode::setState");
}
volSymmTensorField multiMode::tau() const {
    volSymmTensorField f("tau" + name(), *U().mesh);
    gpuContext::get().check(rt_model_get(model_, RT_TAU, f.internal.data()), "tau");
    return f;
}
volSymmTensorField multiMode::theta(label mode) const {
    volSymmTensorField f("theta" + name() + std::to_string(mode), *U().mesh);
    gpuContext::get().check(rt_model_get(modes_.at(mode), RT_THETA, f.internal.data()), "theta");
    return f;
}
volSymmTensorField multiMode::tau(label mode) const {
    volSymmTensorField f("tau" + name(), *U().mesh);
    gpuContext::get().check(rt_model_get(modes_.at(mode), RT_TAU, f.internal.data()), "tau");
    return f;
}
void multiMode::correct(const volScalarField*, const volTensorField* gradU) {
    gpuContext& c = gpuContext::get();
    const fvMesh& mesh = *U().mesh;
    c.check(rt_model_set_flow(model_, U().internal.data(), U().boundary.empty() ? nullptr : U().boundary.data(),
                              phi().values.data()),
            "multiMode::correct (U, phi)");
    const dictionary sd = mesh.solverDict("theta" + name());
    if (!sd.text().empty()) c.check(rt_model_set_solver(model_, sd.text().c_str()), "fvSolution theta");
    std::string ds;
    if (topEntry(mesh.fvSchemes.subDict("divSchemes").text(), "div(phi,theta)", ds))
        c.check(rt_model_set_scheme(model_, ds.c_str()), "fvSchemes div(phi,theta)");
    c.check(rt_model_correct(model_, mesh.time.deltaTValue(), gradU ? gradU->internal.data() : nullptr),
            "multiMode::correct");
}

std::unique_ptr<constitutiveEq> constitutiveEq::New(const word& name, const volVectorField& U,
                                                    const surfaceScalarField& phi, const dictionary& dict) {
    word typeName;
    if (!lookupWord(dict.text(), "type", typeName))
        throw FatalError(RT_E_BAD_ARG, "keyword type is undefined in dictionary parameters");
    auto& table = dictionaryConstructorTable();
    auto it = table.find(typeName);
    if (it == table.end()) {
        std::string msg = "Unknown constitutiveEq type " + typeName + "\n\nValid constitutiveEq types are :\n" +
                          std::to_string(table.size()) + "\n(\n";
        for (auto& kv : table) msg += kv.first + "\n";
        throw FatalError(RT_E_UNKNOWN_TYPE, msg + ")");
    }
    return it->second(name, U, phi, dict);
}

logConformationEq::logConformationEq(const word& typeName, const word& name, const volVectorField& U,
                                     const surfaceScalarField& phi, const dictionary& dict)
    : constitutiveEq(name, U, phi), type_(typeName) {
    gpuContext& c = gpuContext::get();
    c.check(rt_model_create(c.ptr(), name.c_str(), U.mesh->device(), dict.text().c_str(), &model_),
            "constitutiveEq::New");
    lookupScalar(dict.text(), "rho", rho_);
    // theta starts at zero with the identity eigen state until setState()
    volSymmTensorField th("theta" + name, *U.mesh);
    c.check(rt_model_set_state(model_, th.internal.data(), nullptr, nullptr, nullptr), "theta");
}

logConformationEq::~logConformationEq() {
    if (model_) rt_model_destroy(model_);
}

void logConformationEq::setState(const volSymmTensorField& theta, const volTensorField* eigVals,
                                 const volTensorField* eigVecs, const volSymmTensorField* tau) {
    gpuContext& c = gpuContext::get();
    c.check(rt_model_set_state(model_, theta.internal.data(), eigVals ? eigVals->internal.data() : nullptr,
                               eigVecs ? eigVecs->internal.data() : nullptr, tau ? tau->internal.data() : nullptr),
            "setState");
    for (size_t p = 0; p < theta.bcType.size(); ++p)
        if (theta.bcType[p] != zeroGradient && theta.bcType[p] != empty) {
            const patchInfo& P = theta.mesh->boundary[p];
            c.check(rt_model_set_bc(model_, (int)p, theta.bcType[p], theta.boundary.data() + 6 * (size_t)P.start),
                    "theta boundary condition");
        }
}

void logConformationEq::setFluidity(const volScalarField& Phi) {
    gpuContext& c = gpuContext::get();
    rt_model* fl = rt_model_fluidity(model_);
    if (!fl) throw FatalError(RT_E_BAD_ARG, "setFluidity: " + type_ + " has no fluidity equation");
    c.check(rt_model_set_state(fl, Phi.internal.data(), nullptr, nullptr, nullptr), "Phi");
    for (size_t p = 0; p < Phi.bcType.size(); ++p)
        if (Phi.bcType[p] != zeroGradient && Phi.bcType[p] != empty) {
            const patchInfo& P = Phi.mesh->boundary[p];
            c.check(rt_model_set_bc(fl, (int)p, Phi.bcType[p], Phi.boundary.data() + (size_t)P.start),
                    "Phi boundary condition");
        }
}

volScalarField logConformationEq::fluidity() const {
    volScalarField f("Phi" + name(), *U().mesh);
    rt_model* fl = rt_model_fluidity(model_);
    if (fl) gpuContext::get().check(rt_model_get(fl, RT_THETA, f.internal.data()), "get Phi");
    return f;
}

void logConformationEq::correct(const volScalarField*, const volTensorField* gradU) {
    gpuContext& c = gpuContext::get();
    const fvMesh& mesh = *U().mesh;
    if (rt_model* fl = rt_model_fluidity(model_)) {
        // BMPLog's fluidity equation: fvSolution Phi, relaxationFactors, div(phi,Phi)
        const dictionary fd = mesh.solverDict("Phi" + name());
        if (!fd.text().empty()) c.check(rt_model_set_solver(fl, fd.text().c_str()), "fvSolution Phi");
        double fr = 0;
        const dictionary frd = mesh.fvSolution.subDict("relaxationFactors");
        if (!lookupScalar(frd.subDict("equations").text(), "Phi" + name(), fr))
            lookupScalar(frd.text(), "Phi" + name(), fr);
        c.check(rt_model_set_relax(fl, fr), "relax Phi");
        std::string fs;
        if (topEntry(mesh.fvSchemes.subDict("divSchemes").text(), "div(phi,Phi" + name() + ")", fs))
            c.check(rt_model_set_scheme(fl, fs.c_str()), "fvSchemes div(phi,Phi)");
    }
    c.check(rt_model_set_flow(model_, U().internal.data(), U().boundary.empty() ? nullptr : U().boundary.data(),
                              phi().values.data()),
            "constitutiveEq::correct (U, phi)");
    const dictionary sd = mesh.solverDict("theta" + name());
    if (!sd.text().empty()) c.check(rt_model_set_solver(model_, sd.text().c_str()), "fvSolution theta");
    // relaxationFactors { equations { theta f; } } -> fvMatrix::relax()
    double f = 0;
    const dictionary rf = mesh.fvSolution.subDict("relaxationFactors");
    if (!lookupScalar(rf.subDict("equations").text(), "theta" + name(), f)) lookupScalar(rf.text(), "theta" + name(), f);
    c.check(rt_model_set_relax(model_, f), "relax");
    // fvSchemes divSchemes { div(phi,theta) GaussDefCmpw <limiter>; }
    std::string ds;
    if (topEntry(mesh.fvSchemes.subDict("divSchemes").text(), "div(phi,theta" + name() + ")", ds))
        c.check(rt_model_set_scheme(model_, ds.c_str()), "fvSchemes div(phi,theta)");
    c.check(rt_model_correct(model_, mesh.time.deltaTValue(), gradU ? gradU->internal.data() : nullptr),
            "constitutiveEq::correct");
    rt_solver_perf p[6];
    rt_model_solver_perf(model_, p);
    const char* cmpt[6] = {"xx", "xy", "xz", "yy", "yz", "zz"};
    perf_.clear();
    for (int q = 0; q < 6; ++q)
        perf_.push_back(solverPerformance{"theta" + name() + cmpt[q], p[q].initialResidual, p[q].finalResidual,
                                          p[q].nIterations, p[q].nRefinements, p[q].converged != 0,
                                          p[q].solved != 0});
}

volSymmTensorField logConformationEq::getSymm(int which, const word& n) const {
    volSymmTensorField f(n, *U().mesh);
    gpuContext::get().check(rt_model_get(model_, which, f.internal.data()), "get");
    return f;
}
volTensorField logConformationEq::getTensor(int which, const word& n) const {
    volTensorField f(n, *U().mesh);
    gpuContext::get().check(rt_model_get(model_, which, f.internal.data()), "get");
    return f;
}
volSymmTensorField logConformationEq::tau() const { return getSymm(RT_TAU, "tau" + name()); }
volSymmTensorField logConformationEq::theta() const { return getSymm(RT_THETA, "theta" + name()); }
volTensorField logConformationEq::eigVals() const { return getTensor(RT_EIGVALS, "eigVals" + name()); }
volTensorField logConformationEq::eigVecs() const { return getTensor(RT_EIGVECS, "eigVecs" + name()); }
volVectorField logConformationEq::fvcDivTau(const std::vector<scalar>& tauBoundary) const {
    volVectorField f("div(tau" + name() + ")", *U().mesh);
    // laplacianSchemes { laplacian(eta,U) Gauss linear corrected; } (BSD's explicit laplacian)
    std::string ls;
    if (topEntry(U().mesh->fvSchemes.subDict("laplacianSchemes").text(), "laplacian(eta,U)", ls))
        gpuContext::get().check(rt_model_set_laplacian_scheme(model_, ls.c_str()), "fvSchemes laplacian(eta,U)");
    gpuContext::get().check(
        rt_model_div_tau(model_, tauBoundary.empty() ? nullptr : tauBoundary.data(), f.internal.data()),
        "divTau");
    return f;
}
scalar logConformationEq::divTauCoeff() const {
    double nu = 0;
    gpuContext::get().check(rt_model_div_tau_coeff(model_, &nu), "divTau coefficient");
    return nu;
}
volSymmTensorField logConformationEq::tauTotal() const {
    volSymmTensorField f("tauTotal" + name(), *U().mesh);
    gpuContext::get().check(rt_model_tau_total(model_, f.internal.data()), "tauTotal");
    return f;
}

constitutiveModel::constitutiveModel(const volVectorField& U, const surfaceScalarField& phi,
                                     const dictionary& constitutiveProperties) {
    dictionary params = constitutiveProperties.subDict("parameters");
    if (params.text().empty()) params = constitutiveProperties;
    eqPtr_ = constitutiveEq::New(word(), U, phi, params);
}

// ------------------------------------------------------------ solve surface
fvMatrix<symmTensor>::fvMatrix(volSymmTensorField& psi)
    : psi_(&psi), diag_(psi.mesh->nCells, 0.0), lower_(psi.mesh->nInternalFaces, 0.0),
      upper_(psi.mesh->nInternalFaces, 0.0), source_((size_t)6 * psi.mesh->nCells, 0.0),
      internalCoeffs_((size_t)6 * psi.mesh->nBoundaryFaces, 0.0),
      boundaryCoeffs_((size_t)6 * psi.mesh->nBoundaryFaces, 0.0) {}

std::vector<solverPerformance> fvMatrix<symmTensor>::solve(const dictionary& controls) {
    gpuContext& c = gpuContext::get();
    rt_ldu_symm sys{diag_.data(), lower_.data(), upper_.data(), source_.data(),
                    internalCoeffs_.empty() ? nullptr : internalCoeffs_.data(),
                    boundaryCoeffs_.empty() ? nullptr : boundaryCoeffs_.data()};
    rt_solver_perf p[6];
    c.check(rt_symm_solve(c.ptr(), psi_->mesh->device(), &sys, psi_->internal.data(), controls.text().c_str(), p),
            "fvMatrix<symmTensor>::solve");
    const char* cmpt[6] = {"xx", "xy", "xz", "yy", "yz", "zz"};
    std::vector<solverPerformance> out;
    for (int q = 0; q < 6; ++q)
        out.push_back(solverPerformance{psi_->name + cmpt[q], p[q].initialResidual, p[q].finalResidual,
                                        p[q].nIterations, p[q].nRefinements, p[q].converged != 0, p[q].solved != 0});
    return out;
}

std::vector<solverPerformance> fvMatrix<symmTensor>::solve() {
    return solve(psi_->mesh->solverDict(psi_->name));
}

namespace {

// the default: matrix.solve() (openFoamSolver.C:46-53) -- on the device here
class openFoamSolverSymm : public sparseSolver<symmTensor> {
public:
    using sparseSolver::sparseSolver;
    bool isExternalSolver() const override { return false; }
    void solve(fvMatrix<symmTensor>& m) override { perf_ = m.solve(); }
    void solve(fvMatrix<symmTensor>& m, const dictionary& d) override { perf_ = m.solve(d); }
};

// explicit GPU solver entry (solverType gpuSolver)
class gpuSolverSymm : public sparseSolver<symmTensor> {
public:
    using sparseSolver::sparseSolver;
    bool isExternalSolver() const override { return true; }
    void solve(fvMatrix<symmTensor>& m) override { perf_ = m.solve(mesh_.solverDict(m.psi().name)); }
    void solve(fvMatrix<symmTensor>& m, const dictionary& d) override { perf_ = m.solve(d); }
};

}  // namespace

template <>
std::map<word, sparseSolver<symmTensor>::constructorFn>& sparseSolver<symmTensor>::dictionaryConstructorTable() {
    static std::map<word, constructorFn> table{
        {"openFoamSolver",
         [](const volSymmTensorField&, const fvMesh& m, const dictionary& d) {
             return std::unique_ptr<sparseSolver>(new openFoamSolverSymm(m, d));
         }},
        {"gpuSolver", [](const volSymmTensorField&, const fvMesh& m, const dictionary& d) {
             return std::unique_ptr<sparseSolver>(new gpuSolverSymm(m, d));
         }}};
    return table;
}

template <>
std::unique_ptr<sparseSolver<symmTensor>> sparseSolver<symmTensor>::New(const volSymmTensorField& T,
                                                                       const fvMesh& mesh,
                                                                       const dictionary& fvSolution) {
    word type = "openFoamSolver";
    dictionary sd;
    for (auto& kv : topBlocks(fvSolution.subDict("solvers").text())) {
        bool match = kv.first == T.name;
        if (!match) {
            try {
                match = std::regex_match(T.name, std::regex(kv.first));
            } catch (const std::regex_error&) {
            }
        }
        if (match) {
            sd = dictionary(kv.second);
            break;
        }
    }
    lookupWord(sd.text(), "solverType", type);
    auto& table = dictionaryConstructorTable();
    auto it = table.find(type);
    if (it == table.end()) {
        std::string msg = "Unknown sparseSolver type " + type + "\n\nValid sparseSolvers are : \n" +
                          std::to_string(table.size()) + "\n(\n";
        for (auto& kv : table) msg += kv.first + "\n";
        throw FatalError(RT_E_UNKNOWN_TYPE, msg + ")");
    }
    return it->second(T, mesh, fvSolution);
}

}  // namespace rheoGpu
