// C++ host-API test: drives the stress step the way rheoFoam does
// (constitutiveModel constEq(U, phi); constEq.correct();
//  sparseSolver<symmTensor>::New(...)->solve(fvMatrix&)) through the C++
// mirror of the reference interfaces, and checks the results against the CPU
// oracle (oracle/liboracle.so) on the same inputs.  Exit 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"
#include "rheoGpu.H"
#include "rtcase.h"

using namespace rheoGpu;

static int failures = 0;
#define CHECK(cond, ...)                              \
    do {                                              \
        if (!(cond)) {                                \
            std::printf("FAIL %s: ", #cond);          \
            std::printf(__VA_ARGS__);                 \
            std::printf("\n");                        \
            ++failures;                               \
        }                                             \
    } while (0)

static double relL2(const std::vector<double>& a, const std::vector<double>& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return std::sqrt(num / (den > 0 ? den : 1e-300));
}

int main() {
    // ---- case: 64x64x1 channel, seeded fields
    rtc_mesh* cm = rtc_mesh_cartesian(64, 64, 1, 1.0, 1.0, 1.0 / 64, 1, 1, 0.0, 4, 0);
    int32_t sz[4];
    rtc_mesh_sizes(cm, sz);
    const label N = sz[0], F = sz[1], NB = sz[2], NP = sz[3];
    std::vector<int32_t> owner(F + NB), nbr(F), ps(NP), pz(NP), pk(NP);
    std::vector<double> Sf(3 * (size_t)(F + NB)), w(F), V(N);
    rtc_mesh_export(cm, owner.data(), nbr.data(), Sf.data(), nullptr, w.data(), V.data(), nullptr, ps.data(),
                    pz.data(), pk.data(), nullptr);
    std::vector<patchInfo> patches;
    for (int p = 0; p < NP; ++p) patches.push_back(patchInfo{rtc_mesh_patch_name(cm, p), ps[p], pz[p], pk[p]});
    fvMesh mesh(N, F, NB, owner, nbr, Sf, w, V, patches);
    mesh.time.deltaT = 1e-3;
    mesh.fvSolution = dictionary(R"DICT(
        solvers
        {
            "(theta|tau|C)"
            {
                solver           PBiCGStab;
                preconditioner   { preconditioner DILU; }
                tolerance        1e-10;
                relTol           0;
                maxIter          1000;
            }
        }
    )DICT");
    volVectorField U("U", mesh);
    surfaceScalarField phi("phi", mesh);
    volSymmTensorField theta("theta", mesh);
    volTensorField eigVals("eigVals", mesh), eigVecs("eigVecs", mesh), gradU("gradU", mesh);
    rtc_fields(cm, RTC_FIELDS_CHANNEL, 2, 1e-2, -20, 5, U.internal.data(), U.boundary.data(), phi.values.data(),
               gradU.internal.data(), theta.internal.data(), eigVals.internal.data(), eigVecs.internal.data());

    const dictionary constitutiveProperties(R"(
        parameters
        {
            type             Oldroyd-BLog;
            rho              rho [1 -3 0 0 0 0 0] 1.;
            etaS             etaS [1 -1 -1 0 0 0 0] 0.59;
            etaP             etaP [1 -1 -1 0 0 0 0] 0.41;
            lambda           lambda [0 0 1 0 0 0 0] 1.;
            stabilization    coupling;
        }
    )");

    // ---- constitutiveModel::correct() vs oracle
    constitutiveModel constEq(U, phi, constitutiveProperties);
    auto& eq = dynamic_cast<logConformationEq&>(constEq.eq());
    eq.setState(theta, &eigVals, &eigVecs);
    constEq.correct();
    volSymmTensorField tauG = constEq.tau();
    volSymmTensorField thG = eq.theta();
    CHECK(eq.type() == "Oldroyd-BLog", "type %s", eq.type().c_str());
    CHECK(constEq.rho() == 1.0, "rho %g", constEq.rho());
    CHECK(!constEq.isGNF(), "isGNF");

    std::vector<int32_t> empty(NP);
    for (int p = 0; p < NP; ++p) empty[p] = pk[p] == RTC_EMPTY;
    orc_mesh om{N, F, NB, NP, owner.data(), nbr.data(), Sf.data(), w.data(), V.data(), ps.data(), pz.data(),
                empty.data()};
    orc_params prm{ORC_OLDROYDB, 1.0, 0.59, 0.41, 1.0, 0, 0, 0, 0, 0};
    std::vector<int32_t> bc(NP);
    for (int p = 0; p < NP; ++p) bc[p] = pk[p] == RTC_EMPTY ? ORC_BC_EMPTY : ORC_BC_ZEROGRADIENT;
    std::vector<double> bcv(6 * (size_t)NB, 0.0);
    int32_t valid[6];
    {
        auto v = mesh.validComponentsSymmTensor();
        for (int q = 0; q < 6; ++q) valid[q] = v[q];
    }
    CHECK(valid[2] == -1 && valid[4] == -1 && valid[0] == 1, "validComponents");
    orc_controls ctl{ORC_PBICGSTAB, ORC_PC_DILU, 1e-10, 0.0, 1000, 0};
    std::vector<double> thO = theta.internal, evO = eigVals.internal, vecO = eigVecs.internal,
                        tauO(6 * (size_t)N);
    orc_perf perf[6];
    orc_correct(&om, &prm, U.internal.data(), U.boundary.data(), phi.values.data(), nullptr, 1e-3, bc.data(),
                bcv.data(), valid, &ctl, 0.0, thO.data(), evO.data(), vecO.data(), tauO.data(), perf);
    const double eTh = relL2(thG.internal, thO), eTau = relL2(tauG.internal, tauO);
    std::printf("constitutiveModel::correct  theta relL2 %.3e  tau relL2 %.3e\n", eTh, eTau);
    CHECK(eTh < 1e-9, "theta %.3e", eTh);
    CHECK(eTau < 1e-9, "tau %.3e", eTau);
    for (auto& p : eq.solverPerf())
        if (p.solved) CHECK(p.converged && p.finalResidual < 1e-10, "%s", p.fieldName.c_str());

    // ---- run-time selection errors
    try {
        constitutiveEq::New("", U, phi, dictionary("type Maxwell; rho 1; etaS 1; etaP 1; lambda 1;"));
        CHECK(false, "unknown type accepted");
    } catch (const FatalError& e) {
        const std::string m = e.what();
        CHECK(m.find("Unknown constitutiveEq type Maxwell") != std::string::npos, "%s", m.c_str());
        CHECK(m.find("GiesekusLog") != std::string::npos && m.find("PTTLog") != std::string::npos, "%s",
              m.c_str());
        CHECK(e.code == RT_E_UNKNOWN_TYPE, "code %d", e.code);
    }

    // ---- sparseSolver<symmTensor> surface vs the oracle's segregated solve
    std::vector<double> S(6 * (size_t)N), f(N);
    orc_rhs(&prm, N, gradU.internal.data(), theta.internal.data(), eigVals.internal.data(),
            eigVecs.internal.data(), S.data(), f.data());
    volSymmTensorField T("theta", mesh);
    T.internal = theta.internal;
    fvMatrix<symmTensor> M(T);
    orc_assemble(&om, phi.values.data(), theta.internal.data(), S.data(), 1e-3, bc.data(), bcv.data(),
                 M.diag().data(), M.lower().data(), M.upper().data(), M.source().data(),
                 M.internalCoeffs().data(), M.boundaryCoeffs().data());
    std::vector<double> xO = theta.internal;
    orc_solve_segregated(&om, M.diag().data(), M.lower().data(), M.upper().data(), M.source().data(),
                         M.internalCoeffs().data(), M.boundaryCoeffs().data(), bc.data(), valid, &ctl, xO.data(),
                         perf);
    dictionary fvSol(R"(solvers { theta { solverType gpuSolver; solver PBiCGStab; preconditioner DILU;
                        tolerance 1e-12; relTol 0; maxIter 1000; precision mixed; } })");
    auto solver = sparseSolver<symmTensor>::New(T, mesh, fvSol);
    CHECK(solver->isExternalSolver(), "gpuSolver is external");
    solver->solve(M, fvSol.subDict("solvers").subDict("theta"));
    const double eX = relL2(T.internal, xO);
    std::printf("sparseSolver<symmTensor>::solve  psi relL2 %.3e\n", eX);
    CHECK(eX < 1e-9, "psi %.3e", eX);
    try {
        sparseSolver<symmTensor>::New(T, mesh, dictionary("solvers { theta { solverType petscSolver; } }"));
        CHECK(false, "unknown sparseSolver accepted");
    } catch (const FatalError& e) {
        CHECK(std::string(e.what()).find("Unknown sparseSolver type petscSolver") != std::string::npos, "%s",
              e.what());
    }
    rtc_mesh_free(cm);
    std::printf(failures ? "FAILED (%d)\n" : "OK\n", failures);
    return failures ? 1 : 0;
}
