// rt_api.cpp -- host runtime and C-ABI entry points (include/rheo_gpu.h).
//
// Context (device, stream, HBM budget, kernel timing), mesh and model
// handles, the constitutiveEq run-time selection from dictionary text, the
// correct() sequence and the sparseSolver<symmTensor>::solve surface.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <set>

#include "rt_dict.h"
#include "rt_internal.h"

namespace rt {

void buildMesh(Ctx* ctx, const rt_mesh_desc* d, int renumber, Mesh& m);

// ------------------------------------------------------------------ memory
DBuf& DBuf::operator=(DBuf&& o) noexcept {
    if (this != &o) {
        release();
        ctx = o.ctx;
        p = o.p;
        bytes = o.bytes;
        view = o.view;
        o.ctx = nullptr;
        o.p = nullptr;
        o.bytes = 0;
        o.view = false;
    }
    return *this;
}
void DBuf::release() {
    if (p && !view) {
        if (ctx) {
            ctx->inUse -= bytes;
            ctx->recycle(p, bytes);
        } else {
            (void)hipFree(p);
        }
    }
    p = nullptr;
    bytes = 0;
    ctx = nullptr;
    view = false;
}

namespace {
// 256-byte granularity up to 1 MiB, then 8 classes per power of two (at most
// 12.5% slack), so that buffers of similar meshes share a class
size_t sizeClass(size_t b) {
    b = (b + 255) / 256 * 256;
    if (b <= (size_t(1) << 20)) return b;
    const int k = 63 - __builtin_clzll(b);
    const size_t step = size_t(1) << (k - 3);
    return (b + step - 1) / step * step;
}
}  // namespace

DBuf Ctx::alloc(size_t bytes) {
    if (bytes == 0) bytes = 8;
    bytes = sizeClass(bytes);
    if (budget && inUse + bytes > budget)
        throw Error{RT_E_OOM, "HBM budget exceeded: " + std::to_string(inUse + bytes) + " > " +
                                  std::to_string(budget) + " bytes"};
    DBuf b;
    // best fit: the smallest cached block that holds the request, if it
    // wastes at most half the request (the serving loop sees many sizes, and
    // a miss costs a hipMalloc)
    auto it = cache.lower_bound(bytes);
    if (it != cache.end() && it->first <= bytes + bytes / 2 &&
        !(budget && inUse + it->first > budget)) {
        bytes = it->first;
        b.p = it->second;
        cachedBytes -= bytes;
        cache.erase(it);
    } else {
        hipError_t e = hipMalloc(&b.p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            trimCache(0);  // give the cached memory back and try once more
            e = hipMalloc(&b.p, bytes);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error{RT_E_OOM, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e)};
        }
    }
    b.ctx = this;
    b.bytes = bytes;
    inUse += bytes;
    return b;
}

void Ctx::recycle(void* p, size_t bytes) {
    cache.emplace(bytes, p);
    cachedBytes += bytes;
    if (cachedBytes > cacheCap) trimCache(cacheCap / 2);
}

void Ctx::trimCache(size_t keep) {
    // largest first: they are the least likely to be reused exactly
    while (cachedBytes > keep && !cache.empty()) {
        auto it = std::prev(cache.end());
        (void)hipFree(it->second);
        cachedBytes -= it->first;
        cache.erase(it);
    }
}

void* Ctx::takePinnedBlock() {
    if (!pinnedBlocks.empty()) {
        void* p = pinnedBlocks.back();
        pinnedBlocks.pop_back();
        return p;
    }
    void* p = nullptr;
    RT_HIP(hipHostMalloc(&p, kPinnedBlock, hipHostMallocDefault));
    return p;
}

void Ctx::givePinnedBlock(void* p) {
    if (p) pinnedBlocks.push_back(p);
}

void* Ctx::pinnedHost(size_t bytes) {
    if (pinnedBytes < bytes) {
        if (pinned) (void)hipHostFree(pinned);
        pinned = nullptr;
        RT_HIP(hipHostMalloc(&pinned, bytes, hipHostMallocDefault));
        pinnedBytes = bytes;
    }
    return pinned;
}

hipEvent_t Ctx::getEvent() {
    if (eventPool.empty()) {
        hipEvent_t e;
        RT_HIP(hipEventCreate(&e));
        return e;
    }
    hipEvent_t e = eventPool.back();
    eventPool.pop_back();
    return e;
}

void Ctx::flushProfile() {
    if (pending.empty()) return;
    RT_HIP(hipStreamSynchronize(stream));
    for (auto& p : pending) {
        float ms = 0;
        RT_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        auto& s = stats[p.name];
        s.launches += 1;
        s.ms += ms;
        s.bytes += p.bytes;
        eventPool.push_back(p.a);
        eventPool.push_back(p.b);
    }
    pending.clear();
}

// -------------------------------------------------------------- controls
Controls parseControls(const char* text) {
    Controls c;
    c.tolerance = 1e-6;
    if (!text) return c;
    Dict d = parseDict(text);
    // accept either the bare entry or `theta { ... }` / `solvers { theta {...} }`
    const Dict* e = &d;
    if (!d.has("solver")) {
        if (d.subDicts.count("solvers")) e = &d.subDicts.at("solvers");
        for (auto& kv : e->subDicts)
            if (kv.second.has("solver")) {
                e = &kv.second;
                break;
            }
    }
    std::string s;
    if (e->word("solver", s)) {
        if (s != "PBiCGStab" && s != "PBiCG")
            throw Error{RT_E_UNKNOWN_TYPE, "Unknown asymmetric matrix solver " + s +
                                               "\n\nValid asymmetric matrix solvers are :\n2\n(\nPBiCG\nPBiCGStab\n)"};
        c.solverName = s;
        c.solver = s == "PBiCG" ? SV_PBICG : SV_PBICGSTAB;
    }
    std::string pc = "none";
    if (e->subDicts.count("preconditioner")) {
        e->subDicts.at("preconditioner").word("preconditioner", pc);
    } else {
        e->word("preconditioner", pc);
    }
    c.precondName = pc;
    // multicolourILU0 is ILU(0) in a two-colour ordering, run as the
    // red-black Schur-complement solve (a mesh whose cell graph is
    // two-colourable, solver PBiCGStab; otherwise refused at the solve).
    // DILU's sweep is sequential; the device replaces it by that red-black
    // ILU(0) where it can run and by the diagonal (Jacobi) preconditioner
    // elsewhere, so iterates differ from the reference's and only converged
    // fields are compared.  diagonal/Jacobi is the same preconditioner, none
    // is none.  DIC is a symmetric-matrix preconditioner and the theta matrix
    // is asymmetric ([OF9] rejects it there).
    if (pc == "diagonal" || pc == "Jacobi") {
        c.precond = PC_JACOBI;
    } else if (pc == "none") {
        c.precond = PC_NONE;
    } else if (pc == "DILU") {
        c.precond = PC_MCILU0;
    } else if (pc == "multicolourILU0") {
        c.precond = PC_MCILU0;
        c.precondStrict = true;
    } else if (pc == "DIC" || pc == "FDIC" || pc == "GAMG") {
        throw Error{RT_E_UNKNOWN_TYPE, "Unknown asymmetric matrix preconditioner " + pc +
                                           "\n\nValid asymmetric matrix preconditioners are :\n5\n(\nDILU\nJacobi\ndiagonal\nmulticolourILU0\nnone\n)"};
    } else {
        throw Error{RT_E_UNKNOWN_TYPE, "Unknown asymmetric matrix preconditioner " + pc +
                                           "\n\nValid asymmetric matrix preconditioners are :\n5\n(\nDILU\nJacobi\ndiagonal\nmulticolourILU0\nnone\n)"};
    }
    double v;
    if (e->scalar("tolerance", v)) c.tolerance = v;
    if (e->scalar("relTol", v)) c.relTol = v;
    if (e->scalar("maxIter", v)) c.maxIter = (int)v;
    if (e->scalar("minIter", v)) c.minIter = (int)v;
    if (e->scalar("innerRelTol", v)) c.innerRelTol = v;
    if (e->scalar("maxRefine", v)) c.maxRefine = (int)v;
    if (e->word("precision", s)) {
        if (s == "mixed") c.mixed = true;
        else if (s == "double") c.mixed = false;
        else throw Error{RT_E_UNKNOWN_TYPE, "Unknown precision " + s + " (valid: double, mixed)"};
    }
    return c;
}

// ------------------------------------------------------------ model setup
namespace {

const char* kTypes[] = {"FENE-CRLog", "FENE-PLog",          "GiesekusLog", "Oldroyd-BLog", "Oldroyd-BRootk",
                        "PTTLog",     "Rolie-PolyLog", "SaramitoLog", "WhiteMetznerCYLog", "XPomPomLog",
                        "BMPLog",     "multiMode"};

double need(const Dict& d, const char* key, const std::string& dictName) {
    double v;
    if (!d.scalar(key, v))
        throw Error{RT_E_BAD_ARG, std::string("keyword ") + key + " is undefined in dictionary " + dictName};
    return v;
}

void parseModel(Model& M, const Dict& d, const std::string& dn) {
    std::string type;
    if (!d.word("type", type)) throw Error{RT_E_BAD_ARG, "keyword type is undefined in dictionary " + dn};
    ModelParams& p = M.prm;
    if (type == "Oldroyd-BLog") p.kind = M_OLDROYDB;
    else if (type == "Oldroyd-BRootk") p.kind = M_ROOTK;
    else if (type == "GiesekusLog") p.kind = M_GIESEKUS;
    else if (type == "FENE-PLog") p.kind = M_FENEP;
    else if (type == "PTTLog") p.kind = M_PTT_LINEAR;
    else if (type == "FENE-CRLog") p.kind = M_FENECR;
    else if (type == "WhiteMetznerCYLog") p.kind = M_WMCY;
    else if (type == "Rolie-PolyLog") p.kind = M_ROLIEPOLY;
    else if (type == "XPomPomLog") p.kind = M_XPP;
    else if (type == "SaramitoLog") p.kind = M_SARAMITO;
    else if (type == "BMPLog") p.kind = M_BMP;
    else {
        std::string msg = "Unknown constitutiveEq type " + type + "\n\nValid constitutiveEq types are :\n" +
                          std::to_string(sizeof(kTypes) / sizeof(kTypes[0])) + "\n(\n";
        for (const char* t : kTypes) msg += std::string(t) + "\n";
        throw Error{RT_E_UNKNOWN_TYPE, msg + ")"};
    }
    M.typeName = type;
    p.rho = need(d, "rho", dn);
    p.etaS = need(d, "etaS", dn);
    p.etaP = need(d, "etaP", dn);
    if (p.kind == M_PTT_LINEAR || p.kind == M_SARAMITO) {
        p.epsilon = need(d, "epsilon", dn);
        p.zeta = need(d, "zeta", dn);
    }
    if (p.kind == M_SARAMITO) {  // SaramitoLog.C:117-170
        p.tau0 = need(d, "tau0", dn);
        p.sarN = need(d, "n", dn);
        p.sarK = p.sarN == 1. ? need(d, "etaP", dn) : need(d, "k", dn);
        std::vector<double> dims;
        if (!d.vector("dims", dims) || dims.size() != 3)
            throw Error{RT_E_BAD_ARG, "keyword dims is undefined in dictionary " + dn};
        p.dimX = dims[0], p.dimY = dims[1], p.dimZ = dims[2];
        p.sarPTT = 0;
        if (p.sarN == 1.) {
            std::string f;
            if (!d.word("PTTfunction", f))
                throw Error{RT_E_BAD_ARG, "keyword PTTfunction is undefined in dictionary " + dn};
            if (f == "none") p.sarPTT = 0;
            else if (f == "linear") p.sarPTT = 1;
            else if (f == "exponential") p.sarPTT = 2;
            else
                throw Error{RT_E_BAD_ARG, "\nThe PTT function specified does not exist.\n\nAvailable PTT functions "
                                          "are:\n\n. none\n. linear\n. exponential"};
        }
    }
    // the keys each constructor looks up (e.g. RoliePolyLog.C:114-121, XPomPomLog.C:115-122)
    if (p.kind == M_ROLIEPOLY) {
        p.lambdaR = need(d, "lambdaR", dn);
        p.lambdaD = need(d, "lambdaD", dn);
        p.beta = need(d, "beta", dn);
        p.delta = need(d, "delta", dn);
        p.chiMax = need(d, "chiMax", dn);
    } else if (p.kind == M_XPP) {
        p.lambdaS = need(d, "lambdaS", dn);
        p.lambdaB = need(d, "lambdaB", dn);
        p.alpha = need(d, "alpha", dn);
        p.q = need(d, "q", dn);
        p.nXPP = need(d, "n", dn);
    } else {
        p.lambda = need(d, "lambda", dn);
    }
    if (p.kind == M_BMP) {  // BMPLog.C:113-119
        p.G0 = need(d, "G0", dn);
        p.kBMP = need(d, "k", dn);
        p.Phi0 = need(d, "Phi0", dn);
        p.PhiInf = need(d, "PhiInf", dn);
    }
    if (p.kind == M_WMCY) {  // WhiteMetznerCYLog.C:119-124
        p.wmM = need(d, "m", dn);
        p.wmN = need(d, "n", dn);
        p.wmK = need(d, "K", dn);
        p.wmL = need(d, "L", dn);
        p.wmA = need(d, "a", dn);
        p.wmB = need(d, "b", dn);
    }
    if (p.kind == M_GIESEKUS) p.alpha = need(d, "alpha", dn);
    if (p.kind == M_ROOTK) p.rootK = need(d, "k", dn);  // Oldroyd_BRootk.C:121
    if (p.kind == M_FENEP || p.kind == M_FENECR) p.L2 = need(d, "L2", dn);
    if (p.kind == M_PTT_LINEAR) {
        std::string f;
        if (!d.word("destructionFunctionType", f))
            throw Error{RT_E_BAD_ARG, "keyword destructionFunctionType is undefined in dictionary " + dn};
        if (f == "linear") p.kind = M_PTT_LINEAR;
        else if (f == "exponential") p.kind = M_PTT_EXP;
        else if (f == "generalized") {
            // PTTLog.C:133-152: the Mittag-Leffler function's alpha and beta,
            // its tolerance and term limit, and the gamma table of the series
            p.kind = M_PTT_GEN;
            const double a = need(d, "alpha", dn), b = need(d, "beta", dn);
            p.alpha = a;
            p.beta = b;
            double v;
            p.mlRtol = d.scalar("rTolMittagLeffler", v) ? v : 1e-12;
            int maxIter = d.scalar("maxIterMittagLeffler", v) ? (int)v : 200;
            if (a <= 0 || b <= 0)
                throw Error{RT_E_BAD_ARG, "Both alpha and beta should be positive values for the Mittag-Leffler "
                                          "function to converge."};
            std::vector<double> g{std::tgamma(b)};
            int k = 0;
            while (k < maxIter && g.back() < 1e+100) {
                g.push_back(std::tgamma(a * k + b));
                k++;
            }
            p.mlMaxIter = k;
            p.mlGammaBeta = g[0];
            M.mlGammaHost.assign(g.begin() + 1, g.end());  // [j] = gamma(alpha j + beta)
        } else
            throw Error{RT_E_BAD_ARG, f + " is not in enumeration: 3(linear exponential generalized)"};
    }
    // checkForStab (constitutiveEq.C:430-439): the key is mandatory
    std::string stab;
    if (!d.word("stabilization", stab))
        throw Error{RT_E_BAD_ARG, "keyword stabilization is undefined in dictionary " + dn};
    if (stab != "none" && stab != "BSD" && stab != "coupling")
        throw Error{RT_E_BAD_ARG, stab + " is not in enumeration: 3(none BSD coupling)"};
    M.stab = stab == "BSD" ? 1 : stab == "coupling" ? 2 : 0;
    // thermo functions: only the isothermal default (Constant) is on this path
    for (const char* tf : {"thermoLambda", "thermoEta", "thermoLambdaR", "thermoLambdaD", "thermoLambdaS",
                           "thermoLambdaB"}) {
        auto it = d.subDicts.find(tf);
        if (it != d.subDicts.end()) {
            std::string t = "Constant";
            it->second.word("type", t);
            if (t != "Constant")
                throw Error{RT_E_UNSUPPORTED, std::string(tf) + " type " + t +
                                                  " (non-isothermal) is not supported on the device path"};
        }
    }
    double q;
    if (d.scalar("ofInvQuirk", q)) p.ofInvQuirk = q != 0;
}

unsigned nb(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// staging upload of a host AoS array, then transpose/permute on the device
static void uploadAos(Ctx* ctx, DBuf& stage, const double* host, int64_t count) {
    const size_t bytes = (size_t)count * sizeof(double);
    if (stage.bytes < bytes) stage = ctx->alloc(bytes);
    RT_HIP(hipMemcpyAsync(stage.p, host, bytes, hipMemcpyHostToDevice, ctx->stream));
}

struct StagingRef {
    DBuf& buf;
};
static StagingRef staging(Ctx* c) { return StagingRef{c->stage}; }

// Uploads of a model's state and flow.  Each host array goes through the
// context's staging buffer and a reorder kernel; consecutive uploads share
// the buffer safely because the copies and kernels run in stream order.
// wait = true (the public setters): return once the caller's arrays have
// been read, since the caller may change them at once.  wait = false (the
// server): return at once; the server's requests keep their arrays
// unchanged until the ticket finishes, and from page-locked arrays the
// copies then overlap other work.
void modelUploadState(Model& M, const double* theta, const double* eigVals, const double* eigVecs,
                      const double* tau, bool wait) {
    if (!theta) throw Error{RT_E_BAD_ARG, "rt_model_set_state: theta is required (MUST_READ)"};
    Ctx* c = M.ctx;
    Mesh& m = *M.mesh;
    const int32_t N = m.N;
    RT_HIP(hipSetDevice(c->device));
    DBuf& st = staging(c).buf;
    const int32_t* perm = m.permD.as<int32_t>();
    std::vector<double> phi6;
    if (M.scalar()) {  // the fluidity: Phi [N] into component xx
        phi6.assign((size_t)6 * N, 0.0);
        for (int32_t i = 0; i < N; ++i) phi6[6 * (size_t)i] = theta[i];
        theta = phi6.data();
        eigVals = eigVecs = tau = nullptr;
        wait = true;  // phi6 is local
    }
    uploadAos(c, st, theta, (int64_t)6 * N);
    launch_permute_rows(c, N, 6, 6, st.as<double>(), perm, M.theta.as<double>());
    if (eigVals) {
        uploadAos(c, st, eigVals, (int64_t)9 * N);
        launch_eigvals_from_tensor(c, N, st.as<double>(), perm, M.eigVal.as<double>());
    }
    if (eigVecs) {
        uploadAos(c, st, eigVecs, (int64_t)9 * N);
        launch_aos_to_soa(c, N, 9, st.as<double>(), perm, M.eigVec.as<double>());
    }
    if (tau) {
        uploadAos(c, st, tau, (int64_t)6 * N);
        launch_permute_rows(c, N, 6, 6, st.as<double>(), perm, M.tau.as<double>());
    }
    if (wait) RT_HIP(hipStreamSynchronize(c->stream));
    M.hasState = true;
}

void modelUploadFlow(Model& M, const double* U, const double* Ub, const double* phi, bool wait) {
    Mesh& m = *M.mesh;
    if (!U || !phi || (m.NB > 0 && !Ub))
        throw Error{RT_E_BAD_ARG, "rt_model_set_flow: U, Ub and phi are required"};
    Ctx* c = M.ctx;
    RT_HIP(hipSetDevice(c->device));
    DBuf& st = staging(c).buf;
    uploadAos(c, st, U, (int64_t)3 * m.N);
    launch_permute_rows(c, m.N, 3, 4, st.as<double>(), m.permD.as<int32_t>(), M.U.as<double>());
    if (m.NB > 0) {  // [NB][3] -> [3][NB]
        uploadAos(c, st, Ub, (int64_t)3 * m.NB);
        launch_aos_to_soa(c, m.NB, 3, st.as<double>(), nullptr, M.Ub.as<double>());
    }
    RT_HIP(hipMemcpyAsync(M.phi.p, phi, ((size_t)m.F + m.NB) * 8, hipMemcpyHostToDevice, c->stream));
    launch_face_to_entry(c, m, M.phi.as<double>(), M.ePhi.as<double>());
    if (wait) RT_HIP(hipStreamSynchronize(c->stream));
    M.hasFlow = true;
}

int modelInvMask(Model& M) {
    if (!M.prm.ofInvQuirk) return 0;
    const Mesh& m = *M.mesh;
    const int32_t r = m.perm[0];
    double l[3];
    for (int q = 0; q < 3; ++q)
        RT_HIP(hipMemcpyAsync(&l[q], M.eigVal.as<double>() + (int64_t)q * m.N + r, sizeof(double),
                              hipMemcpyDeviceToHost, M.ctx->stream));
    RT_HIP(hipStreamSynchronize(M.ctx->stream));
    const double s = l[0] * l[0] + l[1] * l[1] + l[2] * l[2];
    int mask = 0;
    for (int q = 0; q < 3; ++q)
        if (l[q] * l[q] / s < 1e-15) mask |= 1 << q;
    return mask;
}

// The mesh and flow part of the per-row kernels' arguments.
AssembleArgs meshArgs(Model& M) {
    Mesh& m = *M.mesh;
    AssembleArgs a{};
    a.N = m.N;
    a.ellOff = m.ellOff.as<int64_t>();
    a.ellWidth = m.ellWidth.as<int32_t>();
    a.ellCol = m.ellCol.as<int32_t>();
    a.ellFid = m.ellFid.as<int32_t>();
    a.ellSlot = m.ellSlot.as<uint8_t>();
    a.eSf = m.eSf.as<double>();
    a.eW = m.eW.as<double>();
    a.ePhi = M.ePhi.as<double>();
    a.nEll = std::max<int64_t>(1, m.nEll);
    a.bPtr = m.bPtr.as<int32_t>();
    a.bFace = m.bFace.as<int32_t>();
    a.bPatch = m.bPatch.as<int32_t>();
    a.bcType = M.bcType.as<int32_t>();
    a.bcVal = M.bcVal.as<double>();
    a.patchT = M.patchT.as<double>();
    a.NB = m.NB;
    a.bSf = m.bSf.as<double>();
    a.phiB = M.phi.as<double>() + m.F;
    a.V = m.V.as<double>();
    a.U = M.U.as<double>();
    a.Ub = M.Ub.as<double>();
    return a;
}

void modelCorrectBegin(Model& M, double deltaT, const double* gradU, int iteration) {
    if (!M.hasState) throw Error{RT_E_BAD_ARG, "correct(): theta was never set (rt_model_set_state)"};
    if (!M.hasFlow) throw Error{RT_E_BAD_ARG, "correct(): U/phi were never set (rt_model_set_flow)"};
    if (!(deltaT > 0)) throw Error{RT_E_BAD_ARG, "correct(): deltaT must be positive"};
    Ctx* ctx = M.ctx;
    Mesh& m = *M.mesh;
    const int32_t N = m.N;
    const bool bicg = M.ctl.solver == SV_PBICG;
    M.kry.ensure(m, M.ctl.mixed, bicg);
    if (!M.diag.p) {
        M.diag = ctx->alloc((size_t)N * 8);
        M.aval = ctx->alloc((size_t)std::max<int64_t>(1, m.nEll) * 8);
        M.slotMask = ctx->alloc((size_t)std::max<int32_t>(1, m.nSlices) * 4);
        M.b = ctx->alloc((size_t)6 * N * 8);
        RT_HIP(hipMemsetAsync(M.b.p, 0, (size_t)6 * N * 8, ctx->stream));
    }
    if (modelKeepsFactor(M.prm.kind) && !M.fOld.p) M.fOld = ctx->alloc((size_t)N * 8);
    // inner iterations of a step keep theta.oldTime() of the step's first one
    const double* thetaOld = nullptr;
    if (iteration > 0) {
        if (!M.iterating)
            throw Error{RT_E_BAD_ARG, "correct(): an inner iteration needs a preceding iteration 0 of its step"};
        thetaOld = M.thetaOld.as<double>();
    } else if (iteration == 0) {
        if (!M.thetaOld.p) M.thetaOld = ctx->alloc((size_t)6 * N * 8);
        launch_copy(ctx, M.theta.p, M.thetaOld.p, (size_t)6 * N * 8);
        M.iterating = true;
    } else {
        M.iterating = false;  // a plain step
    }
    if (gradU) {
        if (!M.gradU.p) M.gradU = ctx->alloc((size_t)9 * N * 8);
        DBuf& st = staging(ctx).buf;
        uploadAos(ctx, st, gradU, (int64_t)9 * N);
        launch_aos_to_soa(ctx, N, 9, st.as<double>(), m.permD.as<int32_t>(), M.gradU.as<double>());
    }
    SolveState* st = M.kry.state.as<SolveState>();
    const bool mixed = M.ctl.mixed;
    // gAverage(theta) for the normalisation factor of the initial residual
    launch_mean6(ctx, N, M.theta.as<double>(), M.kry.partials.as<double>(), st->c,
                 M.kry.grid);
    AssembleArgs a = meshArgs(M);
    a.gradU = gradU ? M.gradU.as<double>() : nullptr;
    a.thetaOld = thetaOld;
    a.theta = M.theta.as<double>();
    a.eigVal = M.eigVal.as<double>();
    a.eigVec = M.eigVec.as<double>();
    a.tauOld = M.prm.kind == M_SARAMITO ? M.tau.as<double>()
               : M.scalar()                ? M.parentTau->as<double>()
                                           : nullptr;
    a.aux = M.prm.kind == M_BMP ? M.fluid->impl.theta.as<double>() : nullptr;
    a.rDeltaT = 1.0 / deltaT;
    a.relax = M.relax;
    a.prm = M.prm;
    a.invMask = modelInvMask(M);
    a.aval = M.aval.as<double>();
    a.slotMask = M.slotMask.as<uint32_t>();
    a.diag = M.diag.as<double>();
    a.b = M.b.as<double>();
    a.fOld = modelKeepsFactor(M.prm.kind) ? M.fOld.as<double>() : nullptr;
    a.atval = bicg ? M.kry.atval.as<double>() : nullptr;
    a.partials = M.kry.partials.as<double>();
    a.cs = st->c;
    a.nnzOut = &st->misc[0];
    for (int q = 0; q < 6; ++q) M.valid6[q] = M.scalar() ? (q == 0 ? 1 : -1) : m.valid[q];
    for (int q = 0; q < 6; ++q) a.valid[q] = M.valid6[q];
    a.rD = M.kry.rD.as<double>();
    a.ctl = kryCtl(M.ctl, mixed, m);
    // the red-black solve opens with a sweep that writes r and r0 on the
    // second colour from the eliminated first colour; the assembly's own
    // residual then only feeds the normalisation and the initial residual
    a.r = a.ctl.rb ? nullptr : (mixed ? M.kry.rs.p : M.kry.r.p);
    a.r0 = a.ctl.rb ? nullptr : (mixed ? M.kry.r0s.p : M.kry.r0.p);
    a.nRed = m.nRed;
    a.mixed = mixed ? 1 : 0;
    a.noConv = M.scheme == DS_NONE ? 1 : 0;
    a.bounded = M.bounded ? 1 : 0;
    a.souT = nullptr;
    if (M.scheme != DS_UPWIND && M.scheme != DS_NONE) {
        if (!M.gradTh.p) M.gradTh = ctx->alloc((size_t)18 * N * 8);
        if (!M.souT.p) M.souT = ctx->alloc((size_t)6 * N * 8);
        launch_defc(ctx, a, M.scheme, m.C4.as<double>(), M.gradTh.as<double>(), M.souT.as<double>(), M.kry.grid);
        a.souT = M.souT.as<double>();
    }
    launch_assemble(ctx, a, M.kry.grid);
    if (!M.pendingSolve) M.pendingSolve = new SolveArgs{};
    SolveArgs& sa = *M.pendingSolve;
    sa = SolveArgs{};
    sa.mesh = &m;
    sa.kry = &M.kry;
    sa.diag = M.diag.as<double>();
    sa.aval = M.aval.as<double>();
    sa.slotMask = M.slotMask.as<uint32_t>();
    sa.atval = bicg ? M.kry.atval.as<double>() : nullptr;
    sa.b = M.b.as<double>();
    sa.x = M.theta.as<double>();
    sa.valid = M.valid6;
    sa.ctl = &M.ctl;
    sa.xbarKnown = true;
    sa.initDone = true;
    solve_begin(ctx, sa);
}

void modelCorrectFinish(Model& M) {
    Ctx* ctx = M.ctx;
    solve_finish(M.kry, M.perf);
    if (M.scalar()) return;  // the fluidity has no eigen state or tau
    launch_eig_tau(ctx, M.mesh->N, M.prm, M.theta.as<double>(), M.eigVal.as<double>(),
                   M.eigVec.as<double>(), M.tau.as<double>(),
                   modelKeepsFactor(M.prm.kind) ? M.fOld.as<double>() : nullptr);
}

bool modelCorrectAdvance(Model& M, bool block) {
    if (!solve_advance(M.ctx, M.kry, block)) return false;
    modelCorrectFinish(M);
    return true;
}

void modelCorrect(Model& M, double deltaT, const double* gradU, int iteration) {
    modelCorrectBegin(M, deltaT, gradU, iteration);
    while (!modelCorrectAdvance(M, true)) {
    }
    RT_HIP(hipStreamSynchronize(M.ctx->stream));
}

// device bytes of one model on an existing mesh (state, flow, matrix, Krylov)
size_t modelBytesEstimate(int32_t N, int32_t F, int32_t NB, int64_t nEll, bool mixed, bool gradU, bool download) {
    const double n = N, f = F, nb = NB, ne = (double)nEll;
    double by = 0;
    // model state + flow + matrix: theta tau eigVal eigVec U b diag fOld, Ub, phi, bc, ePhi, aval
    by += 8 * n * (6 + 6 + 3 + 9 + 4 + 6 + 1 + 1) + 8 * (3 * nb + f + nb) + 6 * 8 * nb + 8 * ne + 8 * ne +
          4 * n / 64;
    // Krylov: r r0 p v t (and the mixed solve's correction e) x 6 components,
    // rD, ahat, partials, state
    by += 6 * n * (mixed ? 6 * 4 : 5 * 8) + 8 * n + ne * (mixed ? 4 : 8) + 2048 * 24 * 8 + 4096;
    // staging for AoS transfers (9 comps)
    by += 9 * 8 * n;
    if (gradU) by += 9 * 8 * n;      // the prescribed gradient
    if (download) by += 2 * 6 * 8 * n;  // theta and tau in the caller's order on their way home
    // size classes round each buffer up by at most 1/8, and the cache hands
    // out blocks up to 1.5 times a request; 1.25 covers the usual mix
    return (size_t)(by * 1.25) + (1u << 20);
}

}  // namespace rt

using namespace rt;

// ======================================================================= C-ABI
namespace {

// The wedge patch's faceT ([OF9] wedgePolyPatch::calcGeometry): n the mean of
// the patch's unit face normals; the centre-plane normal cn takes each
// component of n to sign(n_i) (max(|n_i|, 1/2) - 1/2) and is normalised, so
// it is the coordinate axis nearest n; faceT is the rotation taking cn onto n
// ([OF9] transform.H rotationTensor: s I + (1 - s) a a / |a|^2 + n cn - cn n,
// s = cn . n, a = cn x n).  A patch whose cn is no coordinate axis, or whose
// n is one (no angle), is refused as the reference refuses it.
void wedgeFaceT(Ctx* c, const Mesh& m, int patch, double T[9]) {
    const rt_patch& p = m.patches[patch];
    if (p.size <= 0) throw Error{RT_E_BAD_ARG, "rt_model_set_bc: wedge patch " + std::to_string(patch) + " has no faces"};
    std::vector<double> sf((size_t)3 * p.size);
    for (int d = 0; d < 3; ++d)
        RT_HIP(hipMemcpyAsync(sf.data() + (size_t)d * p.size, m.bSf.as<double>() + (int64_t)d * m.NB + p.start,
                              (size_t)p.size * 8, hipMemcpyDeviceToHost, c->stream));
    RT_HIP(hipStreamSynchronize(c->stream));
    double n[3] = {0, 0, 0};
    for (int k = 0; k < p.size; ++k) {
        const double x = sf[k], y = sf[(size_t)p.size + k], z = sf[(size_t)2 * p.size + k];
        const double mag = std::sqrt(x * x + y * y + z * z);
        n[0] += x / mag;
        n[1] += y / mag;
        n[2] += z / mag;
    }
    for (double& v : n) v /= p.size;
    double cn[3];
    for (int d = 0; d < 3; ++d) {
        const double sg = n[d] > 0 ? 1.0 : (n[d] < 0 ? -1.0 : 0.0);
        cn[d] = sg * (std::max(std::fabs(n[d]), 0.5) - 0.5);
    }
    const double mc = std::sqrt(cn[0] * cn[0] + cn[1] * cn[1] + cn[2] * cn[2]);
    if (mc == 0) throw Error{RT_E_BAD_ARG, "rt_model_set_bc: wedge patch " + std::to_string(patch) +
                                               " is not aligned with a coordinate axis"};
    for (double& v : cn) v /= mc;
    if (std::fabs(cn[0] + cn[1] + cn[2]) < 1 - 1e-15)
        throw Error{RT_E_BAD_ARG, "rt_model_set_bc: wedge patch " + std::to_string(patch) +
                                      " is not aligned with a coordinate axis"};
    const double s = cn[0] * n[0] + cn[1] * n[1] + cn[2] * n[2];
    const double ax[3] = {cn[1] * n[2] - cn[2] * n[1], cn[2] * n[0] - cn[0] * n[2], cn[0] * n[1] - cn[1] * n[0]};
    const double ms = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2];
    if (std::sqrt(ms) < 1e-15)
        throw Error{RT_E_BAD_ARG, "rt_model_set_bc: wedge patch " + std::to_string(patch) +
                                      " lies in a coordinate plane (no wedge angle)"};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            T[3 * i + j] = (i == j ? s : 0.0) + (1 - s) * ax[i] * ax[j] / ms + (n[i] * cn[j] - cn[i] * n[j]);
}

template <class F>
int guarded(rt::Ctx* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->lastError.clear();
        return RT_OK;
    } catch (const rt::Error& e) {
        if (ctx) ctx->lastError = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->lastError = e.what();
        return RT_E_BAD_ARG;
    }
}

}  // namespace

namespace rt {
static void ctxTeardown(rt_ctx* ctx) {
    Ctx& x = ctx->impl;
    (void)hipSetDevice(x.device);
    (void)hipStreamSynchronize(x.stream);
    for (auto& p : x.pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : x.eventPool) (void)hipEventDestroy(e);
    if (x.pinned) (void)hipHostFree(x.pinned);
    x.stage.release();
    x.trimCache(0);
    for (void* p : x.pinnedBlocks) (void)hipHostFree(p);
    if (x.stream) (void)hipStreamDestroy(x.stream);
    delete ctx;
}

void ctxUnref(Ctx* c) {
    if (--c->refs == 0) ctxTeardown(c->handle);
}
}  // namespace rt

extern "C" {

int rt_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int rt_ctx_create(int device, size_t hbm_budget, rt_ctx** out) {
    if (!out) return RT_E_BAD_ARG;
    *out = nullptr;
    int n = rt_device_count();
    if (n <= 0 || device < 0 || device >= n) return RT_E_NO_DEVICE;
    auto* c = new rt_ctx;
    c->impl.handle = c;
    int rc = guarded(&c->impl, [&] {
        Ctx& x = c->impl;
        x.device = device;
        RT_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        RT_HIP(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw Error{RT_E_NO_DEVICE, std::string("device ") + std::to_string(device) + " is " +
                                            prop.gcnArchName + "; this library is built for gfx950"};
        x.numCUs = prop.multiProcessorCount;
        x.budget = hbm_budget;
        RT_HIP(hipStreamCreateWithFlags(&x.stream, hipStreamNonBlocking));
        noallocInit(device);
    });
    if (rc != RT_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return RT_OK;
}


void rt_ctx_destroy(rt_ctx* ctx) {
    if (ctx) ctxUnref(&ctx->impl);
}

const char* rt_last_error(const rt_ctx* ctx) { return ctx ? ctx->impl.lastError.c_str() : "NULL context"; }

size_t rt_ctx_bytes_in_use(const rt_ctx* ctx) { return ctx ? ctx->impl.inUse : 0; }

int rt_ctx_set_profiling(rt_ctx* ctx, int enable) {
    if (!ctx) return RT_E_BAD_ARG;
    return guarded(&ctx->impl, [&] {
        ctx->impl.flushProfile();
        ctx->impl.profiling = enable != 0;
    });
}

int rt_ctx_kernel_stats(rt_ctx* ctx, rt_kernel_stat* out, int cap) {
    if (!ctx) return -RT_E_BAD_ARG;
    int n = 0;
    int rc = guarded(&ctx->impl, [&] {
        ctx->impl.flushProfile();
        for (auto& kv : ctx->impl.stats) {
            if (out && n < cap) {
                std::memset(out[n].name, 0, sizeof(out[n].name));
                std::strncpy(out[n].name, kv.first.c_str(), sizeof(out[n].name) - 1);
                out[n].launches = kv.second.launches;
                out[n].total_ms = kv.second.ms;
                out[n].bytes = kv.second.bytes;
            }
            ++n;
        }
    });
    return rc == RT_OK ? n : -rc;
}

int rt_host_alloc(rt_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out || bytes == 0) return RT_E_BAD_ARG;
    *out = nullptr;
    return guarded(&ctx->impl, [&] {
        RT_HIP(hipSetDevice(ctx->impl.device));
        void* p = nullptr;
        if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) {
            (void)hipGetLastError();
            throw Error{RT_E_OOM, "rt_host_alloc: page-locked allocation of " + std::to_string(bytes) +
                                      " bytes failed"};
        }
        *out = p;
    });
}

void rt_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

int rt_ctx_copy_bench(rt_ctx* ctx, size_t bytes, int reps, double* ms) {
    if (!ctx || !ms || bytes < 16 || reps <= 0) return RT_E_BAD_ARG;
    return guarded(&ctx->impl, [&] {
        Ctx& c = ctx->impl;
        RT_HIP(hipSetDevice(c.device));
        bytes = bytes / 16 * 16;
        DBuf a = c.alloc(bytes), b = c.alloc(bytes);
        RT_HIP(hipMemsetAsync(a.p, 0, bytes, c.stream));
        launch_copy(&c, a.p, b.p, bytes);  // warm-up
        hipEvent_t e0 = c.getEvent(), e1 = c.getEvent();
        RT_HIP(hipEventRecord(e0, c.stream));
        for (int i = 0; i < reps; ++i) launch_copy(&c, a.p, b.p, bytes);
        RT_HIP(hipEventRecord(e1, c.stream));
        RT_HIP(hipEventSynchronize(e1));
        float t = 0;
        RT_HIP(hipEventElapsedTime(&t, e0, e1));
        c.eventPool.push_back(e0);
        c.eventPool.push_back(e1);
        *ms = t / reps;
    });
}

void rt_ctx_reset_stats(rt_ctx* ctx) {
    if (!ctx) return;
    try {
        ctx->impl.flushProfile();
    } catch (...) {
    }
    ctx->impl.stats.clear();
}

int rt_mesh_create(rt_ctx* ctx, const rt_mesh_desc* desc, int renumber, rt_mesh** out) {
    if (!ctx || !out) return RT_E_BAD_ARG;
    *out = nullptr;
    auto* m = new rt_mesh;
    int rc = guarded(&ctx->impl, [&] {
        RT_HIP(hipSetDevice(ctx->impl.device));
        buildMesh(&ctx->impl, desc, renumber, m->impl);
    });
    if (rc != RT_OK) {
        delete m;
        return rc;
    }
    ctxRef(&ctx->impl);
    *out = m;
    return RT_OK;
}

void rt_mesh_destroy(rt_mesh* mesh) {
    if (!mesh) return;
    Ctx* c = mesh->impl.ctx;
    delete mesh;
    if (c) ctxUnref(c);
}

int rt_mesh_valid_components(const rt_mesh* mesh, int32_t out[6]) {
    if (!mesh || !out) return RT_E_BAD_ARG;
    for (int q = 0; q < 6; ++q) out[q] = mesh->impl.valid[q];
    return RT_OK;
}

int rt_model_create(rt_ctx* ctx, const char* name, const rt_mesh* mesh, const char* dict,
                    rt_model** out) {
    if (!ctx || !out) return RT_E_BAD_ARG;
    *out = nullptr;
    if (!mesh) {
        ctx->impl.lastError = "rt_model_create: NULL mesh";
        return RT_E_BAD_ARG;
    }
    auto* mo = new rt_model;
    int rc = guarded(&ctx->impl, [&] {
        if (!dict) throw Error{RT_E_BAD_ARG, "rt_model_create: NULL dictionary"};
        RT_HIP(hipSetDevice(ctx->impl.device));
        Dict full = parseDict(dict);
        createModel(mo->impl, &ctx->impl, const_cast<Mesh*>(&mesh->impl), name ? name : "",
                     parametersOf(full), "constitutiveProperties.parameters");
    });
    if (rc != RT_OK) {
        destroyModes(mo->impl);
        delete mo;
        return rc;
    }
    ctxRef(&ctx->impl);
    *out = mo;
    return RT_OK;
}

void rt_model_destroy(rt_model* model) {
    if (!model) return;
    Ctx* c = model->impl.ctx;
    destroyModes(model->impl);
    delete model->impl.pendingSolve;
    delete model;
    if (c) ctxUnref(c);
}

rt_model* rt_model_fluidity(rt_model* model) { return model ? model->impl.fluid : nullptr; }

int rt_model_mode_count(const rt_model* model) {
    return model ? (int)model->impl.modes.size() : 0;
}

rt_model* rt_model_mode(rt_model* model, int i) {
    if (!model || i < 0 || i >= (int)model->impl.modes.size()) return nullptr;
    return model->impl.modes[i];
}

}  // extern "C"

namespace rt {

void destroyModes(Model& M) {
    for (rt_model* x : M.modes) {
        destroyModes(x->impl);
        delete x->impl.pendingSolve;
        delete x;
    }
    M.modes.clear();
    if (M.fluid) {
        delete M.fluid->impl.pendingSolve;
        delete M.fluid;
        M.fluid = nullptr;
    }
}

// A model on ctx/mesh from its parameters dictionary.  multiMode
// (multiMode.C:43-83): one model per entry of `models`, named name + keyword,
// and only the tau sum held here.
void createModel(Model& M, Ctx* c, Mesh* mesh, const std::string& name, const Dict& d, const std::string& dn) {
    M.ctx = c;
    M.mesh = mesh;
    M.name = name;
    const int32_t N = M.mesh->N, NB = M.mesh->NB;
    std::string type;
    if (d.word("type", type) && type == "multiMode") {
        std::vector<std::pair<std::string, Dict>> list;
        if (!d.dictList("models", list) || list.empty())
            throw Error{RT_E_BAD_ARG, "keyword models is undefined in dictionary " + dn};
        M.typeName = "multiMode";
        double etaS = 0, rho = 0;
        for (auto& e : list) {
            auto* sub = new rt_model;
            M.modes.push_back(sub);
            createModel(sub->impl, c, mesh, name + e.first, e.second, dn + ".models." + e.first);
            if (sub->impl.isMulti()) throw Error{RT_E_BAD_ARG, "multiMode: a mode cannot itself be multiMode"};
            etaS += sub->impl.prm.etaS;  // multiMode.C:80
            rho += sub->impl.prm.rho;
        }
        M.prm.etaS = etaS;
        M.prm.rho = rho / list.size();  // multiMode::rho() averages
        M.tau = c->alloc((size_t)6 * N * 8);
        RT_HIP(hipMemsetAsync(M.tau.p, 0, M.tau.bytes, c->stream));
        M.ctl = parseControls(nullptr);
        RT_HIP(hipStreamSynchronize(c->stream));
        return;
    }
    parseModel(M, d, dn);
    allocModelState(M);
    if (M.prm.kind == M_PTT_GEN) {
        M.mlGamma = c->alloc(std::max<size_t>(1, M.mlGammaHost.size()) * 8);
        if (!M.mlGammaHost.empty())
            RT_HIP(hipMemcpyAsync(M.mlGamma.p, M.mlGammaHost.data(), M.mlGammaHost.size() * 8,
                                  hipMemcpyHostToDevice, c->stream));
        RT_HIP(hipStreamSynchronize(c->stream));
        M.prm.mlGamma = M.mlGamma.as<double>();
    }
    if (M.prm.kind == M_BMP) {
        // the fluidity equation Phi<name> (BMPLog.C:98-111, 151-163)
        auto* fl = new rt_model;
        M.fluid = fl;
        Model& F = fl->impl;
        F.ctx = c;
        F.mesh = mesh;
        F.name = name;
        F.typeName = "BMPLog fluidity";
        F.prm = M.prm;
        F.prm.kind = M_FLUIDITY;
        F.parentTau = &M.tau;
        allocModelState(F);
    }
}

// device state of a (non-multiMode) model: fields, flow, boundary conditions,
// with the reference's defaults (zero tau, identity eigen state, zeroGradient)
void allocModelState(Model& M) {
    Ctx* c = M.ctx;
    const int32_t N = M.mesh->N, NB = M.mesh->NB;
    {
        M.theta = c->alloc((size_t)6 * N * 8);
        M.tau = c->alloc((size_t)6 * N * 8);
        M.eigVal = c->alloc((size_t)3 * N * 8);
        M.eigVec = c->alloc((size_t)9 * N * 8);
        M.U = c->alloc((size_t)4 * N * 8);
        M.Ub = c->alloc((size_t)3 * std::max(1, NB) * 8);
        M.phi = c->alloc(((size_t)M.mesh->F + NB) * 8);
        M.ePhi = c->alloc((size_t)std::max<int64_t>(1, M.mesh->nEll) * 8);
        M.bcType = c->alloc((size_t)std::max(1, M.mesh->nPatches) * 4);
        M.bcVal = c->alloc((size_t)6 * std::max(1, NB) * 8);
        M.patchT = c->alloc((size_t)9 * std::max(1, M.mesh->nPatches) * 8);
        RT_HIP(hipMemsetAsync(M.bcVal.p, 0, M.bcVal.bytes, c->stream));
        RT_HIP(hipMemsetAsync(M.patchT.p, 0, M.patchT.bytes, c->stream));
        RT_HIP(hipMemsetAsync(M.tau.p, 0, M.tau.bytes, c->stream));
        M.hBcType.resize(M.mesh->nPatches);
        for (int p = 0; p < M.mesh->nPatches; ++p)
            M.hBcType[p] = M.mesh->patches[p].kind == RT_PATCH_EMPTY ? RT_BC_EMPTY : RT_BC_ZEROGRADIENT;
        if (!M.hBcType.empty())
            RT_HIP(hipMemcpyAsync(M.bcType.p, M.hBcType.data(), M.hBcType.size() * 4,
                                  hipMemcpyHostToDevice, c->stream));
        // READ_IF_PRESENT default of eigVals/eigVecs: identity tensor
        launch_fill(c, M.eigVal.as<double>(), (int64_t)3 * N, 1.0);
        launch_fill(c, M.eigVec.as<double>(), (int64_t)9 * N, 0.0);
        for (int q : {0, 4, 8}) launch_fill(c, M.eigVec.as<double>() + (int64_t)q * N, N, 1.0);
        M.ctl = parseControls(nullptr);
        RT_HIP(hipStreamSynchronize(c->stream));
    }
}

// multiMode models whose modes' solves may run as one group: two or more
// modes on the model's own context, no fluidity equation (BMP solves it
// before theta), and no profiling (a profiled run charges each launch to
// one solve)
bool multiGroupable(const Model& M) {
    if (M.modes.size() < 2 || M.ctx->profiling) return false;
    for (const rt_model* mode : M.modes)
        if (mode->impl.ctx != M.ctx || mode->impl.fluid) return false;
    return true;
}

}  // namespace rt

extern "C" {

const char* rt_model_type(const rt_model* model) { return model ? model->impl.typeName.c_str() : ""; }

// multiMode forwards a setter to every mode (first failure wins)
#define RT_FORWARD_TO_MODES(call)                                   \
    if (model->impl.isMulti()) {                                    \
        for (rt_model* mode_ : model->impl.modes) {                 \
            rt_model* model = mode_;                                \
            int rc_ = (call);                                       \
            if (rc_ != RT_OK) {                                     \
                mode_->impl.ctx->lastError = "mode " + mode_->impl.name + ": " + mode_->impl.ctx->lastError; \
                return rc_;                                         \
            }                                                       \
        }                                                           \
        return RT_OK;                                               \
    }

int rt_model_set_state(rt_model* model, const double* theta, const double* eigVals,
                       const double* eigVecs, const double* tau) {
    if (!model) return RT_E_BAD_ARG;
    Model& M = model->impl;
    if (M.isMulti()) {
        M.ctx->lastError = "rt_model_set_state: a multiMode model has no state of its own; set each mode's "
                           "(rt_model_mode)";
        return RT_E_BAD_ARG;
    }
    return guarded(M.ctx, [&] { modelUploadState(M, theta, eigVals, eigVecs, tau, true); });
}
int rt_model_set_bc(rt_model* model, int patch, int bcType, const double* values) {
    if (!model) return RT_E_BAD_ARG;
    RT_FORWARD_TO_MODES(rt_model_set_bc(model, patch, bcType, values));
    Model& M = model->impl;
    return guarded(M.ctx, [&] {
        Mesh& m = *M.mesh;
        if (patch < 0 || patch >= m.nPatches) throw Error{RT_E_BAD_ARG, "rt_model_set_bc: no such patch"};
        if (bcType < RT_BC_ZEROGRADIENT || bcType > RT_BC_WEDGE)
            throw Error{RT_E_UNKNOWN_TYPE, "rt_model_set_bc: unknown patchField type"};
        const bool emptyPatch = m.patches[patch].kind == RT_PATCH_EMPTY;
        if (emptyPatch != (bcType == RT_BC_EMPTY))
            throw Error{RT_E_BAD_ARG, "rt_model_set_bc: empty patches take exactly the empty condition"};
        if (bcType == RT_BC_FIXEDVALUE && !values)
            throw Error{RT_E_BAD_ARG, "rt_model_set_bc: fixedValue needs values"};
        // a scalar field's symmetry or wedge value is its cell value: the
        // fluidity takes the zeroGradient form (the same coefficients while the
        // patch flux is zero)
        const bool transform = bcType == RT_BC_SYMMETRY || bcType == RT_BC_WEDGE;
        // copies on the model's stream, ordered with the work already queued
        // there; the host arrays are read before the call returns
        Ctx* c = M.ctx;
        if (bcType == RT_BC_WEDGE && !M.scalar()) {
            double T[9];
            wedgeFaceT(c, m, patch, T);
            RT_HIP(hipMemcpyAsync(M.patchT.as<double>() + 9 * (int64_t)patch, T, sizeof T, hipMemcpyHostToDevice,
                                  c->stream));
            RT_HIP(hipStreamSynchronize(c->stream));
        }
        M.hBcType[patch] = transform && M.scalar() ? RT_BC_ZEROGRADIENT : bcType;
        RT_HIP(hipMemcpyAsync(M.bcType.as<int32_t>() + patch, &M.hBcType[patch], 4, hipMemcpyHostToDevice,
                              c->stream));
        std::vector<double> cols;
        if (bcType == RT_BC_FIXEDVALUE) {
            const rt_patch& p = m.patches[patch];
            cols.resize((size_t)6 * p.size);
            const int nc = M.scalar() ? 1 : 6;  // the fluidity takes one value per face
            for (int q = 0; q < 6; ++q) {
                double* col = cols.data() + (size_t)q * p.size;
                for (int k = 0; k < p.size; ++k) col[k] = q < nc ? values[nc * (size_t)k + q] : 0.0;
                RT_HIP(hipMemcpyAsync(M.bcVal.as<double>() + (int64_t)q * m.NB + p.start, col,
                                      (size_t)p.size * 8, hipMemcpyHostToDevice, c->stream));
            }
        }
        RT_HIP(hipStreamSynchronize(c->stream));
    });
}

int rt_model_set_flow(rt_model* model, const double* U, const double* Ub, const double* phi) {
    if (!model) return RT_E_BAD_ARG;
    RT_FORWARD_TO_MODES(rt_model_set_flow(model, U, Ub, phi));
    Model& M = model->impl;
    return guarded(M.ctx, [&] {
        modelUploadFlow(M, U, Ub, phi, true);
        if (M.fluid) {
            const int rc = rt_model_set_flow(M.fluid, U, Ub, phi);
            if (rc != RT_OK) throw Error{rc, "fluidity: " + M.ctx->lastError};
        }
    });
}

int rt_model_set_solver(rt_model* model, const char* controls) {
    if (!model) return RT_E_BAD_ARG;
    RT_FORWARD_TO_MODES(rt_model_set_solver(model, controls));
    Model& M = model->impl;
    return guarded(M.ctx, [&] { M.ctl = parseControls(controls); });
}

int rt_model_set_relax(rt_model* model, double factor) {
    if (!model) return RT_E_BAD_ARG;
    RT_FORWARD_TO_MODES(rt_model_set_relax(model, factor));
    model->impl.relax = factor;
    return RT_OK;
}

int rt_model_set_scheme(rt_model* model, const char* divScheme) {
    if (!model) return RT_E_BAD_ARG;
    RT_FORWARD_TO_MODES(rt_model_set_scheme(model, divScheme));
    Model& M = model->impl;
    std::vector<std::string> tok;
    {
        std::string cur;
        for (const char* c = divScheme ? divScheme : ""; *c; ++c) {
            if (std::isspace((unsigned char)*c) || *c == ';') {
                if (!cur.empty()) tok.push_back(cur);
                cur.clear();
            } else {
                cur += *c;
            }
        }
        if (!cur.empty()) tok.push_back(cur);
    }
    // [OF9] convectionScheme::New: an optional `bounded` wrapper, then the
    // scheme; "Gauss upwind" builds the same matrix as "GaussDefCmpw upwind"
    const bool bounded = !tok.empty() && tok[0] == "bounded";
    if (bounded) tok.erase(tok.begin());
    if (tok.size() == 2 && tok[0] == "Gauss" && tok[1] == "upwind") tok.erase(tok.begin());
    if (tok.size() == 2 && tok[0] == "GaussDefCmpw") tok.erase(tok.begin());
    if (tok.size() != 1) {
        M.ctx->lastError = std::string("rt_model_set_scheme: expected \"[bounded] GaussDefCmpw <limiter>\", got \"") +
                           (divScheme ? divScheme : "") + "\" (only the GaussDefCmpw convection scheme is supported)";
        return tok.size() == 2 ? RT_E_UNSUPPORTED : RT_E_BAD_ARG;
    }
    static const std::pair<const char*, int> names[] = {{"upwind", DS_UPWIND},   {"cubista", DS_CUBISTA},
                                                        {"minmod", DS_MINMOD},   {"smart", DS_SMART},
                                                        {"waceb", DS_WACEB},     {"superbee", DS_SUPERBEE},
                                                        {"none", DS_NONE}};
    int scheme = -1;
    for (const auto& n : names)
        if (tok[0] == n.first) scheme = n.second;
    if (scheme < 0) {
        // limiters.H:88-97
        M.ctx->lastError = "Error in div(phi, theta" + M.name +
                           ")\n\nThe deferred limited scheme is not specified or does not exist.\n\n"
                           "Valid schemes are:\n\n. upwind\n. cubista\n. minmod\n. smart\n. waceb\n"
                           ". superbee\n. none (no convection)";
        return RT_E_UNKNOWN_TYPE;
    }
    if (scheme != DS_UPWIND && scheme != DS_NONE && !M.mesh->C4.p) {
        M.ctx->lastError = "rt_model_set_scheme: the " + tok[0] +
                           " correction needs cell centres; the mesh was created without them";
        return RT_E_BAD_ARG;
    }
    M.scheme = scheme;
    M.bounded = bounded;
    return RT_OK;
}

int rt_model_correct(rt_model* model, double deltaT, const double* gradU) {
    return rt_model_correct_iter(model, deltaT, gradU, -1);
}

int rt_model_correct_iter(rt_model* model, double deltaT, const double* gradU, int iteration) {
    if (!model) return RT_E_BAD_ARG;
    Model& M = model->impl;
    if (M.isMulti() && multiGroupable(M)) {
        // multiMode::correct (multiMode.C:250-260) with the modes' Krylov
        // rounds batched: every mode's assembly is queued, then their solves
        // run as one group (one launch per kernel for all modes; each mode's
        // reductions stay its own, so the fields are bit-identical to solving
        // the modes in turn), and each mode's eigen/tau follows its own solve;
        // then tau = the sum of the modes' tau (multiMode.C:219-227)
        const int rc = guarded(M.ctx, [&] {
            RT_HIP(hipSetDevice(M.ctx->device));
            for (rt_model* mode : M.modes) {
                try {
                    modelCorrectBegin(mode->impl, deltaT, gradU, iteration);
                } catch (const Error& e) {
                    throw Error{e.code, "mode " + mode->impl.name + ": " + e.msg};
                }
            }
            bool same = true, mixed0 = false, rb0 = false;
            for (size_t i = 0; i < M.modes.size(); ++i) {
                bool mx = false, rb = false;
                if (!solve_groupable(M.modes[i]->impl.kry, &mx, &rb)) same = false;
                if (i == 0) mixed0 = mx, rb0 = rb;
                else if (mx != mixed0 || rb != rb0) same = false;
            }
            if (!same) {  // PBiCG, or modes with different solve forms: one after another
                for (rt_model* mode : M.modes)
                    while (!modelCorrectAdvance(mode->impl, true)) {
                    }
            } else {
                SolveGroup G;
                G.ctx = M.ctx;
                G.mixed = mixed0;
                G.rb = rb0;
                for (rt_model* mode : M.modes) group_join(G, mode->impl.kry, M.ctx->stream);
                std::vector<Krylov*> done;
                while (!G.members.empty()) {
                    RT_HIP(hipStreamSynchronize(M.ctx->stream));  // the queued rounds' states are back
                    done.clear();
                    group_advance(G, done);
                    for (Krylov* K : done)
                        for (rt_model* mode : M.modes)
                            if (&mode->impl.kry == K) modelCorrectFinish(mode->impl);
                }
                RT_HIP(hipStreamSynchronize(M.ctx->stream));  // the group's tables are read no more
            }
            const int64_t n = (int64_t)6 * M.mesh->N;
            launch_fill(M.ctx, M.tau.as<double>(), n, 0.0);
            for (rt_model* mode : M.modes) launch_add(M.ctx, n, mode->impl.tau.as<double>(), M.tau.as<double>());
            RT_HIP(hipStreamSynchronize(M.ctx->stream));
        });
        if (rc != RT_OK) return rc;
        for (rt_model* mode : M.modes)
            for (int q = 0; q < 6; ++q)
                if (mode->impl.perf[q].solved && !mode->impl.perf[q].converged) return RT_W_NOT_CONVERGED;
        return RT_OK;
    }
    if (M.isMulti()) {
        // multiMode::correct (multiMode.C:250-260): each mode in turn, then
        // tau = sum of the modes' tau (multiMode.C:219-227)
        int worst = RT_OK;
        for (rt_model* mode : M.modes) {
            const int rc = rt_model_correct_iter(mode, deltaT, gradU, iteration);
            if (rc != RT_OK && rc != RT_W_NOT_CONVERGED) {
                M.ctx->lastError = "mode " + mode->impl.name + ": " + mode->impl.ctx->lastError;
                return rc;
            }
            if (rc == RT_W_NOT_CONVERGED) worst = rc;
        }
        const int rc = guarded(M.ctx, [&] {
            const int64_t n = (int64_t)6 * M.mesh->N;
            launch_fill(M.ctx, M.tau.as<double>(), n, 0.0);
            for (rt_model* mode : M.modes) launch_add(M.ctx, n, mode->impl.tau.as<double>(), M.tau.as<double>());
        });
        return rc != RT_OK ? rc : worst;
    }
    int rc = guarded(M.ctx, [&] {
        RT_HIP(hipSetDevice(M.ctx->device));
        // BMPLog::correct (BMPLog.C:151-191): the fluidity, then theta with it
        if (M.fluid) modelCorrect(M.fluid->impl, deltaT, gradU, iteration);
        modelCorrect(M, deltaT, gradU, iteration);
    });
    if (rc != RT_OK) return rc;
    for (int q = 0; q < 6; ++q)
        if (M.perf[q].solved && !M.perf[q].converged) return RT_W_NOT_CONVERGED;
    if (M.fluid && M.fluid->impl.perf[0].solved && !M.fluid->impl.perf[0].converged) return RT_W_NOT_CONVERGED;
    return RT_OK;
}

int rt_model_get(const rt_model* model, int field, double* out) {
    if (!model || !out) return RT_E_BAD_ARG;
    Model& M = const_cast<Model&>(model->impl);
    return guarded(M.ctx, [&] {
        Ctx* c = M.ctx;
        Mesh& m = *M.mesh;
        const int32_t N = m.N;
        RT_HIP(hipSetDevice(c->device));
        DBuf& st = staging(c).buf;
        int nc = field == RT_THETA || field == RT_TAU ? 6 : 9;
        if (field < RT_THETA || field > RT_EIGVECS) throw Error{RT_E_BAD_ARG, "rt_model_get: unknown field"};
        if (M.isMulti() && field != RT_TAU)
            throw Error{RT_E_BAD_ARG, "rt_model_get: a multiMode model holds only the summed tau; read the "
                                      "modes' fields (rt_model_mode)"};
        if (M.scalar() && field != RT_THETA)
            throw Error{RT_E_BAD_ARG, "rt_model_get: the fluidity equation holds only Phi (RT_THETA, [N])"};
        const size_t bytes = (size_t)nc * N * 8;
        if (st.bytes < bytes) st = c->alloc(bytes);
        const int32_t* perm = m.permD.as<int32_t>();
        if (field == RT_THETA) launch_unpermute_rows(c, N, 6, 6, M.theta.as<double>(), perm, st.as<double>());
        else if (field == RT_TAU) launch_unpermute_rows(c, N, 6, 6, M.tau.as<double>(), perm, st.as<double>());
        else if (field == RT_EIGVALS) launch_eigvals_to_tensor(c, N, M.eigVal.as<double>(), perm, st.as<double>());
        else launch_soa_to_aos(c, N, 9, M.eigVec.as<double>(), perm, st.as<double>());
        if (M.scalar()) {  // Phi is component xx
            std::vector<double> all((size_t)6 * N);
            RT_HIP(hipMemcpyAsync(all.data(), st.p, bytes, hipMemcpyDeviceToHost, c->stream));
            RT_HIP(hipStreamSynchronize(c->stream));
            for (int32_t i = 0; i < N; ++i) out[i] = all[6 * (size_t)i];
            return;
        }
        RT_HIP(hipMemcpyAsync(out, st.p, bytes, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(hipStreamSynchronize(c->stream));
    });
}

int rt_model_div_tau(rt_model* model, const double* tauB, double* out) {
    if (!model || !out) return RT_E_BAD_ARG;
    Model& M = model->impl;
    if (M.isMulti()) {
        // multiMode::divTau sums the modes' terms (multiMode.C:141-151)
        const size_t n = (size_t)3 * M.mesh->N;
        std::vector<double> part(n);
        for (size_t k = 0; k < n; ++k) out[k] = 0.0;
        for (rt_model* mode : M.modes) {
            const int rc = rt_model_div_tau(mode, tauB, part.data());
            if (rc != RT_OK) {
                M.ctx->lastError = "mode " + mode->impl.name + ": " + mode->impl.ctx->lastError;
                return rc;
            }
            for (size_t k = 0; k < n; ++k) out[k] = out[k] + part[k];
        }
        return RT_OK;
    }
    return guarded(M.ctx, [&] {
        Ctx* c = M.ctx;
        Mesh& m = *M.mesh;
        const int32_t N = m.N;
        RT_HIP(hipSetDevice(c->device));
        if (M.stab != 0) {
            if (!M.hasFlow)
                throw Error{RT_E_BAD_ARG, "rt_model_div_tau: the stabilization term needs U (rt_model_set_flow)"};
            if (!m.C4.p || (m.NB > 0 && !m.bDc.p))
                throw Error{RT_E_BAD_ARG, std::string("rt_model_div_tau: stabilization ") +
                                              (M.stab == 1 ? "BSD" : "coupling") +
                                              " needs the mesh's cell and boundary face centres (rt_mesh_desc C, Cb)"};
        }
        DBuf tb;
        if (tauB && m.NB > 0) {
            std::vector<double> soa((size_t)6 * m.NB);
            for (int32_t b = 0; b < m.NB; ++b)
                for (int q = 0; q < 6; ++q) soa[(size_t)q * m.NB + b] = tauB[6 * (size_t)b + q];
            tb = c->alloc(soa.size() * 8);
            RT_HIP(hipMemcpyAsync(tb.p, soa.data(), soa.size() * 8, hipMemcpyHostToDevice, c->stream));
        }
        DBuf d = c->alloc((size_t)3 * N * 8);
        DBuf G;
        if (M.stab != 0) G = c->alloc((size_t)9 * N * 8);
        launch_div_tau(c, meshArgs(M), M.tau.as<double>(), tb.p ? tb.as<double>() : nullptr, M.prm.rho, M.stab,
                       M.lapScheme, M.prm.etaP / M.prm.rho, G.as<double>(), m.C4.as<double>(), m.bDc.as<double>(),
                       d.as<double>(), grid_for(c, N));
        DBuf& st = staging(c).buf;
        if (st.bytes < (size_t)3 * N * 8) st = c->alloc((size_t)3 * N * 8);
        launch_unpermute_rows(c, N, 3, 3, d.as<double>(), m.permD.as<int32_t>(), st.as<double>());
        RT_HIP(hipMemcpyAsync(out, st.p, (size_t)3 * N * 8, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(hipStreamSynchronize(c->stream));
    });
}

int rt_model_div_tau_coeff(const rt_model* model, double* nu) {
    if (!model || !nu) return RT_E_BAD_ARG;
    const Model& M = model->impl;
    if (M.isMulti()) {  // the modes' matrices are summed (multiMode.C:141-151)
        double s = 0;
        for (const rt_model* mode : M.modes) {
            double v = 0;
            const int rc = rt_model_div_tau_coeff(mode, &v);
            if (rc != RT_OK) return rc;
            s = s + v;
        }
        *nu = s;
        return RT_OK;
    }
    // fvm::laplacian(etaS/rho) (none) or ((etaP + etaS)/rho) (BSD, coupling)
    *nu = M.stab == 0 ? M.prm.etaS / M.prm.rho : (M.prm.etaP + M.prm.etaS) / M.prm.rho;
    return RT_OK;
}

int rt_model_set_laplacian_scheme(rt_model* model, const char* text) {
    if (!model) return RT_E_BAD_ARG;
    RT_FORWARD_TO_MODES(rt_model_set_laplacian_scheme(model, text));
    Model& M = model->impl;
    std::vector<std::string> tok;
    {
        std::string cur;
        for (const char* c = text ? text : ""; *c; ++c) {
            if (std::isspace((unsigned char)*c) || *c == ';') {
                if (!cur.empty()) tok.push_back(cur);
                cur.clear();
            } else {
                cur += *c;
            }
        }
        if (!cur.empty()) tok.push_back(cur);
    }
    if (tok.size() != 3 || tok[0] != "Gauss" || tok[1] != "linear" ||
        (tok[2] != "corrected" && tok[2] != "orthogonal" && tok[2] != "uncorrected")) {
        M.ctx->lastError = std::string("rt_model_set_laplacian_scheme: expected \"Gauss linear "
                                       "corrected|orthogonal|uncorrected\", got \"") + (text ? text : "") + "\"";
        return RT_E_UNSUPPORTED;
    }
    M.lapScheme = tok[2] == "corrected" ? 0 : tok[2] == "orthogonal" ? 1 : 2;
    return RT_OK;
}

int rt_model_tau_total(rt_model* model, double* out) {
    if (!model || !out) return RT_E_BAD_ARG;
    Model& M = model->impl;
    return guarded(M.ctx, [&] {
        Ctx* c = M.ctx;
        Mesh& m = *M.mesh;
        const int32_t N = m.N;
        RT_HIP(hipSetDevice(c->device));
        DBuf d = c->alloc((size_t)6 * N * 8);
        // multiMode: sum of the modes' tau and of their etaS; U from the first mode
        Model& F = M.isMulti() ? M.modes.front()->impl : M;
        launch_tau_total(c, meshArgs(F), M.tau.as<double>(), M.prm.etaS, d.as<double>(), grid_for(c, N));
        DBuf& st = staging(c).buf;
        if (st.bytes < (size_t)6 * N * 8) st = c->alloc((size_t)6 * N * 8);
        launch_unpermute_rows(c, N, 6, 6, d.as<double>(), m.permD.as<int32_t>(), st.as<double>());
        RT_HIP(hipMemcpyAsync(out, st.p, (size_t)6 * N * 8, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(hipStreamSynchronize(c->stream));
    });
}

int rt_model_solver_perf(const rt_model* model, rt_solver_perf perf[6]) {
    if (!model || !perf) return RT_E_BAD_ARG;
    const Model& M = model->impl;
    if (M.isMulti()) {  // the worst mode per component: most iterations, largest residuals
        for (int q = 0; q < 6; ++q) perf[q] = rt_solver_perf{0, 0, 0, 0, 1, 0};
        for (const rt_model* mode : M.modes)
            for (int q = 0; q < 6; ++q) {
                const rt_solver_perf& s = mode->impl.perf[q];
                rt_solver_perf& t = perf[q];
                t.initialResidual = std::max(t.initialResidual, s.initialResidual);
                t.finalResidual = std::max(t.finalResidual, s.finalResidual);
                t.nIterations = std::max(t.nIterations, s.nIterations);
                t.nRefinements = std::max(t.nRefinements, s.nRefinements);
                t.converged = t.converged && (s.converged || !s.solved);
                t.solved = t.solved || s.solved;
            }
        for (int q = 0; q < 6; ++q)
            if (!perf[q].solved) perf[q].converged = 0;
        return RT_OK;
    }
    for (int q = 0; q < 6; ++q) perf[q] = M.perf[q];
    return RT_OK;
}

int rt_symm_solve(rt_ctx* ctx, const rt_mesh* mesh, const rt_ldu_symm* sys, double* psi,
                  const char* controls, rt_solver_perf perf[6]) {
    if (!ctx) return RT_E_BAD_ARG;
    return guarded(&ctx->impl, [&] {
        if (!mesh || !sys || !psi || !perf || !sys->diag || !sys->source ||
            (mesh->impl.F > 0 && (!sys->lower || !sys->upper)))
            throw Error{RT_E_BAD_ARG, "rt_symm_solve: incomplete system"};
        Ctx* c = &ctx->impl;
        Mesh& m = const_cast<Mesh&>(mesh->impl);
        RT_HIP(hipSetDevice(c->device));
        Controls ctl = parseControls(controls);
        const int32_t N = m.N, F = m.F, NB = m.NB;
        // boundary coefficients (sparseSolver.C:72-127): iC -> diag, bC -> source
        std::vector<double> d(sys->diag, sys->diag + N);
        std::vector<double> src(sys->source, sys->source + (size_t)6 * N);
        for (int32_t p = 0; p < m.nPatches; ++p) {
            const rt_patch& P = m.patches[p];
            if (P.kind == RT_PATCH_EMPTY) continue;
            for (int32_t k = 0; k < P.size; ++k) {
                const int32_t b = P.start + k;
                const int32_t cell = m.hOwner[(size_t)F + b];
                if (sys->internalCoeffs) {
                    const double* ic = sys->internalCoeffs + 6 * (size_t)b;
                    for (int q = 1; q < 6; ++q)
                        if (ic[q] != ic[0] && m.valid[q] > 0)
                            throw Error{RT_E_UNSUPPORTED,
                                        "rt_symm_solve: component-dependent internalCoeffs (transform patches) are not supported"};
                    d[cell] += ic[0];
                }
                if (sys->boundaryCoeffs)
                    for (int q = 0; q < 6; ++q) src[6 * (size_t)cell + q] += sys->boundaryCoeffs[6 * (size_t)b + q];
            }
        }
        std::unique_ptr<Krylov> Kp(new Krylov);  // workspace of this call
        Krylov& K = *Kp;
        const bool bicg = ctl.solver == SV_PBICG;
        K.ensure(m, ctl.mixed, bicg);
        DBuf dD = c->alloc((size_t)N * 8), dA = c->alloc((size_t)std::max<int64_t>(1, m.nEll) * 8),
             dB = c->alloc((size_t)6 * N * 8), dX = c->alloc((size_t)6 * N * 8),
             dLo = c->alloc((size_t)std::max(1, F) * 8), dUp = c->alloc((size_t)std::max(1, F) * 8);
        std::vector<double> dperm(N);
        for (int32_t r = 0; r < N; ++r) dperm[r] = d[m.iperm[r]];
        RT_HIP(hipMemcpyAsync(dD.p, dperm.data(), (size_t)N * 8, hipMemcpyHostToDevice, c->stream));
        if (F > 0) {
            RT_HIP(hipMemcpyAsync(dLo.p, sys->lower, (size_t)F * 8, hipMemcpyHostToDevice, c->stream));
            RT_HIP(hipMemcpyAsync(dUp.p, sys->upper, (size_t)F * 8, hipMemcpyHostToDevice, c->stream));
        }
        launch_ldu_to_ell(c, m, dLo.as<double>(), dUp.as<double>(), dA.as<double>());
        if (bicg) launch_ldu_to_ell(c, m, dUp.as<double>(), dLo.as<double>(), K.atval.as<double>());  // A^T
        DBuf& st = staging(c).buf;
        uploadAos(c, st, src.data(), (int64_t)6 * N);
        launch_permute_rows(c, N, 6, 6, st.as<double>(), m.permD.as<int32_t>(), dB.as<double>());
        RT_HIP(hipStreamSynchronize(c->stream));
        uploadAos(c, st, psi, (int64_t)6 * N);
        launch_permute_rows(c, N, 6, 6, st.as<double>(), m.permD.as<int32_t>(), dX.as<double>());
        // gAverage(psi) per component for the normalisation factor
        SolveState* ss = K.state.as<SolveState>();
        for (int q = 0; q < 6; ++q) {
            double s = 0;
            for (int32_t i = 0; i < N; ++i) s += psi[6 * (size_t)i + q];
            const double xbar = s / N;
            RT_HIP(hipMemcpyAsync(&ss->c[q].xbar, &xbar, 8, hipMemcpyHostToDevice, c->stream));
            RT_HIP(hipStreamSynchronize(c->stream));
        }
        SolveArgs sa{};
        sa.mesh = &m;
        sa.kry = &K;
        sa.diag = dD.as<double>();
        sa.aval = dA.as<double>();
        sa.atval = bicg ? K.atval.as<double>() : nullptr;
        sa.b = dB.as<double>();
        sa.x = dX.as<double>();
        sa.valid = m.valid;
        sa.ctl = &ctl;
        sa.xbarKnown = true;
        sa.initDone = false;
        DBuf dS;
        if (kryCtl(ctl, ctl.mixed, m).rb) {  // the Schur pivots of a general LDU matrix
            dS = c->alloc((size_t)2 * N * 8);
            launch_rb_schur_diag(c, m, dLo.as<double>(), dUp.as<double>(), dD.as<double>(), dS.as<double>());
            sa.rDs = dS.as<double>();
        }
        solve_lockstep(c, sa, perf);
        launch_unpermute_rows(c, N, 6, 6, dX.as<double>(), m.permD.as<int32_t>(), st.as<double>());
        RT_HIP(hipMemcpyAsync(psi, st.p, (size_t)6 * N * 8, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
