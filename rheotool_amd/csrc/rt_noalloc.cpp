// rt_noalloc.cpp -- rt_symm_solve_noalloc: the segregated symmTensor solve of
// rt_symm_solve without any device or page-locked allocation, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).
//
// Every buffer the solve needs beyond the caller's arrays -- the device copies
// of the LDU system, the ELL matrix values, the Krylov vectors, the reduction
// partials, the solver state and the upload/download staging -- is carved out
// of one file-scope device array, `scratch`, of RT_NOALLOC_SCRATCH_BYTES.  Each
// device holds its own copy of that array from the first context created on it
// until the process ends; its device address is looked up once per device
// (noallocInit, called by rt_ctx_create with the device current) and only that
// address is handed to kernels and copies.
//
// The array keeps the last call's contents: the solver state is cleared by a
// fill kernel on every call.  Calls that overlap on one device race on the
// array (the library's contexts do not serialise against each other).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "rt_internal.h"

namespace rt {
namespace {

constexpr size_t kScratchBytes = RT_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct Slot {
    std::once_flag once;
    hipError_t err = hipErrorNotInitialized;
    void* ptr = nullptr;  // device address of `scratch` on this device
    Krylov kry;           // views into the array, rebound on every call
    SolveState status;    // pageable status read-back of the solve
};
Slot g_slots[kMaxDevices];

// Bump allocator over the scratch array: 256-byte aligned sub-buffers.
struct Carve {
    char* base;
    size_t off = 0;
    void* take(size_t bytes) {
        off = (off + 255) / 256 * 256;
        void* p = base + off;
        off += bytes == 0 ? 8 : bytes;
        return p;
    }
};

constexpr int BLKN = 256;
inline unsigned blocks(int64_t n) { return (unsigned)((n + BLKN - 1) / BLKN); }

// The system in device row order with the boundary coefficients applied, as
// rt_symm_solve applies them on the host (sparseSolver.C:72-127): per row,
// diag += internalCoeffs and source += boundaryCoeffs of its non-empty patch
// faces, in increasing boundary face order.
__global__ void __launch_bounds__(BLKN) k_gather_system(int32_t N, const int32_t* __restrict__ permD,
                                                        const int32_t* __restrict__ bPtr,
                                                        const int32_t* __restrict__ bFace,
                                                        const double* __restrict__ diagExt,
                                                        const double* __restrict__ srcExt,
                                                        const double* __restrict__ iC,
                                                        const double* __restrict__ bC, double* __restrict__ diag,
                                                        double* __restrict__ b) {
    const int32_t r = blockIdx.x * BLKN + threadIdx.x;
    if (r >= N) return;
    const int32_t cell = permD[r];
    double d = diagExt[cell];
    double s[6];
    for (int q = 0; q < 6; ++q) s[q] = srcExt[6 * (int64_t)cell + q];
    for (int32_t k = bPtr[r]; k < bPtr[r + 1]; ++k) {
        const int64_t f = bFace[k];
        if (iC) d += iC[6 * f];
        if (bC)
            for (int q = 0; q < 6; ++q) s[q] += bC[6 * f + q];
    }
    diag[r] = d;
    for (int q = 0; q < 6; ++q) b[6 * (int64_t)r + q] = s[q];
}

// gAverage(psi) of the normalisation factor, computed by the host from the
// caller's psi, into the component records
__global__ void k_set_xbar(CompState* cs, double x0, double x1, double x2, double x3, double x4, double x5) {
    if (threadIdx.x == 0) {
        cs[0].xbar = x0;
        cs[1].xbar = x1;
        cs[2].xbar = x2;
        cs[3].xbar = x3;
        cs[4].xbar = x4;
        cs[5].xbar = x5;
    }
}

}  // namespace

void noallocInit(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    Slot& s = g_slots[device];
    std::call_once(s.once, [&] {
        void* p = nullptr;
        s.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (s.err == hipSuccess) s.ptr = p;
        else (void)hipGetLastError();
    });
}

namespace {

int solveNoAlloc(Ctx* c, const Mesh& mc, const rt_ldu_symm* sys, double* psi, const char* controls,
                 rt_solver_perf perf[6]) {
    Mesh& m = const_cast<Mesh&>(mc);
    RT_HIP(hipSetDevice(c->device));
    if (c->device < 0 || c->device >= kMaxDevices)
        throw Error{RT_E_UNSUPPORTED, "rt_symm_solve_noalloc: device index beyond the library's table"};
    Slot& slot = g_slots[c->device];
    if (!slot.ptr)
        throw Error{RT_E_HIP, std::string("rt_symm_solve_noalloc: the scratch array's device address is unknown: ") +
                                  hipGetErrorString(slot.err)};
    Controls ctl = parseControls(controls);
    const int32_t N = m.N, F = m.F, NB = m.NB;
    // component-dependent internalCoeffs (transform patches) are refused as in rt_symm_solve
    if (sys->internalCoeffs)
        for (int32_t p = 0; p < m.nPatches; ++p) {
            const rt_patch& P = m.patches[p];
            if (P.kind == RT_PATCH_EMPTY) continue;
            for (int32_t k = 0; k < P.size; ++k) {
                const double* ic = sys->internalCoeffs + 6 * (size_t)(P.start + k);
                for (int q = 1; q < 6; ++q)
                    if (ic[q] != ic[0] && m.valid[q] > 0)
                        throw Error{RT_E_UNSUPPORTED,
                                    "rt_symm_solve: component-dependent internalCoeffs (transform patches) are not supported"};
            }
        }
    const bool mixed = ctl.mixed, bicg = ctl.solver == SV_PBICG;
    const size_t n6 = (size_t)6 * N, ne = (size_t)std::max<int64_t>(1, m.nEll);
    const int grid = grid_for(c, N);

    // carve the array; sizes first, so that an oversized mesh is refused
    // before anything is written
    Carve cv{static_cast<char*>(slot.ptr)};
    double* diagExt = (double*)cv.take((size_t)N * 8);
    double* stage = (double*)cv.take(n6 * 8);  // source upload, then psi in and out
    double* iC = sys->internalCoeffs ? (double*)cv.take((size_t)6 * NB * 8) : nullptr;
    double* bC = sys->boundaryCoeffs ? (double*)cv.take((size_t)6 * NB * 8) : nullptr;
    double* lo = (double*)cv.take((size_t)std::max(1, F) * 8);
    double* up = (double*)cv.take((size_t)std::max(1, F) * 8);
    double* dD = (double*)cv.take((size_t)N * 8);
    double* dA = (double*)cv.take(ne * 8);
    double* dB = (double*)cv.take(n6 * 8);
    double* dX = (double*)cv.take(n6 * 8);
    Krylov& K = slot.kry;
    K.noAlloc = false;
    K.r = DBuf(); K.r0 = DBuf(); K.p = DBuf(); K.v = DBuf(); K.t = DBuf();
    K.rs = DBuf(); K.r0s = DBuf(); K.ps = DBuf(); K.vs = DBuf(); K.ts = DBuf(); K.es = DBuf();
    K.s64 = DBuf(); K.s32 = DBuf(); K.atval = DBuf(); K.a32 = DBuf(); K.at32 = DBuf();
    K.ahat32 = DBuf(); K.ahat64 = DBuf();
    auto view = [&](size_t bytes) { return DBuf::viewOf(cv.take(bytes), bytes); };
    if (mixed) {
        K.rs = view(n6 * 4); K.r0s = view(n6 * 4); K.ps = view(n6 * 4); K.vs = view(n6 * 4);
        K.ts = view(n6 * 4); K.es = view(n6 * 4); K.ahat32 = view(ne * 4);
    } else {
        K.r = view(n6 * 8); K.r0 = view(n6 * 8); K.p = view(n6 * 8); K.v = view(n6 * 8); K.t = view(n6 * 8);
        K.ahat64 = view(ne * 8);
    }
    if (bicg) {
        K.atval = view(ne * 8);
        if (mixed) {
            K.s32 = view(n6 * 4); K.a32 = view(ne * 4); K.at32 = view(ne * 4);
        } else {
            K.s64 = view(n6 * 8);
        }
    }
    const bool rb = kryCtl(ctl, mixed, m).rb != 0;
    double* dS = rb ? (double*)cv.take((size_t)2 * std::max(1, N) * 8) : nullptr;
    K.rD = view((size_t)std::max(1, N) * 8);
    K.partials = view((size_t)grid * 24 * sizeof(double));
    K.state = view(sizeof(SolveState));
    if (cv.off > kScratchBytes) {
        K.r = DBuf(); K.r0 = DBuf(); K.p = DBuf(); K.v = DBuf(); K.t = DBuf();
        K.rs = DBuf(); K.r0s = DBuf(); K.ps = DBuf(); K.vs = DBuf(); K.ts = DBuf(); K.es = DBuf();
        K.s64 = DBuf(); K.s32 = DBuf(); K.atval = DBuf(); K.a32 = DBuf(); K.at32 = DBuf();
        K.ahat32 = DBuf(); K.ahat64 = DBuf(); K.rD = DBuf(); K.partials = DBuf(); K.state = DBuf();
        throw Error{RT_E_BAD_ARG, "rt_symm_solve_noalloc: this mesh and solver need " + std::to_string(cv.off) +
                                      " bytes of scratch memory; the limit is " + std::to_string(kScratchBytes) +
                                      " bytes (RT_NOALLOC_SCRATCH_BYTES): use rt_symm_solve"};
    }
    K.ctx = c;
    K.N = N;
    K.grid = grid;
    K.noAlloc = true;

    // uploads (the caller's pageable arrays) and the system in device order
    hipStream_t st = c->stream;
    RT_HIP(hipMemcpyAsync(diagExt, sys->diag, (size_t)N * 8, hipMemcpyHostToDevice, st));
    RT_HIP(hipMemcpyAsync(stage, sys->source, n6 * 8, hipMemcpyHostToDevice, st));
    if (iC) RT_HIP(hipMemcpyAsync(iC, sys->internalCoeffs, (size_t)6 * NB * 8, hipMemcpyHostToDevice, st));
    if (bC) RT_HIP(hipMemcpyAsync(bC, sys->boundaryCoeffs, (size_t)6 * NB * 8, hipMemcpyHostToDevice, st));
    if (F > 0) {
        RT_HIP(hipMemcpyAsync(lo, sys->lower, (size_t)F * 8, hipMemcpyHostToDevice, st));
        RT_HIP(hipMemcpyAsync(up, sys->upper, (size_t)F * 8, hipMemcpyHostToDevice, st));
    }
    if (N > 0)
        hipLaunchKernelGGL(k_gather_system, dim3(blocks(N)), dim3(BLKN), 0, st, N, m.permD.as<int32_t>(),
                           m.bPtr.as<int32_t>(), m.bFace.as<int32_t>(), diagExt, stage, iC, bC, dD, dB);
    RT_HIP(hipGetLastError());
    launch_ldu_to_ell(c, m, lo, up, dA);
    if (rb) launch_rb_schur_diag(c, m, lo, up, dD, dS);  // the Schur pivots of a general LDU matrix
    if (bicg) launch_ldu_to_ell(c, m, up, lo, K.atval.as<double>());  // A^T
    RT_HIP(hipMemcpyAsync(stage, psi, n6 * 8, hipMemcpyHostToDevice, st));
    launch_permute_rows(c, N, 6, 6, stage, m.permD.as<int32_t>(), dX);
    // the solver state starts from zero (the array holds the last call's)
    launch_fill(c, K.state.as<double>(), (int64_t)(sizeof(SolveState) / sizeof(double)), 0.0);
    double xbar[6];
    for (int q = 0; q < 6; ++q) {
        double s = 0;
        for (int32_t i = 0; i < N; ++i) s += psi[6 * (size_t)i + q];
        xbar[q] = N > 0 ? s / N : 0.0;
    }
    SolveState* ss = K.state.as<SolveState>();
    hipLaunchKernelGGL(k_set_xbar, dim3(1), dim3(64), 0, st, ss->c, xbar[0], xbar[1], xbar[2], xbar[3], xbar[4],
                       xbar[5]);
    RT_HIP(hipGetLastError());

    SolveArgs sa{};
    sa.mesh = &m;
    sa.kry = &K;
    sa.diag = dD;
    sa.aval = dA;
    sa.atval = bicg ? K.atval.as<double>() : nullptr;
    sa.b = dB;
    sa.x = dX;
    sa.valid = m.valid;
    sa.ctl = &ctl;
    sa.xbarKnown = true;
    sa.initDone = false;
    sa.hostStatus = &slot.status;
    sa.rDs = dS;
    solve_lockstep(c, sa, perf);
    launch_unpermute_rows(c, N, 6, 6, dX, m.permD.as<int32_t>(), stage);
    RT_HIP(hipMemcpyAsync(psi, stage, n6 * 8, hipMemcpyDeviceToHost, st));
    RT_HIP(hipStreamSynchronize(st));
    return RT_OK;
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" int rt_symm_solve_noalloc(rt_ctx* ctx, const rt_mesh* mesh, const rt_ldu_symm* sys, double* psi,
                                     const char* controls, rt_solver_perf perf[6]) {
    if (!ctx) return RT_E_BAD_ARG;
    Ctx* c = &ctx->impl;
    try {
        if (!mesh || !sys || !psi || !perf || !sys->diag || !sys->source ||
            (mesh->impl.F > 0 && (!sys->lower || !sys->upper)))
            throw Error{RT_E_BAD_ARG, "rt_symm_solve_noalloc: incomplete system"};
        if (!mesh->impl.ctx || mesh->impl.ctx->device != c->device)
            throw Error{RT_E_BAD_ARG, "rt_symm_solve_noalloc: the mesh lives on another device"};
        const int rc = solveNoAlloc(c, mesh->impl, sys, psi, controls, perf);
        c->lastError.clear();
        return rc;
    } catch (const Error& e) {
        c->lastError = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        c->lastError = e.what();
        return RT_E_BAD_ARG;
    }
}
