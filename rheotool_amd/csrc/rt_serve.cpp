// rt_serve.cpp -- continuous-batching request loop for independent stress steps.
//
// Requests (a mesh, a model dictionary, the state and the flow of one
// stress step) are admitted while a slot and the HBM budget allow, then
// advanced together: every progress() call moves each in-flight request
// further (assembly, Krylov rounds, eigen/tau, download) without blocking on
// any of them (results go home with asynchronous copies, which overlap other
// requests' kernels when the output arrays are page-locked, rt_host_alloc).
// Each slot owns a context with its own stream for a request's own work
// (uploads, assembly, eigen/tau, downloads).  The Krylov solves of all
// in-flight requests are batched: one launch per kernel runs the rounds of
// every request of the same precision and preconditioner form (grouped
// rounds over the requests' matrices, SolveGroup), on a stream of the
// server's own, so small requests no longer pay a launch per kernel each.
// Requests that arrive while others run are admitted as soon as a slot and
// memory free up and join the batch at its next round.  A request's kernels
// depend only on its own data (its own workgroups, per-request reductions in
// a fixed order), so its result does not depend on the mix.
#include <chrono>
#include <deque>
#include <map>

#include "rt_internal.h"

namespace rt {

struct Request {
    int64_t ticket = 0;
    // the submission: strings copied, host arrays referenced (the caller keeps
    // them alive until the ticket is finished), the mesh is an uploaded handle
    const rt_mesh* mesh = nullptr;
    int32_t nCells = 0;
    std::string dict, controls;
    const double *theta = nullptr, *eigVals = nullptr, *eigVecs = nullptr, *U = nullptr, *Ub = nullptr,
                 *phi = nullptr, *gradU = nullptr;
    double deltaT = 0;
    double *thetaOut = nullptr, *tauOut = nullptr;
    size_t estBytes = 0;
    // runtime
    int slot = -1;
    rt_model* model = nullptr;
    enum { QUEUED, RUNNING, FINISHING, DONE, FAILED } state = QUEUED;
    hipEvent_t done = nullptr;
    DBuf dlTheta, dlTau;  // theta and tau in the caller's row order, on their way to the host
    std::chrono::steady_clock::time_point tSubmit, tDone;
    std::vector<double> ownTheta, ownTau;  // results when the caller gave no output arrays
    rt_solver_perf perf[6];
    int rc = RT_OK;
    std::string err;
    int group = -1;  // the batch its Krylov solve runs in, -1: its own rounds
};

}  // namespace rt

using namespace rt;

struct rt_server {
    rt_ctx* parent = nullptr;
    size_t budget = 0, inflightBytes = 0;
    std::vector<rt_ctx*> slots;
    std::vector<Request*> slotReq;
    std::deque<Request*> queue;
    std::deque<int64_t> finished;  // tickets done or failed, not yet reported by rt_server_finished
    std::map<int64_t, std::unique_ptr<Request>> all;
    int64_t nextTicket = 1;
    std::string defaultControls = "solver PBiCGStab; preconditioner DILU; tolerance 1e-10; relTol 0; "
                                  "maxIter 1000; precision mixed;";
    // batched Krylov rounds: one group per (precision, preconditioner form),
    // on the stream of a context of the server's own
    bool batching = true;
    rt_ctx* gctx = nullptr;
    SolveGroup groups[4];
    std::map<Krylov*, Request*> byKrylov;
};

namespace {

void releaseRequest(rt_server* S, Request* r) {
    if (r->group >= 0 && r->model) {
        group_remove(S->groups[r->group], r->model->impl.kry);
        S->byKrylov.erase(&r->model->impl.kry);
        r->group = -1;
    }
    if (r->model) rt_model_destroy(r->model);
    r->model = nullptr;
    if (r->done) (void)hipEventDestroy(r->done);
    r->done = nullptr;
    r->dlTheta.release();
    r->dlTau.release();
    if (r->slot >= 0) {
        S->slotReq[r->slot] = nullptr;
        S->inflightBytes -= r->estBytes;
        r->slot = -1;
    }
    S->finished.push_back(r->ticket);
    // the caller's input arrays are no longer read
    r->theta = r->eigVals = r->eigVecs = r->U = r->Ub = r->phi = r->gradU = nullptr;
}

void fail(rt_server* S, Request* r, int rc, const std::string& msg) {
    r->rc = rc;
    r->err = msg;
    r->state = Request::FAILED;
    r->tDone = std::chrono::steady_clock::now();
    // work already queued for the request (kernels, copies into the caller's
    // arrays) ends before its buffers and the caller's arrays are let go
    if (r->slot >= 0) (void)hipStreamSynchronize(S->slots[r->slot]->impl.stream);
    releaseRequest(S, r);
}

// creates the device objects of a request in a free slot and launches its first work
void admit(rt_server* S, Request* r, int slot) {
    rt_ctx* c = S->slots[slot];
    r->slot = slot;
    S->slotReq[slot] = r;
    S->inflightBytes += r->estBytes;
    int rc = rt_model_create(c, "", r->mesh, r->dict.c_str(), &r->model);
    if (rc == RT_OK && rt_model_mode_count(r->model) > 0) {
        c->impl.lastError = "rt_server: multiMode requests are not supported (submit one request per mode)";
        rc = RT_E_UNSUPPORTED;
    }
    if (rc == RT_OK && (rt_model_fluidity(r->model) || r->model->impl.prm.kind == M_SARAMITO)) {
        // their extra state (BMP's fluidity, Saramito's previous tau) is not part of a request
        c->impl.lastError = "rt_server: BMPLog and SaramitoLog requests are not supported";
        rc = RT_E_UNSUPPORTED;
    }
    if (rc == RT_OK) rc = rt_model_set_solver(r->model, r->controls.c_str());
    if (rc != RT_OK) {
        fail(S, r, rc, rt_last_error(c));
        return;
    }
    try {
        RT_HIP(hipSetDevice(c->impl.device));
        // the request's arrays stay unchanged until its ticket finishes, so
        // the uploads are queued without waiting for them
        Model& M = r->model->impl;
        modelUploadState(M, r->theta, r->eigVals, r->eigVecs, nullptr, false);
        if (!r->eigVals && !r->eigVecs)  // the eigen state of theta, computed here instead of uploaded
            launch_eig_tau(M.ctx, M.mesh->N, M.prm, M.theta.as<double>(), M.eigVal.as<double>(),
                           M.eigVec.as<double>(), nullptr, nullptr);
        modelUploadFlow(M, r->U, r->Ub, r->phi, false);
        modelCorrectBegin(r->model->impl, r->deltaT, r->gradU);
        RT_HIP(hipEventCreateWithFlags(&r->done, hipEventDisableTiming));
        r->state = Request::RUNNING;
        bool mixed = false, rb = false;
        if (S->batching && solve_groupable(M.kry, &mixed, &rb)) {
            const int g = (mixed ? 2 : 0) + (rb ? 1 : 0);
            SolveGroup& G = S->groups[g];
            G.mixed = mixed;
            G.rb = rb;
            group_join(G, M.kry, c->impl.stream);
            r->group = g;
            S->byKrylov[&M.kry] = r;
        }
    } catch (const Error& e) {
        fail(S, r, e.code, e.msg);
    }
}

// Enqueues the results' way home on the slot's stream: theta and tau back
// to the caller's row order (what rt_model_get does), then device-to-host
// copies into the output arrays.  Nothing here waits: from page-locked
// output arrays the copies run while other requests compute.
void beginDownload(rt_ctx* c, Request* r) {
    Model& M = r->model->impl;
    const Mesh& m = *M.mesh;
    Ctx* x = &c->impl;
    const size_t N = (size_t)m.N, bytes = 6 * N * sizeof(double);
    double* th = r->thetaOut;
    double* ta = r->tauOut;
    if (!th) {
        r->ownTheta.resize(6 * N);
        th = r->ownTheta.data();
    }
    if (!ta) {
        r->ownTau.resize(6 * N);
        ta = r->ownTau.data();
    }
    r->dlTheta = x->alloc(bytes);
    r->dlTau = x->alloc(bytes);
    const int32_t* perm = m.permD.as<int32_t>();
    launch_unpermute_rows(x, m.N, 6, 6, M.theta.as<double>(), perm, r->dlTheta.as<double>());
    launch_unpermute_rows(x, m.N, 6, 6, M.tau.as<double>(), perm, r->dlTau.as<double>());
    RT_HIP(hipMemcpyAsync(th, r->dlTheta.p, bytes, hipMemcpyDeviceToHost, x->stream));
    RT_HIP(hipMemcpyAsync(ta, r->dlTau.p, bytes, hipMemcpyDeviceToHost, x->stream));
}

// a request whose solve has finished: eigen/tau and the way home, on its slot's stream
void finishSolve(rt_server* S, Request* r, bool solveDone) {
    rt_ctx* c = S->slots[r->slot];
    if (!solveDone) modelCorrectFinish(r->model->impl);
    beginDownload(c, r);
    RT_HIP(hipEventRecord(r->done, c->impl.stream));
    r->state = Request::FINISHING;
}

void advance(rt_server* S, Request* r) {
    rt_ctx* c = S->slots[r->slot];
    try {
        RT_HIP(hipSetDevice(c->impl.device));
        if (r->state == Request::RUNNING && r->group < 0) {
            if (modelCorrectAdvance(r->model->impl, false)) finishSolve(S, r, true);
        }
        if (r->state == Request::FINISHING) {
            hipError_t e = hipEventQuery(r->done);
            if (e == hipErrorNotReady) return;
            RT_HIP(e);
            rt_model_solver_perf(r->model, r->perf);
            r->rc = RT_OK;
            for (int q = 0; q < 6; ++q)
                if (r->perf[q].solved && !r->perf[q].converged) r->rc = RT_W_NOT_CONVERGED;
            r->state = Request::DONE;
            r->tDone = std::chrono::steady_clock::now();
            releaseRequest(S, r);
        }
    } catch (const Error& e) {
        fail(S, r, e.code, e.msg);
    }
}

}  // namespace

extern "C" {

int rt_server_create(rt_ctx* ctx, size_t hbm_budget, int max_inflight, rt_server** out) {
    if (!ctx || !out) return RT_E_BAD_ARG;
    *out = nullptr;
    if (max_inflight <= 0 || max_inflight > 64) {
        ctx->impl.lastError = "rt_server_create: max_inflight must be in 1..64";
        return RT_E_BAD_ARG;
    }
    auto* S = new rt_server;
    S->parent = ctx;
    S->budget = hbm_budget;
    if (rt_ctx_create(ctx->impl.device, 0, &S->gctx) != RT_OK) {
        delete S;
        ctx->impl.lastError = "rt_server_create: context creation failed";
        return RT_E_HIP;
    }
    for (auto& g : S->groups) g.ctx = &S->gctx->impl;
    for (int i = 0; i < max_inflight; ++i) {
        rt_ctx* c = nullptr;
        int rc = rt_ctx_create(ctx->impl.device, 0, &c);
        if (rc != RT_OK) {
            for (auto* x : S->slots) rt_ctx_destroy(x);
            rt_ctx_destroy(S->gctx);
            delete S;
            ctx->impl.lastError = "rt_server_create: slot context creation failed";
            return rc;
        }
        S->slots.push_back(c);
        S->slotReq.push_back(nullptr);
    }
    ctxRef(&ctx->impl);
    *out = S;
    return RT_OK;
}

void rt_server_destroy(rt_server* S) {
    if (!S) return;
    (void)hipStreamSynchronize(S->gctx->impl.stream);
    for (auto& kv : S->all)
        if (kv.second->slot >= 0) {
            (void)hipStreamSynchronize(S->slots[kv.second->slot]->impl.stream);
            releaseRequest(S, kv.second.get());
        }
    for (auto* c : S->slots) rt_ctx_destroy(c);
    for (auto& g : S->groups) g.tables.release();
    rt_ctx_destroy(S->gctx);
    Ctx* parent = &S->parent->impl;
    delete S;
    ctxUnref(parent);
}

size_t rt_request_bytes_estimate(const rt_mesh* mesh, int mixed, int gradU) {
    if (!mesh) return 0;
    const Mesh& m = mesh->impl;
    return modelBytesEstimate(m.N, m.F, m.NB, m.nEll, mixed != 0, gradU != 0, true);
}

int rt_server_set_batching(rt_server* S, int enable) {
    if (!S) return RT_E_BAD_ARG;
    if (!S->all.empty() && S->batching != (enable != 0)) {
        for (auto& kv : S->all)
            if (kv.second->state == Request::QUEUED || kv.second->state == Request::RUNNING ||
                kv.second->state == Request::FINISHING) {
                S->parent->impl.lastError = "rt_server_set_batching: requests are in flight";
                return RT_E_BAD_ARG;
            }
    }
    S->batching = enable != 0;
    return RT_OK;
}

int64_t rt_server_batched_rounds(const rt_server* S) {
    if (!S) return 0;
    int64_t n = 0;
    for (auto& g : S->groups) n += g.launches;
    return n;
}

int rt_server_submit(rt_server* S, const rt_request* q, int64_t* ticket) {
    if (!S || !q || !ticket) return RT_E_BAD_ARG;
    rt_ctx* P = S->parent;
    if (!q->mesh || !q->dict || !q->theta || !q->U || !q->phi || !(q->deltaT > 0) ||
        (q->mesh->impl.NB > 0 && !q->Ub)) {
        P->impl.lastError = "rt_server_submit: incomplete request";
        return RT_E_BAD_ARG;
    }
    auto r = std::make_unique<Request>();
    const Mesh& mm = q->mesh->impl;
    const size_t N = mm.N, F = mm.F, NB = mm.NB;
    r->mesh = q->mesh;
    r->nCells = mm.N;
    r->dict = q->dict;
    r->controls = q->controls ? q->controls : S->defaultControls;
    r->theta = q->theta;
    r->eigVals = q->eigVals;
    r->eigVecs = q->eigVecs;
    r->U = q->U;
    r->Ub = NB ? q->Ub : nullptr;
    r->phi = q->phi;
    r->deltaT = q->deltaT;
    r->thetaOut = q->thetaOut;
    r->tauOut = q->tauOut;
    r->gradU = q->gradU;
    const bool mixed = r->controls.find("mixed") != std::string::npos;
    r->estBytes = modelBytesEstimate((int32_t)N, (int32_t)F, (int32_t)NB, mm.nEll, mixed, q->gradU != nullptr, true);
    if (S->budget && r->estBytes > S->budget) {
        P->impl.lastError = "rt_server_submit: request needs " + std::to_string(r->estBytes) +
                            " bytes, more than the whole HBM budget";
        return RT_E_OOM;
    }
    r->ticket = S->nextTicket++;
    r->tSubmit = std::chrono::steady_clock::now();
    *ticket = r->ticket;
    S->queue.push_back(r.get());
    S->all[r->ticket] = std::move(r);
    return RT_OK;
}

int rt_server_progress(rt_server* S) {
    if (!S) return -RT_E_BAD_ARG;
    // admission: first fit over the queue while a slot and memory are free
    for (auto it = S->queue.begin(); it != S->queue.end();) {
        int slot = -1;
        for (size_t i = 0; i < S->slots.size(); ++i)
            if (!S->slotReq[i]) {
                slot = (int)i;
                break;
            }
        if (slot < 0) break;
        Request* r = *it;
        if (S->budget && S->inflightBytes + r->estBytes > S->budget) {
            ++it;
            continue;
        }
        it = S->queue.erase(it);
        admit(S, r, slot);
    }
    // batched Krylov rounds; requests whose solve finished go on with eigen/tau
    std::vector<Krylov*> done;
    for (auto& g : S->groups) {
        if (g.members.empty()) continue;
        done.clear();
        try {
            RT_HIP(hipSetDevice(S->gctx->impl.device));
            group_advance(g, done);
        } catch (const Error& e) {
            // the batch's launch failed: every member fails with it
            std::vector<Request*> members;
            for (auto& mb : g.members) members.push_back(S->byKrylov[mb.K]);
            for (Request* r : members) fail(S, r, e.code, e.msg);
            continue;
        }
        for (Krylov* K : done) {
            Request* r = S->byKrylov[K];
            S->byKrylov.erase(K);
            r->group = -1;
            try {
                RT_HIP(hipSetDevice(S->slots[r->slot]->impl.device));
                finishSolve(S, r, false);
            } catch (const Error& e) {
                fail(S, r, e.code, e.msg);
            }
        }
    }
    int live = (int)S->queue.size();
    for (size_t i = 0; i < S->slots.size(); ++i) {
        Request* r = S->slotReq[i];
        if (!r) continue;
        advance(S, r);
        if (S->slotReq[i]) ++live;
    }
    return live;
}

int rt_server_finished(rt_server* S, int64_t* tickets, int max) {
    if (!S || (max > 0 && !tickets)) return -RT_E_BAD_ARG;
    int n = 0;
    while (n < max && !S->finished.empty()) {
        tickets[n++] = S->finished.front();
        S->finished.pop_front();
    }
    return n;
}

int rt_server_poll(rt_server* S, int64_t ticket, double* theta, double* tau, rt_solver_perf perf[6],
                   double* latency_ms) {
    if (!S) return -RT_E_BAD_ARG;
    auto it = S->all.find(ticket);
    if (it == S->all.end()) {
        S->parent->impl.lastError = "rt_server_poll: unknown ticket";
        return -RT_E_BAD_ARG;
    }
    Request* r = it->second.get();
    if (r->state == Request::FAILED) {
        S->parent->impl.lastError = r->err;
        return -r->rc;
    }
    if (r->state != Request::DONE) return 0;
    const size_t n = (size_t)6 * r->nCells;
    const double* th = r->thetaOut ? r->thetaOut : r->ownTheta.data();
    const double* ta = r->tauOut ? r->tauOut : r->ownTau.data();
    if (theta && theta != th) std::copy(th, th + n, theta);
    if (tau && tau != ta) std::copy(ta, ta + n, tau);
    if (perf)
        for (int q = 0; q < 6; ++q) perf[q] = r->perf[q];
    if (latency_ms)
        *latency_ms = std::chrono::duration<double, std::milli>(r->tDone - r->tSubmit).count();
    return 1;
}

}  // extern "C"
