"""The continuous-batching server: every request's result is the same alone
and in a mix (bit-exact: a request's kernels depend only on its own data),
requests arriving while others run are admitted as slots free up, the HBM
budget gates admission, and results agree with the FP64 oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from rheotool_amd import abi, cases

pytestmark = pytest.mark.gpu


def make_request(i):
    """Mixed sizes, 2D/3D, three models, Wi log-uniform on [0.1, 20]."""
    rng = np.random.default_rng(100 + i)
    twoD = bool(rng.integers(2))
    ncell = int(np.exp(rng.uniform(np.log(1e3), np.log(3e4))))
    if twoD:
        n = max(8, int(np.sqrt(ncell)))
        m = cases.cartesian(n, n, 1, twoD=True, channel=True)
    else:
        n = max(6, int(round(ncell ** (1 / 3))))
        m = cases.cartesian(n, n, n, channel=True)
    f = cases.make_fields(m, cases.FIELDS_CHANNEL, seed=1000 + i)
    model = ["Oldroyd-BLog", "GiesekusLog", "PTTLog"][i % 3]
    lam = float(np.exp(rng.uniform(np.log(0.1), np.log(20.0))))
    kw = {"GiesekusLog": dict(alpha=0.2), "PTTLog": dict(epsilon=0.25, zeta=0.0)}.get(model, {})
    dt = 0.1 / n
    return m, f, model, lam, kw, dt


def run_alone(ctx, req, controls):
    m, f, model, lam, kw, dt = req
    dm = ctx.mesh(m)
    eq = ctx.constitutiveEq("", dm, abi.model_dict(model, lam=lam, **kw))
    eq.set_solver(controls)
    eq.set_state(f.theta, f.eigVals, f.eigVecs)
    eq.set_flow(f.U, f.Ub, f.phi)
    eq.correct(dt)
    return eq.theta(), eq.tau()


@pytest.mark.parametrize("batching", [True, False])
@pytest.mark.parametrize("precision,precond", [("mixed", "diagonal"), ("double", "diagonal"),
                                               ("mixed", "DILU"), ("double", "multicolourILU0")])
def test_same_result_alone_and_in_mix(precision, precond, batching):
    """Each request's result is the same bits alone (a blocking step) and in a
    mix of staggered arrivals, with the requests' Krylov rounds batched into
    one launch per kernel or each request launching its own."""
    controls = abi.solver_dict(1e-10, precision=precision, precond=precond)
    reqs = [make_request(i) for i in range(10)]
    print(f"[case] serve mix {precision} {precond} batching={batching}", flush=True)
    with abi.Context(0) as ctx:
        alone = [run_alone(ctx, r, controls) for r in reqs]
        srv = abi.Server(ctx, hbm_budget=0, max_inflight=4, batching=batching)
        tickets = {}
        meshes = [ctx.mesh(r[0]) for r in reqs]
        # staggered arrivals: submit some, progress, submit more while running
        for i, r in enumerate(reqs):
            m, f, model, lam, kw, dt = r
            tickets[i] = srv.submit(meshes[i], f, abi.model_dict(model, lam=lam, **kw), dt, controls)
            if i % 3 == 2:
                for _ in range(3):
                    srv.progress()
        spins = 0
        while srv.progress() > 0:
            spins += 1
            assert spins < 2_000_000
        for i, r in enumerate(reqs):
            out = srv.poll(tickets[i])
            assert out is not None
            assert np.array_equal(out["theta"], alone[i][0]), i
            assert np.array_equal(out["tau"], alone[i][1]), i
            assert out["latency_ms"] > 0
        assert (srv.batched_rounds > 0) == batching
        srv.close()



def test_page_locked_requests_give_the_same_bits():
    """Requests whose inputs and result arrays are page-locked
    (Context.pinned_empty, rt_host_alloc): the uploads and the asynchronous
    downloads land the same bits as the pageable path, written in place into
    the caller's arrays."""
    import dataclasses
    controls = abi.solver_dict(1e-10, precision="mixed")
    reqs = [make_request(i) for i in range(6)]
    with abi.Context(0) as ctx:
        alone = [run_alone(ctx, r, controls) for r in reqs]

        def pin(a):
            p = ctx.pinned_empty(a.shape)
            p[...] = a
            return p

        srv = abi.Server(ctx, hbm_budget=0, max_inflight=3)
        meshes = [ctx.mesh(r[0]) for r in reqs]
        outs, tickets = [], []
        for i, r in enumerate(reqs):
            m, f, model, lam, kw, dt = r
            pf = dataclasses.replace(f, **{k: pin(getattr(f, k)) for k in
                                           ("theta", "eigVals", "eigVecs", "U", "Ub", "phi")})
            o = (ctx.pinned_empty((m.nCells, 6)), ctx.pinned_empty((m.nCells, 6)))
            outs.append(o)
            tickets.append(srv.submit(meshes[i], pf, abi.model_dict(model, lam=lam, **kw), dt, controls,
                                      out=o))
            srv.progress()
        with pytest.raises(ValueError):  # result arrays of the wrong shape are refused
            srv.submit(meshes[0], reqs[0][1], abi.model_dict("Oldroyd-BLog"), 1e-3, controls,
                       out=(np.empty((3, 6)), np.empty((3, 6))))
        spins = 0
        while srv.progress() > 0:
            spins += 1
            assert spins < 2_000_000
        for i in range(len(reqs)):
            out = srv.poll(tickets[i])
            assert out["theta"] is outs[i][0] and out["tau"] is outs[i][1]
            assert np.array_equal(outs[i][0], alone[i][0]), i
            assert np.array_equal(outs[i][1], alone[i][1]), i
        srv.close()
        del outs


def test_requests_without_eigen_state():
    """A request may leave out eigVals/eigVecs: the server computes them from
    theta on the device (calcEig) before the step.  The results match the
    FP64 oracle fed the host's eigen state of the same theta, and are the
    same bits alone and in a mix."""
    controls = abi.solver_dict(1e-10, precision="mixed")
    reqs = [make_request(i) for i in range(5)]
    with abi.Context(0) as ctx:
        meshes = [ctx.mesh(r[0]) for r in reqs]

        def serve(idx, inflight):
            srv = abi.Server(ctx, max_inflight=inflight)
            t = {i: srv.submit(meshes[i], reqs[i][1], abi.model_dict(reqs[i][2], lam=reqs[i][3], **reqs[i][4]),
                               reqs[i][5], controls, eigen=False) for i in idx}
            while srv.progress() > 0:
                pass
            out = {i: srv.poll(k) for i, k in t.items()}
            srv.close()
            return out

        mix = serve(range(len(reqs)), 3)
        for i, r in enumerate(reqs):
            alone = serve([i], 1)[i]
            assert np.array_equal(alone["theta"], mix[i]["theta"]), i
            assert np.array_equal(alone["tau"], mix[i]["tau"]), i
            m, f, model, lam, kw, dt = r
            p = O.params(model, etaS=0.59, etaP=0.41, lam=lam, alpha=kw.get("alpha", 0.0),
                         epsilon=kw.get("epsilon", 0.0), zeta=kw.get("zeta", 0.0))
            th, _, _, tau, _ = O.correct(m, p, f, dt)
            v = m.validComponents > 0
            assert np.linalg.norm(mix[i]["theta"][:, v] - th[:, v]) / np.linalg.norm(th[:, v]) < 1e-9, i
            assert np.linalg.norm(mix[i]["tau"] - tau) / np.linalg.norm(tau) < 1e-9, i

def test_server_results_match_oracle():
    with abi.Context(0) as ctx:
        srv = abi.Server(ctx, max_inflight=3)
        reqs = [make_request(i) for i in range(4)]
        meshes = [ctx.mesh(r[0]) for r in reqs]
        t = [srv.submit(dm, r[1], abi.model_dict(r[2], lam=r[3], **r[4]), r[5])
             for dm, r in zip(meshes, reqs)]
        while srv.progress() > 0:
            pass
        for k, r in zip(t, reqs):
            m, f, model, lam, kw, dt = r
            out = srv.poll(k)
            p = O.params(model, etaS=0.59, etaP=0.41, lam=lam, alpha=kw.get("alpha", 0.0),
                         epsilon=kw.get("epsilon", 0.0), zeta=kw.get("zeta", 0.0))
            th, _, _, tau, _ = O.correct(m, p, f, dt)
            v = m.validComponents > 0
            assert np.linalg.norm(out["theta"][:, v] - th[:, v]) / np.linalg.norm(th[:, v]) < 1e-9
            assert np.linalg.norm(out["tau"] - tau) / np.linalg.norm(tau) < 1e-9
        srv.close()


def test_budget_gates_admission():
    with abi.Context(0) as ctx:
        reqs = [make_request(i) for i in range(3)]
        srv = abi.Server(ctx, hbm_budget=1 << 20, max_inflight=2)  # 1 MiB: nothing fits
        dm = ctx.mesh(reqs[0][0])
        with pytest.raises(abi.RheoError) as e:
            srv.submit(dm, reqs[0][1], abi.model_dict(reqs[0][2], lam=reqs[0][3], **reqs[0][4]),
                       reqs[0][5])
        assert e.value.code == abi.RT_E_OOM
        srv.close()


def test_shared_mesh_requests_and_queueing():
    """Several requests on one uploaded mesh; two slots admit two at a time,
    the rest queue and still finish with unchanged results."""
    with abi.Context(0) as ctx:
        m, f, model, lam, kw, dt = make_request(1)
        dm = ctx.mesh(m)
        d = abi.model_dict(model, lam=lam, **kw)
        alone = run_alone(ctx, (m, f, model, lam, kw, dt), abi.solver_dict(1e-10, precision="mixed"))
        probe = abi.Server(ctx, max_inflight=1)
        t0 = probe.submit(dm, f, d, dt)
        while probe.progress() > 0:
            pass
        assert probe.poll(t0) is not None
        probe.close()
        srv = abi.Server(ctx, hbm_budget=200 << 20, max_inflight=2)
        ts = [srv.submit(dm, f, d, dt) for _ in range(5)]
        seen = []
        while srv.progress() > 0:
            seen += srv.finished()
        seen += srv.finished()
        assert sorted(seen) == sorted(ts) and srv.finished() == []  # each finish reported once
        for t in ts:
            out = srv.poll(t)
            assert np.array_equal(out["theta"], alone[0])


def test_objects_outlive_their_context():
    """A mesh, a model and a server released after rt_ctx_destroy still free
    their memory safely (the context is torn down at its last reference), and
    a model keeps working until it is destroyed."""
    m, f, model, lam, kw, dt = make_request(2)
    ref = None
    with abi.Context(0) as ctx0:
        ref = run_alone(ctx0, (m, f, model, lam, kw, dt), abi.solver_dict(1e-10, precision="mixed"))
    ctx = abi.Context(0)
    dm = ctx.mesh(m)
    eq = ctx.constitutiveEq("", dm, abi.model_dict(model, lam=lam, **kw))
    eq.set_solver(abi.solver_dict(1e-10, precision="mixed"))
    eq.set_state(f.theta, f.eigVals, f.eigVecs)
    eq.set_flow(f.U, f.Ub, f.phi)
    srv = abi.Server(ctx, max_inflight=2)
    ctx.close()
    eq.correct(dt)
    assert np.array_equal(eq.theta(), ref[0])
    srv.close()
    eq.close()
    dm.close()


def test_random_graph_requests_with_prescribed_gradu():
    """Random-face-graph requests with an analytic grad U (the gradU argument
    of correct()) mixed with channel requests: each result equals the model
    run alone and the graph ones agree with the oracle to 1e-9.  With
    eigenvalues of C up to 1e6 (|theta| up to 14), tau = exp(theta)-based
    amplifies a theta difference by |theta|, so both sides solve to 1e-12
    here to keep the solver tolerance itself out of the comparison."""
    controls = abi.solver_dict(1e-12, precision="mixed")
    with abi.Context(0) as ctx:
        graphs = []
        for i in range(3):
            m = cases.random_graph(4000 + 1500 * i, kNear=5, sigma=0.5, seed=20 + i)
            f = cases.make_fields(m, cases.FIELDS_TG, seed=30 + i, lnLo=-9.0, lnHi=float(np.log(1e6)))
            graphs.append((m, f))
        d = abi.model_dict("GiesekusLog", lam=50.0, alpha=0.01)
        alone = []
        for m, f in graphs:
            dm = ctx.mesh(m)
            eq = ctx.constitutiveEq("", dm, d)
            eq.set_solver(controls)
            eq.set_state(f.theta, f.eigVals, f.eigVecs)
            eq.set_flow(f.U, f.Ub, f.phi)
            eq.correct(1e-3, f.gradU)
            alone.append((eq.theta(), eq.tau()))
        srv = abi.Server(ctx, max_inflight=3)
        meshes = [ctx.mesh(m) for m, _ in graphs]
        others = [make_request(i) for i in range(2)]
        ts = []
        for k, ((m, f), dm) in enumerate(zip(graphs, meshes)):
            ts.append(srv.submit(dm, f, d, 1e-3, controls, gradU=f.gradU))
            if k < len(others):
                r = others[k]
                srv.submit(ctx.mesh(r[0]), r[1], abi.model_dict(r[2], lam=r[3], **r[4]), r[5], controls)
        while srv.progress() > 0:
            pass
        p = O.params("GiesekusLog", etaS=0.59, etaP=0.41, lam=50.0, alpha=0.01)
        for t, (m, f), a in zip(ts, graphs, alone):
            out = srv.poll(t)
            assert np.array_equal(out["theta"], a[0]) and np.array_equal(out["tau"], a[1])
            th, _, _, tau, _ = O.correct(m, p, f, 1e-3, gradU=f.gradU, ctl=O.controls(tol=1e-12))
            assert np.linalg.norm(out["theta"] - th) / np.linalg.norm(th) < 1e-9
            assert np.linalg.norm(out["tau"] - tau) / np.linalg.norm(tau) < 1e-9
        srv.close()


@pytest.mark.parametrize("precision,with_grad", [("mixed", False), ("double", False), ("mixed", True),
                                                  ("double", True)])
def test_admission_estimate_covers_what_a_request_allocates(precision, with_grad):
    """rt_request_bytes_estimate, the amount admission reserves against the HBM
    budget, is at least what a request's step allocates in a fresh context
    (the model, its solve, the prescribed gradient, the staging and the two
    result buffers the server downloads through) and not grossly more."""
    from rheotool_amd.abi import lib
    m = cases.cartesian(40, 36, 30, channel=True)
    f = cases.make_fields(m, cases.FIELDS_TG, seed=3, lnLo=-2.0, lnHi=2.0)
    with abi.Context(0) as ctx:
        dm = ctx.mesh(m)
        base = ctx.bytes_in_use
        eq = ctx.constitutiveEq("", dm, abi.model_dict("GiesekusLog", lam=1.0, alpha=0.2))
        eq.set_solver(abi.solver_dict(1e-10, precision=precision))
        eq.set_state(f.theta, f.eigVals, f.eigVecs)
        eq.set_flow(f.U, f.Ub, f.phi)
        eq.correct(1e-3, f.gradU if with_grad else None)
        used = ctx.bytes_in_use - base + 2 * 6 * 8 * m.nCells  # + the server's result buffers
        est = lib().rt_request_bytes_estimate(dm.h, int(precision == "mixed"), int(with_grad))
        print(f"[case] estimate {est} used {used}", flush=True)
        assert used <= est <= 1.6 * used
        eq.close()
