"""The stress step at the benchmark configurations' full sizes, against the
oracle driven to convergence.

- 128^3 drop-in (2.1 M cells, Oldroyd-B and Giesekus, FP64): the whole step
  against the oracle at the same 1e-10 control, within 1e-9.
- 256^3 precision form (16.8 M cells, permuted, jittered, FENE-P): the mixed
  precision and the FP64 steps at the reference's 1e-10 control, each within
  1e-9 per field of the oracle's step solved to 1e-14 (the decomposed oracle,
  one thread per slab, takes about ten seconds at this size); the explicit
  (phi = 0) step against the oracle's source on a sample of cells.
- 4 M-cell random face graph (Giesekus alpha = 0.01, Wi = 50, prescribed
  grad U, conformation eigenvalues up to 1e6): at this conditioning the
  normalised 1e-10 control does not fix the fields to 1e-9 -- the oracle's
  own step at 1e-10 is 5.5e-10 (theta) and 7.5e-9 (tau) from its step at
  1e-14, because tau = G (exp(theta) - I) turns the absolute theta error of
  the cells with theta ~ 14 into the same relative tau error.  So the 1e-9
  bar is checked at the control that reaches it (1e-12: both device paths
  within 1e-9 of the converged oracle per field), and at 1e-10 each device
  path must be as close to the converged oracle as the oracle's own 1e-10
  step is, to the factor 5 by which the preconditioners' stopping residuals
  differ.  theta finite, C = exp(theta) positive definite.
- The literal eigenvalue floor of that configuration (ln lambda down to
  ln 1e-12) on a 50 k-cell graph: the explicit source dt e^27.6 / Wi
  overflows exp() in the oracle and on the device alike (theta, and where tau
  is finite, agree).
- The graph as a serving request beside a 256^3 request, with the device
  memory filled to about 80 %: both results the same bits as the same steps
  run directly.
"""
import numpy as np
import pytest

from oracle import oracle as O
from rheotool_amd import abi, cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = abi.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def box256():
    """The 256^3 precision form (built once: about half a minute on the host)."""
    return cases.box3d(256, lnLo=-8.0, lnHi=float(np.log(30.0)), perm_seed=3, jitter=0.1)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def step(ctx, dm, f, model, lam, dt, kw, precision, gradU=None, tol=1e-10):
    eq = ctx.constitutiveEq("", dm, abi.model_dict(model, lam=lam, **kw))
    eq.set_solver(abi.solver_dict(tol, precision=precision))
    eq.set_state(f.theta, f.eigVals, f.eigVecs)
    eq.set_flow(f.U, f.Ub, f.phi)
    assert eq.correct(dt, gradU) == abi.RT_OK, eq.solver_perf()
    out = dict(theta=eq.theta(), tau=eq.tau(), ev=eq.eigVals(), perf=eq.solver_perf())
    eq.close()
    return out


@pytest.mark.parametrize("model,kw", [("Oldroyd-BLog", {}), ("GiesekusLog", dict(alpha=0.1))])
def test_dropin_128_fp64_matches_oracle(ctx, model, kw):
    print(f"[case] 128^3 drop-in {model} FP64 vs oracle", flush=True)
    m, f = cases.box3d(128, lnLo=-9.0, lnHi=5.0)
    dm = ctx.mesh(m)
    g = step(ctx, dm, f, model, 10.0, 5e-4, kw, "double")
    p = O.params(model, etaS=0.59, etaP=0.41, lam=10.0, alpha=kw.get("alpha", 0.0))
    th, ev, vec, tau, perf = O.correct(m, p, f, 5e-4)
    assert rel(g["theta"], th) < 1e-9 and rel(g["tau"], tau) < 1e-9
    assert all(q.converged for q in g["perf"] if q.solved)


def oracle_threads():
    """Host threads for the decomposed oracle: OMP_NUM_THREADS when set (the
    cores assigned to this GPU), else the CPUs this process may use; at most 16."""
    import os
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))
    return max(1, min(16, n))


def test_precision_256_matches_converged_oracle(ctx, box256):
    """Both device paths at the reference's 1e-10 control against the oracle's
    step solved to 1e-14 (decomposed into slabs, one thread each; at 1e-14 the
    decomposition changes nothing): within 1e-9 per field."""
    m, f = box256
    nd = oracle_threads()
    print(f"[case] 256^3 precision form vs the converged oracle ({nd} slabs)", flush=True)
    p = O.params("FENE-PLog", etaS=0.59, etaP=0.41, lam=5.0, L2=100.0)
    th_o, _, _, tau_o, perf_o = O.decomposed_correct(m, p, f, 5e-4, nd, ctl=O.controls(tol=1e-14))
    assert all(q.finalResidual < 1e-13 for q in perf_o)
    dm = ctx.mesh(m)
    kw = dict(L2=100.0)
    for precision in ("double", "mixed"):
        g = step(ctx, dm, f, "FENE-PLog", 5.0, 5e-4, kw, precision)
        for k, want in (("theta", th_o), ("tau", tau_o)):
            assert np.all(np.isfinite(g[k]))
            assert rel(g[k], want) < 1e-9, (precision, k, rel(g[k], want))
        assert all(q.converged for q in g["perf"] if q.solved)
        if precision == "mixed":
            assert all(q.nRefinements >= 1 for q in g["perf"] if q.solved)


def test_precision_256_explicit_step_matches_oracle_source(ctx, box256):
    """phi = 0 and a prescribed gradient: every cell's step is
    theta + dt S(theta, grad U, eigen state) exactly; 20000 sampled cells
    against the oracle's source."""
    import dataclasses
    print("[case] 256^3 explicit step vs oracle source sample", flush=True)
    m, f = box256
    f = dataclasses.replace(f, phi=np.zeros_like(f.phi))
    dm = ctx.mesh(m)
    dt = 1e-3
    g = step(ctx, dm, f, "FENE-PLog", 5.0, dt, dict(L2=100.0), "double", gradU=f.gradU)
    idx = np.random.default_rng(0).choice(m.nCells, 20000, replace=False)
    p = O.params("FENE-PLog", etaS=0.59, etaP=0.41, lam=5.0, L2=100.0)
    S, _ = O.rhs(p, f.gradU[idx], f.theta[idx], f.eigVals[idx], f.eigVecs[idx])
    want = f.theta[idx] + dt * S
    err = np.abs(g["theta"][idx] - want).max() / np.abs(want).max()
    assert err < 1e-13, err


@pytest.fixture(scope="module")
def graph4m():
    """The 4 M-cell random face graph, its fields, and the oracle's step
    solved to 1e-14 (about 20 s on one host thread)."""
    m = cases.random_graph(4_000_000, kNear=5, sigma=0.5, seed=3)
    f = cases.make_fields(m, cases.FIELDS_TG, seed=11, lnLo=-9.0, lnHi=float(np.log(1e6)))
    p = O.params("GiesekusLog", etaS=0.59, etaP=0.41, lam=50.0, alpha=0.01)
    th, ev, vec, tau, perf = O.correct(m, p, f, 1e-3, ctl=O.controls(tol=1e-14), gradU=f.gradU)
    assert all(q.finalResidual < 1e-13 for q in perf)
    return m, f, p, dict(theta=th, tau=tau)


def check_graph_fields(out):
    assert np.all(np.isfinite(out["theta"]))
    lam = np.diagonal(out["ev"], axis1=1, axis2=2)
    assert np.all(lam > 0)  # C = V diag(lam) V^T positive definite


def test_random_graph_4m_matches_converged_oracle(ctx, graph4m):
    """Solved to 1e-12, both device paths are within 1e-9 of the converged
    oracle per field."""
    m, f, _, conv = graph4m
    print("[case] 4M-cell random face graph at 1e-12 vs the converged oracle", flush=True)
    dm = ctx.mesh(m)
    for precision in ("double", "mixed"):
        g = step(ctx, dm, f, "GiesekusLog", 50.0, 1e-3, dict(alpha=0.01), precision, gradU=f.gradU, tol=1e-12)
        check_graph_fields(g)
        for k in ("theta", "tau"):
            assert rel(g[k], conv[k]) < 1e-9, (precision, k, rel(g[k], conv[k]))


def test_random_graph_4m_reference_control(ctx, graph4m):
    """At the reference's 1e-10 control: the oracle's own step misses the
    converged one by more than 1e-9 in tau (the reference algorithm cannot
    meet that bar here), and each device path lands as close to the converged
    step as the oracle's 1e-10 step, to a factor 5."""
    m, f, p, conv = graph4m
    print("[case] 4M-cell random face graph at the 1e-10 control", flush=True)
    th, _, _, tau, _ = O.correct(m, p, f, 1e-3, ctl=O.controls(tol=1e-10), gradU=f.gradU)
    e_o = dict(theta=rel(th, conv["theta"]), tau=rel(tau, conv["tau"]))
    print(f"[case] oracle at 1e-10 vs 1e-14: {e_o}", flush=True)
    assert e_o["tau"] > 1e-9
    dm = ctx.mesh(m)
    for precision in ("double", "mixed"):
        g = step(ctx, dm, f, "GiesekusLog", 50.0, 1e-3, dict(alpha=0.01), precision, gradU=f.gradU)
        check_graph_fields(g)
        for k in ("theta", "tau"):
            bar = max(1e-9, 5 * e_o[k])
            assert rel(g[k], conv[k]) < bar, (precision, k, rel(g[k], conv[k]), bar)


def test_random_graph_literal_eigenvalue_floor_overflows_like_oracle(ctx):
    """ln lambda on [ln 1e-12, ln 1e6] at Wi = 50, dt = 1e-3: the explicit
    relaxation source dt e^27.6 / Wi ~ 2e7 sends theta to ~2e7 in one step and
    exp() overflows in most cells -- in the reference algorithm as the oracle
    restates it and on the device alike.  theta agrees; tau and the
    conformation eigenvalues are non-finite in the same cells; where theta
    stays moderate tau agrees to the level the normalised tolerance allows."""
    m = cases.random_graph(50_000, kNear=5, sigma=0.5, seed=3)
    f = cases.make_fields(m, cases.FIELDS_TG, seed=11, lnLo=float(np.log(1e-12)), lnHi=float(np.log(1e6)))
    lam0 = np.diagonal(f.eigVals, axis1=1, axis2=2)
    assert lam0.min() < 2e-12 and lam0.max() > 5e5
    print("[case] 50k-cell graph, eigenvalues 1e-12..1e6", flush=True)
    dm = ctx.mesh(m)
    g = step(ctx, dm, f, "GiesekusLog", 50.0, 1e-3, dict(alpha=0.01), "double", gradU=f.gradU)
    p = O.params("GiesekusLog", etaS=0.59, etaP=0.41, lam=50.0, alpha=0.01)
    th, ev, _, tau, _ = O.correct(m, p, f, 1e-3, gradU=f.gradU)
    assert rel(g["theta"], th) < 1e-9
    assert np.abs(th).max() > 1e6  # the overshoot happened
    fin_g, fin_o = np.isfinite(g["tau"]), np.isfinite(tau)
    assert (~fin_o).any() and np.array_equal(fin_g, fin_o)
    lg, lo = np.diagonal(g["ev"], axis1=1, axis2=2), np.diagonal(ev, axis1=1, axis2=2)
    assert np.array_equal(np.isfinite(lg), np.isfinite(lo))
    sane = np.abs(th).max(1) < 30.0
    assert sane.sum() > 100
    dth = np.abs(g["theta"] - th).max()
    assert rel(g["tau"][sane], tau[sane]) < max(10 * dth, 1e-9)


def test_random_graph_4m_in_the_server_beside_256_at_80_percent_hbm(ctx, box256):
    """The 4 M-cell graph as a serving request, next to a 256^3 precision-form
    request, with the device's memory filled to about 80 % by other
    allocations: no failure for want of memory, the graph's fields within
    the bars of the direct mixed-vs-FP64 test above, and the 256^3 result the
    same bits as the same step run directly."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")  # the runtime the library already loaded

    def mem_info():
        free, total = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
        return free.value, total.value

    print("[case] 4M graph + 256^3 in the server at 80% HBM", flush=True)
    g = cases.random_graph(4_000_000, kNear=5, sigma=0.5, seed=3)
    fg = cases.make_fields(g, cases.FIELDS_TG, seed=11, lnLo=-9.0, lnHi=float(np.log(1e6)))
    mb, fb = box256
    dg, db = ctx.mesh(g), ctx.mesh(mb)
    kwg, kwb = dict(alpha=0.01), dict(L2=100.0)
    ref_g = step(ctx, dg, fg, "GiesekusLog", 50.0, 1e-3, kwg, "mixed", gradU=fg.gradU)
    ref_b = step(ctx, db, fb, "FENE-PLog", 5.0, 5e-4, kwb, "mixed")
    free, total = mem_info()
    # both requests need about 20 GB; the rest of 80 % is taken by ballast
    ballast_bytes = int(0.8 * total) - (total - free) - (20 << 30)
    assert ballast_bytes > 0
    ballast = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ballast), ctypes.c_size_t(ballast_bytes)) == 0
    try:
        srv = abi.Server(ctx, hbm_budget=0, max_inflight=2)
        controls = abi.solver_dict(1e-10, precision="mixed")
        tg = srv.submit(dg, fg, abi.model_dict("GiesekusLog", lam=50.0, **kwg), 1e-3, controls, gradU=fg.gradU)
        tb = srv.submit(db, fb, abi.model_dict("FENE-PLog", lam=5.0, **kwb), 5e-4, controls)
        peak = 0.0
        while srv.progress() > 0:
            f_now, _ = mem_info()
            peak = max(peak, 1 - f_now / total)
        out_g, out_b = srv.poll(tg), srv.poll(tb)
        srv.close()
    finally:
        assert hip.hipFree(ballast) == 0
    print(f"[case] peak device memory in use {peak:.2f} of {total / 2**30:.0f} GiB", flush=True)
    assert peak > 0.75
    for k in ("theta", "tau"):
        assert np.all(np.isfinite(out_g[k]))
        assert np.array_equal(out_g[k], ref_g[k]), k
        assert np.array_equal(out_b[k], ref_b[k]), k
    assert all(q.converged for q in out_g["perf"] if q.solved)
