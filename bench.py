#!/usr/bin/env python3
"""Stress-step throughput of the log-conformation step on one MI355X.

Default workload (the headline): the 16.8M-cell precision form -- a 256^3
box with cells permuted by seed and nodes jittered +-10%, FENE-P-Log
(L2=100, Wi=5, beta=0.59, dt=5e-4), FP32 Krylov + FP64 refinement to the
reference's 1e-10 normalised residual, FP64 eigen/log/exp.  One step = one
constitutiveEq::correct() (gradient, decomposition, assembly, six-component
solve, eigen-decomposition, tau) with every input resident in HBM.

Also reported on the same inputs: the plain FP64 drop-in (time and per-field
error of the mixed result against it), the dominant kernel's HBM roofline
fraction from HIP events, per-step p50/p99 latency, and the CPU oracle
(reference algorithm, PBiCGStab + DILU, one thread) timed on a bounded
sample of the same workload.

python bench.py [--steps K] [--warmup W] [--workload precision|dropin|channel]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "stress-step cell-updates/s at 1e-9 rel err vs FP64; HBM roofline %; p99 ms"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (mesh kwargs, model, model kwargs, lam, dt, ln range, precision)
    "precision": dict(n=256, perm_seed=3, jitter=0.1, model="FENE-PLog", kw=dict(L2=100.0),
                      lam=5.0, dt=5e-4, ln=(-8.0, float(np.log(30.0))), precision="mixed",
                      desc="256^3 box, cells permuted (seed 3), nodes jittered +-10%, FENE-P-Log "
                           "L2=100 Wi=5 beta=0.59 dt=5e-4; FP32 Krylov + FP64 refinement to "
                           "1e-10; FP64 eigen/log/exp"),
    "precision_natural": dict(n=256, perm_seed=0, jitter=0.1, model="FENE-PLog", kw=dict(L2=100.0),
                              lam=5.0, dt=5e-4, ln=(-8.0, float(np.log(30.0))), precision="mixed",
                              desc="256^3 box in natural (lexicographic) cell order, nodes jittered "
                                   "+-10%, FENE-P-Log L2=100 Wi=5 beta=0.59 dt=5e-4; FP32 Krylov + "
                                   "FP64 refinement to 1e-10; FP64 eigen/log/exp"),
    "dropin": dict(n=128, perm_seed=0, jitter=0.0, model="GiesekusLog", kw=dict(alpha=0.1),
                   lam=10.0, dt=5e-4, ln=(-9.0, 5.0), precision="double",
                   desc="128^3 box, Taylor-Green + shear, Giesekus-Log alpha=0.1 Wi=10 dt=5e-4; "
                        "plain FP64 PBiCGStab to 1e-10"),
    "channel": dict(n=128, model="Oldroyd-BLog", kw={}, lam=1.0, dt=1e-3, precision="double",
                    desc="128x128x1 channel, Oldroyd-B-Log Wi=1 beta=0.59 dt=1e-3; FP64"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make_case(w, n=None):
    from rheotool_amd import cases
    n = n or w["n"]
    if w["model"] == "Oldroyd-BLog" and "perm_seed" not in w:
        return cases.channel2d(n)
    return cases.box3d(n, lnLo=w["ln"][0], lnHi=w["ln"][1], perm_seed=w.get("perm_seed", 0),
                       jitter=w.get("jitter", 0.0))


PRECOND = "DILU"  # the fvSolution preconditioner keyword the bench's solves name
INNER_REL_TOL = None  # mixed precision: the inner solves' relative target (None: the library default)


def new_model(ctx, dm, w, f, precision):
    from rheotool_amd import abi
    eq = ctx.constitutiveEq("", dm, abi.model_dict(w["model"], etaS=0.59, etaP=0.41, lam=w["lam"],
                                                   **w["kw"]))
    eq.set_solver(abi.solver_dict(1e-10, precision=precision, precond=PRECOND, innerRelTol=INNER_REL_TOL))
    eq.set_state(f.theta, f.eigVals, f.eigVecs)
    eq.set_flow(f.U, f.Ub, f.phi)
    return eq


def min_pass_bytes(N, F, perf, mixed):
    """Bytes of one step under the minimum-pass accounting of a fully fused
    PBiCGStab: the shared matrix read twice per lockstep iteration plus the
    Jacobi diagonal, 17 vector passes per component-iteration, one FP64
    residual-and-correction pass per refinement sweep, and the per-cell pre-
    and post-solve passes (assembly inputs and outputs; eigen state and tau)."""
    w = 4.0 if mixed else 8.0
    solved = [p for p in perf if p.solved]
    nc = len(solved)
    K = max((p.nIterations for p in solved), default=0)
    iters = sum(p.nIterations for p in solved)
    nref = max((p.nRefinements for p in solved), default=0)
    b_a = lambda wa: N * wa + 2 * F * wa + 8 * F
    b_pre = N * (48 + 24 + 72 + 24 + 8 + 6 * w + w) + F * (8 + 24 + 8 + 8 + 2 * w)
    b_post = N * (48 + 24 + 72 + 48)
    b_ref = b_a(8.0) + nc * N * (8 + 8 + 8 + 4 + 8)
    return b_pre + K * (2 * b_a(w) + N * w) + 17 * N * w * iters + nref * b_ref + b_post


def cpu_threads():
    """Threads of the all-cores baseline: OMP_NUM_THREADS when set (the host
    cores assigned to this GPU), else the CPUs this process may use, at most
    16."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_step(run, reps):
    """Median of reps calls after one warm-up call; run() returns seconds."""
    ts = [run() for _ in range(reps + 1)]
    return float(np.median(ts[1:]))


def cpu_baseline(w, n_single, n_all, reps=5):
    """The oracle (the reference's algorithm: PBiCGStab + DILU, FP64) on
    bounded samples of the same workload, each the median of reps steps
    after one warm-up, every step from the same state:
    (i) one thread, as one rank of the reference runs, on the n_single^3
        version;
    (ii) all cores on the n_all^3 version, decomposed into one "simple" slab
        per thread as decomposePar does, DILU per slab and the processor
        interfaces inside Amul.
    Returns the single-thread sample's mesh, fields and result (the accuracy
    check runs on it) and the cpu_baseline record."""
    from oracle import oracle as O
    p = O.params(w["model"], etaS=0.59, etaP=0.41, lam=w["lam"], alpha=w["kw"].get("alpha", 0.0),
                 L2=w["kw"].get("L2", 100.0))
    m, f = make_case(w, n_single)
    out = {}

    def serial():
        t0 = time.perf_counter()
        out["r"] = O.correct(m, p, f, w["dt"])
        return time.perf_counter() - t0

    def label(n, mesh):
        return ("the 2D case itself" if mesh.emptyDir >= 0 else
                f"the {n}^3 version of the same workload") + f" ({mesh.nCells} cells)"

    t1 = median_step(serial, reps)
    th, ev, vec, tau, perf = out["r"]
    single = dict(value=m.nCells / t1, unit="cell-updates/s", cores=1, kind="port",
                  sample=f"one stress step (oracle, PBiCGStab+DILU, FP64) on {label(n_single, m)}, "
                         f"median of {reps} after a warm-up: {t1:.3f} s",
                  iterations=[int(q.nIterations) for q in perf])
    nt = cpu_threads()
    mA, fA = (m, f) if n_all == n_single else make_case(w, n_all)
    ext = np.ptp(mA.C, axis=0)
    axis = 2 if ext[2] >= 0.99 * ext.max() else int(np.argmax(ext))
    dd = O.Decomposed(mA, O.simple_domains(mA, nt, axis))

    def decomposed():
        dd.set_fields(fA)
        t0 = time.perf_counter()
        out["p"] = dd.correct(p, w["dt"])
        return time.perf_counter() - t0

    try:
        tP = median_step(decomposed, reps)
    finally:
        dd.close()
    cpu = dict(value=mA.nCells / tP, unit="cell-updates/s", cores=nt, kind="port",
               sample=f"one stress step (oracle, PBiCGStab+DILU, FP64) on {label(n_all, mA)} decomposed "
                      f"into {nt} slabs along axis {axis}, one thread each, median of {reps} after a "
                      f"warm-up: {tP:.3f} s",
               iterations=[int(q.nIterations) for q in out["p"]],
               cpu=cpu_model(), omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
               single_thread=single)
    return m, f, (th, tau), cpu


SERVE_DESC = ("serving stream: {n} stress-step requests, Poisson arrivals at {rate}/s, each on its own "
              "2D or 3D Cartesian channel (coin flip) with cells log-uniform on [1e3, {mx:.0e}], model uniform "
              "over Oldroyd-B / Giesekus alpha=0.2 / linear PTT eps=0.25, Wi log-uniform [0.1, 20], "
              "fields from seed {seed}+i, dt = 0.1 dx; FP32 Krylov + FP64 refinement; HBM budget 200 GB")


def serve_specs(nreq, max_cells, seed=4):
    """Per request: its own near-square (2D) or near-cubic (3D) Cartesian
    shape for a cell count log-uniform on [1e3, max_cells], the model and Wi."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(nreq):
        twoD = bool(rng.integers(2))
        ncell = float(np.exp(rng.uniform(np.log(1e3), np.log(max_cells))))
        if twoD:
            nx = max(8, int(round(np.sqrt(ncell))))
            shape = (nx, max(8, int(round(ncell / nx))), 1)
        else:
            nx = max(6, int(round(ncell ** (1 / 3))))
            ny = max(6, int(round(np.sqrt(ncell / nx))))
            shape = (nx, ny, max(6, int(round(ncell / (nx * ny)))))
        model = ["Oldroyd-BLog", "GiesekusLog", "PTTLog"][int(rng.integers(3))]
        lam = float(np.exp(rng.uniform(np.log(0.1), np.log(20.0))))
        specs.append((shape, model, lam))
    return specs


def serve_stream(ctx, nreq, rate, max_cells, pageable=False, upload_eigen=False, batching=True,
                 seed_base=1000, max_inflight=16):
    """Continuous-batching stream: requests arrive as a Poisson process, each
    on its own mesh with its own fields (seed seed_base + i); they are
    admitted by the HBM budget and run overlapped, their Krylov rounds batched.
    Meshes are uploaded before the stream (a serving system keeps its meshes
    resident) and fields are generated beforehand; the timed stream covers
    submit to result of every request.  Returns the measured figures."""
    import dataclasses
    from rheotool_amd import abi, cases
    specs = serve_specs(nreq, max_cells)
    t0 = time.perf_counter()
    reqs = []
    for i, (shape, model, lam) in enumerate(specs):
        nx, ny, nz = shape
        m = cases.cartesian(nx, ny, nz, twoD=nz == 1, channel=True)
        f = cases.make_fields(m, cases.FIELDS_CHANNEL, seed=seed_base + i)
        reqs.append([m, f, ctx.mesh(m), 0.1 / nx])
        if (i + 1) % 32 == 0:
            log(f"[serve] {i + 1} of {nreq} meshes and field sets prepared "
                f"({time.perf_counter() - t0:.1f} s)")
    log(f"[serve] {nreq} meshes and field sets prepared in {time.perf_counter() - t0:.1f} s")
    outs = [None] * nreq
    if not pageable:
        # a server's receive and send buffers: page-locked, filled before the
        # requests arrive
        t1 = time.perf_counter()

        def pin(a):
            p = ctx.pinned_empty(a.shape)
            p[...] = a
            return p

        for r in reqs:
            m, f = r[0], r[1]
            r[1] = dataclasses.replace(f, **{k: pin(getattr(f, k)) for k in
                                             ("theta", "eigVals", "eigVecs", "U", "Ub", "phi")})
        for i, r in enumerate(reqs):
            N = r[0].nCells
            outs[i] = (ctx.pinned_empty((N, 6)), ctx.pinned_empty((N, 6)))
        log(f"[serve] page-locked request buffers ready in {time.perf_counter() - t1:.1f} s")
    kwmap = {"GiesekusLog": dict(alpha=0.2), "PTTLog": dict(epsilon=0.25, zeta=0.0)}
    controls = abi.solver_dict(1e-10, precision="mixed", precond=PRECOND, innerRelTol=INNER_REL_TOL)
    srv = abi.Server(ctx, hbm_budget=200 << 30, max_inflight=max_inflight, batching=batching)
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    index_of = {}
    done = [None] * nreq
    cells = 0
    t_start = time.perf_counter()
    i = 0
    while True:
        now = time.perf_counter() - t_start
        while i < nreq and arrivals[i] <= now:
            _, model, lam = specs[i]
            m, f, dm, dt = reqs[i]
            t = srv.submit(dm, f, abi.model_dict(model, lam=lam, **kwmap.get(model, {})), dt, controls,
                           out=outs[i], eigen=upload_eigen)
            index_of[t] = i
            i += 1
        live = srv.progress()
        for t in srv.finished():  # poll only the requests that finished
            j = index_of[t]
            out = srv.poll(t)
            done[j] = out["latency_ms"]
            cells += reqs[j][0].nCells
        if i == nreq and live == 0 and all(d is not None for d in done):
            break
    elapsed = time.perf_counter() - t_start
    lat = np.array(done)
    rounds = srv.batched_rounds
    srv.close()
    for r in reqs:
        r[2].close()
    return dict(requests=nreq, rate=rate, total_cells=int(cells), value=round(cells / elapsed, 1),
                p50_ms=round(float(np.percentile(lat, 50)), 3), p99_ms=round(float(np.percentile(lat, 99)), 3),
                max_ms=round(float(lat.max()), 3), wall_s=round(elapsed, 3), batched_rounds=rounds,
                batching=batching, host_buffers="pageable" if pageable else "page-locked",
                eigen_state="uploaded" if upload_eigen else "computed from theta on the device",
                workload=SERVE_DESC.format(n=nreq, rate=rate, mx=max_cells, seed=seed_base))


MM_SPEC = [("M1", "Oldroyd-BLog", dict(lam=0.8, etaS=0.1, etaP=0.3)),
           ("M2", "GiesekusLog", dict(lam=2.5, alpha=0.2, etaS=0.05, etaP=0.55)),
           ("M3", "PTTLog", dict(lam=0.3, epsilon=0.25, zeta=0.0, etaS=0.02, etaP=0.2)),
           ("M4", "GiesekusLog", dict(lam=0.05, alpha=0.3, etaS=0.02, etaP=0.1))]


def multimode_bench(args):
    """A four-mode multiMode step (--workload multimode): its modes' solves
    batched, against the same four modes stepped as models of their own one
    after another (what a step without batching launches)."""
    from rheotool_amd import abi, cases
    n = args.n or 48
    ctx = abi.Context(0)
    m = cases.cartesian(n, n, n, jitter=0.1)
    base = cases.make_fields(m, cases.FIELDS_TG, seed=31, lnLo=-2.0, lnHi=2.0)
    states = [cases.make_fields(m, cases.FIELDS_TG, seed=40 + k, lnLo=-2.0, lnHi=2.0) for k in range(len(MM_SPEC))]
    dm = ctx.mesh(m)
    controls = abi.solver_dict(1e-10, precond=PRECOND, precision="mixed")

    def entry(name, model, **kw):
        body = abi.model_dict(model, **kw).split("{", 1)[1].rsplit("}", 1)[0]
        return f"{name} {{{body}}}"
    text = ("parameters\n{\n    type multiMode;\n    models\n    (\n" +
            "\n".join(entry(a, b, **c) for a, b, c in MM_SPEC) + "\n    );\n}\n")
    eq = ctx.constitutiveEq("", dm, text)
    for md, g in zip(eq.modes(), states):
        md.set_state(g.theta, g.eigVals, g.eigVecs)
    eq.set_solver(controls)
    eq.set_flow(base.U, base.Ub, base.phi)
    alone = []
    for (a, b, c), g in zip(MM_SPEC, states):
        x = ctx.constitutiveEq("", dm, abi.model_dict(b, **c))
        x.set_state(g.theta, g.eigVals, g.eigVecs)
        x.set_solver(controls)
        x.set_flow(base.U, base.Ub, base.phi)
        alone.append(x)
    dt = 1e-3

    def timed(step):
        for _ in range(args.warmup):
            step()
        ts = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            step()
            ts.append(1e3 * (time.perf_counter() - t0))
        return float(np.median(ts)), float(np.percentile(ts, 99))

    def in_turn():
        for x in alone:
            x.correct(dt)
    batched, b99 = timed(lambda: eq.correct(dt))
    turn, t99 = timed(in_turn)
    for md, x in zip(eq.modes(), alone):  # the same bits either way
        if not np.array_equal(md.theta(), x.theta()):
            raise SystemExit("multimode bench: batched and one-by-one modes differ")
    nm = len(MM_SPEC)
    line = dict(metric=METRIC, value=round(m.nCells * nm / (batched * 1e-3), 1), unit="cell-updates/s",
                n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=round(batched, 3),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32+f64", data="synthetic",
                config=dict(workload=f"multiMode of {nm} modes (Oldroyd-B, Giesekus, PTT, Giesekus) on a {n}^3 "
                                     "box, nodes jittered +-10%; FP32 Krylov + FP64 refinement to 1e-10; "
                                     "value counts each mode's cells; median step without kernel events",
                            cells=m.nCells, modes=nm, preconditioner=PRECOND),
                p99_ms=round(b99, 3), modes_in_turn_ms_per_step=round(turn, 3), modes_in_turn_p99_ms=round(t99, 3),
                speedup_vs_modes_in_turn=round(turn / batched, 3))
    print(json.dumps(line), flush=True)
    ctx.close()


def serve_bench(args):
    """The serving stream as the bench line (--workload serve)."""
    from rheotool_amd import abi
    ctx = abi.Context(0)
    r = serve_stream(ctx, args.requests, args.rate, args.max_cells, pageable=args.pageable,
                     upload_eigen=args.upload_eigen, batching=not args.no_batching,
                     max_inflight=args.max_inflight)
    nreq = r["requests"]
    line = dict(metric=METRIC, value=r["value"], unit="cell-updates/s", n_gpus=1,
                steps=nreq, warmup=0, ms_per_step=round(1e3 * r["wall_s"] / nreq, 3),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32+f64",
                data="synthetic",
                config=dict(workload=r["workload"], requests=nreq, total_cells=r["total_cells"],
                            host_buffers=r["host_buffers"], eigen_state=r["eigen_state"],
                            batching=r["batching"], preconditioner=PRECOND),
                p50_ms=r["p50_ms"], p99_ms=r["p99_ms"], max_ms=r["max_ms"], wall_s=r["wall_s"],
                batched_rounds=r["batched_rounds"])
    print(json.dumps(line), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="precision",
                    choices=sorted(WORKLOADS) + ["copy", "serve", "multimode"])
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--rate", type=float, default=100.0, help="serve: Poisson arrival rate, 1/s")
    ap.add_argument("--max-cells", type=float, default=2e6)
    ap.add_argument("--max-inflight", type=int, default=16, help="serve: requests on the device at once")
    ap.add_argument("--pageable", action="store_true",
                    help="serve: requests in pageable host arrays instead of page-locked ones")
    ap.add_argument("--upload-eigen", action="store_true",
                    help="serve: send each request's eigen state instead of computing it from theta")
    ap.add_argument("--no-batching", action="store_true",
                    help="serve: each request launches its own Krylov rounds")
    ap.add_argument("--no-serve-sample", action="store_true",
                    help="leave the serving-stream sample out of the default line")
    ap.add_argument("--serve-requests", type=int, default=128, help="serving sample: requests")
    ap.add_argument("--serve-rate", type=float, default=1000.0, help="serving sample: arrivals per second")
    ap.add_argument("--serve-max-cells", type=float, default=2e5, help="serving sample: largest request")
    ap.add_argument("--no-transfers", action="store_true",
                    help="leave the transfer-inclusive drop-in measurement out")
    ap.add_argument("--n", type=int, default=0, help="override the mesh size")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--renumber", default="auto", choices=["auto", "none", "rcm", "sfc"],
                    help="device cell order (auto: RCM when the input order has poor locality)")
    ap.add_argument("--no-colour", action="store_true",
                    help="keep a two-colourable mesh in one block (no colour-by-colour row order)")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--inner-rel-tol", type=float, default=None,
                    help="mixed precision: relative target of each FP32 inner solve (innerRelTol)")
    ap.add_argument("--precond", default="DILU",
                    help="preconditioner keyword of the solves (DILU, multicolourILU0, diagonal, none)")
    ap.add_argument("--cpu-sample", type=int, default=64,
                    help="cells per side of the one-thread CPU sample")
    ap.add_argument("--cpu-sample-all", type=int, default=128,
                    help="cells per side of the all-cores (decomposed) CPU sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the stress step "
              "would shard by cell partition, one rank per GPU, with halo exchanges per SpMV; "
              "only the one-GPU path is built)", file=sys.stderr)
        sys.exit(2)

    global PRECOND, INNER_REL_TOL
    PRECOND = args.precond
    INNER_REL_TOL = args.inner_rel_tol
    from rheotool_amd import abi
    if args.workload == "copy":
        # HBM calibration: a plain copy of 2 GiB each way
        with abi.Context(0) as ctx:
            nbytes = 2 << 30
            ms = ctx.copy_bench(nbytes, reps=max(args.steps, 5))
            print(json.dumps(dict(workload="copy", bytes_read=nbytes, bytes_written=nbytes,
                                  ms=round(ms, 4), GBps=round(2 * nbytes / (ms * 1e-3) / 1e9, 1))))
        return
    if args.workload == "serve":
        return serve_bench(args)
    if args.workload == "multimode":
        return multimode_bench(args)
    w = WORKLOADS[args.workload]
    n = args.n or w["n"]
    t0 = time.perf_counter()
    m, f = make_case(w, n)
    log(f"[bench] case {args.workload}: {m.nCells} cells, {m.nInternal} faces, "
        f"generated in {time.perf_counter() - t0:.1f} s")
    ctx = abi.Context(0)
    t0 = time.perf_counter()
    dm = ctx.mesh(m, {"auto": abi.RT_RENUMBER_AUTO, "none": abi.RT_RENUMBER_NONE,
                      "rcm": abi.RT_RENUMBER_RCM, "sfc": abi.RT_RENUMBER_SFC}[args.renumber]
                  | (abi.RT_RENUMBER_NO_COLOUR if args.no_colour else 0))
    log(f"[bench] mesh uploaded in {time.perf_counter() - t0:.1f} s")
    eq = new_model(ctx, dm, w, f, w["precision"])
    log(f"[bench] device bytes in use {ctx.bytes_in_use / 1e9:.2f} GB")

    # one step from the initial state, also through the FP64 path: error table
    errors = None
    fp64_ms = None
    eq.correct(w["dt"])
    first = dict(theta=eq.theta(), tau=eq.tau())
    log(f"[bench] first step perf {eq.solver_perf()}")
    if w["precision"] == "mixed" and not args.no_fp64_compare:
        ref = new_model(ctx, dm, w, f, "double")
        ref.correct(w["dt"])
        errors = dict(theta=rel_l2(first["theta"], ref.theta()), tau=rel_l2(first["tau"], ref.tau()))
        log(f"[bench] mixed vs FP64 drop-in after one step: {errors}")
        for _ in range(max(0, args.warmup - 1)):
            ref.correct(w["dt"])
        ts = []
        for _ in range(min(args.steps, 5)):
            t1 = time.perf_counter()
            ref.correct(w["dt"])
            ts.append(time.perf_counter() - t1)
        fp64_ms = 1e3 * float(np.median(ts))
        log(f"[bench] FP64 drop-in median step {fp64_ms:.2f} ms, perf {ref.solver_perf()}")
        ref.close()
    for _ in range(max(0, args.warmup - 1)):
        eq.correct(w["dt"])

    # timed region: K steps, inputs resident in HBM, kernel events on
    ctx.reset_stats()
    ctx.set_profiling(True)
    step_ms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        eq.correct(w["dt"])
        step_ms.append(1e3 * (time.perf_counter() - t1))
    elapsed = time.perf_counter() - t_start
    stats = ctx.kernel_stats()
    ctx.set_profiling(False)
    iters = [p.nIterations for p in eq.solver_perf()]
    mp_bytes = min_pass_bytes(m.nCells, m.nInternal, eq.solver_perf(), w["precision"] == "mixed")
    log(f"[bench] last step perf {eq.solver_perf()}")
    value = m.nCells * args.steps / elapsed
    dom = max(stats.items(), key=lambda kv: kv[1]["ms"])
    name, st = dom
    avg_ms = st["ms"] / st["launches"]
    achieved = st["bytes"] / st["launches"] / (avg_ms * 1e-3) / 1e9
    step_gpu_ms = sum(s["ms"] for s in stats.values()) / args.steps
    step_bytes = sum(s["bytes"] for s in stats.values()) / args.steps
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("kernels", {}).get(name)
        except Exception:
            traffic = None
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel=name,
                    avg_launch_ms=round(avg_ms, 4), launches_per_step=st["launches"] / args.steps)
    kernels = {k: dict(ms_per_step=round(v["ms"] / args.steps, 3),
                       GBps=round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None)
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"])}

    # the drop-in as a rheoFoam caller drives it:
    # per step, U and phi go up, correct() runs, theta and tau come back
    # (pageable host arrays, the eigen state stays on the device)
    transfers = None
    if not args.no_transfers:
        th_h, tau_h = np.empty((m.nCells, 6)), np.empty((m.nCells, 6))
        ts = []
        for k in range(max(2, min(args.steps, 5)) + 1):
            t1 = time.perf_counter()
            eq.set_flow(f.U, f.Ub, f.phi)
            eq.correct(w["dt"])
            eq.get_into(abi.RT_THETA, th_h)
            eq.get_into(abi.RT_TAU, tau_h)
            if k:
                ts.append(time.perf_counter() - t1)
        t_tr = float(np.median(ts))
        up = m.nCells * 24 + m.nBoundary * 24 + (m.nInternal + m.nBoundary) * 8
        transfers = dict(value=round(m.nCells / t_tr, 1), unit="cell-updates/s", ms_per_step=round(1e3 * t_tr, 3),
                         host_to_device_bytes=int(up), device_to_host_bytes=int(m.nCells * 96),
                         steps=len(ts), note="per step: set_flow(U, Ub, phi) upload, correct(), theta and tau "
                                             "downloaded into pageable arrays; median step")
        log(f"[bench] drop-in with transfers {transfers['ms_per_step']} ms per step")

    cpu = None
    err_sample = None
    if not args.no_cpu_baseline and w["model"] != "Oldroyd-BLog" or (
            not args.no_cpu_baseline and args.workload == "channel"):
        ns = args.cpu_sample if args.workload != "channel" else n
        na = args.cpu_sample_all if args.workload != "channel" else n
        ms, fs, (th_o, tau_o), cpu = cpu_baseline(w, ns, na)
        dms = ctx.mesh(ms)
        es = new_model(ctx, dms, w, fs, w["precision"])
        es.correct(w["dt"])
        err_sample = dict(theta=rel_l2(es.theta(), th_o), tau=rel_l2(es.tau(), tau_o))
        es.close()
        dms.close()
        log(f"[bench] cpu baseline {cpu['value']:.3e} cell-updates/s on {cpu['cores']} threads, "
            f"{cpu['single_thread']['value']:.3e} on one; GPU vs oracle on the sample "
            f"{err_sample}")

    line = dict(
        metric=METRIC, value=round(value, 1), unit="cell-updates/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=round(1e3 * elapsed / args.steps, 3),
        higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f32+f64" if w["precision"] == "mixed" else "f64", data="synthetic",
        config=dict(workload=w["desc"], cells=m.nCells, internal_faces=m.nInternal,
                    model=w["model"], precision=w["precision"], preconditioner=PRECOND),
        roofline=roofline, cpu_baseline=cpu,
        p50_ms=round(float(np.percentile(step_ms, 50)), 3),
        p99_ms=round(float(np.percentile(step_ms, 99)), 3),
        gpu_ms_per_step=round(step_gpu_ms, 3),
        step_algorithmic_GB=round(step_bytes / 1e9, 3),
        step_roofline_frac=round(step_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
        step_min_pass_GB=round(mp_bytes / 1e9, 3),
        step_min_pass_roofline_frac=round(mp_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
        iterations=iters,
        fp64_dropin_ms_per_step=None if fp64_ms is None else round(fp64_ms, 3),
        speedup_vs_fp64_dropin=None if fp64_ms is None else round(fp64_ms / (1e3 * elapsed / args.steps), 3),
        err_vs_fp64=errors, err_vs_oracle_sample=err_sample, kernels=kernels,
        dropin_with_transfers=transfers,
    )
    eq.close()
    dm.close()
    if not args.no_serve_sample and args.workload == "precision":
        # the serving configuration, at a size the default run affords: its own
        # p50/p99 and aggregate rate beside the single-step figures
        line["serve"] = serve_stream(ctx, args.serve_requests, args.serve_rate, args.serve_max_cells)
        log(f"[bench] serve sample {line['serve']}")
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
