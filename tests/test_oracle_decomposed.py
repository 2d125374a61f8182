"""The decomposed form of the oracle's stress step (orc_decomp_*), the
all-cores CPU baseline: a decomposePar'd run with one thread per domain,
processor-patch interfaces inside Amul and sumA, DILU on each domain's own
matrix and gSum reductions ([OF9] lduMatrix::Amul / updateMatrixInterfaces,
DILUPreconditioner, processorFvPatchField).

Pinned by what the decomposition must not change:
- one domain is the serial step operation for operation (bit-equal);
- any domain assignment (slabs, or cells scattered at random, where almost
  every face is an interface) solves the same linear system, so with the
  solver driven to round-off the result equals the serial one;
- a dropped or mis-signed interface term moves the result by O(dt phi / V),
  which the second check sees at once (the planted-mistake test below).
"""
import numpy as np
import pytest

from oracle import oracle as O
from rheotool_amd import cases

TIGHT = dict(tol=1e-14, maxIter=300)


def box(seed=3):
    return cases.box3d(12, lnLo=-3, lnHi=2, perm_seed=seed, jitter=0.1)


def inlet_channel(seed=5):
    m = cases.cartesian(20, 10, 1, twoD=True, channel=True, jitter=0.1)
    f = cases.make_fields(m, cases.FIELDS_CHANNEL, seed=seed)
    rng = np.random.default_rng(seed)
    bc = O.default_bc(m)
    bc[0] = O.BC_FIXEDVALUE
    bcval = np.zeros((m.nBoundary, 6))
    s, n = m.patchStart[0], m.patchSize[0]
    bcval[s:s + n] = f.theta[m.owner[m.nInternal + s:m.nInternal + s + n]] + 0.05 * rng.normal(size=(n, 6))
    return m, f, bc, bcval


def run(m, f, p, dt, domains, ctl=None, bc=None, bcval=None, tau_old=None, scheme="upwind"):
    d = O.Decomposed(m, domains)
    try:
        d.set_fields(f, tau_old=tau_old, bcval=bcval)
        perf = d.correct(p, dt, ctl=ctl, bc=bc, scheme=scheme)
        return (*d.get(), perf)
    finally:
        d.close()


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("model,kw", [("GiesekusLog", dict(alpha=0.1)), ("FENE-PLog", dict(L2=100.0)),
                                      ("SaramitoLog", dict(tau0=0.3))])
def test_one_domain_is_the_serial_step_bit_for_bit(model, kw):
    m, f = box()
    p = O.params(model, lam=2.0, **kw)
    tau0 = 0.5 * np.random.default_rng(1).normal(size=(m.nCells, 6)) if model == "SaramitoLog" else None
    ref = O.correct(m, p, f, 5e-3, tau_old=tau0)
    got = run(m, f, p, 5e-3, np.zeros(m.nCells, np.int32), tau_old=tau0)
    for i in range(4):
        assert np.array_equal(got[i], ref[i]), i
    assert [q.nIterations for q in got[4]] == [q.nIterations for q in ref[4]]
    assert [q.finalResidual for q in got[4]] == [q.finalResidual for q in ref[4]]


def test_one_domain_with_a_fixed_value_inlet_and_bounded_scheme():
    m, f, bc, bcval = inlet_channel()
    p = O.params("Oldroyd-BLog", lam=1.0)
    for scheme in ("upwind", "bounded upwind"):
        ref = O.correct(m, p, f, 1e-3, bc=bc, bcval=bcval, scheme=scheme)
        got = run(m, f, p, 1e-3, np.zeros(m.nCells, np.int32), bc=bc, bcval=bcval, scheme=scheme)
        for i in range(4):
            assert np.array_equal(got[i], ref[i]), (scheme, i)


@pytest.mark.parametrize("n", [2, 3, 7])
def test_slabs_solve_the_undecomposed_system(n):
    m, f = box()
    p = O.params("GiesekusLog", lam=2.0, alpha=0.1)
    ctl = O.controls(**TIGHT)
    ref = O.correct(m, p, f, 5e-3, ctl=ctl)
    got = run(m, f, p, 5e-3, O.simple_domains(m, n), ctl=ctl)
    assert rel(got[0], ref[0]) < 1e-10 and rel(got[3], ref[3]) < 1e-10
    assert all(q.converged for q in got[4])


def test_scattered_domains_and_inlet_solve_the_undecomposed_system():
    """Cells dealt to 5 domains at random: nearly every face is a processor
    face, the DILU of each domain sees almost no couplings, and the solve
    still reaches the serial solution."""
    m, f, bc, bcval = inlet_channel()
    p = O.params("Oldroyd-BLog", lam=1.0)
    ctl = O.controls(**TIGHT)
    dom = np.random.default_rng(0).integers(0, 5, m.nCells).astype(np.int32)
    dom[:5] = np.arange(5)
    for scheme in ("upwind", "bounded upwind"):
        ref = O.correct(m, p, f, 1e-3, ctl=ctl, bc=bc, bcval=bcval, scheme=scheme)
        got = run(m, f, p, 1e-3, dom, ctl=ctl, bc=bc, bcval=bcval, scheme=scheme)
        assert rel(got[0], ref[0]) < 1e-10 and rel(got[3], ref[3]) < 1e-10, scheme


def test_a_dropped_interface_term_is_seen():
    """Planted mistake: zero the flux through the processor faces (as if the
    interface coefficients were lost).  The decomposed result then leaves
    the serial solution by far more than the bar above."""
    m, f = box()
    p = O.params("GiesekusLog", lam=2.0, alpha=0.1)
    ctl = O.controls(**TIGHT)
    dom = O.simple_domains(m, 3)
    ref = O.correct(m, p, f, 5e-3, ctl=ctl)
    cut = dom[m.owner[:m.nInternal]] != dom[m.neighbour]
    assert cut.any()
    g = cases.Fields(**{k: getattr(f, k) for k in f.__dataclass_fields__})
    g.phi = f.phi.copy()
    g.phi[:m.nInternal][cut] = 0.0
    got = run(m, g, p, 5e-3, dom, ctl=ctl)
    assert rel(got[0], ref[0]) > 1e-6


def test_simple_domains_cut_equal_counts_along_the_axis():
    m, f = box()
    dom = O.simple_domains(m, 4, axis=2)
    assert np.bincount(dom).tolist() == [m.nCells // 4] * 4
    z = m.C[:, 2]
    assert all(z[dom == k].max() <= z[dom == k + 1].min() for k in range(3))


def test_unsupported_requests_raise():
    m, f = box()
    p = O.params("GiesekusLog", lam=2.0, alpha=0.1)
    with pytest.raises(ValueError):
        O.Decomposed(m, np.full(m.nCells, -1, np.int32))
    d = O.Decomposed(m, O.simple_domains(m, 2))
    d.set_fields(f)
    with pytest.raises(ValueError):
        d.correct(p, 5e-3, scheme="cubista")
    with pytest.raises(ValueError):
        d.correct(p, 5e-3, ctl=O.controls(solver=O.PBICG))
    d.close()
