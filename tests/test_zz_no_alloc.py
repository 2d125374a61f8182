"""rt_symm_solve_noalloc, the segregated solve without device allocation,
against rt_symm_solve on the same systems: the same kernels on the same
inputs, so psi and the solver performance must agree bit for bit.  Two calls
in a row, the second on a smaller mesh, for each solver and precision; a
mesh beyond the scratch array is refused and leaves the next call intact."""
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


def system(m, seed):
    f = cases.make_fields(m, cases.FIELDS_TG, seed=seed, lnLo=-3.0, lnHi=2.0)
    S = np.random.default_rng(seed).standard_normal((m.nCells, 6))
    sysm = O.assemble(m, f.phi, f.theta, S, 1e-3)
    return sysm, f.theta


def both(ctx, dm, m, sysm, psi, controls):
    kw = dict(internalCoeffs=sysm["iC"][:m.nBoundary], boundaryCoeffs=sysm["bC"][:m.nBoundary])
    a = ctx.symm_solve(dm, sysm["diag"], sysm["lower"], sysm["upper"], sysm["source"], psi, controls, **kw)
    used = ctx.bytes_in_use
    b = ctx.symm_solve(dm, sysm["diag"], sysm["lower"], sysm["upper"], sysm["source"], psi, controls,
                       noalloc=True, **kw)
    assert ctx.bytes_in_use == used  # nothing taken from the context's allocator
    return a, b


@pytest.mark.parametrize("solver,precision", [("PBiCGStab", "double"), ("PBiCGStab", "mixed"),
                                              ("PBiCG", "double"), ("PBiCG", "mixed")])
def test_noalloc_solve_equals_symm_solve(ctx, solver, precision):
    big = cases.cartesian(36, 36, 36, jitter=0.1, perm_seed=4)
    small = cases.cartesian(20, 20, 1, twoD=True, channel=True)
    controls = abi.solver_dict(1e-10, precision=precision, solver=solver)
    for m, seed in ((big, 3), (small, 5)):
        print(f"[case] noalloc {solver} {precision} N={m.nCells}", flush=True)
        dm = ctx.mesh(m)
        sysm, psi = system(m, seed)
        (xa, pa), (xb, pb) = both(ctx, dm, m, sysm, psi, controls)
        assert np.array_equal(xa, xb)
        for q in range(6):
            assert (pa[q].nIterations, pa[q].nRefinements, pa[q].converged, pa[q].solved) == \
                   (pb[q].nIterations, pb[q].nRefinements, pb[q].converged, pb[q].solved)
            assert pa[q].finalResidual == pb[q].finalResidual
            if m.validComponents[q] > 0:
                assert pb[q].converged == 1 and pb[q].finalResidual < 1e-10
        dm.close()


def test_noalloc_refuses_meshes_beyond_the_scratch_array(ctx):
    m = cases.cartesian(72, 72, 72)
    dm = ctx.mesh(m)
    sysm, psi = system(m, 7)
    print(f"[case] noalloc oversize N={m.nCells}", flush=True)
    with pytest.raises(abi.RheoError) as e:
        ctx.symm_solve(dm, sysm["diag"], sysm["lower"], sysm["upper"], sysm["source"], psi,
                       abi.solver_dict(1e-10), noalloc=True)
    assert e.value.code == abi.RT_E_BAD_ARG and "RT_NOALLOC_SCRATCH_BYTES" in e.value.msg
    dm.close()
    small = cases.cartesian(12, 12, 12)
    ds = ctx.mesh(small)
    sysm, psi = system(small, 8)
    (xa, _), (xb, _) = both(ctx, ds, small, sysm, psi, abi.solver_dict(1e-10))
    assert np.array_equal(xa, xb)
    ds.close()
