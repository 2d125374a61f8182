"""Pinning of the oracle's restatement of [OF9] fvMatrix::relax (the
equation relaxation factor of Oldroyd_BLog.C:166 and the other models), by the
properties under-relaxation must have, not by retyping its formula:

- the exact solution of the unrelaxed system still solves the relaxed one
  when it is the field the relaxation is taken about (the diagonal and the
  source move together), so relaxation changes the iteration, not the answer;
- the relaxed matrix is diagonally dominant by the factor 1/alpha;
- alpha = 1 on a diagonally dominant matrix changes nothing in the cells
  with no inflow through a patch (an inflowing zeroGradient patch face has a
  negative coefficient, which the relaxation replaces by its magnitude).
A relaxation that dropped the source term, flipped a sign or forgot the
boundary coefficients fails the first property."""
import numpy as np
import pytest
import scipy.sparse
import scipy.sparse.linalg

from oracle import oracle as O
from rheotool_amd import cases


def system(seed=3, inlet=True):
    m = cases.cartesian(8, 6, 1, twoD=True, channel=True, jitter=0.1)
    f = cases.make_fields(m, cases.FIELDS_CHANNEL, seed=seed)
    rng = np.random.default_rng(seed)
    S = rng.normal(size=(m.nCells, 6))
    bc = O.default_bc(m)
    bcval = np.zeros((m.nBoundary, 6))
    if inlet:
        bc[0] = O.BC_FIXEDVALUE
        s, n = m.patchStart[0], m.patchSize[0]
        bcval[s:s + n] = rng.normal(size=(n, 6))
    sysm = O.assemble(m, f.phi, f.theta, S, 0.05, bc=bc, bcval=bcval)
    return m, f, bc, sysm


def dense(m, D, sysm, q):
    """Component q's matrix and source with the patch coefficients folded in."""
    N, F = m.nCells, m.nInternal
    diag = D.copy()
    src = sysm["source"][:, q].copy()
    for b in range(m.nBoundary):
        c = m.owner[F + b]
        diag[c] += sysm["iC"][b, q]
        src[c] += sysm["bC"][b, q]
    A = scipy.sparse.coo_matrix((np.concatenate([diag, sysm["upper"], sysm["lower"]]),
                                 (np.concatenate([np.arange(N), m.owner[:F], m.neighbour]),
                                  np.concatenate([np.arange(N), m.neighbour, m.owner[:F]]))),
                                shape=(N, N)).tocsc()
    return A, src


@pytest.mark.parametrize("alpha", [0.3, 0.7, 0.95])
@pytest.mark.parametrize("inlet", [False, True])
def test_relaxation_keeps_the_solution(alpha, inlet):
    m, f, bc, sysm = system(inlet=inlet)
    valid = m.validComponents
    star = np.zeros((m.nCells, 6))
    for q in range(6):
        if valid[q] > 0:
            A, b = dense(m, sysm["diag"], sysm, q)
            star[:, q] = scipy.sparse.linalg.spsolve(A, b)
    D, S = O.relax(m, sysm, alpha, star, bc=bc)
    assert not np.allclose(D, sysm["diag"])  # it did relax
    relaxed = dict(sysm, diag=D, source=S)
    for q in range(6):
        if valid[q] > 0:
            A, b = dense(m, D, relaxed, q)
            r = A @ star[:, q] - b
            assert np.abs(r).max() < 1e-11 * np.abs(b).max(), (q, np.abs(r).max())
    # about another field the relaxed system has another solution: the source term matters
    D2, S2 = O.relax(m, sysm, alpha, star + 0.1, bc=bc)
    A, b = dense(m, D2, dict(sysm, diag=D2, source=S2), 0)
    assert np.abs(A @ star[:, 0] - b).max() > 1e-6


@pytest.mark.parametrize("alpha", [0.5, 0.8])
def test_relaxed_matrix_is_diagonally_dominant(alpha):
    m, f, bc, sysm = system()
    D, _ = O.relax(m, sysm, alpha, f.theta, bc=bc)
    F = m.nInternal
    off = np.zeros(m.nCells)
    np.add.at(off, m.owner[:F], np.abs(sysm["upper"]))
    np.add.at(off, m.neighbour, np.abs(sysm["lower"]))
    Dq = D.copy()
    for b in range(m.nBoundary):  # the smallest patch coefficient over the components
        Dq[m.owner[F + b]] += sysm["iC"][b].min()
    assert np.all(alpha * Dq >= off * (1 - 1e-14))


def test_unit_factor_leaves_a_dominant_matrix_alone():
    m, f, bc, sysm = system(inlet=False)  # zeroGradient: the same patch coefficient in every component
    D, S = O.relax(m, sysm, 1.0, f.theta, bc=bc)
    F = m.nInternal
    inflow = np.zeros(m.nCells, bool)
    inflow[m.owner[F:][sysm["iC"][:m.nBoundary, 0] < 0]] = True
    assert inflow.any() and (~inflow).sum() > m.nCells // 2
    # to rounding: a patch coefficient is added and subtracted again
    assert np.allclose(D[~inflow], sysm["diag"][~inflow], rtol=1e-15, atol=0)
    assert np.allclose(S[~inflow], sysm["source"][~inflow], rtol=1e-14, atol=1e-15)
    assert np.all(D[inflow] >= sysm["diag"][inflow]) and np.any(D[inflow] > 1.5 * sysm["diag"][inflow])
