"""The CPU FP64 oracle pinned against independent numpy/scipy results.

Every expectation comes from tests/golden (eigh, Daleckii-Krein log
derivative, fsolve steady states, dense assembly + spsolve) or from closed
forms; none of them re-types the oracle's own formulas."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from rheotool_amd import cases
from tests.golden.make_golden import MODELS, full

G = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(G, name))


def _params(model, prm, lam=None):
    return O.params(model, etaS=0.59, etaP=0.41, lam=prm["lambda_"] if lam is None else lam,
                    alpha=prm["alpha"], L2=prm["L2"], epsilon=prm["epsilon"], zeta=prm["zeta"])


# ------------------------------------------------------------ calcEig
@pytest.mark.parametrize("method", [0, 1])
def test_calc_eig_vs_lapack(method):
    d = load("eig.npz")
    ev, vec = O.calc_eig(d["theta"], method)
    lam = np.log(np.diagonal(ev, axis1=1, axis2=2))
    scale = np.maximum(1.0, np.abs(d["evals"]).max(1, keepdims=True))
    assert np.all(np.diff(lam, axis=1) >= 0), "eigenvalues must be ascending"
    assert np.max(np.abs(lam - d["evals"]) / scale) < 1e-13
    # reconstructed conformation tensor, sign/order invariant
    C = np.einsum("nik,nk,njk->nij", vec, np.diagonal(ev, axis1=1, axis2=2), vec)
    relC = np.abs(C - d["C"]).max((1, 2)) / np.abs(d["C"]).max((1, 2))
    assert relC.max() < 1e-12
    # orthonormal eigenvectors, reconstruction of theta itself
    VtV = np.einsum("nki,nkj->nij", vec, vec)
    assert np.abs(VtV - np.eye(3)).max() < 1e-13
    th = np.einsum("nik,nk,njk->nij", vec, lam, vec)
    assert np.abs(th - full(d["theta"])).max() < 1e-13 * 30
    # off-diagonal of eigVals zero (calcEig: vals *= 0 then diagonal)
    off = ev.copy()
    for i in range(3):
        off[:, i, i] = 0
    assert not off.any()


def test_eig_solvers_agree_on_projectors():
    d = load("eig.npz")
    e0, v0 = O.calc_eig(d["theta"], 0)
    e1, v1 = O.calc_eig(d["theta"], 1)
    lam = d["evals"]
    gaps = np.diff(lam, axis=1)
    for k in range(3):
        ok = np.ones(len(lam), bool)
        if k > 0:
            ok &= gaps[:, k - 1] > 1e-6
        if k < 2:
            ok &= gaps[:, k] > 1e-6
        P0 = np.einsum("ni,nj->nij", v0[:, :, k], v0[:, :, k])[ok]
        P1 = np.einsum("ni,nj->nij", v1[:, :, k], v1[:, :, k])[ok]
        assert np.abs(P0 - P1).max() < 1e-9


# ---------------------------------------------------- sources (a5-a7)
@pytest.mark.parametrize("model", list(MODELS))
def test_rhs_matches_daleckii_krein(model):
    d = load("rhs.npz")
    k = model.replace("-", "_")
    p = _params(model, MODELS[model])
    S, _ = O.rhs(p, d[k + "_gradU"], d[k + "_theta"], d[k + "_eigVals"], d[k + "_eigVecs"])
    ref = d[k + "_S"]
    err = np.abs(S - ref).max(1) / (1 + np.abs(ref).max(1))
    assert err.max() < 1e-10, err.max()


def test_rhs_planted_mistake_is_caught():
    """A wrong relaxation time (or a dropped term) must fail the identity."""
    d = load("rhs.npz")
    k = "Oldroyd_BLog"
    prm = dict(MODELS["Oldroyd-BLog"])
    p = _params("Oldroyd-BLog", prm, lam=prm["lambda_"] * 1.01)
    S, _ = O.rhs(p, d[k + "_gradU"], d[k + "_theta"], d[k + "_eigVals"], d[k + "_eigVecs"])
    assert np.abs(S - d[k + "_S"]).max() > 1e-4
    # transposed velocity gradient
    p = _params("Oldroyd-BLog", prm)
    S, _ = O.rhs(p, np.swapaxes(d[k + "_gradU"], 1, 2), d[k + "_theta"], d[k + "_eigVals"],
                 d[k + "_eigVecs"])
    assert np.abs(S - d[k + "_S"]).max() > 1e-3


@pytest.mark.parametrize("model", list(MODELS))
def test_rhs_vanishes_at_steady_homogeneous_flow(model):
    d = load("steady.npz")
    p = _params(model, MODELS[model])
    key = model.replace("-", "_") + "__"
    cases_ = sorted({k.rsplit("__", 1)[0] for k in d.files if k.startswith(key)})
    assert cases_
    for c in cases_:
        C = d[c + "__C"]
        w, V = np.linalg.eigh(C)
        th = V @ np.diag(np.log(w)) @ V.T
        th6 = np.array([[th[0, 0], th[0, 1], th[0, 2], th[1, 1], th[1, 2], th[2, 2]]])
        S, _ = O.rhs(p, d[c + "__gradU"][None], th6, np.diag(w)[None], V[None])
        assert np.abs(S).max() < 1e-9 * max(1.0, np.abs(np.log(w)).max()), (c, S)


def test_oldroyd_b_shear_closed_form():
    """OB steady shear: C = [[1+2Wi^2, Wi, 0],[Wi, 1, 0],[0,0,1]], tau_xy = etaP gd."""
    lam, gd, etaP = 1.3, 0.7, 0.41
    Wi = lam * gd
    C = np.array([[1 + 2 * Wi ** 2, Wi, 0], [Wi, 1, 0], [0, 0, 1.0]])
    w, V = np.linalg.eigh(C)
    th = V @ np.diag(np.log(w)) @ V.T
    th6 = np.array([[th[0, 0], th[0, 1], th[0, 2], th[1, 1], th[1, 2], th[2, 2]]])
    p = O.params("Oldroyd-BLog", etaP=etaP, lam=lam)
    gU = np.zeros((1, 3, 3))
    gU[0, 1, 0] = gd
    S, _ = O.rhs(p, gU, th6, np.diag(w)[None], V[None])
    assert np.abs(S).max() < 1e-12
    tau = O.tau(p, np.diag(w)[None], V[None])[0]
    assert abs(tau[1] - etaP * gd) < 1e-13
    assert abs((tau[0] - tau[3]) - 2 * etaP * lam * gd ** 2) < 1e-12


# ---------------------------------------------------- assembly (a8-a9)
def _mesh_from_spec(spec):
    nx, ny, nz, twoD, channel, jitter, perm, kind = spec
    return cases.cartesian(int(nx), int(ny), int(nz), twoD=bool(twoD), channel=bool(channel),
                           jitter=float(jitter), perm_seed=int(perm))


LDU = ["2x2x1", "3x3x1", "3x3x3", "4x4x4j", "5x4x1p"]


def _to_dense(m, sys):
    N, F = m.nCells, m.nInternal
    A = np.diag(sys["diag"].copy())
    for f in range(F):
        A[m.owner[f], m.neighbour[f]] += sys["upper"][f]
        A[m.neighbour[f], m.owner[f]] += sys["lower"][f]
    for p in range(m.nPatches):
        if m.patchKind[p] == cases.EMPTY:
            continue
        for k in range(m.patchSize[p]):
            b = m.patchStart[p] + k
            A[m.owner[F + b], m.owner[F + b]] += sys["iC"][b, 0]
    return A


@pytest.mark.parametrize("name", LDU)
def test_assembly_matches_dense(name):
    d = load(f"ldu_{name}.npz")
    m = _mesh_from_spec(d["spec"])
    sys = O.assemble(m, d["phi"], d["theta"], d["S"], float(d["dt"]))
    A = _to_dense(m, sys)
    assert np.abs(A - d["A"]).max() < 1e-13 * np.abs(d["A"]).max()
    assert np.abs(sys["source"] - d["b"]).max() < 1e-13 * np.abs(d["b"]).max()


@pytest.mark.parametrize("name", LDU)
@pytest.mark.parametrize("precond", [O.PC_DILU, O.PC_DIAGONAL, O.PC_NONE])
@pytest.mark.parametrize("solver", [O.PBICGSTAB, O.PBICG])
def test_segregated_solve_matches_spsolve(name, precond, solver):
    d = load(f"ldu_{name}.npz")
    m = _mesh_from_spec(d["spec"])
    sys = O.assemble(m, d["phi"], d["theta"], d["S"], float(d["dt"]))
    ctl = O.controls(solver=solver, precond=precond, tol=1e-12)
    x, perf = O.solve_segregated(m, sys, d["theta"], ctl)
    valid = m.validComponents
    for q in range(6):
        if valid[q] < 0:
            assert np.array_equal(x[:, q], d["theta"][:, q])  # frozen in 2D
            assert perf[q].solved == 0
            continue
        ref = d["x"][:, q]
        assert np.linalg.norm(x[:, q] - ref) / np.linalg.norm(ref) < 1e-10
        assert perf[q].converged == 1


def test_row_sum_invariant():
    """A.1 = V/dt + sum of boundary outflow + net face flux (upwind, Euler)."""
    m, f = cases.box3d(6, lnLo=-2, lnHi=2)
    S = np.zeros((m.nCells, 6))
    dt = 0.01
    sys = O.assemble(m, f.phi, f.theta, S, dt)
    A = _to_dense(m, sys)
    net = np.zeros(m.nCells)
    np.add.at(net, m.owner[:m.nInternal], f.phi[:m.nInternal])
    np.add.at(net, m.neighbour, -f.phi[:m.nInternal])
    np.add.at(net, m.owner[m.nInternal:], f.phi[m.nInternal:])
    assert np.allclose(A.sum(1), m.V / dt + net, rtol=1e-12, atol=1e-14)
    # column sums: conservation, V/dt + boundary outflow of that cell
    bnd = np.zeros(m.nCells)
    np.add.at(bnd, m.owner[m.nInternal:], f.phi[m.nInternal:])
    assert np.allclose(A.sum(0), m.V / dt + bnd, rtol=1e-12, atol=1e-14)


def test_foam_residual_properties():
    d = load("ldu_3x3x3.npz")
    m = _mesh_from_spec(d["spec"])
    sys = O.assemble(m, d["phi"], d["theta"], d["S"], float(d["dt"]))
    l, u = np.ascontiguousarray(m.owner[:m.nInternal]), m.neighbour
    b = np.ascontiguousarray(d["b"][:, 0])
    x = np.ascontiguousarray(d["x"][:, 0])
    diag = sys["diag"].copy()  # addBoundaryDiag
    np.add.at(diag, m.owner[m.nInternal:], sys["iC"][:, 0])
    sys["diag"] = diag
    r0 = O.foam_residual(m.nCells, l, u, sys["diag"], sys["lower"], sys["upper"], b, x)
    assert r0 < 1e-14
    x1 = x + 1e-3
    r1 = O.foam_residual(m.nCells, l, u, sys["diag"], sys["lower"], sys["upper"], b, x1)
    r2 = O.foam_residual(m.nCells, l, u, 3 * sys["diag"], 3 * sys["lower"], 3 * sys["upper"],
                         3 * b, x1)
    assert r1 > 1e-6 and abs(r1 - r2) < 1e-12 * r1
    # by hand from the sparseSolver.C:129-180 definition
    A = _to_dense(m, dict(sys, iC=0 * sys["iC"]))
    Ax = A @ x1
    pA = A.sum(1) * x1.mean()
    ref = np.abs(b - Ax).sum() / (np.abs(Ax - pA).sum() + np.abs(b - pA).sum() + 1e-20)
    assert abs(r1 - ref) < 1e-12 * ref


# ------------------------------------------------ gradient, end to end
def test_gauss_gradient_exact_for_linear_field():
    m = cases.cartesian(5, 4, 3, jitter=0.2, perm_seed=9)
    Gm = np.array([[0.3, -1.2, 0.5], [2.0, 0.1, -0.7], [0.4, 0.9, -0.4]])
    U = m.C @ Gm
    Ub = m.Cf[m.nInternal:] @ Gm
    g = O.grad(m, U, Ub)
    # Gauss linear is exact for linear fields only on meshes where the face
    # interpolation is exact; on the jittered mesh check the divergence
    # identity sum_f Sf (x) U_f = V grad instead, on the Cartesian mesh exactness.
    m2 = cases.cartesian(5, 4, 3)
    g2 = O.grad(m2, m2.C @ Gm, m2.Cf[m2.nInternal:] @ Gm)
    assert np.abs(g2 - Gm).max() < 1e-12
    assert np.abs(g - Gm).max() < 0.5  # consistent to first order on skewed cells


def test_one_cell_homogeneous_shear_converges_to_closed_form():
    """rheoTestFoam-style: phi = 0, one cell, prescribed gradU; many steps reach
    the Oldroyd-B steady shear state."""
    m = cases.cartesian(1, 1, 1)
    f = cases.make_fields(m, cases.FIELDS_TG, seed=3, lnLo=-0.1, lnHi=0.1)
    f.phi[:] = 0.0
    lam, gd = 0.5, 1.2
    gU = np.zeros((1, 3, 3))
    gU[0, 1, 0] = gd
    p = O.params("Oldroyd-BLog", etaP=0.41, lam=lam)
    for _ in range(3000):
        th, ev, vec, tau, perf = O.correct(m, p, f, 5e-3, gradU=gU)
        f.theta, f.eigVals, f.eigVecs = th, ev, vec
    Wi = lam * gd
    C = vec[0] @ ev[0] @ vec[0].T
    Cref = np.array([[1 + 2 * Wi ** 2, Wi, 0], [Wi, 1, 0], [0, 0, 1.0]])
    assert np.abs(C - Cref).max() < 1e-8
    assert abs(tau[0, 1] - 0.41 * gd) < 1e-8


def test_one_step_explicit_limit():
    """phi = 0: the solve is diagonal, theta_new = theta_old + dt S exactly."""
    m = cases.cartesian(2, 2, 2)
    f = cases.make_fields(m, cases.FIELDS_TG, seed=3, lnLo=-2, lnHi=2)
    f.phi[:] = 0.0
    p = O.params("GiesekusLog", etaP=0.41, lam=0.7, alpha=0.1)
    S, _ = O.rhs(p, f.gradU, f.theta, f.eigVals, f.eigVecs)
    dt = 1e-3
    th, *_ = O.correct(m, p, f, dt, gradU=f.gradU)
    assert np.abs(th - (f.theta + dt * S)).max() < 1e-12


def test_inverse_quirk_closed_form():
    """[OF9] inv(tensorField) adds 1 to the diagonal components that are zero
    in the field's first tensor before inverting every tensor, and subtracts
    it after (its treatment of 2D fields).  With V = I and no velocity
    gradient the Oldroyd-B source is (inv(Lambda) - I)/lambda, so the quirk's
    inverse is read off directly: the first cell's diag(0, b, c) inverts to
    diag(0, 1/b, 1/c), and every other cell's xx to 1/(a + 1) - 1.  Without
    the quirk (the default) the other cells invert plainly."""
    lam = 2.0
    ev = np.zeros((3, 3, 3))
    ev[0] = np.diag([0.0, 2.0, 4.0])
    ev[1] = np.diag([0.5, 2.0, 4.0])
    ev[2] = np.diag([3.0, 0.25, 1.0])
    vec = np.broadcast_to(np.eye(3), (3, 3, 3)).copy()
    th = np.zeros((3, 6))
    g = np.zeros((3, 3, 3))
    S, _ = O.rhs(O.params("Oldroyd-BLog", lam=lam, of_inv_quirk=True), g, th, ev, vec)
    inv = np.array([[0.0, 0.5, 0.25], [1 / 1.5 - 1, 0.5, 0.25], [1 / 4.0 - 1, 4.0, 1.0]])
    want = (inv - 1.0) / lam
    assert np.allclose(S[:, [0, 3, 5]], want, rtol=1e-15, atol=1e-16)
    assert np.all(S[:, [1, 2, 4]] == 0)
    S0, _ = O.rhs(O.params("Oldroyd-BLog", lam=lam), g[1:], th[1:], ev[1:], vec[1:])
    assert np.allclose(S0[:, [0, 3, 5]], (1 / np.diagonal(ev[1:], axis1=1, axis2=2) - 1) / lam, rtol=1e-15)
