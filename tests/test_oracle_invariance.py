"""Invariances the log-conformation source must have, checked on the oracle
for all ten models (constitutiveEq.C:323-358 decomposeGradU, the model
sources of the *Log classes).

- Frame indifference: rotating the state and the flow by R (theta -> R theta
  R^T, grad U -> R grad U R^T, eigenvectors -> R V, the previous tau -> R tau
  R^T) rotates the source: S -> R S R^T, and leaves the FENE-type factor f
  unchanged.
- The eigen state is a basis choice: flipping the sign of eigenvectors or
  reordering the eigenpairs (values with their vectors) changes nothing.

Neither follows from retyping the formulas: decomposeGradU works in the
eigenbasis (M = V^T L^T V, Omega~ from the eigenvalue differences), and the
identities hold only if every product is formed on the right side.  The
last test shows the check is not vacuous: an unrotated gradient fails it.
"""
import numpy as np
import pytest

from oracle import oracle as O

MODELS = {
    "Oldroyd-BLog": dict(lam=1.7),
    "GiesekusLog": dict(lam=0.8, alpha=0.2),
    "FENE-PLog": dict(lam=2.3, L2=100.0),
    "PTTLog": dict(lam=1.1, epsilon=0.25, zeta=0.1),
    "PTTLog-exp": dict(lam=1.1, epsilon=0.25, zeta=0.1, ptt_exp=True),
    "FENE-CRLog": dict(lam=1.5, L2=50.0),
    "WhiteMetznerCYLog": dict(lam=1.2, m=0.6, wm_n=0.5, K=0.8, Lc=1.1, a=2.0, b=1.5),
    "Rolie-PolyLog": dict(lambdaR=0.3, lambdaD=2.0, beta=0.5, delta=-0.5, chiMax=3.0),
    "XPomPomLog": dict(lambdaS=0.4, lambdaB=1.6, alpha=0.1, q=3.0, n=2.0),
    "SaramitoLog": dict(lam=1.3, tau0=0.3, sar_n=1.0, dims=(1, 1, 1), ptt_function="linear",
                        epsilon=0.2, zeta=0.1),
    "SaramitoLog-hb": dict(lam=0.8, tau0=0.25, sar_n=0.75, sar_k=1.5, dims=(1, 1, 1)),
    "BMPLog": dict(lam=1.0, G0=0.05, k_bmp=1.0, Phi0=2.0, PhiInf=200.0),
}

IJ = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def full(s6):
    t = np.empty(s6.shape[:-1] + (3, 3))
    for k, (i, j) in enumerate(IJ):
        t[..., i, j] = t[..., j, i] = s6[..., k]
    return t


def six(t):
    return np.stack([t[..., i, j] for i, j in IJ], axis=-1)


def rotations(n, rng):
    q, r = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    q = q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]
    q[np.linalg.det(q) < 0, :, 0] *= -1
    return q


def state(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, 3, 3))
    th = 0.6 * (A + A.transpose(0, 2, 1))
    w, V = np.linalg.eigh(th)
    ev = np.zeros((n, 3, 3))
    ev[:, [0, 1, 2], [0, 1, 2]] = np.exp(w)
    L = 0.7 * rng.normal(size=(n, 3, 3))
    T = rng.normal(size=(n, 3, 3))
    tau = 0.8 * (T + T.transpose(0, 2, 1))
    aux = np.zeros((n, 6))
    aux[:, 0] = rng.uniform(1.0, 50.0, n)  # BMP's fluidity
    return rng, th, ev, V, L, tau, aux


def source(key, L, th, ev, V, tau, aux):
    model = key.split("-exp")[0].split("-hb")[0]
    p = O.params(model, **MODELS[key])
    S, f = O.rhs(p, np.ascontiguousarray(L), six(th), np.ascontiguousarray(ev), np.ascontiguousarray(V),
                 tau_old=six(tau), aux=aux)
    return full(S), f


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("key", sorted(MODELS))
def test_source_is_frame_indifferent(key):
    rng, th, ev, V, L, tau, aux = state(400, 7)
    R = rotations(len(th), rng)
    Rt = R.transpose(0, 2, 1)
    S, f = source(key, L, th, ev, V, tau, aux)
    S2, f2 = source(key, R @ L @ Rt, R @ th @ Rt, ev, R @ V, R @ tau @ Rt, aux)
    assert rel(S2, R @ S @ Rt) < 1e-11, rel(S2, R @ S @ Rt)
    assert np.allclose(f2, f, rtol=1e-12, atol=0)


@pytest.mark.parametrize("key", sorted(MODELS))
def test_source_ignores_eigenvector_signs_and_order(key):
    rng, th, ev, V, L, tau, aux = state(400, 8)
    n = len(th)
    S, f = source(key, L, th, ev, V, tau, aux)
    perm = np.array([rng.permutation(3) for _ in range(n)])
    signs = rng.choice([-1.0, 1.0], size=(n, 3))
    V2 = np.take_along_axis(V, perm[:, None, :], axis=2) * signs[:, None, :]
    w = np.diagonal(ev, axis1=1, axis2=2)
    ev2 = np.zeros_like(ev)
    ev2[:, [0, 1, 2], [0, 1, 2]] = np.take_along_axis(w, perm, axis=1)
    S2, f2 = source(key, L, th, ev2, V2, tau, aux)
    assert rel(S2, S) < 1e-11, rel(S2, S)
    assert np.allclose(f2, f, rtol=1e-12, atol=0)


def test_an_unrotated_gradient_breaks_the_identity():
    rng, th, ev, V, L, tau, aux = state(200, 9)
    R = rotations(len(th), rng)
    Rt = R.transpose(0, 2, 1)
    S, _ = source("GiesekusLog", L, th, ev, V, tau, aux)
    S2, _ = source("GiesekusLog", L, R @ th @ Rt, ev, R @ V, tau, aux)
    assert rel(S2, R @ S @ Rt) > 1e-3
