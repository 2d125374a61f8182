"""Oracle pinning for Oldroyd-BRootk (Oldroyd_BRootk.C:129-206, the
conformation in the root-k form theta = C^(1/k)) and PTTLog's generalized
destruction function (PTTLog.C:135-152,202-230, the Mittag-Leffler function
E_{alpha,beta} of the trace).

The fixture (tests/golden/rhs_kernel.npz, written by make_golden.py) holds
right-hand sides from independent formulations: d(C^(1/k))/dt along the
Oldroyd-B conformation equation through the Daleckii-Krein derivative of
x^(1/k), and the log right-hand side of PTT with Gamma(beta) E_{alpha,beta}
taken from the closed forms E_{1,1} = exp, E_{1,2} = (e^z - 1)/z,
E_{2,1} = cosh(sqrt z)."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.golden.make_golden import PTT_GEN, ROOTK, mittag_leffler_closed

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load():
    return np.load(os.path.join(HERE, "rhs_kernel.npz"))


def rootk_params(key, **over):
    prm = dict(ROOTK[key])
    prm.update(over)
    return O.params("Oldroyd-BRootk", lam=prm["lam"], root_k=prm["k"])


def pttgen_params(key, **over):
    prm = dict(PTT_GEN[key])
    prm.update(over)
    return O.params("PTTLog", lam=prm["lam"], epsilon=prm["epsilon"], zeta=prm["zeta"], ptt_gen=prm["ab"])


def rhs_err(p, d, pre):
    S, _ = O.rhs(p, d[pre + "_gradU"], d[pre + "_theta"], d[pre + "_eigVals"], d[pre + "_eigVecs"])
    ref = d[pre + "_S"]
    return (np.abs(S - ref).max(1) / (1 + np.abs(ref).max(1))).max()


@pytest.mark.parametrize("key", list(ROOTK))
def test_rootk_rhs_matches_daleckii_krein(key):
    assert rhs_err(rootk_params(key), load(), f"rootk_{key}") < 1e-10


@pytest.mark.parametrize("key", list(PTT_GEN))
def test_ptt_generalized_rhs_matches_closed_forms(key):
    assert rhs_err(pttgen_params(key), load(), f"pttgen_{key}") < 1e-10


@pytest.mark.parametrize("kind,key,over", [("rootk", "k2", dict(k=2.02)), ("rootk", "k3p5", dict(lam=0.71)),
                                           ("pttgen", "ml12", dict(ab=(1, 1))),
                                           ("pttgen", "ml21", dict(ab=(1, 2))),
                                           ("pttgen", "ml11", dict(zeta=0.0))])
def test_planted_mistake_is_caught(kind, key, over):
    p = rootk_params(key, **over) if kind == "rootk" else pttgen_params(key, **over)
    assert rhs_err(p, load(), f"{kind}_{key}") > 1e-6


def test_series_tolerance_and_term_limit_are_honoured():
    """rTolMittagLeffler and maxIterMittagLeffler (PTTLog.C:135-136): a loose
    tolerance or a short series leaves a truncation error the size of the
    first dropped terms; the default tolerance lands within 1e-12 of exp."""
    d = load()
    pre = "pttgen_ml11"
    tight = rhs_err(pttgen_params("ml11"), d, pre)
    assert tight < 1e-12
    loose = rhs_err(O.params("PTTLog", lam=1.1, epsilon=0.25, zeta=0.1, ptt_gen=(1, 1), ml_rtol=1e-3), d, pre)
    short = rhs_err(O.params("PTTLog", lam=1.1, epsilon=0.25, zeta=0.1, ptt_gen=(1, 1), ml_max_iter=2), d, pre)
    assert loose > 1e-6 and short > 1e-3


def test_ptt_generalized_at_rest_equals_linear():
    """z = 0 (tr C = 3): the series stops after its first term, Gamma(beta) E(0)
    = 1, so the generalized function equals the linear one bit for bit."""
    d = load()
    pre = "pttgen_ml21"
    R = np.eye(3)
    th = np.zeros((4, 6))
    ev = np.broadcast_to(np.eye(3), (4, 3, 3)).copy()
    vec = np.broadcast_to(R, (4, 3, 3)).copy()
    L = d[pre + "_gradU"][:4]
    S_gen, _ = O.rhs(pttgen_params("ml21"), L, th, ev, vec)
    S_lin, _ = O.rhs(O.params("PTTLog", lam=1.3, epsilon=0.2, zeta=0.15), L, th, ev, vec)
    assert np.array_equal(S_gen, S_lin)
    assert math.isclose(math.gamma(1) * mittag_leffler_closed(2, 1, 0.0), 1.0)


def test_rootk_step_eigen_state_and_tau():
    """One phi = 0 step on a tiny mesh: theta_new = theta + dt S; the eigen
    state after calcEig holds the conformation eigenvalues (eig theta_new)^k
    (Oldroyd_BRootk.C:194-200) and tau = etaP/lambda (theta_new^k - I) for k = 2
    (:204)."""
    from rheotool_amd import cases

    d = load()
    pre = "rootk_k2"
    m = cases.cartesian(2, 2, 2)
    f = cases.make_fields(m, cases.FIELDS_TG, seed=3, lnLo=-1.0, lnHi=1.0)
    n = m.nCells
    f.theta[:] = d[pre + "_theta"][:n]
    f.eigVals[:] = d[pre + "_eigVals"][:n]
    f.eigVecs[:] = d[pre + "_eigVecs"][:n]
    f.phi[:] = 0.0
    p = rootk_params("k2")
    dt = 1e-3
    th, ev, vec, tau, _ = O.correct(m, p, f, dt, gradU=d[pre + "_gradU"][:n])
    assert np.abs(th - (f.theta + dt * d[pre + "_S"][:n])).max() < 1e-12
    full = np.array([[[t[0], t[1], t[2]], [t[1], t[3], t[4]], [t[2], t[4], t[5]]] for t in th])
    w = np.linalg.eigvalsh(full)
    assert np.allclose(np.sort(np.diagonal(ev, axis1=1, axis2=2), 1), np.sort(w, 1) ** 2, rtol=1e-12)
    want = (0.41 / 1.4) * (full @ full - np.eye(3))
    got = np.array([[[t[0], t[1], t[2]], [t[1], t[3], t[4]], [t[2], t[4], t[5]]] for t in tau])
    assert np.abs(got - want).max() < 1e-12 * (1 + np.abs(want).max())
