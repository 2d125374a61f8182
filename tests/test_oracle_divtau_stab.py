"""Oracle pinning for the stabilization terms of constitutiveEq::divTau
(constitutiveEq.C:88-125): BSD subtracts fvc::laplacian(etaP/rho, U) and
coupling subtracts fvc::div((etaP/rho) fvc::grad(U)), both beside
fvc::div(tau/rho).  Pinned by closed forms on Cartesian meshes: both terms
vanish for a linear velocity with exact patch values, and for U_x = y^2 with
walls at y = 0 and 1 the Gauss-linear operators give 2 in the interior and the
hand-derived wall-cell values (BSD: 3/2 in the wall cell; coupling: 13/8 in the
wall cell, 15/8 in the next one, from the patch gradient
G_P + n (snGrad - n.G_P)).  The non-orthogonal correction of the corrected
laplacian is pinned by a property: on a jittered mesh it removes most of the
orthogonal scheme's error for a linear field."""
import numpy as np
import pytest

from oracle import oracle as O
from rheotool_amd import cases


def _lin(m, rng):
    A = rng.normal(size=(3, 3))
    u0 = rng.normal(size=3)
    return u0[None, :] + m.C @ A.T, u0[None, :] + m.Cf[m.nInternal:] @ A.T


@pytest.mark.parametrize("stab", ["BSD", "coupling"])
@pytest.mark.parametrize("lap", ["corrected", "orthogonal", "uncorrected"])
def test_stab_terms_vanish_for_linear_velocity(stab, lap):
    m = cases.cartesian(5, 4, 3, L=(1.0, 0.7, 1.3))
    U, Ub = _lin(m, np.random.default_rng(1))
    tau = np.zeros((m.nCells, 6))
    out = O.div_tau_stab(m, U, Ub, tau, None, rho=1.3, etaP=0.41, stab=stab, lap=lap)
    assert np.abs(out).max() < 1e-11


def test_none_is_div_tau():
    m = cases.cartesian(4, 4, 3, jitter=0.1)
    rng = np.random.default_rng(2)
    U, Ub = rng.normal(size=(m.nCells, 3)), rng.normal(size=(m.nBoundary, 3))
    tau, tauB = rng.normal(size=(m.nCells, 6)), rng.normal(size=(m.nBoundary, 6))
    assert np.array_equal(O.div_tau_stab(m, U, Ub, tau, tauB, 1.7, 0.4, "none"), O.div_tau(m, tau, tauB, 1.7))


@pytest.mark.parametrize("stab", ["BSD", "coupling"])
def test_stab_quadratic_profile_closed_form(stab):
    ny = 8
    m = cases.cartesian(4, ny, 3)
    h = 1.0 / ny
    U = np.zeros((m.nCells, 3))
    U[:, 0] = m.C[:, 1] ** 2
    Ub = np.zeros((m.nBoundary, 3))
    Ub[:, 0] = m.Cf[m.nInternal:, 1] ** 2
    etaP, rho = 0.41, 1.3
    gam = etaP / rho
    rng = np.random.default_rng(3)
    tau = rng.normal(size=(m.nCells, 6))
    base = O.div_tau(m, tau, None, rho)
    out = O.div_tau_stab(m, U, Ub, tau, None, rho=rho, etaP=etaP, stab=stab)
    term = (base - out)[:, 0] / gam  # the operator applied to U_x
    j = np.rint(m.C[:, 1] / h - 0.5).astype(int)
    dist = np.minimum(j, ny - 1 - j)  # cells from the nearest wall
    if stab == "BSD":
        want = np.where(dist == 0, 1.5, 2.0)
    else:
        want = np.select([dist == 0, dist == 1], [13.0 / 8.0, 15.0 / 8.0], 2.0)
    assert np.allclose(term, want, rtol=0, atol=1e-10)
    assert np.abs((base - out)[:, 1:]).max() < 1e-11


def test_corrected_laplacian_removes_nonorthogonal_error():
    m = cases.cartesian(7, 6, 5, jitter=0.2)
    U, Ub = _lin(m, np.random.default_rng(4))
    tau = np.zeros((m.nCells, 6))
    interior = np.ones(m.nCells, bool)
    interior[m.owner[m.nInternal:]] = False
    err = {lap: np.abs(O.div_tau_stab(m, U, Ub, tau, None, 1.0, 1.0, "BSD", lap)[interior]).max()
           for lap in ("corrected", "orthogonal", "uncorrected")}
    assert err["corrected"] < 0.2 * err["orthogonal"]
    assert err["corrected"] < 0.2 * err["uncorrected"]
