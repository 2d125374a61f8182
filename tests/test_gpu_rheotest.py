"""The rheoTestFoam protocol (solvers/rheoTestFoam) on the device path against
the same protocol through the oracle: ramp steady states and the transient
multi-mode report with inner iterations, within 1e-9 relative."""
import os

import numpy as np
import pytest

import oracle_backends
from rheotool_amd import rheotest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rows(text):
    return np.array([[float(x) for x in l.split("\t")[:7]] for l in text.strip().splitlines()[2:]])


def test_ramp_fenecr_device_matches_oracle():
    case = os.path.join(GOLDEN, "rheotest_fenecr")
    print("[case] rheotest ramp FENE-CR 2 rates", flush=True)
    dev = rheotest.run(case, rates=2)
    orc = oracle_backends.rheotest_run(case, rates=2)
    a, b = rows(dev), rows(orc)
    assert a.shape == b.shape == (2, 7)
    assert np.allclose(a, b, rtol=1e-9, atol=1e-12), (a, b)
    assert all("Converged" in l for l in dev.strip().splitlines()[2:])


def test_transient_multimode_xpp_inner_iterations_device_matches_oracle():
    case = os.path.join(GOLDEN, "rheotest_xpp")
    print("[case] rheotest transient XPP 2 modes 20 steps x 5 inner", flush=True)
    dev = rheotest.run(case, max_steps=20)
    orc = oracle_backends.rheotest_run(case, max_steps=20)
    a, b = rows(dev), rows(orc)
    assert a.shape == b.shape == (20, 7)
    assert np.array_equal(a[:, 0], b[:, 0])
    err = np.abs(a[:, 1:] - b[:, 1:]).max() / np.abs(b[:, 1:]).max()
    assert err < 1e-9, err
