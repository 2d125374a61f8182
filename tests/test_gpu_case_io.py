"""Stress-only runs of OpenFOAM case directories (blockMesh or polyMesh, 0/
fields, fvSchemes/fvSolution/controlDict) on the device path against the same
runs through the oracle: theta and tau of every mode within 1e-9 relative L2
after ten steps, and the files written at each write time read back equal to
the device state."""
import os

import numpy as np
import pytest

from rheotool_amd import polymesh, stressfoam

from test_case_io import cavity_velocity, channel_velocity, prepare

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name,vel", [("stresscase_cavity", cavity_velocity),
                                      ("stresscase_channel", channel_velocity)])
def test_case_run_device_matches_oracle(tmp_path, name, vel):
    print(f"[case] stress-only case run {name}, 10 steps", flush=True)
    dev_case = prepare(tmp_path / "dev", name, vel)
    orc_case = prepare(tmp_path / "orc", name, vel)
    dev = stressfoam.StressCase(dev_case)
    written = dev.run(steps=10)
    orc = stressfoam.StressCase(orc_case, oracle=True)
    orc.run(steps=10, write=False)
    d, dtau = dev.state()
    o, otau = orc.state()
    for k in o:
        assert rel(d[k]["theta"], o[k]["theta"]) < 1e-9, k
        assert rel(d[k]["tau"], o[k]["tau"]) < 1e-9, k
    assert rel(dtau, otau) < 1e-9
    last = written[-1]
    for md in dev.modes:
        f = polymesh.read_field(os.path.join(last, md.field), dev.pm)
        assert np.array_equal(f.internal, d[md.name]["theta"])


def test_written_polymesh_gives_the_same_run(tmp_path):
    """The blockMesh subset's mesh written as constant/polyMesh and read back
    runs the same step bit for bit."""
    print("[case] polyMesh round trip run", flush=True)
    a = prepare(tmp_path / "a", "stresscase_cavity", cavity_velocity)
    b = prepare(tmp_path / "b", "stresscase_cavity", cavity_velocity)
    polymesh.write_polymesh(b, polymesh.read_block_mesh(b))
    ra = stressfoam.StressCase(a)
    ra.run(steps=3, write=False)
    rb = stressfoam.StressCase(b)
    rb.run(steps=3, write=False)
    assert np.array_equal(ra.state()[1], rb.state()[1])
