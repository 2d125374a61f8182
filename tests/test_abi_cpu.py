"""CPU-side checks of the C-ABI library: it loads without a GPU and exports
every function include/rheo_gpu.h declares."""
import ctypes as C
import os
import re

import pytest

from rheotool_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_no_device_reports_cleanly():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    assert abi.device_count() == 0
    with pytest.raises(abi.RheoError) as e:
        abi.Context(0)
    assert e.value.code == abi.RT_E_NO_DEVICE


def test_arrays_of_the_wrong_size_are_refused_before_the_library_reads_them():
    """The C entry points read a fixed number of values from each pointer
    (the sizes in include/rheo_gpu.h); the Python side checks them first."""
    import numpy as np

    assert abi._f64(np.arange(6.0), (2, 3)).shape == (2, 3)
    assert abi._f64([[1, 2, 3]], (3,)).dtype == np.float64
    with pytest.raises(ValueError):
        abi._f64(np.zeros(5), (2, 3))
    with pytest.raises(ValueError):
        abi._f64(np.zeros((4, 6)), (3, 6))
