"""Host-side checks that need no GPU: the C-ABI library loads and exports
every symbol include/kcds.h declares, the generator library works, and the
drop-in CLI rejects bad arguments the way the reference does (exit 1)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "kcds.h")).read()
    return sorted(set(re.findall(r"\b(kcds_[a-z_]+)\s*\(", src)))


def test_header_and_python_exports_agree():
    from test_amd import kcds
    assert sorted(kcds.EXPORTS) == declared_symbols()


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "test_amd", "libkcds.so"))
    for sym in declared_symbols():
        assert hasattr(lib, sym), sym


def test_create_without_gpu_fails_cleanly():
    from test_amd import kcds
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    assert kcds.lib().kcds_create(ctypes.byref(h), 0, 0) != 0


def test_cli_usage_errors():
    exe = os.path.join(ROOT, "test_amd", "kcds")
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "usage" in r.stdout
    r = subprocess.run([exe, "/nonexistent/graph.txt", "3"], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "Open Failed" in r.stdout
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "ref_data_test.txt"), "9"],
                       stdout=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "2..8" in r.stdout


def test_concat_graphs():
    from test_amd.kcds import concat_graphs
    a = (np.array([0, 1, 2], np.uint32), np.array([1, 0], np.uint32))
    b = (np.array([0, 2, 3, 4], np.uint32), np.array([1, 2, 0, 0], np.uint32))
    off, nbr, voff = concat_graphs([a, b])
    assert voff.tolist() == [0, 2, 5]
    assert off.tolist() == [0, 1, 2, 4, 5, 6]
    assert nbr.tolist() == [1, 0, 3, 4, 2, 2]
