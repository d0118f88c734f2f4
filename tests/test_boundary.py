"""The C-ABI boundary on the CPU: the library loads, exports every entry point
declared in include/chemtensor_hip.h, rejects bad arguments with LAPACK-style
negative statuses before touching a device, and its host-side PCG32 input
generator matches the oracle bit for bit."""
import ctypes
import os
import re

import numpy as np
import pytest

import chemtensor_mojo_amd as ct
from chemtensor_mojo_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "chemtensor_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ct_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, f"{n} has no ctypes signature"


def test_version_string():
    assert b"gfx950" in _lib.load().ct_version()


def test_argument_errors_without_device():
    lib = _lib.load()
    # NULL handle -> -1 (first argument)
    assert lib.ct_synchronize(None) == -1
    assert lib.ct_heff_apply(None, 0, 0, None, None, None, None, None, None, 1, 1, 2, 1, 1, 1, 1, 1) == -1
    # bad dims on the host-only generator
    dims = (ctypes.c_int64 * 4)()
    assert lib.ct_random_mps_c_compatible(0, 2, 4, 1, dims, None, None) == -1
    assert lib.ct_random_mps_c_compatible(3, 2, 0, 1, dims, None, None) == -3
    assert b"chi_max" in lib.ct_last_error()


@pytest.mark.parametrize("args", [(16, 2, 8, 42), (9, 3, 5, 7), (6, 2, 2, 5), (12, 2, 64, 1)])
def test_host_random_mps_matches_oracle(ora, args):
    mine = ct.random_mps_c_compatible_tensors(*args)
    ref = ora.random_mps(*args)
    assert len(mine) == len(ref)
    for a, b in zip(mine, ref):
        assert a.shape == b.shape
        assert np.array_equal(a, b)


def test_gpu_entry_points_fail_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError):
        ct.DeviceContext(0)
