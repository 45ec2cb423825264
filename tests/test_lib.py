"""The C-ABI library: it builds, loads without a GPU and exports every entry
point include/tse_hip.h declares; the product has no CPU fallback."""
import ctypes
import os
import re

import pytest

from helpers import ROOT

HEADER = os.path.join(ROOT, "include", "tse_hip.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tse_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _declared()
    for n in ("tse_compute_p", "tse_compute_q", "tse_flow_error",
              "tse_compute_c", "tse_compute_t", "tse_l2_norm"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from transportse_amd import _lib
    lib = _lib.load()
    for n in _declared():
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, n
    assert lib.tse_abi_version() == _lib.ABI_VERSION


def test_host_sequential_norm_matches_oracle(oracle_mod):
    import numpy as np
    from oracle import Oracle
    from transportse_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(0)
    o = Oracle(1, 1, 1, np.ones(1), np.ones(1), np.ones(1), 9.8, 0.0)
    dp = ctypes.POINTER(ctypes.c_double)
    flag = ctypes.c_int32()
    for n in (1000, 3 << 20):   # the second sum runs on a thread above 2^20
        a, b = rng.standard_normal(n), rng.standard_normal(n)
        v = lib.tse_host_l2_norm_seq(a.ctypes.data_as(dp), b.ctypes.data_as(dp),
                                     a.size, ctypes.byref(flag))
        assert v == o.l2_norm(a, b) and flag.value == 0


def test_no_cpu_fallback_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from transportse_amd.device import HipContext
    import numpy as np
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        HipContext(2, 2, 2, np.ones(2), np.ones(2), np.ones(2), 9.8, 0.0)


def test_c_host_example_compiles_as_c99(tmp_path):
    """include/tse_hip.h is plain C: a C99 host program using it compiles
    with -Werror and links against the library."""
    from transportse_amd.build import build_c_example
    out = build_c_example(str(tmp_path / "c_abi_example"))
    assert os.path.exists(out)
