"""The C-ABI library loads and exports every entry point declared in include/rrtmgsw.h
(no GPU call is made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rrtmg_sw_amd", "librrtmgsw_hip.so")


def declared():
    src = open(os.path.join(ROOT, "include", "rrtmgsw.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rrtmgsw_[a-z_]+)\s*\(", src)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 9
    lib = ctypes.CDLL(LIB)
    for n in names:
        assert hasattr(lib, n), n
    from rrtmg_sw_amd import _lib
    assert sorted(_lib.EXPORTED) == names


def test_autograd_wrapper_refuses_other_forms():
    """The adjoint is the derivative of the exact-exponential forward: the torch.autograd.Function
    refuses the table and mixed forms before anything runs."""
    import torch
    from rrtmg_sw_amd.torch_adjoint import McicaShortwave
    x = torch.zeros(2, 3, dtype=torch.float64)
    for p in ("fp64", "mixed"):
        with pytest.raises(ValueError):
            McicaShortwave.apply(x, x, x, x, x, x, None, {}, p)
