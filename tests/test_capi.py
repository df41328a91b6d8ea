"""CPU checks of the C ABI: the library loads without a GPU, exports every
function include/muspin_hip.h declares, and the struct layouts match."""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "muspin_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msp_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from muspinsim_amd import _native as N
    lib = N.load_library()
    names = _declared()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(N.SIGNATURES)


def test_struct_layouts():
    from muspinsim_amd import _native as N
    assert N.SYSTEM_DTYPE.itemsize == 224
    assert N.SYSTEM_DTYPE.fields["gammas"][1] == 96
    assert N.TERM_DTYPE.itemsize == 88
    assert N.TERM_DTYPE.fields["t"][1] == 16


def test_argument_errors_need_no_gpu():
    """Validation happens before any HIP call, so it is testable on the CPU."""
    from muspinsim_amd import _native as N
    lib = N.load_library()
    assert lib.msp_eigh(0, 1, None, None, None, None, 0, None, 0, None) == N.MSP_ERR_VALUE
    assert "out of range" in N.last_error()
    assert lib.msp_integral(4, 1, None, None, 1, -1.0, 1.0, None, 0, None, None) == N.MSP_ERR_VALUE
    assert "tau" in N.last_error()
    with pytest.raises(ValueError):
        N.check(N.MSP_ERR_VALUE)
    with pytest.raises(np.linalg.LinAlgError):
        N.check(N.MSP_ERR_NOCONV)
    assert lib.msp_eigh_workspace(64, 10, N.MSP_EIGH_JACOBI) == 0
    assert lib.msp_eigh_workspace(192, 2, N.MSP_EIGH_DC) > 2 * 3 * 192 * 192 * 8
    # zero-size batches are no-ops
    assert lib.msp_gemm_cn(4, 0, None, None, None, None) == N.MSP_OK
