"""ctypes binding of libmuspin_hip.so (include/muspin_hip.h).

The library is the only compute path: if it is missing or the device is not
an MI355X (gfx950) every call raises instead of falling back to anything.
"""

import ctypes
import os

import numpy as np
# torch's HIP runtime must be the one in the process before the library is
# loaded: both resolve libamdhip64 by soname, and the first one loaded wins.
import torch  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmuspin_hip.so")

MSP_OK = 0
MSP_ERR_VALUE = 1
MSP_ERR_NOCONV = 2
MSP_ERR_HIP = 3
MSP_ERR_UNSUPPORTED = 4

MSP_EIGH_AUTO = 0
MSP_EIGH_JACOBI = 1
MSP_EIGH_DC = 2
MSP_EIGH_DC2 = 3

MSP_MAX_SPINS = 16
MSP_NOALLOC_SCRATCH_BYTES = 64 << 20     # static scratch of msp_eigh_noalloc, per device

# numpy mirrors of the C structs (natural alignment, no packing)
SYSTEM_DTYPE = np.dtype([
    ("n_spins", np.int32), ("n", np.int32), ("muon_index", np.int32),
    ("term_offset", np.int32), ("n_terms", np.int32), ("pad", np.int32, 3),
    ("dims", np.int32, MSP_MAX_SPINS), ("gammas", np.float64, MSP_MAX_SPINS),
])
TERM_DTYPE = np.dtype([
    ("kind", np.int32), ("i", np.int32), ("j", np.int32), ("pad", np.int32),
    ("t", np.float64, 9),
])
assert SYSTEM_DTYPE.itemsize == 224 and TERM_DTYPE.itemsize == 88

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_SZ = ctypes.c_size_t

# name -> (restype, argtypes); every symbol of include/muspin_hip.h
SIGNATURES = {
    "msp_init": (_I, []),
    "msp_version": (_I, []),
    "msp_last_error": (_I, [ctypes.c_char_p, _SZ]),
    "msp_rotate_fields": (_I, [_I, _P, _P, _P, _P, _P, _P, _P]),
    "msp_build_h": (_I, [_I, _I, _P, _P, _P, _P, _P, _I, _P, _P]),
    "msp_build_h_lower": (_I, [_I, _I, _P, _P, _P, _P, _P, _I, _P, _P]),
    "msp_eigh_workspace": (_SZ, [_I, _I, _I]),
    "msp_eigh": (_I, [_I, _I, _P, _P, _P, _P, _I, _P, _SZ, _P]),
    "msp_eigh_noalloc": (_I, [_I, _I, _P, _P, _P, _P, _I, _P]),
    "msp_eigh_mixed_workspace": (_SZ, [_I, _I]),
    "msp_timesum_mixed_workspace": (_SZ, [_I, _I, _I]),
    "msp_timesum_mixed": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _SZ, _P]),
    "msp_timesum_mfma64_workspace": (_SZ, [_I, _I, _I]),
    "msp_timesum_mfma64": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _SZ, _P]),
    "msp_eigh_mixed": (_I, [_I, _I, _P, _P, _P, _P, _P, _SZ, _P]),
    "msp_apply_muon_op": (_I, [_I, _I, _P, _P, _P, _I, _P, _P, _P]),
    "msp_apply_kron": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P]),
    "msp_gemm_cn": (_I, [_I, _I, _P, _P, _P, _P]),
    "msp_form_c": (_I, [_I, _I, _P, _P, _I, _D, _D, _P, _P]),
    "msp_timesum": (_I, [_I, _I, _I, _P, _P, _I, _P, _P, _I, _P, _P]),
    "msp_integral": (_I, [_I, _I, _P, _P, _I, _D, _D, _P, _I, _P, _P]),
    "msp_phase_evolve": (_I, [_I, _I, _P, _P, _P, _P, _P]),
    "msp_liouvillian": (_I, [_I, _I, _P, _I, _P, _P, _P, _P]),
    "msp_lindblad_workspace": (_SZ, [_I, _I]),
    "msp_lindblad_modes": (_I, [_I, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _SZ, _P]),
    "msp_mode_sum": (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _I, _P, _P]),
    "msp_mode_integral": (_I, [_I, _I, _I, _P, _P, _D, _D, _P, _I, _P, _P]),
    "msp_mode_states": (_I, [_I, _I, _P, _P, _P, _P, _P, _P]),
    "msp_adjoint_freq": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _D, _P, _P, _P]),
    "msp_hadamard_real": (_I, [_I, _I, _P, _P, _P, _P]),
    "msp_adjoint_gprime": (_I, [_I, _I, _P, _P, _P, _P, _P, _D, _D, _P, _P]),
    "msp_batch_sum": (_I, [_I, _I, _P, _P, _P]),
    "msp_trace_products": (_I, [_I, _I, _P, _P, _P, _P]),
    "msp_accumulate": (_I, [_I, _I, _P, _P, _P, _D, _P, _P]),
    "msp_gemm_cn_upper": (_I, [_I, _I, _P, _P, _P, _P]),
    "msp_gemm_muon_upper": (_I, [_I, _I, _P, _P, _P, _I, _P, _P, _P]),
    "msp_integral_obs": (_I, [_I, _I, _P, _P, _D, _D, _D, _D, _P, _P, _P]),
    "msp_celio_workspace": (_SZ, [_I, _I]),
    "msp_celio_evolve": (_I, [_I, _I, _P, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P,
                              _P, _P, _SZ, _P]),
    "msp_graph_begin": (_I, [_P]),
    "msp_graph_end": (_I, [_P, ctypes.POINTER(_P)]),
    "msp_graph_launch": (_I, [_P, _P]),
    "msp_graph_destroy": (_I, [_P]),
    "msp_profile_enable": (_I, [_I]),
    "msp_profile_report": (_I, [ctypes.c_char_p, _SZ]),
}

_lib = None
_initialised = False


def load_library(path=LIB_PATH):
    """Load and bind the library without touching the GPU."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `python -m muspinsim_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    buf = ctypes.create_string_buffer(512)
    load_library().msp_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(status):
    if status == MSP_OK:
        return
    msg = last_error()
    if status == MSP_ERR_VALUE:
        raise ValueError(msg)
    if status == MSP_ERR_NOCONV:
        raise np.linalg.LinAlgError(msg)
    if status == MSP_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def lib():
    """The bound library, initialised on the current HIP device (gfx950 only)."""
    global _initialised
    L = load_library()
    if not _initialised:
        check(L.msp_init())
        _initialised = True
    return L


def profile_report():
    """{kernel: (launches, total_ms)} since the last report (msp_profile_report)."""
    buf = ctypes.create_string_buffer(1 << 16)
    check(load_library().msp_profile_report(buf, 1 << 16))
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split()
        out[name] = (int(cnt), float(ms))
    return out
