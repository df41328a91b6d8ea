"""MI355X-native engine for muspinsim's spin-dynamics hot path.

Host classes mirror the reference API (Hamiltonian, Lindbladian,
ExperimentRunner, SpinSystem, operators, the input-file reader, fitting and
Celio's method); every numerical kernel of the path runs in
libmuspin_hip.so (hand-written gfx950 HIP, C ABI in include/muspin_hip.h).
The top-level names follow muspinsim/__init__.py; they import lazily, so
that importing the package does not load torch or the library.
"""

from .version import __version__  # noqa: F401

_EXPORTS = {
    "MuSpinInput": ("input", "MuSpinInput"),
    "MuSpinConfig": ("config", "MuSpinConfig"),
    "ExperimentRunner": ("experiment", "ExperimentRunner"),
    "FittingRunner": ("fitting", "FittingRunner"),
}


def __getattr__(name):
    if name in _EXPORTS:
        import importlib
        mod, attr = _EXPORTS[name]
        return getattr(importlib.import_module(f"{__name__}.{mod}"), attr)
    if name == "constants":
        import importlib
        return importlib.import_module(f"{__name__}.constants")
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
