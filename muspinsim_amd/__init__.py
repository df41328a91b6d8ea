"""MI355X-native engine for muspinsim's spin-dynamics hot path.

Host classes mirror the reference API (Hamiltonian, Lindbladian,
ExperimentRunner, SpinSystem, operators); every numerical kernel of the path
runs in libmuspin_hip.so (hand-written gfx950 HIP, C ABI in
include/muspin_hip.h).
"""

__version__ = "0.1.0"
