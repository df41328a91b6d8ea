"""Physical constants and isotope data (muspinsim/constants.py:12-108).

The reference reads isotope data from soprano, which is not available here.
This module carries its own table of the common nuclei (gyromagnetic ratio in
MHz/T, spin, quadrupole moment in millibarn, the unit soprano returns and
EFG_2_MHZ assumes).  1H and 2H agree with the reference's own tests
(tests/test_constants.py:14-40); the other entries are standard NMR table
values and their agreement with soprano's is not pinned.
"""

import numpy as np
from scipy import constants as cnst

ELEC_GAMMA = -28024.9514242                 # MHz/T (constants.py:12)
MU_GAMMA = -(ELEC_GAMMA / 206.7669883)      # MHz/T (constants.py:13)
MU_TAU = 2.19703                            # us (constants.py:14)
EFG_2_MHZ = (                               # constants.py:20-25
    cnst.physical_constants["atomic unit of electric field gradient"][0] * cnst.e * 1e-37 / cnst.h
)

# element -> (default mass number, {A: (gamma MHz/T, I, Q mb)})
_ISOTOPES = {
    "H": (1, {1: (42.577478518, 0.5, 0.0), 2: (6.535903, 1.0, 2.86), 3: (45.414835, 0.5, 0.0)}),
    "Li": (7, {6: (6.265884, 1.0, -0.808), 7: (16.548177, 1.5, -40.0)}),
    "B": (11, {10: (4.575079, 3.0, 84.59), 11: (13.662979, 1.5, 40.59)}),
    "C": (12, {12: (0.0, 0.0, 0.0), 13: (10.708395, 0.5, 0.0)}),
    "N": (14, {14: (3.077706, 1.0, 20.44), 15: (-4.317265, 0.5, 0.0)}),
    "O": (16, {16: (0.0, 0.0, 0.0), 17: (-5.774236, 2.5, -25.58)}),
    "F": (19, {19: (40.078, 0.5, 0.0)}),
    "Na": (23, {23: (11.268878, 1.5, 104.0)}),
    "Mg": (24, {24: (0.0, 0.0, 0.0), 25: (-2.6083, 2.5, 199.4)}),
    "Al": (27, {27: (11.103083, 2.5, 146.6)}),
    "Si": (28, {28: (0.0, 0.0, 0.0), 29: (-8.465499, 0.5, 0.0)}),
    "P": (31, {31: (17.251, 0.5, 0.0)}),
    "Cl": (35, {35: (4.1765, 1.5, -81.65), 37: (3.4766, 1.5, -64.35)}),
    "V": (51, {50: (4.2483, 6.0, 210.0), 51: (11.2133, 3.5, -52.0)}),
    "Cu": (63, {63: (11.319, 1.5, -220.0), 65: (12.103, 1.5, -204.0)}),
}


def _lookup(elem, iso):
    try:
        default, table = _ISOTOPES[elem]
    except KeyError as exc:
        raise ValueError(f"Invalid isotope {iso} for element {elem}") from exc
    A = default if iso is None else int(iso)
    if A not in table:
        raise ValueError(f"Invalid isotope {iso} for element {elem}")
    return table[A]


def gyromagnetic_ratio(elem="mu", iso=None):
    """MHz/T (constants.py:28-53)."""
    if elem == "e":
        return ELEC_GAMMA
    if elem == "mu":
        return MU_GAMMA
    return _lookup(elem, iso)[0]


def quadrupole_moment(elem="mu", iso=None):
    """Quadrupole moment in millibarn (constants.py:56-77)."""
    if elem in ("e", "mu"):
        return 0
    return _lookup(elem, iso)[2]


def spin(elem="mu", iso=None):
    """Spin in units of hbar (constants.py:80-108)."""
    if elem == "mu":
        return 0.5
    if elem == "e":
        iso = iso or 1
        if iso < 1 or int(iso) != iso:
            raise ValueError(f"Invalid multiplicity {iso} for electron")
        return 0.5 * int(iso)
    return _lookup(elem, iso)[1]


def dipolar_tensor(g_i, g_j, r):
    """Dipolar coupling tensor (MHz) for r in Angstrom (spinsys.py:400-409)."""
    r = np.asarray(r, dtype=float)
    rn = np.linalg.norm(r)
    D = -(np.eye(3) - 3.0 / rn ** 2 * r[:, None] * r[None, :])
    return D * (-(cnst.mu_0 * cnst.hbar * (g_i * g_j * 1e6)) / (2 * (rn * 1e-10) ** 3))


def quadrupolar_tensor(Q, I, efg):
    """EFG -> coupling tensor in MHz (spinsys.py:433-438)."""
    if I == 0.5:
        raise ValueError("Can not set up quadrupolar coupling for spin 1/2 particle")
    return EFG_2_MHZ * Q / (2 * I * (2 * I - 1)) * np.asarray(efg, dtype=float)
