"""Orientation sets and quaternion conventions (oracle).

Restates:
  muspinsim/utils.py:34-37     zcw_gen (soprano ZCW 'sphere')
  muspinsim/utils.py:40-50     eulrange_gen
  muspinsim/utils.py:53-68     quat_from_polar
  muspinsim/simconfig.py:152-159, 596-613   orientation weights, stored conjugate
  muspinsim/experiment.py:384-411           B' = q.rotate(B) + B_intr, p' = q.rotate(p)

Quaternions are numpy arrays [w, x, y, z] with the Hamilton product.  The
Euler convention (zyz: q = Qz(c) Qy(b) Qz(a)) is the one the reference's
tests pin with explicit axis-angle products (tests/test_config.py:214-286).

ZCW: soprano is not available, so the point set is generated from the
published Zaremba-Conroy-Wolfsberg recipe (Eden & Levitt, J. Magn. Reson.
132, 220 (1998), 'sphere' constants c = (1, 2, 1), N = F_{m+2}, g = F_m).
Parity against soprano's exact point set is unpinned.
"""

import numpy as np


def qmul(p, q):
    """Hamilton product p * q."""
    w1, x1, y1, z1 = p
    w2, x2, y2, z2 = q
    return np.array([
        w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
        w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
        w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
        w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2,
    ])


def qconj(q):
    return np.array([q[0], -q[1], -q[2], -q[3]])


def from_axis_angle(n, theta):
    n = np.asarray(n, dtype=float)
    n = n / np.linalg.norm(n)
    return np.concatenate([[np.cos(theta / 2.0)], np.sin(theta / 2.0) * n])


def from_euler(a, b, c, mode="zyz"):
    """q = Q_z(c) Q_m(b) Q_z(a), m = y for 'zyz' and x for 'zxz'."""
    mid = {"zyz": [0, 1, 0], "zxz": [1, 0, 0]}[mode]
    q = from_axis_angle([0, 0, 1], c)
    q = qmul(q, from_axis_angle(mid, b))
    q = qmul(q, from_axis_angle([0, 0, 1], a))
    return q


def quat_from_polar(theta, phi):
    """utils.py:53-68: Euler (phi, theta, phi) zyz."""
    return from_euler(phi, theta, phi, "zyz")


def rotation_matrix(q):
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def rotate(q, v):
    """q v q* for a 3-vector v."""
    vq = np.concatenate([[0.0], np.asarray(v, dtype=float)])
    return qmul(qmul(q, vq), qconj(q))[1:]


def fibonacci_at_least(N):
    F = [1, 1]
    while F[-1] < N:
        F.append(F[-1] + F[-2])
    return F


def zcw_angles(N):
    """ZCW 'sphere' polar angle pairs (theta, phi), at least N of them."""
    F = fibonacci_at_least(max(int(N), 1))
    Nm = F[-1]
    g = F[-3] if len(F) >= 3 else 1
    j = np.arange(Nm)
    theta = np.arccos(1.0 * (2.0 * np.mod(j / Nm, 1.0) - 1.0))
    phi = 2.0 * np.pi / 1.0 * np.mod(j * g / Nm, 1.0)
    return np.stack([theta, phi], axis=1)


def eulrange(N):
    """N^3 Euler triplets with sin(beta) weights (utils.py:40-50)."""
    N = int(N)
    a = np.linspace(0, 2 * np.pi, N)
    b = np.linspace(0, np.pi, N + 2)[1:-1]
    c = np.linspace(0, 2 * np.pi, N)
    a, b, c = np.array(np.meshgrid(a, b, c)).reshape((3, -1))
    w = np.sin(b)
    return np.array([a, b, c, w]).T


def orientation_entry(v, mode="zyz"):
    """(conjugated quaternion, weight) for one orientation line
    (simconfig.py:596-613)."""
    v = list(v)
    w = 1.0
    if len(v) == 2:
        q = quat_from_polar(*v)
    elif len(v) == 3:
        q = from_euler(*v, mode=mode)
    elif len(v) == 4:
        q = from_euler(*v[:3], mode=mode)
        w = v[3]
    else:
        raise ValueError("orientation needs 2, 3 or 4 values")
    return qconj(q), w


def normalise_weights(weights, averaged):
    """w * (N / sum w) when averaged over orientation, else 1.0
    (simconfig.py:152-159)."""
    weights = np.asarray(weights, dtype=float)
    if averaged:
        norm = len(weights) / np.sum(weights)
        return np.array([w * norm for w in weights])
    return np.ones(len(weights))
