"""Crystallite orientations: quaternions, Euler / polar conventions and the
powder sets of the reference (muspinsim/utils.py:34-68,
muspinsim/simconfig.py:596-613).

Quaternions are float64 arrays (..., 4) = (w, x, y, z) with the Hamilton
product, vectorised over leading axes.  Euler angles (a, b, c) in mode 'zyz'
(or 'zxz') give q = Qz(c) Qm(b) Qz(a); polar angles (theta, phi) are the zyz
triplet (phi, theta, phi).  A configuration stores the CONJUGATE of each
crystallite rotation, so the lab-frame field and polarisation are rotated
into the crystal frame instead of rotating every tensor.

ZCW sets follow the published Zaremba-Conroy-Wolfsberg construction
(Eden & Levitt, J. Magn. Reson. 132, 220 (1998), 'sphere': N = F_{m+2}
points, generator F_m); the reference takes them from soprano, which is not
available here, so the exact point set is not pinned to soprano's.
"""

import numpy as np


def qmul(p, q):
    """Hamilton product, broadcast over leading axes."""
    p = np.asarray(p, dtype=float)
    q = np.asarray(q, dtype=float)
    w1, v1 = p[..., :1], p[..., 1:]
    w2, v2 = q[..., :1], q[..., 1:]
    w = w1 * w2 - np.sum(v1 * v2, axis=-1, keepdims=True)
    v = w1 * v2 + w2 * v1 + np.cross(v1, v2)
    return np.concatenate([w, v], axis=-1)


def conj(q):
    q = np.array(q, dtype=float)
    q[..., 1:] *= -1.0
    return q


def axis_angle(axis, angle):
    """Quaternions of rotations by `angle` (array) about a fixed unit axis."""
    angle = np.asarray(angle, dtype=float)
    ax = np.asarray(axis, dtype=float) / np.linalg.norm(axis)
    out = np.empty(angle.shape + (4,))
    out[..., 0] = np.cos(angle / 2)
    out[..., 1:] = np.sin(angle / 2)[..., None] * ax
    return out


def from_euler(a, b, c, mode="zyz"):
    """q = Qz(c) Qm(b) Qz(a), m = y ('zyz') or x ('zxz')."""
    if mode not in ("zyz", "zxz"):
        raise ValueError(f"Invalid Euler angle convention {mode}")
    mid = (0, 1, 0) if mode == "zyz" else (1, 0, 0)
    z = (0, 0, 1)
    return qmul(qmul(axis_angle(z, c), axis_angle(mid, b)), axis_angle(z, a))


def from_polar(theta, phi):
    """Rotation taking z to the direction (theta, phi) (utils.py:53-68)."""
    return from_euler(phi, theta, phi, "zyz")


def euler_zyz(q):
    """(a, b, c) with from_euler(a, b, c, 'zyz') = +-q: from the rotation
    matrix R = Rz(c) Ry(b) Rz(a), b = arccos R22, a = atan2(R21, -R20),
    c = atan2(R12, R02); at b = 0 or pi the split of a and c is not
    determined and c = 0 is taken."""
    w, x, y, z = np.asarray(q, dtype=float) / np.linalg.norm(q)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    b = np.arccos(np.clip(R[2, 2], -1.0, 1.0))
    if np.hypot(R[2, 0], R[2, 1]) > 1e-12:
        return np.array([np.arctan2(R[2, 1], -R[2, 0]), b, np.arctan2(R[1, 2], R[0, 2])])
    if R[2, 2] > 0:
        return np.array([np.arctan2(R[1, 0], R[0, 0]), b, 0.0])
    return np.array([np.arctan2(R[1, 0], R[1, 1]), b, 0.0])


def rotate(q, v):
    """q v q^* for quaternions (N, 4) and vectors (N, 3) (or broadcastable)."""
    q = np.asarray(q, dtype=float)
    v = np.asarray(v, dtype=float)
    w = q[..., :1]
    u = q[..., 1:]
    c = np.cross(u, v)
    return v + 2.0 * (w * c + np.cross(u, c))


def zcw(N):
    """ZCW polar angle pairs (theta, phi), the smallest Fibonacci count >= N."""
    F = [1, 1]
    while F[-1] < max(int(N), 1):
        F.append(F[-1] + F[-2])
    Nm = F[-1]
    g = F[-3] if len(F) >= 3 else 1
    j = np.arange(Nm)
    theta = np.arccos(2.0 * np.mod(j / Nm, 1.0) - 1.0)
    phi = 2.0 * np.pi * np.mod(j * g / Nm, 1.0)
    return np.stack([theta, phi], axis=1)


def eulrange(N):
    """N^3 zyz triplets over a full grid with sin(b) weights: rows
    (a, b, c, w) (utils.py:40-50)."""
    N = int(N)
    a = np.linspace(0, 2 * np.pi, N)
    b = np.linspace(0, np.pi, N + 2)[1:-1]
    c = np.linspace(0, 2 * np.pi, N)
    A, Bb, C = (x.reshape(-1) for x in np.meshgrid(a, b, c))
    return np.stack([A, Bb, C, np.sin(Bb)], axis=1)


def stored_orientations(entries, mode="zyz"):
    """(conjugated quaternions (N, 4), raw weights (N,)) of orientation
    entries: (theta, phi) polar pairs, (a, b, c) Euler triplets or
    (a, b, c, w) weighted triplets (simconfig.py:596-613)."""
    qs, ws = [], []
    for e in entries:
        e = np.atleast_1d(np.asarray(e, dtype=float))
        if e.size == 2:
            q, w = from_polar(e[0], e[1]), 1.0
        elif e.size == 3:
            q, w = from_euler(e[0], e[1], e[2], mode), 1.0
        elif e.size == 4:
            q, w = from_euler(e[0], e[1], e[2], mode), e[3]
        else:
            raise ValueError("an orientation has 2 (polar), 3 (Euler) or 4 (Euler + weight) values")
        qs.append(conj(q))
        ws.append(w)
    return np.array(qs).reshape(-1, 4), np.array(ws, dtype=float)


def normalised_weights(weights, averaged):
    """w * N / sum(w) when orientations are averaged, else all 1.0
    (simconfig.py:152-159)."""
    weights = np.asarray(weights, dtype=float)
    if averaged:
        return weights * (len(weights) / np.sum(weights))
    return np.ones(len(weights))
