"""Utilities of the drop-in surface (muspinsim/utils.py:13-93): Clonable,
deepmap, the powder generators, polar-angle quaternions and get_xy.

The reference takes its quaternions from ase; `Quaternion` here is a small
class of our own with the methods the package and its callers use
(from_euler_angles, from_axis_angle, rotate, conjugate, rotation_matrix,
euler_angles, product), on (w, x, y, z) arrays.
"""

from collections.abc import Iterable
from copy import deepcopy

import numpy as np

from . import orientations as _o


class Clonable:
    """Objects with a .clone() that deep-copies their attributes."""

    def clone(self):
        cls = self.__class__
        copy = cls.__new__(cls)
        copy.__dict__ = deepcopy(self.__dict__)
        return copy


def deepmap(func, obj):
    """func applied to every non-iterable leaf of nested iterables."""
    if isinstance(obj, Iterable):
        return [deepmap(func, x) for x in obj]
    return func(obj)


class Quaternion:
    """Unit quaternion (w, x, y, z); rotate(v) = q v q^*."""

    def __init__(self, qs=(1.0, 0.0, 0.0, 0.0)):
        self.q = np.array(qs, dtype=float)

    @classmethod
    def from_euler_angles(cls, a, b, c, mode="zyz"):
        return cls(_o.from_euler(a, b, c, mode))

    @classmethod
    def from_axis_angle(cls, n, theta):
        return cls(_o.axis_angle(n, theta))

    def __mul__(self, other):
        return Quaternion(_o.qmul(self.q, other.q))

    def conjugate(self):
        return Quaternion(_o.conj(self.q))

    def rotate(self, v):
        return _o.rotate(self.q, np.asarray(v, dtype=float))

    def rotation_matrix(self):
        return np.stack([self.rotate(e) for e in np.eye(3)], axis=1)

    def euler_angles(self, mode="zyz"):
        if mode != "zyz":
            raise ValueError("only the zyz convention is provided")
        return _o.euler_zyz(self.q)

    def __repr__(self):
        return f"Quaternion({self.q.tolist()})"


def zcw_gen(N, mode="sphere"):
    """(theta, phi) pairs of a ZCW powder set of at least N points
    (utils.py:34-37; the sphere set)."""
    if mode != "sphere":
        raise ValueError("only the 'sphere' ZCW set is provided")
    return _o.zcw(int(N))


def eulrange_gen(N):
    """N^3 (a, b, c, weight) rows over a zyz Euler grid (utils.py:40-50)."""
    return _o.eulrange(int(N))


def quat_from_polar(theta, phi):
    """Rotation taking z to the direction (theta, phi) (utils.py:53-68)."""
    return Quaternion(_o.from_polar(theta, phi))


def get_xy(z):
    """Two unit axes x, y with (x, y, z) right handed (utils.py:71-93)."""
    zn = np.linalg.norm(z)
    if zn == 0:
        raise ValueError("Can not find X and Y for null Z vector")
    z = np.asarray(z, dtype=float) / zn
    if z[0] == 0 and z[1] == 0:
        return np.array([1.0, 0, 0]), np.array([0, z[2], 0])
    x = np.array([z[1], -z[0], 0.0]) / np.hypot(z[0], z[1])
    return x, np.cross(z, x)
