"""Minimal forward-mode dual numbers over numpy arrays.

Used by the host table producer to get exact Jacobians (the reference obtains
them with ``jax.jacfwd`` / ``lineax.JacobianLinearOperator``) while the values
themselves may be complex (complex-step differentiation with respect to the
cosmological parameters).  A ``Dual`` holds ``v`` (shape S) and ``d`` (shape
(m,) + S): m tangent directions.  Branch decisions use real parts only.
"""
from __future__ import annotations

import numpy as np


class Dual:
    __slots__ = ("v", "d")
    __array_priority__ = 1000

    def __init__(self, v, d):
        self.v = v
        self.d = d

    # ---- helpers -------------------------------------------------------------
    @staticmethod
    def _split(o):
        if isinstance(o, Dual):
            return o.v, o.d
        return o, None

    def _bin(self, o, fv, fd):
        ov, od = Dual._split(o)
        v = fv(self.v, ov)
        return Dual(v, fd(self.v, self.d, ov, od))

    # ---- arithmetic -----------------------------------------------------------
    def __add__(self, o):
        ov, od = Dual._split(o)
        return Dual(self.v + ov, self.d if od is None else self.d + od)

    __radd__ = __add__

    def __sub__(self, o):
        ov, od = Dual._split(o)
        return Dual(self.v - ov, self.d if od is None else self.d - od)

    def __rsub__(self, o):
        return Dual(o - self.v, -self.d)

    def __neg__(self):
        return Dual(-self.v, -self.d)

    def __mul__(self, o):
        ov, od = Dual._split(o)
        if od is None:
            return Dual(self.v * ov, self.d * ov)
        return Dual(self.v * ov, self.d * ov + self.v * od)

    __rmul__ = __mul__

    def __truediv__(self, o):
        ov, od = Dual._split(o)
        if od is None:
            return Dual(self.v / ov, self.d / ov)
        q = self.v / ov
        return Dual(q, (self.d - q * od) / ov)

    def __rtruediv__(self, o):
        q = o / self.v
        return Dual(q, -q * self.d / self.v)

    def __pow__(self, p):
        if isinstance(p, Dual):
            return exp(p * log(self))
        v = self.v ** p
        return Dual(v, p * self.v ** (p - 1) * self.d)

    def __rpow__(self, b):
        v = b ** self.v
        return Dual(v, v * np.log(b) * self.d)

    # ---- comparisons (real part) -------------------------------------------------
    def __lt__(self, o):
        return np.real(self.v) < np.real(Dual._split(o)[0])

    def __le__(self, o):
        return np.real(self.v) <= np.real(Dual._split(o)[0])

    def __gt__(self, o):
        return np.real(self.v) > np.real(Dual._split(o)[0])

    def __ge__(self, o):
        return np.real(self.v) >= np.real(Dual._split(o)[0])


def val(x):
    return x.v if isinstance(x, Dual) else x


def exp(x):
    if isinstance(x, Dual):
        e = np.exp(x.v)
        return Dual(e, e * x.d)
    return np.exp(x)


def log(x):
    if isinstance(x, Dual):
        return Dual(np.log(x.v), x.d / x.v)
    return np.log(x)


def sqrt(x):
    if isinstance(x, Dual):
        s = np.sqrt(x.v)
        return Dual(s, 0.5 * x.d / s)
    return np.sqrt(x)


def rabs(x):
    """|x| with the sign taken from the real part (complex-step safe)."""
    if isinstance(x, Dual):
        s = np.where(np.real(x.v) < 0, -1.0, 1.0)
        return Dual(x.v * s, x.d * s)
    return np.where(np.real(x) < 0, -x, x)


def where(c, a, b):
    """Elementwise select; tangents are selected along with the values."""
    av, ad = Dual._split(a)
    bv, bd = Dual._split(b)
    v = np.where(c, av, bv)
    if ad is None and bd is None:
        return v
    m = (ad if ad is not None else bd).shape[0]
    shp = (m,) + np.shape(v)
    ad = np.zeros(shp, dtype=np.result_type(v, np.float64)) if ad is None else ad
    bd = np.zeros(shp, dtype=np.result_type(v, np.float64)) if bd is None else bd
    return Dual(v, np.where(c, ad, bd))


def minimum(a, b):
    return where(np.real(val(a)) <= np.real(val(b)), a, b)


def seed(v, m, k):
    """Dual with value ``v`` and unit tangent in direction ``k`` of ``m``."""
    v = np.asarray(v)
    d = np.zeros((m,) + v.shape, dtype=np.result_type(v, np.float64))
    d[k] = 1.0
    return Dual(v, d)
