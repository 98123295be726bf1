"""Magnetic anisotropy: predefined uniaxial and cubic energies, their
arithmetic, and the polynomial form the device evaluates.

Mirrors interface/nmag/anisotropy.py (PredefinedAnisotropy :21-90,
uniaxial_anisotropy :92-136, cubic_anisotropy :138-201).  An energy density
E(m) of a given order is represented, as in src/nsim_anisotropy.ml:119-133,
by the homogeneous polynomial of that degree in (mx, my, mz) that equals E on
the unit sphere.  That polynomial is unique (a homogeneous polynomial that
vanishes on the sphere vanishes everywhere), so an E that is such a
polynomial on the sphere -- uniaxial, cubic, their sums -- is represented
exactly.  The device evaluates H_anis = -(g - m (m.g)) / (mu0 Ms), g = grad
E (src/nsim_anisotropy.ml:167-197,216-234).
"""
import numpy as np

from .units import to_su


class NmagUserError(Exception):
    """User error, as raised by the reference front end."""


class PredefinedAnisotropy:
    """An anisotropy energy function with the polynomial order used to
    represent it (interface/nmag/anisotropy.py:21-90)."""

    def __init__(self, function=None, order=None, anis_type="functional", axis1=None,
                 axis2=None, axis3=None, K1=None, K2=None, K3=None, stringifier=None):
        if function is None and order is None:
            raise NmagUserError("Please specify a sampling order.")
        self.anis_type = anis_type
        self.function = function
        self.order = order
        self.K1, self.K2, self.K3 = K1, K2, K3
        self.axis1, self.axis2, self.axis3 = axis1, axis2, axis3
        self._str_extra = stringifier

    def has_function(self):
        return self.function is not None

    def __str__(self):
        s = self.anis_type
        if self._str_extra:
            s += ", %s" % self._str_extra(self)
        return "<PredefinedAnisotropy:%s>" % s

    def _other(self, y):
        if not isinstance(y, PredefinedAnisotropy):
            raise NmagUserError("%s <-- this is not a PredefinedAnisotropy object." % (y,))
        if not (self.has_function() and y.has_function()):
            raise NmagUserError("you cannot operate on an anisotropy object if this has not "
                                "an associated energy function!")
        return y

    def __neg__(self):
        return PredefinedAnisotropy(lambda m: -self.function(m), self.order)

    def __pos__(self):
        return PredefinedAnisotropy(self.function, self.order, self.anis_type, self.axis1,
                                    self.axis2, self.axis3, self.K1, self.K2, self.K3,
                                    self._str_extra)

    def __add__(self, y):
        y = self._other(y)
        return PredefinedAnisotropy(lambda m: self.function(m) + y.function(m),
                                    max(self.order, y.order))

    def __sub__(self, y):
        y = self._other(y)
        return PredefinedAnisotropy(lambda m: self.function(m) - y.function(m),
                                    max(self.order, y.order))


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def uniaxial_anisotropy(axis, K1, K2=0):
    """E_anis = -K1 <axis,m>^2 - K2 <axis,m>^4 (interface/nmag/anisotropy.py:92-136)."""
    axis = _unit(axis)

    def f(m):
        a = float(np.dot(axis, m))
        e = -K1 * a ** 2
        return e - K2 * a ** 4 if K2 else e

    def stringifier(a):
        s = "axis=[%s, %s, %s], K1=%s" % (tuple(a.axis1) + (a.K1,))
        return s + (", %s" % a.K2 if a.K2 else "")

    return PredefinedAnisotropy(f, 4 if K2 else 2, "uniaxial", axis1=axis, K1=K1, K2=K2,
                                stringifier=stringifier)


def cubic_anisotropy(axis1, axis2, K1, K2=0, K3=0):
    """E_anis = K1 (a1^2 a2^2 + a1^2 a3^2 + a2^2 a3^2) + K2 a1^2 a2^2 a3^2
    + K3 (a1^4 a2^4 + a1^4 a3^4 + a2^4 a3^4), a_i = <axis_i, m>, with axis2
    orthogonalised against axis1 and axis3 = axis1 x axis2
    (interface/nmag/anisotropy.py:138-201)."""
    a1 = np.asarray(axis1, dtype=np.float64)
    a3 = np.cross(a1, np.asarray(axis2, dtype=np.float64))
    a2 = np.cross(a3, a1)
    a1, a2, a3 = _unit(a1), _unit(a2), _unit(a3)

    def f(m):
        x, y, z = float(np.dot(a1, m)), float(np.dot(a2, m)), float(np.dot(a3, m))
        e = K1 * ((x * y) ** 2 + (x * z) ** 2 + (y * z) ** 2)
        if K2:
            e = e + K2 * (x * y * z) ** 2
        if K3:
            e = e + K3 * ((x * y) ** 4 + (x * z) ** 4 + (y * z) ** 4)
        return e

    def stringifier(a):
        s = "axis1=[%s, %s, %s], axis2=[%s, %s, %s], K1=%s" % (
            tuple(a.axis1) + tuple(a.axis2) + (a.K1,))
        s += ", %s" % a.K2 if a.K2 else ""
        return s + (", %s" % a.K3 if a.K3 else "")

    order = 8 if K3 else 6 if K2 else 4
    return PredefinedAnisotropy(f, order, "cubic", axis1=a1, axis2=a2, axis3=a3, K1=K1,
                                K2=K2, K3=K3, stringifier=stringifier)


def monomials(order):
    """Exponent triples (px, py, pz) with px + py + pz = order, px descending
    (the terms of src/nsim_anisotropy.ml:27-28)."""
    return [(a, b, order - a - b) for a in range(order, -1, -1) for b in range(order - a, -1, -1)]


def sphere_points(count):
    """count points spread over the unit sphere (Fibonacci lattice)."""
    k = np.arange(count) + 0.5
    z = 1.0 - 2.0 * k / count
    r = np.sqrt(1.0 - z * z)
    phi = np.pi * (1.0 + 5 ** 0.5) * k
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def fit_polynomial(energy_su, order):
    """(powers (k,3) int32, coeffs (k,)) of the homogeneous polynomial of
    degree `order` equal to energy_su (a function of the unit vector m, in
    simulation units) on the unit sphere: least squares over 4x as many
    sphere points as terms.  The reference solves the square system at
    randomly drawn, well-separated points (src/nsim_anisotropy.ml:51-133); for
    an energy that is such a polynomial both give that polynomial.  Terms
    below 1e-13 of the largest are dropped as rounding (the reference drops
    |c| <= 1e-10)."""
    if order is None or int(order) != order or not 1 <= order <= 16:
        raise NmagUserError("anisotropy_order must be an integer from 1 to 16, not %r" % order)
    order = int(order)
    P = np.array(monomials(order), dtype=np.int32)
    pts = sphere_points(4 * len(P))
    A = np.prod(pts[:, None, :] ** P[None, :, :], axis=2)
    e = np.array([float(energy_su(p)) for p in pts])
    c, *_ = np.linalg.lstsq(A, e, rcond=None)
    keep = np.abs(c) > 1e-13 * max(np.abs(c).max(), 1e-300)
    return np.ascontiguousarray(P[keep]), np.ascontiguousarray(c[keep])


def poly_energy(powers, coeffs, m):
    """sum_j c_j m^p_j at the rows of m (n,3)."""
    m = np.asarray(m, dtype=np.float64).reshape(-1, 3)
    return np.prod(m[:, None, :] ** np.asarray(powers)[None, :, :], axis=2) @ np.asarray(coeffs)


def su_energy(function):
    """The energy function in simulation units (energy densities in J/m^3)."""
    return lambda m: float(to_su(function(m), "J/m^3"))


__all__ = ["PredefinedAnisotropy", "uniaxial_anisotropy", "cubic_anisotropy", "fit_polynomial",
           "monomials", "poly_energy", "NmagUserError"]
