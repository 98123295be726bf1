"""SI values and nmag's simulation units (su).

The reference's default simulation units are A=1e-3, kg=1e-27, m=1e-9, s=1e-12
(interface/nmag/main.py:188-196; conversion interface/nsim/su_units.py:213-254).
A quantity is held as (value, dimension vector over m, kg, s, A); its su value
is value / prod(unit_d ** power_d).
"""
import math
import re

import numpy as np

_BASE = ("m", "kg", "s", "A")
SU_BASE = {"m": 1e-9, "kg": 1e-27, "s": 1e-12, "A": 1e-3}

# derived units in terms of (m, kg, s, A)
_DERIVED = {
    "": (0, 0, 0, 0),
    "m": (1, 0, 0, 0), "kg": (0, 1, 0, 0), "s": (0, 0, 1, 0), "A": (0, 0, 0, 1),
    "J": (2, 1, -2, 0), "T": (0, 1, -2, -1), "N": (1, 1, -2, 0), "W": (2, 1, -3, 0),
    "V": (2, 1, -3, -1), "C": (0, 0, 1, 1), "Hz": (0, 0, -1, 0),
}


def _parse(units):
    """Parse 'A/m', 'J/m^3', 'm/A s', '1/s', 's^(-1)' into a dimension tuple."""
    dims = np.zeros(4, dtype=int)
    if units is None:
        return tuple(dims)
    s = units.replace("(", "").replace(")", "").strip()
    sign = 1
    for tok in re.findall(r"/|[A-Za-z]+(?:\^-?\d+)?|1", s):
        if tok == "/":
            sign = -1
            continue
        if tok == "1":
            continue
        name, _, pw = tok.partition("^")
        p = int(pw) if pw else 1
        if name not in _DERIVED:
            raise ValueError("unknown unit %r in %r" % (name, units))
        dims += sign * p * np.array(_DERIVED[name])
    return tuple(int(v) for v in dims)


class SI:
    """A dimensioned SI value: SI(8.6e5, "A/m").  Supports * / + - and **."""

    def __init__(self, value, units=""):
        if isinstance(units, tuple):
            self.dims = units
        else:
            self.dims = _parse(units)
        self.value = value

    def _combine(self, other, op):
        if isinstance(other, SI):
            return other
        return SI(other, "")

    def __mul__(self, o):
        o = o if isinstance(o, SI) else SI(o)
        return SI(self.value * o.value, tuple(a + b for a, b in zip(self.dims, o.dims)))

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = o if isinstance(o, SI) else SI(o)
        return SI(self.value / o.value, tuple(a - b for a, b in zip(self.dims, o.dims)))

    def __rtruediv__(self, o):
        return SI(o) / self

    def __add__(self, o):
        o = o if isinstance(o, SI) else SI(o)
        if o.dims != self.dims:
            raise ValueError("adding incompatible SI quantities")
        return SI(self.value + o.value, self.dims)

    __radd__ = __add__

    def __sub__(self, o):
        return self + (-1.0) * o

    def __neg__(self):
        return SI(-self.value, self.dims)

    def __pow__(self, p):
        return SI(self.value ** p, tuple(a * p for a in self.dims))

    def is_compatible_with(self, other):
        other = other if isinstance(other, SI) else SI(1.0, other)
        return self.dims == other.dims

    def in_units_of(self, other):
        other = other if isinstance(other, SI) else SI(1.0, other)
        if not self.is_compatible_with(other):
            raise ValueError("incompatible units")
        return self.value / other.value

    def __repr__(self):
        parts = ["%s^%d" % (b, p) if p != 1 else b for b, p in zip(_BASE, self.dims) if p]
        return "SI(%r, '%s')" % (self.value, " ".join(parts))


mu0 = SI(4e-7 * math.pi, "N/A^2")


def su_factor(dims):
    """SI value of one simulation unit of the given dimension."""
    f = 1.0
    for b, p in zip(_BASE, dims):
        f *= SU_BASE[b] ** p
    return f


def to_su(q, compatible_with=None):
    """Simulation-unit number of an SI quantity (plain numbers pass through)."""
    if not isinstance(q, SI):
        return q
    if compatible_with is not None:
        ref = compatible_with if isinstance(compatible_with, SI) else SI(1.0, compatible_with)
        if q.dims != ref.dims:
            raise ValueError("expected units compatible with %r, got %r" % (ref, q))
    return np.asarray(q.value, dtype=np.float64) / su_factor(q.dims)


def from_su(x, dims):
    return np.asarray(x) * su_factor(dims if isinstance(dims, tuple) else _parse(dims))
