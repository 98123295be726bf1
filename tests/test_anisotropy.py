"""Polynomial anisotropy (cubic, sums, custom functions): the oracle's field
(oracle.poly_anis, src/nsim_anisotropy.ml:167-234) pinned by finite
differences of the energy along great circles, by the uniaxial closed form,
and by H.m = 0; the front end's fit (interface/nmag/anisotropy.py,
src/nsim_anisotropy.ml:119-133) pinned by the energies it must reproduce and
by the exact coefficients of a degree-2 uniaxial energy."""
import numpy as np
import pytest

import oracle
from helpers import random_m
from nmag import (MagMaterial, NmagUserError, PredefinedAnisotropy, SI, cubic_anisotropy,
                  uniaxial_anisotropy)
from nmag.anisotropy import fit_polynomial, monomials, poly_energy, su_energy

S = 0.37  # 1 / (mu0 Ms) of the field, any positive number for these checks


def tangents(m):
    """Two unit tangents of the sphere at each row of m."""
    ref = np.where(np.abs(m[:, :1]) < 0.9, [[1.0, 0, 0]], [[0, 1.0, 0]])
    t1 = np.cross(m, ref)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    return t1, np.cross(m, t1)


def great_circle_fd(E, m, t, h=1e-4):
    """dE/dtheta of E along m cos(theta) + t sin(theta) at theta = 0
    (fourth-order central differences)."""
    def at(th):
        return np.array([E(x) for x in m * np.cos(th) + t * np.sin(th)])
    return (8 * (at(h) - at(-h)) - (at(2 * h) - at(-2 * h))) / (12 * h)


@pytest.mark.parametrize("K", [(3.1, 0.0, 0.0), (2.0, -1.5, 0.0), (1.0, 0.7, -2.2)])
def test_cubic_field_is_minus_tangential_energy_gradient(K):
    """H.t = -S dE/dtheta along both tangents, H.m = 0: against the cubic
    energy formula itself, not the fitted polynomial."""
    an = cubic_anisotropy([1.0, 2.0, 0.5], [0.3, -1.0, 2.0], *K)
    P, c = fit_polynomial(an.function, an.order)
    m = random_m(40, 3)
    H = oracle.poly_anis(m, P, c, S)
    assert np.max(np.abs((H * m).sum(1))) < 1e-12 * np.abs(H).max()
    for t in tangents(m):
        fd = great_circle_fd(an.function, m, t)
        assert np.allclose((H * t).sum(1), -S * fd, rtol=0, atol=1e-8 * max(1.0, np.abs(fd).max()))


def test_uniaxial_polynomial_equals_the_closed_form():
    """-K1 (a.m)^2 as a degree-2 polynomial (the reference's representation of
    uniaxial_anisotropy with K2 = 0): the same field as the oracle's uniaxial
    closed form, q (a - u m)."""
    a = np.array([0.3, -0.4, 0.866])
    a /= np.linalg.norm(a)
    K1 = 2.5
    P, c = fit_polynomial(uniaxial_anisotropy(a, K1).function, 2)
    m = random_m(30, 8)
    u = m @ a
    closed = S * (2 * K1 * u)[:, None] * (a[None, :] - u[:, None] * m)
    assert np.allclose(oracle.poly_anis(m, P, c, S), closed, rtol=0, atol=1e-13 * K1)
    # the fitted coefficients are those of -K1 sum_ij a_i a_j m_i m_j
    want = {}
    for i in range(3):
        for j in range(3):
            p = [0, 0, 0]
            p[i] += 1
            p[j] += 1
            want[tuple(p)] = want.get(tuple(p), 0.0) - K1 * a[i] * a[j]
    got = {tuple(p): v for p, v in zip(P.tolist(), c)}
    assert set(got) <= set(want)
    for p, v in want.items():
        assert abs(got.get(p, 0.0) - v) < 1e-13 * K1


def test_fit_reproduces_energies_of_sums_and_orders():
    cub = cubic_anisotropy([0, 0, 1], [1, 1, 0], 4.5e-3, 1e-3, 2e-4)
    uni = uniaxial_anisotropy([1, 0, 1], 3e-3, 1e-3)
    for an in (cub, uni, cub + uni, cub - uni, -cub, +uni):
        P, c = fit_polynomial(an.function, an.order)
        assert all(sum(p) == an.order for p in P.tolist())
        m = random_m(50, 12)
        e = np.array([an.function(x) for x in m])
        assert np.max(np.abs(poly_energy(P, c, m) - e)) < 1e-12 * max(1e-3, np.abs(e).max())
    assert (cub + uni).order == 8 and (cub - uni).order == 8
    assert len(monomials(8)) == 45 and len(monomials(4)) == 15


def test_custom_function_material_and_errors():
    def e(m):
        return SI(1e4, "J/m^3") * (m[0] ** 2 * m[1] ** 2)
    mat = MagMaterial("X", anisotropy=e, anisotropy_order=4)
    P, c = mat.su_anisotropy_poly()
    m = random_m(20, 2)
    want = np.array([su_energy(e)(x) for x in m])
    assert np.max(np.abs(poly_energy(P, c, m) - want)) < 1e-12 * np.abs(want).max()
    assert mat.su_anisotropy() is None
    cubic = MagMaterial("C", anisotropy=cubic_anisotropy([1, 0, 0], [0, 1, 0], SI(4e4, "J/m^3")))
    assert cubic.su_anisotropy_poly() is not None and cubic.anisotropy_order == 4
    uni = MagMaterial("U", anisotropy=uniaxial_anisotropy([0, 0, 1], SI(4e4, "J/m^3")))
    assert uni.su_anisotropy() is not None and uni.su_anisotropy_poly() is None
    with pytest.raises(NmagUserError):
        MagMaterial("X", anisotropy=e)  # no order
    with pytest.raises(NmagUserError):
        MagMaterial("X", anisotropy=cubic_anisotropy([1, 0, 0], [0, 1, 0], 1.0),
                    anisotropy_order=4)
    with pytest.raises(NmagUserError):
        PredefinedAnisotropy()
    with pytest.raises(NmagUserError):
        fit_polynomial(lambda m: 0.0, 40)
