"""Polynomial anisotropy on the device (nsim_set_anisotropy): the cubic
energy and custom polynomials, alone and beside the uniaxial K1/K2 fields,
against the oracle; per-magnet polynomials of a batch; the adjoint's
(dH_anis/dm)^T against central differences; the Simulation front end with
cubic_anisotropy (interface/nmag/anisotropy.py:138-201)."""
import math

import numpy as np
import pytest

import oracle
from helpers import PERMALLOY, material, orc_params, random_m, rel_err
from nmag import SI, MagMaterial, Simulation, cubic_anisotropy, kuhn_cuboid
from nmag._native import NativeBatch, NativeMagnet, NsimError
from nmag.anisotropy import fit_polynomial

pytestmark = pytest.mark.gpu


def _note(*a):
    print("[case]", *a, flush=True)


def _cubic(K=(0.03, -0.02, 0.01)):
    an = cubic_anisotropy([1.0, 2.0, 0.5], [0.3, -1.0, 2.0], *K)
    return fit_polynomial(an.function, an.order)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("mixed", 1e-6)])
def test_cubic_and_uniaxial_fields_match_the_oracle(prec, tol):
    _note("cubic + uniaxial", prec)
    coords, tets = kuhn_cuboid(8, 6, 3, 2.5)
    mat = material()
    g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=prec, nbc_rtol=1e-13,
                     dbc_rtol=1e-13, nbc_maxits=20000, dbc_maxits=20000)
    P, c = _cubic()
    n = g.n
    m = random_m(n, 21)
    rng = np.random.default_rng(22)
    H = np.tile([0.01, -0.02, 0.015], (n, 1))
    k1 = rng.uniform(0.01, 0.03, n)
    ax = np.tile([0.0, 0.6, 0.8], (n, 1))
    g.set_anisotropy(P, c)
    g.set_field("H_ext", H)
    g.set_field("axis", ax)
    g.set_field("K1", k1)
    g.set_field("m", m)
    g.execute()
    ref = oracle.OracleMesh(coords, tets).fields(orc_params(mat), m, H_ext=H, k1=k1, axis=ax,
                                                 rtol=1e-13, poly=(P, c))
    # H_anis is computed in FP64 in every mode
    assert rel_err(g.get_field("H_anis"), ref["H_anis"]) < 1e-13
    for key in ("H_total", "dmdt"):
        assert rel_err(g.get_field(key), ref[key]) < tol, key
    # clearing leaves the uniaxial part alone
    g.set_anisotropy(np.zeros((0, 3)), np.zeros(0))
    g.execute()
    only = oracle.OracleMesh(coords, tets).fields(orc_params(mat), m, H_ext=H, k1=k1, axis=ax,
                                                  rtol=1e-13)
    assert rel_err(g.get_field("H_anis"), only["H_anis"]) < 1e-13


def test_batch_magnets_take_their_own_polynomials():
    _note("batch of 3 with per-magnet polynomials")
    dims = [(6, 5, 2, 2.5), (5, 5, 3, 2.5), (7, 4, 2, 2.0)]
    meshes = [kuhn_cuboid(*d) for d in dims]
    mats = [material(alpha=a) for a in (0.5, 0.3, 0.2)]
    b = NativeBatch(meshes, [mt.su_coeffs() for mt in mats], nbc_rtol=1e-13, dbc_rtol=1e-13,
                    nbc_maxits=20000, dbc_maxits=20000)
    polys = [_cubic(), None, fit_polynomial(lambda m: 0.02 * m[0] * m[2] - 0.01 * m[1] ** 2, 2)]
    for k, p in enumerate(polys):
        if p is not None:
            b.set_anisotropy(*p, magnet=k)
    ms = [random_m(len(cd), 30 + k) for k, (cd, _) in enumerate(meshes)]
    b.set_field("m", np.concatenate(ms))
    b.execute()
    Ha, dm = b.split(b.get_field("H_anis")), b.split(b.get_field("dmdt"))
    for k, (cd, t) in enumerate(meshes):
        ref = oracle.OracleMesh(cd, t).fields(orc_params(mats[k]), ms[k], rtol=1e-13,
                                              poly=polys[k])
        if polys[k] is None:
            assert np.max(np.abs(Ha[k])) == 0.0
        else:
            assert rel_err(Ha[k], ref["H_anis"]) < 1e-13, k
        assert rel_err(dm[k], ref["dmdt"]) < 1e-10, k


def test_vjp_of_the_polynomial_field_against_central_differences():
    _note("rhs_vjp with a cubic polynomial")
    coords, tets = kuhn_cuboid(6, 4, 2, 2.5)
    mat = material()
    g = NativeMagnet(coords, tets, mat.su_coeffs(), nbc_rtol=1e-13, dbc_rtol=1e-13,
                     nbc_maxits=20000, dbc_maxits=20000)
    g.set_anisotropy(*_cubic((0.3, -0.2, 0.1)))  # strong, so it dominates the products
    m = random_m(g.n, 5).reshape(-1)
    rng = np.random.default_rng(6)
    lam = rng.standard_normal(3 * g.n)
    r = g.rhs_vjp(m, lam)
    eps = 1e-5
    for seed in (7, 8):
        u = np.random.default_rng(seed).standard_normal(3 * g.n)
        fd = (g.rhs(m + eps * u) - g.rhs(m - eps * u)) / (2 * eps)
        assert abs(r["y"] @ u - lam @ fd) <= 1e-6 * abs(lam @ fd)


def test_arguments_are_checked():
    coords, tets = kuhn_cuboid(2, 2, 1, 2.5)
    g = NativeMagnet(coords, tets, material().su_coeffs())
    for P, c, k in (([[-1, 0, 2]], [1.0], -1), ([[17, 0, 0]], [1.0], -1),
                    ([[2, 0, 0]], [float("nan")], -1), ([[2, 0, 0]], [1.0], 3),
                    (np.zeros((257, 3)), np.ones(257), -1)):
        with pytest.raises(NsimError):
            g.set_anisotropy(P, c, magnet=k)


def test_simulation_with_cubic_anisotropy_relaxes_onto_a_cube_axis(tmp_path, monkeypatch):
    """A 10 nm permalloy-like cube (demag nearly isotropic) with K1 > 0 cubic
    anisotropy along rotated axes: relaxed m lies along the axis nearest the
    start, E_anis -> 0 there."""
    monkeypatch.chdir(tmp_path)
    _note("Simulation + cubic_anisotropy relax")
    a1 = np.array([1.0, 1.0, 0.0]) / math.sqrt(2)
    an = cubic_anisotropy(a1, [-1.0, 1.0, 0.0], SI(3e5, "J/m^3"))
    mat = MagMaterial("Py", llg_damping=0.5, anisotropy=an, **PERMALLOY)
    sim = Simulation()
    sim.load_mesh(kuhn_cuboid(4, 4, 4, 2.5), [("Py", mat)], unit_length=SI(1e-9, "m"))
    sim.set_m((a1 + [0.2, -0.3, 0.25]).tolist())
    sim.relax()
    m = sim.get_subfield_average_siv("m", "Py")
    assert float(np.dot(m, a1)) > 0.995
    # the anisotropy field is the device's, as the oracle has it at the final m
    mm = sim.get_subfield("m_Py")
    P, c = mat.su_anisotropy_poly()
    ref = oracle.poly_anis(mm, P, c, 1.0 / (mat.su_mu0 * mat.su_Ms))
    assert rel_err(sim._magnet.get_field("H_anis"), ref) < 1e-12
