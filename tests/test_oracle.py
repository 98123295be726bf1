"""Pins the CPU oracle against closed forms, invariants and independent
computations (numerical quadrature, scipy direct solves, finite differences).

The reference ships no golden vectors for this path, so every function of the
oracle is pinned here by a check that a plausible mistake (dropped term, wrong
sign, transposed index, wrong scale) would fail.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from helpers import material, orc_params, random_m, rel_err
from nmag import kuhn_cuboid, cuboid_sizes


def aharoni_nz(lx, ly, lz):
    """Demagnetising factor along z of a uniformly magnetised rectangular prism
    of sides lx, ly, lz (A. Aharoni, J. Appl. Phys. 83, 3432 (1998), eq. 1)."""
    a, b, c = lx / 2.0, ly / 2.0, lz / 2.0
    abc = math.sqrt(a * a + b * b + c * c)
    ab, bc, ac = math.sqrt(a * a + b * b), math.sqrt(b * b + c * c), math.sqrt(a * a + c * c)
    t = ((b * b - c * c) / (2 * b * c) * math.log((abc - a) / (abc + a))
         + (a * a - c * c) / (2 * a * c) * math.log((abc - b) / (abc + b))
         + b / (2 * c) * math.log((ab + a) / (ab - a))
         + a / (2 * c) * math.log((ab + b) / (ab - b))
         + c / (2 * a) * math.log((bc - b) / (bc + b))
         + c / (2 * b) * math.log((ac - a) / (ac + a))
         + 2 * math.atan(a * b / (c * abc))
         + (a ** 3 + b ** 3 - 2 * c ** 3) / (3 * a * b * c)
         + (a * a + b * b - 2 * c * c) / (3 * a * b * c) * abc
         + c / (a * b) * (ac + bc)
         - (ab ** 3 + bc ** 3 + ac ** 3) / (3 * a * b * c))
    return t / math.pi


def test_aharoni_formula_itself():
    assert abs(aharoni_nz(1, 1, 1) - 1.0 / 3.0) < 1e-12
    s = aharoni_nz(3, 2, 1) + aharoni_nz(1, 3, 2) + aharoni_nz(2, 1, 3)
    assert abs(s - 1.0) < 1e-12


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 2), (12, 12, 4), (40, 20, 2)])
def test_kuhn_mesh_matches_product_generator_and_sizes(dims):
    c1, t1 = oracle.kuhn_cuboid(*dims, 2.5)
    c2, t2 = kuhn_cuboid(*dims, 2.5)
    assert np.array_equal(c1, c2) and np.array_equal(t1, t2)
    M = oracle.OracleMesh(c1, t1, build_bem=False)
    n, T, nb, ntri = cuboid_sizes(*dims)
    assert (M.n, M.ntet, M.nb, M.ntri) == (n, T, nb, ntri)


def test_config_sizes():
    assert cuboid_sizes(12, 12, 4) == (845, 3456, 482, 960)
    assert cuboid_sizes(80, 40, 5) == (19926, 96000, 7602, 15200)
    assert cuboid_sizes(40, 20, 2) == (2583, 9600, 1842, 3680)


def test_volumes_and_stiffness_closed_forms():
    h = 2.5
    M = oracle.OracleMesh(*oracle.kuhn_cuboid(6, 5, 4, h), build_bem=False)
    V = M.volumes()
    assert abs(V.sum() - 6 * 5 * 4 * h ** 3) < 1e-9
    i = 2 + 7 * (2 + 6 * 2)  # an interior node
    assert abs(V[i] - h ** 3) < 1e-12
    rp, ci, K = M.stiffness()
    A = sp.csr_matrix((K, ci, rp), shape=(M.n, M.n))
    assert abs(A - A.T).max() < 1e-13
    assert np.abs(A @ np.ones(M.n)).max() < 1e-12
    row = A.getrow(i).toarray().ravel()
    nz = {int(j - i): row[j] for j in np.nonzero(np.abs(row) > 1e-12)[0]}
    assert nz == pytest.approx({0: 6 * h, 1: -h, -1: -h, 7: -h, -7: -h, 42: -h, -42: -h})
    # energy identity of a linear field: u^T K u = |g|^2 * volume
    g = np.array([0.3, -1.1, 0.7])
    u = M.coords @ g
    assert abs(u @ (A @ u) - g @ g * V.sum()) < 1e-9 * (g @ g) * V.sum()


def test_exchange_spin_wave_closed_form():
    h, k = 2.0, 0.3
    M = oracle.OracleMesh(*oracle.kuhn_cuboid(20, 4, 4, h), build_bem=False)
    mat = material()
    P = orc_params(mat)
    x = M.coords[:, 0]
    m = np.stack([np.cos(k * x), np.sin(k * x), 0 * x], axis=1)
    f = M.fields(P, m, do_demag=False)
    expect = P.negJ * 2 * (1 - math.cos(k * h)) / h ** 2 * m
    ix = np.nonzero((x > 0) & (x < 40) & (M.coords[:, 1] > 0) & (M.coords[:, 1] < 8)
                    & (M.coords[:, 2] > 0) & (M.coords[:, 2] < 8))[0]
    assert rel_err(f["H_exch"][ix], expect[ix]) < 1e-12


def test_divergence_and_gradient_identities():
    h = 2.5
    M = oracle.OracleMesh(*oracle.kuhn_cuboid(6, 5, 4, h), build_bem=False)
    Ms = 0.86
    rho = M.div(Ms, random_m(M.n, 3))
    assert abs(rho.sum()) < 1e-12 * np.abs(rho).max() * M.n
    rz = M.div(Ms, np.tile([0.0, 0.0, 1.0], (M.n, 1)))
    z = M.coords[:, 2]
    face = np.nonzero((z == 4 * h) & (M.coords[:, 0] > 0) & (M.coords[:, 0] < 6 * h)
                      & (M.coords[:, 1] > 0) & (M.coords[:, 1] < 5 * h))[0]
    interior = np.nonzero((z > 0) & (z < 4 * h) & (M.coords[:, 0] > 0) & (M.coords[:, 0] < 6 * h)
                          & (M.coords[:, 1] > 0) & (M.coords[:, 1] < 5 * h))[0]
    # uniform m: no volume charge, surface charge Ms (m.n) * integral of the hat = Ms h^2
    assert np.abs(rz[interior]).max() < 1e-12
    assert np.allclose(rz[face], Ms * h * h, rtol=1e-12)
    g = np.array([0.2, -0.5, 1.3])
    H = M.grad(M.coords @ g)
    assert np.allclose(H, -g, atol=1e-12)


def test_cg_matches_direct_solves():
    M = oracle.OracleMesh(*oracle.kuhn_cuboid(7, 5, 4, 2.0), build_bem=False)
    rp, ci, K = M.stiffness()
    A = sp.csr_matrix((K, ci, rp), shape=(M.n, M.n)).tocsc()
    rng = np.random.default_rng(0)
    b = rng.standard_normal(M.n)
    b -= b.mean()
    x, it, res = M.cg(b, np.zeros(M.n), rtol=1e-14)
    # Neumann problem: compare after removing the constant null space
    ref = spla.lsqr(A, b, atol=1e-15, btol=1e-15, iter_lim=20000)[0]
    assert rel_err(x - x.mean(), ref - ref.mean()) < 1e-10
    # Dirichlet problem on the interior with boundary values fixed
    bn, _, _, _ = M.boundary()
    fixed = np.zeros(M.n, dtype=bool)
    fixed[bn] = True
    x0 = np.zeros(M.n)
    x0[bn] = rng.standard_normal(len(bn))
    xd, _, _ = M.cg(None, x0, fixed=fixed, rtol=1e-14)
    I = np.nonzero(~fixed)[0]
    Aii = A[I][:, I]
    ref_i = spla.spsolve(Aii, -(A[I][:, bn] @ x0[bn]))
    assert rel_err(xd[I], ref_i) < 1e-10
    assert np.array_equal(xd[bn], x0[bn])
    # a linear function is harmonic: reproduced exactly from its boundary values
    g = np.array([1.0, -2.0, 0.5])
    xl = np.zeros(M.n)
    xl[bn] = (M.coords @ g)[bn]
    xs, _, _ = M.cg(None, xl, fixed=fixed, rtol=1e-14)
    assert rel_err(xs, M.coords @ g) < 1e-11


def _quad_double_layer(x, p0, p1, p2, n, order=60):
    g, w = np.polynomial.legendre.leggauss(order)
    g, w = (g + 1) / 2, w / 2
    area2 = np.linalg.norm(np.cross(p1 - p0, p2 - p0))
    tot = np.zeros(3)
    for gi, wi in zip(g, w):
        for gj, wj in zip(g, w):
            s, t = gi, gj * (1 - gi)
            lam = np.array([1 - s - t, s, t])
            y = lam[0] * p0 + lam[1] * p1 + lam[2] * p2
            d = x - y
            tot += wi * wj * (1 - gi) * area2 * (n @ d) / np.linalg.norm(d) ** 3 * lam
    return tot / (4 * math.pi)


@pytest.mark.parametrize("obs", [[0.5, 0.4, 1.0], [3.0, 2.0, -0.7], [0.7, 0.5, 0.3],
                                 [-4.0, 6.0, 2.5]])
def test_lindholm_vs_quadrature(obs):
    """Lindholm weights = (1/4pi) int lambda_a n.(x-y)/|x-y|^3 dS (double layer)."""
    rng = np.random.default_rng(5)
    p0, p1, p2 = np.zeros(3), np.array([2.0, 0.1, 0.0]), np.array([0.3, 1.7, 0.2])
    R, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    p0, p1, p2, x = (R @ p0, R @ p1, R @ p2, R @ np.asarray(obs))
    n = np.cross(p1 - p0, p2 - p0)
    n /= np.linalg.norm(n)
    got = oracle.lindholm(x, p0, p1, p2, n)
    ref = _quad_double_layer(x, p0, p1, p2, n)
    assert np.allclose(got, ref, rtol=1e-8, atol=1e-12)
    # the outward normal decides the sign
    assert np.allclose(oracle.lindholm(x, p0, p1, p2, -n), -got, rtol=1e-14)
    assert np.allclose(oracle.lindholm(x, p0, p2, p1, n), got[[0, 2, 1]], rtol=1e-12)


def test_vertex_solid_angles_of_a_cuboid():
    M = oracle.OracleMesh(*oracle.kuhn_cuboid(4, 3, 2, 2.5), build_bem=False)
    bn, _, _, sa = M.boundary()
    x = M.coords[bn]
    on = [(np.isclose(x[:, d], 0) | np.isclose(x[:, d], L)).astype(int)
          for d, L in enumerate((10.0, 7.5, 5.0))]
    k = on[0] + on[1] + on[2]  # 1 face, 2 edge, 3 corner
    expect = np.select([k == 1, k == 2, k == 3], [2 * math.pi, math.pi, math.pi / 2])
    assert np.allclose(sa, expect, rtol=1e-12)


@pytest.mark.parametrize("dims", [(1, 1, 1, 2.5), (4, 3, 2, 2.5), (12, 12, 4, 2.5)])
def test_bem_row_sums(dims):
    M = oracle.OracleMesh(*oracle.kuhn_cuboid(*dims))
    B = M.bem()
    assert np.allclose(B.sum(axis=1), -1.0, atol=1e-10)


def test_atan2_solid_angle_form_is_the_reference_acos_form_without_its_noise():
    """The literal acos form of the reference and the atan2 form used here are
    the same function; the acos form carries ~1e-9 absolute noise on the
    40x20x2 mesh (acos near 1), the atan2 form keeps B 1 = -1 to ~1e-14."""
    c, t = oracle.kuhn_cuboid(40, 20, 2, 2.5)
    A = oracle.OracleMesh(c, t).bem()
    R = oracle.OracleMesh(c, t, acos_form=True).bem()
    assert np.abs(A - R).max() < 5e-9
    assert np.abs(A.sum(axis=1) + 1).max() < 1e-13
    x, p0, p1, p2 = (np.array([5.0, 3.0, 40.0]), np.zeros(3), np.array([2.0, 0.0, 0.0]),
                     np.array([0.0, 2.0, 0.0]))
    n = np.array([0.0, 0.0, 1.0])
    ref = _quad_double_layer(x, p0, p1, p2, n)
    assert np.allclose(oracle.lindholm(x, p0, p1, p2, n), ref, rtol=1e-10)


def test_uniform_cuboid_demag_factor_against_aharoni():
    M = oracle.OracleMesh(*oracle.kuhn_cuboid(12, 12, 4, 2.5))
    mat = material()
    P = orc_params(mat)
    V = M.volumes()
    for d, (lx, ly, lz) in enumerate([(30.0, 10.0, 30.0), (10.0, 30.0, 30.0),
                                      (30.0, 30.0, 10.0)]):
        m = np.zeros((M.n, 3))
        m[:, d] = 1.0
        f = M.fields(P, m)
        N = -(f["H_demag"][:, d] * V).sum() / V.sum() / P.Ms
        assert abs(N - aharoni_nz(lx, ly, lz)) < 0.02


def test_demag_field_is_odd_and_gauge_free():
    M = oracle.OracleMesh(*oracle.kuhn_cuboid(5, 4, 3, 2.5))
    P = orc_params(material())
    m = random_m(M.n, 2)
    a = M.fields(P, m)
    b = M.fields(P, -m)
    assert rel_err(-b["H_demag"], a["H_demag"]) < 1e-11
    # adding a constant to the Neumann warm start leaves H_demag unchanged
    c = M.fields(P, m, phi1=np.full(M.n, 3.7))
    assert rel_err(c["H_demag"], a["H_demag"]) < 1e-10


def test_llg_precession_and_damping_directions():
    M = oracle.OracleMesh(*oracle.kuhn_cuboid(1, 1, 1, 2.5), build_bem=False)
    alpha = 0.1
    mat = material(alpha=alpha)
    P = orc_params(mat)
    m = np.tile([1.0, 0.0, 0.0], (M.n, 1))
    H = np.tile([0.0, 0.0, 0.05], (M.n, 1))
    f = M.fields(P, m, H_ext=H, do_demag=False)
    gp = mat.su_llg_gamma_G / (1 + alpha ** 2)
    # precession: m moves from +x towards +y about +z; damping pulls m towards H
    assert np.allclose(f["dmdt"], [0.0, gp * 0.05, gp * alpha * 0.05], rtol=1e-12)
    # |m| = 1 is preserved to first order, and the c3 term restores it otherwise
    m2 = random_m(M.n, 1)
    f2 = M.fields(P, m2, H_ext=H, do_demag=False)
    assert np.abs((f2["dmdt"] * m2).sum(axis=1)).max() < 1e-14
    f3 = M.fields(P, 1.01 * m2, H_ext=H, do_demag=False)
    radial = (f3["dmdt"] * m2).sum(axis=1)
    assert np.allclose(radial, P.c3 * (1 - 1.01 ** 2) * 1.01, rtol=1e-10)


def test_anisotropy_is_negative_tangential_energy_gradient():
    M = oracle.OracleMesh(*oracle.kuhn_cuboid(2, 2, 1, 2.5), build_bem=False)
    mat = material()
    P = orc_params(mat)
    m = random_m(M.n, 4)
    rng = np.random.default_rng(1)
    k1 = rng.uniform(0.0, 0.05, M.n)
    k2 = rng.uniform(0.0, 0.01, M.n)
    ax = random_m(M.n, 8)
    f = M.fields(P, m, k1=k1, k2=k2, axis=ax, do_demag=False)
    eps = 1e-6
    for i in range(0, M.n, 3):
        E = lambda v: -k1[i] * (ax[i] @ v) ** 2 - k2[i] * (ax[i] @ v) ** 4
        g = np.array([(E(m[i] + eps * e) - E(m[i] - eps * e)) / (2 * eps) for e in np.eye(3)])
        gt = g - m[i] * (m[i] @ g)
        assert np.allclose(f["H_anis"][i], -gt / (P.mu0 * P.Ms), rtol=1e-6, atol=1e-12)


def test_jtimes_against_finite_differences():
    M = oracle.OracleMesh(*oracle.kuhn_cuboid(4, 3, 3, 2.5), build_bem=False)
    P = orc_params(material(alpha=0.3))
    m = random_m(M.n, 6)
    Hfix = np.tile([0.02, 0.01, -0.03], (M.n, 1))
    v = random_m(M.n, 7)

    def f(mm):
        # H frozen except for its exchange part: dmdt with H = Hfix + H_exch(mm)
        out = M.fields(P, mm, H_ext=Hfix, do_demag=False)
        return out["dmdt"] - P.c3 * (1 - (mm * mm).sum(1))[:, None] * mm

    eps = 1e-6
    fd = (f(m + eps * v) - f(m - eps * v)) / (2 * eps)
    H = M.fields(P, m, H_ext=Hfix, do_demag=False)["H_total"]
    assert rel_err(M.jtimes(P, m, H, v), fd) < 1e-7
