"""Differentiation form: the discrete adjoint against finite differences.

The vector-Jacobian product of the right-hand side is checked by dot-product
tests against central differences of the forward right-hand side, both of
this library and of the CPU oracle (which pins it to the reference's field
algorithm).  Trajectory gradients of the loss <m_x>(T) with respect to the
per-element anisotropy, Ms, the pulse amplitudes and m(0) are checked against
central differences of the same forward run (relative step 1e-3, bar 1e-4),
across checkpoint intervals, and mixed precision against FP64.
"""
import numpy as np
import pytest

import oracle
from helpers import material, orc_params, random_m
from nmag import SI, kuhn_cuboid, to_su
from nmag._native import NativeMagnet
from nmag.adjoint import TrajectoryProblem, trajectory_loss

pytestmark = pytest.mark.gpu

K_SU = float(to_su(SI(1.0, "J/m^3"), "J/m^3"))
MT_SU = float(to_su(SI(1e-3 / (4e-7 * np.pi), "A/m"), "A/m"))  # 1 mT / mu0 in su


def _magnet(dims=(6, 4, 2, 2.5), precision="fp64", alpha=0.5, rtol=1e-13, potential="auto"):
    coords, tets = kuhn_cuboid(*dims)
    mat = material(alpha=alpha)
    g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=precision, nbc_rtol=rtol,
                     dbc_rtol=rtol, nbc_maxits=20000, dbc_maxits=20000, potential=potential)
    return coords, tets, mat, g


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    m = random_m(n, seed)
    H = np.tile(rng.uniform(-0.03, 0.03, 3), (n, 1))
    k1 = rng.uniform(0.01, 0.04, n)
    ax = rng.standard_normal(3)
    ax = np.tile(ax / np.linalg.norm(ax), (n, 1))
    return m, H, k1, ax


def _setup(g, H, k1, ax):
    g.set_field("H_ext", H)
    g.set_field("axis", ax)
    g.set_field("K1", k1)


def _f(g, m):
    return g.rhs(m.reshape(-1))


@pytest.mark.parametrize("route", ["iterative", "dense"])
def test_rhs_vjp_dot_products_against_central_differences(route):
    coords, tets, mat, g = _magnet(potential=route)
    assert g.dense_potential == (route == "dense")
    m, H, k1, ax = _inputs(g.n, 1)
    _setup(g, H, k1, ax)
    rng = np.random.default_rng(2)
    lam = rng.standard_normal(3 * g.n)
    r = g.rhs_vjp(m.reshape(-1), lam)
    eps = 1e-5
    # m
    u = rng.standard_normal(3 * g.n)
    fd = (_f(g, m.reshape(-1) + eps * u) - _f(g, m.reshape(-1) - eps * u)) / (2 * eps)
    assert abs(r["y"] @ u - lam @ fd) <= 1e-6 * abs(lam @ fd)
    # H_ext
    dH = rng.standard_normal((g.n, 3)) * 0.01
    g.set_field("H_ext", H + eps * dH)
    fp = _f(g, m)
    g.set_field("H_ext", H - eps * dH)
    fm = _f(g, m)
    g.set_field("H_ext", H)
    ref = lam @ (fp - fm) / (2 * eps)
    assert abs(r["H_ext"] @ dH.reshape(-1) - ref) <= 1e-6 * abs(ref)
    # K1 per node
    dk = rng.standard_normal(g.n) * 0.01
    g.set_field("K1", k1 + eps * dk)
    fp = _f(g, m)
    g.set_field("K1", k1 - eps * dk)
    fm = _f(g, m)
    g.set_field("K1", k1)
    ref = lam @ (fp - fm) / (2 * eps)
    assert abs(r["K1"] @ dk - ref) <= 1e-6 * abs(ref)
    # Ms (exchange constant A held fixed)
    c = mat.su_coeffs()
    A = -c["negJ"] * c["mu0"] * c["Ms"] / 2

    def with_ms(ms):
        cc = dict(c, Ms=ms, negJ=-2 * A / (c["mu0"] * ms))
        g.set_material(cc)
        return _f(g, m)

    h = 1e-6 * c["Ms"]
    ref = lam @ (with_ms(c["Ms"] + h) - with_ms(c["Ms"] - h)) / (2 * h)
    g.set_material(c)
    assert abs(r["Ms"] - ref) <= 1e-6 * abs(ref)


def test_rhs_vjp_against_oracle_jacobian():
    """<J^T lam, u> of the GPU adjoint = <lam, J u> with J u from central
    differences of the oracle's right-hand side."""
    coords, tets, mat, g = _magnet(dims=(5, 3, 2, 2.0))
    m, H, k1, ax = _inputs(g.n, 5)
    _setup(g, H, k1, ax)
    orc = oracle.OracleMesh(coords, tets)
    P = orc_params(mat)
    rng = np.random.default_rng(6)
    lam = rng.standard_normal(3 * g.n)
    u = rng.standard_normal((g.n, 3))
    r = g.rhs_vjp(m.reshape(-1), lam)
    eps = 1e-5
    fo = lambda mm: orc.fields(P, mm, H_ext=H, k1=k1, axis=ax, rtol=1e-13)["dmdt"].reshape(-1)
    ref = lam @ (fo(m + eps * u) - fo(m - eps * u)) / (2 * eps)
    assert abs(r["y"] @ u.reshape(-1) - ref) <= 1e-6 * abs(ref)


@pytest.mark.parametrize("prec,tol", [("mixed", 1e-5), ("mixed24", 1e-4)])
def test_rhs_vjp_mixed_precision_close_to_fp64(prec, tol):
    """The vector-Jacobian product with FP32 (24-bit) operator storage, whose
    demag part applies the transposed potential operator, against FP64."""
    out = {}
    for p in ("fp64", prec):
        _, _, _, g = _magnet(precision=p, rtol=1e-12)
        m, H, k1, ax = _inputs(g.n, 3)
        _setup(g, H, k1, ax)
        lam = np.random.default_rng(4).standard_normal(3 * g.n)
        out[p] = g.rhs_vjp(m.reshape(-1), lam)
    for key in ("y", "K1", "H_ext"):
        a, b = out[prec][key], out["fp64"][key]
        assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(b)), key
    assert abs(out[prec]["Ms"] - out["fp64"]["Ms"]) <= tol * abs(out["fp64"]["Ms"])


def _problem(precision="fp64", checkpoint_every=10, n_steps=100, n_pulses=4, seed=7,
             rtol=1e-12, dims=(8, 4, 1, 2.5), potential="auto"):
    coords, tets = kuhn_cuboid(*dims)
    mat = material(alpha=0.02)
    g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=precision, nbc_rtol=rtol,
                     dbc_rtol=rtol, nbc_maxits=20000, dbc_maxits=20000, potential=potential)
    rng = np.random.default_rng(seed)
    n = len(coords)
    m0 = np.tile([1.0, 0.3, 0.1], (n, 1)) + 0.05 * rng.standard_normal((n, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    dt = 0.2
    prob = TrajectoryProblem(g, mat.su_coeffs(), coords, tets, m0,
                             np.tile([0.0, 0.0, 1.0], (n, 1)), dt=dt, n_steps=n_steps,
                             n_pulses=n_pulses, pulse_len=dt * n_steps / n_pulses,
                             checkpoint_every=checkpoint_every)
    K = rng.normal(2e4, 5e3, len(tets)) * K_SU
    amps = rng.normal(0.0, 5.0, n_pulses) * MT_SU
    return prob, K, mat.su_coeffs()["Ms"], amps


@pytest.mark.parametrize("route", ["iterative", "dense"])
def test_trajectory_gradient_matches_central_differences(route):
    prob, K, Ms, amps = _problem(potential=route)
    prob.run(K, Ms, amps)
    gK, gMs, gA, gm0 = prob.gradient(amps)
    rel = 1e-3

    def fd(fun, x):
        h = rel * max(abs(x), 1e-12)
        return (fun(x + h) - fun(x - h)) / (2 * h)

    ref = fd(lambda v: prob.run(K, v, amps), Ms)
    assert abs(gMs - ref) <= 1e-4 * abs(ref)
    for k in range(len(amps)):
        def f(v, k=k):
            a = amps.copy()
            a[k] = v
            return prob.run(K, Ms, a)
        h = rel * np.max(np.abs(amps))
        ref = (f(amps[k] + h) - f(amps[k] - h)) / (2 * h)
        assert abs(gA[k] - ref) <= 1e-4 * np.max(np.abs(gA)), k
    # per-element K along a random direction (relative perturbation)
    d = np.random.default_rng(9).standard_normal(len(K)) * K
    ref = (prob.run(K + rel * d, Ms, amps) - prob.run(K - rel * d, Ms, amps)) / (2 * rel)
    assert abs(gK @ d - ref) <= 1e-4 * abs(ref)
    # initial magnetisation along a random direction
    u = np.random.default_rng(10).standard_normal(prob.m0.shape) * 1e-2
    m0 = prob.m0.copy()
    prob.m0 = m0 + rel * u
    lp = prob.run(K, Ms, amps)
    prob.m0 = m0 - rel * u
    lm = prob.run(K, Ms, amps)
    prob.m0 = m0
    ref = (lp - lm) / (2 * rel)
    assert abs(np.sum(gm0 * u) - ref) <= 1e-4 * abs(ref)


def test_checkpoint_interval_does_not_change_the_gradient():
    grads = []
    for k in (1, 7, 100):
        prob, K, Ms, amps = _problem(checkpoint_every=k)
        prob.run(K, Ms, amps)
        grads.append(prob.gradient(amps))
    for other in grads[1:]:
        for a, b in zip(other[:3], grads[0][:3]):
            a, b = np.atleast_1d(a), np.atleast_1d(b)
            assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b))


def test_mixed_precision_gradient_close_to_fp64_adjoint():
    res = {}
    for prec in ("fp64", "mixed"):
        prob, K, Ms, amps = _problem(precision=prec, rtol=1e-10)
        prob.run(K, Ms, amps)
        res[prec] = prob.gradient(amps)
    for a, b in zip(res["mixed"][:3], res["fp64"][:3]):
        a, b = np.atleast_1d(a), np.atleast_1d(b)
        assert np.max(np.abs(a - b)) <= 1e-4 * np.max(np.abs(b))


def test_autograd_function_gives_the_adjoint_gradient():
    import torch
    prob, K, Ms, amps = _problem(n_steps=50, n_pulses=2)
    tK = torch.tensor(K, dtype=torch.float64, requires_grad=True)
    tMs = torch.tensor(Ms, dtype=torch.float64, requires_grad=True)
    tA = torch.tensor(amps, dtype=torch.float64, requires_grad=True)
    loss = trajectory_loss(tK, tMs, tA, prob)
    (2.0 * loss).backward()
    assert abs(loss.item() - prob.run(K, Ms, amps)) <= 1e-12
    gK, gMs, gA, _ = prob.gradient(amps)
    np.testing.assert_allclose(tK.grad.numpy(), 2 * gK, rtol=1e-9, atol=1e-20)
    np.testing.assert_allclose(float(tMs.grad), 2 * gMs, rtol=1e-9)
    np.testing.assert_allclose(tA.grad.numpy(), 2 * gA, rtol=1e-9)
