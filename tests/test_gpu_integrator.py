"""The device BDF/Newton-GMRES integrator: closed-form macrospin dynamics,
trajectories against the oracle's rhs integrated by SciPy (DOP853, tight
tolerances) at BDF orders up to 2 and up to 5, CVODE-like stop semantics, and
the psetup/psolve blocks."""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import oracle
from helpers import material, orc_params, random_m, rel_err
from nmag import kuhn_cuboid
from nmag._native import NativeMagnet

pytestmark = pytest.mark.gpu


def _magnet(nx, ny, nz, h, mat, rtol=1e-11, do_demag=True):
    coords, tets = kuhn_cuboid(nx, ny, nz, h)
    g = NativeMagnet(coords, tets, mat.su_coeffs(), do_demag=do_demag, nbc_rtol=rtol,
                     dbc_rtol=rtol, nbc_maxits=5000, dbc_maxits=5000)
    return coords, tets, g


def test_macrospin_precession_and_damping_closed_form():
    """phi = phi0 + g' H t, tan(theta/2) = tan(theta0/2) exp(-alpha g' H t),
    g' = gamma/(1+alpha^2); the global error shrinks with the tolerance."""
    alpha, H = 0.1, 0.1
    mat = material(alpha=alpha)
    _, _, g = _magnet(1, 1, 1, 2.5, mat, do_demag=False)
    th0, ph0 = 1.2, 0.3
    m0 = [math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)]
    gp = mat.su_llg_gamma_G / (1 + alpha ** 2)
    errs = {}
    for tol in (1e-7, 1e-10):
        g.set_field("m", np.tile(m0, (g.n, 1)))
        g.set_field("H_ext", np.tile([0.0, 0.0, H], (g.n, 1)))
        g.ts_init(0.0, rtol=tol, atol=tol)
        worst = 0.0
        for T in (20.0, 100.0):
            assert g.advance(T) == pytest.approx(T, abs=1e-12)
            m = g.get_field("m")
            th = 2 * math.atan(math.tan(th0 / 2) * math.exp(-alpha * gp * H * T))
            ph = ph0 + gp * H * T
            ref = [math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)]
            worst = max(worst, np.abs(m - ref).max())
        errs[tol] = worst
        assert g.stat("ts_steps") > 10 and g.stat("ts_order") == 2
    assert errs[1e-10] < 1e-6
    assert errs[1e-10] < errs[1e-7] / 30


def test_trajectory_matches_oracle_rhs_integrated_by_scipy():
    mat = material(alpha=0.3)
    coords, tets, g = _magnet(6, 5, 3, 2.5, mat)
    x, y = coords[:, 0], coords[:, 1]
    m0 = np.stack([np.ones_like(x), 0.4 * np.sin(x / 7.0), 0.3 * np.cos(y / 5.0)], axis=1)
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    Hx = np.tile([0.0, 0.02, 0.0], (len(coords), 1))
    g.set_field("m", m0)
    g.set_field("H_ext", Hx)
    g.ts_init(0.0, rtol=1e-9, atol=1e-9)
    orc = oracle.OracleMesh(coords, tets)
    P = orc_params(mat)
    state = {"p1": None, "p2": None}

    def f(t, yv):
        r = orc.fields(P, yv.reshape(-1, 3), H_ext=Hx, rtol=1e-13, phi1=state["p1"],
                       phi2=state["p2"])
        state["p1"], state["p2"] = r["phi1"], r["phi2"]
        return r["dmdt"].reshape(-1)

    T = 10.0
    sol = solve_ivp(f, (0.0, T), m0.reshape(-1), method="DOP853", rtol=1e-10, atol=1e-12)
    assert sol.success
    g.advance(T)
    assert g.stat("cg_unconverged") == 0
    m = g.get_field("m")
    ref = sol.y[:, -1].reshape(-1, 3)
    assert np.abs(m - ref).max() < 2e-6
    V = orc.volumes()
    avg = lambda a: (a * V[:, None]).sum(0) / V.sum()
    assert np.abs(avg(m) - avg(ref)).max() < 1e-6


def test_bdf_order_from_step_counts():
    """Local-error control at order q takes h ~ tol^(1/(q+1)), so cutting the
    tolerance 100-fold multiplies the step count by about 100^(1/2) = 10 at
    q = 1 and 100^(1/3) = 4.6 at q = 2; the global error of the macrospin
    closed form shrinks with the tolerance at the matching rate."""
    alpha, H = 0.1, 0.1
    mat = material(alpha=alpha)
    _, _, g = _magnet(1, 1, 1, 2.5, mat, do_demag=False)
    th0, ph0 = 1.2, 0.3
    m0 = [math.sin(th0) * math.cos(ph0), math.sin(th0) * math.sin(ph0), math.cos(th0)]
    gp = mat.su_llg_gamma_G / (1 + alpha ** 2)
    T = 200.0
    th = 2 * math.atan(math.tan(th0 / 2) * math.exp(-alpha * gp * H * T))
    ph = ph0 + gp * H * T
    ref = [math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)]
    ratio, err = {}, {}
    for q in (1, 2):
        steps = {}
        for tol in (1e-6, 1e-8):
            g.set_field("m", np.tile(m0, (g.n, 1)))
            g.set_field("H_ext", np.tile([0.0, 0.0, H], (g.n, 1)))
            g.ts_init(0.0, rtol=tol, atol=tol, max_order=q)
            g.advance(T)
            steps[tol] = g.stat("ts_steps")
            err[(q, tol)] = np.abs(g.get_field("m") - ref).max()
            assert g.stat("ts_order") == q
        ratio[q] = steps[1e-8] / steps[1e-6]
    assert 7.0 < ratio[1] < 14.0, ratio
    assert 3.3 < ratio[2] < 6.5, ratio
    # the global error scales like tol^(q/(q+1)): 100^(1/2) = 10 at q = 1 and
    # 100^(2/3) = 21.5 at q = 2
    assert err[(1, 1e-8)] < err[(1, 1e-6)] / 5
    assert err[(2, 1e-8)] < err[(2, 1e-6)] / 10


def test_stop_semantics_and_statistics():
    mat = material(alpha=0.5)
    _, _, g = _magnet(4, 3, 2, 2.5, mat, rtol=1e-8)
    g.set_field("m", random_m(g.n, 3) * 0.2 + np.array([1.0, 0.0, 0.0]))
    g.ts_init(0.0, rtol=1e-6, atol=1e-6)
    t = g.advance(5.0, exact_tstop=True)
    assert t == pytest.approx(5.0, abs=1e-12)
    s0 = g.stat("ts_steps")
    t2 = g.advance(1e6, max_steps=3)
    assert g.stat("ts_steps") - s0 == 3 and 5.0 < t2 < 1e6
    # non-exact stop: interpolated back to the requested time
    t3 = g.advance(t2 + 0.37, exact_tstop=False)
    assert t3 == pytest.approx(t2 + 0.37, abs=1e-12)
    assert g.stat("ts_max_dm") > 0
    for k in ("ts_rhs", "ts_newton", "ts_linear"):
        assert g.stat(k) > 0
    # orders above 5 and Krylov bases above 500 are refused, not clamped
    with pytest.raises(Exception, match="max_order"):
        g.ts_init(0.0, rtol=1e-6, atol=1e-6, max_order=6)
    with pytest.raises(Exception, match="krylov_max"):
        g.ts_init(0.0, rtol=1e-6, atol=1e-6, krylov_max=501)


def test_relaxation_decreases_energy_and_converges():
    mat = material(alpha=0.5)
    coords, tets, g = _magnet(6, 5, 3, 2.5, mat, rtol=1e-8)
    g.set_field("m", np.tile([1.0, 1.0, 0.3], (g.n, 1)) / math.sqrt(2.09))
    g.ts_init(0.0, rtol=1e-6, atol=1e-6)
    V = g.get_field("volumes")

    def energy():
        g.execute()
        m = g.get_field("m")
        return -(V[:, None] * m * (0.5 * (g.get_field("H_exch") + g.get_field("H_demag")))).sum()

    e = [energy()]
    t = 0.0
    for _ in range(6):
        t = g.advance(t + 50.0)
        e.append(energy())
    assert all(b <= a + 1e-12 * abs(a) for a, b in zip(e, e[1:]))
    # relaxed along the long (x) axis of the 15 x 12.5 x 7.5 nm particle
    mx = (g.get_field("m")[:, 0] * V).sum() / V.sum()
    assert mx > 0.98


def test_psetup_psolve_blocks_match_the_jacobian_diagonal():
    """psolve with "pc_full_solve" 0 applies only the 3x3 node blocks of
    M = I - gamma J (the preconditioner of the full solve)."""
    mat = material(alpha=0.2)
    coords, tets, g = _magnet(2, 2, 1, 2.5, mat)
    g.tune("pc_full_solve", 0)
    m = random_m(g.n, 31)
    H = np.tile([0.03, -0.01, 0.02], (g.n, 1))
    g.set_field("H_ext", H)
    g.rhs(m.reshape(-1))  # H_total at m
    Htot = g.get_field("H_total")
    gamma = 0.7
    g.psetup(m.reshape(-1), gamma)
    r = random_m(g.n, 32)
    z = g.psolve(r.reshape(-1), gamma).reshape(-1, 3)
    orc = oracle.OracleMesh(coords, tets, build_bem=False)
    P = orc_params(mat)
    for i in range(g.n):
        J = np.zeros((3, 3))
        for c in range(3):
            v = np.zeros((g.n, 3))
            v[i, c] = 1.0
            J[:, c] = orc.jtimes(P, m, Htot, v)[i]
        M = np.eye(3) - gamma * J
        assert np.allclose(M @ z[i], r[i], rtol=1e-12, atol=1e-12)
    with pytest.raises(Exception):
        g.psolve(r.reshape(-1), gamma, lr=2)


@pytest.mark.parametrize("precision", ["mixed", "mixed24"])
def test_reduced_storage_trajectory_close_to_fp64(precision):
    """<m>(t) with the potential operator in FP32 (24 bits) against FP64 over
    50 ps of precession and relaxation at the working tolerance 1e-6: within
    the 1e-5 of the precision target."""
    mat = material(alpha=0.1)
    coords, tets = kuhn_cuboid(10, 6, 3, 2.5)
    n = len(coords)
    m0 = random_m(n, 12) * 0.3 + np.array([1.0, 0.5, 0.0])
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    avg = {}
    for p in ("fp64", precision):
        g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=p, potential="dense",
                         nbc_rtol=1e-10, dbc_rtol=1e-10)
        g.set_field("H_ext", np.tile([0.0, 0.02, 0.01], (n, 1)))
        g.set_field("m", m0)
        g.ts_init(0.0, rtol=1e-6, atol=1e-6)
        V = g.get_field("volumes")
        series = []
        for t in np.linspace(5.0, 50.0, 10):
            assert g.advance(t) == pytest.approx(t)
            series.append((g.get_field("m") * V[:, None]).sum(0) / V.sum())
        avg[p] = np.array(series)
        g.close()
    assert np.abs(avg[precision] - avg["fp64"]).max() < 1e-5


def test_config3_mean_m_within_1e5_of_fp64_over_1ns():
    """Config 3: the 845-node cuboid (12 x 12 x 4 cubes of 2.5 nm) relaxing from
    seeded i.i.d. Gaussian directions at alpha 0.5, no applied field, 1 ns.
    With both runs at rtol = atol = 1e-8, where the integration error no
    longer hides the storage error, <m>(t) of the FP32-mixed handle stays
    within 1e-5 of FP64 at every 10 ps checkpoint (about 1e-6 measured).  At
    the working tolerance 1e-6 the two runs take different step sequences and
    the deviation is of the order of the integration error (about 1e-5, while
    FP64 alone moves <m> by 3e-4 between rtol 1e-6 and 1e-8)."""
    mat = material(alpha=0.5)
    coords, tets = kuhn_cuboid(12, 12, 4, 2.5)
    n = len(coords)
    assert n == 845
    m0 = random_m(n, 1)
    avg = {}
    for p in ("fp64", "mixed"):
        g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=p, potential="dense",
                         nbc_rtol=1e-10, dbc_rtol=1e-10)
        g.set_field("m", m0)
        g.ts_init(0.0, rtol=1e-8, atol=1e-8)
        V = g.get_field("volumes")
        series = []
        for t in np.arange(10.0, 1000.0 + 1e-9, 10.0):
            assert g.advance(t) == pytest.approx(t)
            series.append((g.get_field("m") * V[:, None]).sum(0) / V.sum())
        avg[p] = np.array(series)
        g.close()
    assert np.abs(avg["mixed"] - avg["fp64"]).max() < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_persistent_gmres_matches_host_driven_gmres(precision):
    """The one-launch GMRES and the host-driven Arnoldi loop solve the same
    Newton systems to the same tolerance: the trajectories agree to far below
    the integration tolerance and take the same numbers of steps."""
    mat = material(alpha=0.1)
    coords, tets = kuhn_cuboid(10, 6, 3, 2.5)
    g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=precision, potential="dense")
    n = len(coords)
    m0 = random_m(n, 11)
    g.set_field("H_ext", np.tile([0.0, 0.02, 0.01], (n, 1)))
    res = {}
    # host-driven; persistent with the 3x3 block-Jacobi preconditioner; and
    # persistent with two local Schwarz sweeps (fewer Krylov iterations)
    for key, dev, sweeps in (("host", 0, 0), ("dev", 1, 0), ("dev_sweeps", 1, 2)):
        g.tune("gmres_device", dev)
        g.tune("gmres_pc_sweeps", sweeps)
        g.set_field("m", m0)
        g.ts_init(0.0, rtol=1e-7, atol=1e-7)
        assert g.advance(20.0) == pytest.approx(20.0)
        res[key] = (g.get_field("m"),
                    {k: g.stat("ts_" + k) for k in ("steps", "newton", "linear")})
    s0 = res["host"][1]
    assert s0["steps"] > 20
    for key in ("dev", "dev_sweeps"):
        s1 = res[key][1]
        assert abs(s0["steps"] - s1["steps"]) <= max(2, 0.02 * s0["steps"]), key
        assert np.abs(res["host"][0] - res[key][0]).max() < 1e-6, key
    assert abs(s0["linear"] - res["dev"][1]["linear"]) <= 0.05 * s0["linear"] + 5
    assert res["dev_sweeps"][1]["linear"] < s0["linear"]


@pytest.mark.parametrize("device", [1, 0])
def test_orders_up_to_five_against_scipy(device):
    """max_order 5 (order selection) on the persistent and the host-driven
    GMRES: the trajectory matches the oracle's rhs integrated by SciPy, orders
    3 to 5 are taken, and at rtol 1e-9 it needs fewer steps than order 2."""
    mat = material(alpha=0.3)
    coords, tets, g = _magnet(6, 5, 3, 2.5, mat)
    x, y = coords[:, 0], coords[:, 1]
    m0 = np.stack([np.ones_like(x), 0.4 * np.sin(x / 7.0), 0.3 * np.cos(y / 5.0)], axis=1)
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    Hx = np.tile([0.0, 0.02, 0.0], (len(coords), 1))
    g.set_field("H_ext", Hx)
    g.tune("gmres_device", device)
    orc = oracle.OracleMesh(coords, tets)
    P = orc_params(mat)
    state = {"p1": None, "p2": None}

    def f(t, yv):
        r = orc.fields(P, yv.reshape(-1, 3), H_ext=Hx, rtol=1e-13, phi1=state["p1"],
                       phi2=state["p2"])
        state["p1"], state["p2"] = r["phi1"], r["phi2"]
        return r["dmdt"].reshape(-1)

    T = 10.0
    sol = solve_ivp(f, (0.0, T), m0.reshape(-1), method="DOP853", rtol=1e-10, atol=1e-12)
    ref = sol.y[:, -1].reshape(-1, 3)
    steps = {}
    for q in (2, 5):
        g.set_field("m", m0)
        g.ts_init(0.0, rtol=1e-9, atol=1e-9, max_order=q)
        assert g.advance(T) == pytest.approx(T, abs=1e-12)
        assert np.abs(g.get_field("m") - ref).max() < 2e-6, q
        steps[q] = g.stat("ts_steps")
        by_order = [g.stat("ts_order_steps_%d" % k) for k in range(1, 6)]
        assert sum(by_order) == steps[q]
        if q == 5:
            assert sum(by_order[2:]) > 0.5 * steps[q], by_order
    assert steps[5] < 0.8 * steps[2], steps


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_one_exchange_gmres_matches_two_exchange_gmres(precision):
    """The persistent GMRES's one-exchange Arnoldi steps (second Gram-Schmidt
    pass delayed to the next step, Newton-basis directions, look-ahead stop)
    and its two-exchange steps build the same Krylov spaces.  Whether every
    step takes two exchanges (0), or the first 1, 2, 8 or all of them (32)
    take one, switching to two-exchange steps mid-solve, the trajectories agree
    far below the integration tolerance, with the same step counts and nearly
    the same Krylov iteration counts."""
    # a smooth state, damped, at rtol 1e-5: steps long enough for up to about
    # 20 Krylov iterations per Newton solve
    mat = material(alpha=0.5)
    coords, tets = kuhn_cuboid(10, 6, 3, 2.5)
    g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=precision, potential="dense")
    n = len(coords)
    x, y = coords[:, 0], coords[:, 1]
    m0 = np.stack([np.ones_like(x), 0.3 * np.sin(x / 7.0), 0.2 * np.cos(y / 5.0)], axis=1)
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    g.set_field("H_ext", np.tile([0.0, 0.02, 0.01], (n, 1)))
    res = {}
    for steps in (0, 1, 2, 8, 32):
        g.tune("gmres_onex_steps", steps)
        g.set_field("m", m0)
        g.ts_init(0.0, rtol=1e-5, atol=1e-5)
        assert g.advance(100.0) == pytest.approx(100.0)
        res[steps] = (g.get_field("m"), {k: g.stat("ts_" + k) for k in (
            "steps", "newton", "linear", "linear_max")})
    s0 = res[0][1]
    assert s0["steps"] > 10 and s0["linear_max"] >= 10, s0
    for steps in (1, 2, 8, 32):
        s1 = res[steps][1]
        assert abs(s1["steps"] - s0["steps"]) <= max(2, 0.02 * s0["steps"]), (steps, s1, s0)
        assert abs(s1["linear"] - s0["linear"]) <= 0.05 * s0["linear"] + 5, (steps, s1, s0)
        assert np.abs(res[steps][0] - res[0][0]).max() < 1e-6, steps
    with pytest.raises(Exception):
        g.tune("gmres_onex_steps", 33)
    g.close()
