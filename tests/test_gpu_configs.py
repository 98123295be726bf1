"""The benchmark configurations at their own sizes.

* The 64-cuboid batch (seed 4, 520-8,832 nodes per magnet, 194,603 nodes in
  all): sampled magnets against the oracle for that magnet alone, and the
  2 ps relaxation window of one integrator over the whole batch, whose
  persistent GMRES grid exceeds 512 workgroups, against each sampled magnet
  integrated alone.
* The differentiation slab (40 x 20 x 2 cubes, per-element K, 16 pulses) at
  the benchmark's checkpoint spacing over a shorter window: gradients against
  central differences of the same forward map, and mixed against FP64.
* The 20k-node slab: <m>(t) of FP32-mixed against FP64 over 50 ps of the
  benchmark's BDF integration.

Each test prints its case before the GPU work starts.
"""
import numpy as np
import pytest

import oracle
from helpers import orc_params, rel_err
from nmag._native import NativeBatch, NativeMagnet
from nmag.adjoint import TrajectoryProblem
from nmag.workloads import cuboid_batch, differentiation_slab, slab

pytestmark = pytest.mark.gpu


def _note(*a):
    print("[case]", *a, flush=True)


def _set(mag, fields):
    for key in ("m", "H_ext", "axis", "K1"):
        mag.set_field(key, fields[key])


def _cat(mags, key):
    return np.concatenate([g["fields"][key] for g in mags])


def _sample(mags):
    """The three smallest magnets and one of median size."""
    sizes = np.array([len(g["coords"]) for g in mags])
    order = np.argsort(sizes, kind="stable")
    return [int(k) for k in order[:3]] + [int(order[len(order) // 2])]


@pytest.fixture(scope="module")
def c4():
    mags = cuboid_batch()
    sizes = [len(g["coords"]) for g in mags]
    assert len(mags) == 64 and sum(sizes) == 194603
    assert min(sizes) == 520 and max(sizes) == 8832
    return mags


def _c4_batch(mags, precision="fp64"):
    return NativeBatch([(g["coords"], g["tets"]) for g in mags],
                       [g["mat"].su_coeffs() for g in mags], precision=precision,
                       nbc_rtol=1e-10, dbc_rtol=1e-10)


def test_c4_batch_fields_match_the_oracle_on_sampled_magnets(c4):
    _note("C4 fields: 64 cuboids, 194603 nodes, fp64")
    b = _c4_batch(c4)
    assert b.dense_potential
    for key in ("m", "H_ext", "axis", "K1"):
        b.set_field(key, _cat(c4, key))
    b.execute("update_dmdt")
    got = {k: b.split(b.get_field(k)) for k in ("H_exch", "H_anis", "H_demag", "dmdt")}
    b.close()
    for k in _sample(c4):
        g = c4[k]
        f = g["fields"]
        _note("oracle for magnet %d (%d nodes)" % (k, len(g["coords"])))
        ref = oracle.OracleMesh(g["coords"], g["tets"]).fields(
            orc_params(g["mat"]), f["m"], H_ext=f["H_ext"], k1=f["K1"], axis=f["axis"],
            rtol=1e-13)
        assert rel_err(got["H_exch"][k], ref["H_exch"]) < 1e-12, k
        assert rel_err(got["H_anis"][k], ref["H_anis"]) < 1e-13, k
        assert rel_err(got["H_demag"][k], ref["H_demag"]) < 1e-10, k
        assert rel_err(got["dmdt"][k], ref["dmdt"]) < 1e-10, k


def test_c4_relax_window_on_a_grid_above_512_workgroups(c4):
    """The relaxation window of the batch benchmark: one BDF integrator over
    all 64 magnets.  Its persistent Newton-GMRES grid (one workgroup per few
    SELL slices of 64 nodes) is larger than two workgroups per CU can hold, so
    it needs the third workgroup per CU that the kernel's LDS budget leaves."""
    t_end, tol = 2.0, 1e-8
    _note("C4 relax window: 64 cuboids, %.1f ps, rtol %g" % (t_end, tol))
    b = _c4_batch(c4)
    for key in ("m", "H_ext", "axis", "K1"):
        b.set_field(key, _cat(c4, key))
    b.ts_init(0.0, rtol=tol, atol=tol)
    grid = b.stat("ts_gmres_grid")
    _note("persistent GMRES grid: %d workgroups" % grid)
    assert grid > 512
    assert abs(b.advance(t_end) - t_end) < 1e-12
    assert b.stat("ts_steps") > 0
    mb = b.split(b.get_field("m"))
    b.close()
    for k in _sample(c4):
        g = c4[k]
        s = NativeMagnet(g["coords"], g["tets"], g["mat"].su_coeffs(), nbc_rtol=1e-10,
                         dbc_rtol=1e-10)
        _set(s, g["fields"])
        s.ts_init(0.0, rtol=tol, atol=tol)
        s.advance(t_end)
        ms = s.get_field("m")
        V = s.get_field("volumes")
        s.close()
        avg_b = (V[:, None] * mb[k]).sum(0) / V.sum()
        avg_s = (V[:, None] * ms).sum(0) / V.sum()
        assert np.max(np.abs(avg_b - avg_s)) < 1e-5, k
        assert np.max(np.abs(mb[k] - ms)) < 1e-4, k


def test_c4_batch_at_the_benchmark_tolerance_completes():
    """The benchmark's own window settings (rtol = atol = 1e-6), mixed storage."""
    mags = cuboid_batch()
    _note("C4 relax window at rtol 1e-6, mixed")
    b = _c4_batch(mags, "mixed")
    for key in ("m", "H_ext", "axis", "K1"):
        b.set_field(key, _cat(mags, key))
    b.ts_init(0.0, rtol=1e-6, atol=1e-6)
    assert abs(b.advance(2.0) - 2.0) < 1e-12
    m = b.get_field("m")
    b.close()
    assert np.all(np.isfinite(m))
    # the integrator state is not renormalised (nor is the reference's); the
    # c3 term pulls |m| back, and at rtol 1e-6 from random directions it
    # stays within a few 1e-3 of 1
    assert np.max(np.abs(np.linalg.norm(m, axis=1) - 1.0)) < 1e-2


def _c5_problem(precision, n_steps=960, rtol=1e-10):
    coords, tets, mat, K, amps, m0 = differentiation_slab()
    assert len(coords) == 2583 and len(tets) == 9600 and len(amps) == 16
    g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=precision, nbc_rtol=rtol,
                     dbc_rtol=rtol)
    dt = 0.2
    prob = TrajectoryProblem(g, mat.su_coeffs(), coords, tets, m0,
                             np.tile([0.0, 0.0, 1.0], (len(coords), 1)), dt=dt,
                             n_steps=n_steps, n_pulses=len(amps),
                             pulse_len=dt * n_steps / len(amps), checkpoint_every=50)
    return prob, K, mat.su_coeffs()["Ms"], amps


def test_c5_gradients_match_central_differences():
    """192 ps (960 RK4 steps of 0.2 ps, 16 pulses of 12 ps), checkpoints
    every 50 steps as in the 2 ns benchmark; central differences at relative
    step 1e-3 of the same forward map, bar 1e-4."""
    _note("C5 gradients vs FD: 2583 nodes, 960 steps, fp64")
    prob, K, Ms, amps = _c5_problem("fp64")
    prob.run(K, Ms, amps)
    gK, gMs, gA, _ = prob.gradient(amps)
    rel = 1e-3
    h = rel * Ms
    ref = (prob.run(K, Ms + h, amps) - prob.run(K, Ms - h, amps)) / (2 * h)
    assert abs(gMs - ref) <= 1e-4 * abs(ref)
    # pulse amplitudes and per-element K along seeded random directions
    u = np.random.default_rng(11).standard_normal(len(amps)) * np.max(np.abs(amps))
    ref = (prob.run(K, Ms, amps + rel * u) - prob.run(K, Ms, amps - rel * u)) / (2 * rel)
    assert abs(gA @ u - ref) <= 1e-4 * abs(ref)
    d = np.random.default_rng(12).standard_normal(len(K)) * K
    ref = (prob.run(K + rel * d, Ms, amps) - prob.run(K - rel * d, Ms, amps)) / (2 * rel)
    assert abs(gK @ d - ref) <= 1e-4 * abs(ref)
    prob.magnet.close()


def test_c5_mixed_gradient_close_to_the_fp64_adjoint():
    _note("C5 mixed vs fp64 adjoint: 2583 nodes, 960 steps")
    res = {}
    for prec in ("fp64", "mixed"):
        prob, K, Ms, amps = _c5_problem(prec)
        prob.run(K, Ms, amps)
        res[prec] = prob.gradient(amps)[:3]
        prob.magnet.close()
    for a, b in zip(res["mixed"], res["fp64"]):
        a, b = np.atleast_1d(a), np.atleast_1d(b)
        assert np.max(np.abs(a - b)) <= 1e-4 * np.max(np.abs(b))


def test_c2_mixed_mean_m_follows_fp64_over_50_ps():
    """The slab's BDF integration (rtol = atol = 1e-6) from the seeded S-state:
    <m>(t) of FP32 storage within 1e-5 of FP64 at every 5 ps checkpoint."""
    coords, tets, mat, Hext, m0 = slab()
    avgs = {}
    for prec in ("fp64", "mixed"):
        _note("C2 dynamics %s: 19926 nodes, 50 ps" % prec)
        g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=prec, nbc_rtol=1e-6,
                         dbc_rtol=1e-6)
        g.set_field("H_ext", Hext)
        g.set_field("m", m0)
        g.ts_init(0.0, 1e-6, 1e-6)
        V = g.get_field("volumes")
        out, t = [], 0.0
        while t < 50.0 - 1e-9:
            t = g.advance(t + 5.0)
            out.append((g.get_field("m") * V[:, None]).sum(0) / V.sum())
        g.close()
        avgs[prec] = np.array(out)
    assert len(avgs["fp64"]) == 10
    assert np.max(np.abs(avgs["mixed"] - avgs["fp64"])) < 1e-5
