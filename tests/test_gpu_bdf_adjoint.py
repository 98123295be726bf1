"""Differentiation of the BDF integrator (nsim_bdf_forward / _replay /
_adjoint): the discrete adjoint of the recorded step sequence against central
differences of the replayed trajectory (the same steps, orders and
coefficients, each step's equation solved to 1e-13), for the initial state,
the per-node anisotropy K1, Ms and the pulse amplitudes, at BDF order 2 and
with order selection up to 3; and the replay against the recorded run."""
import numpy as np
import pytest

from helpers import material, random_m
from nmag import kuhn_cuboid
from nmag._native import NativeMagnet

pytestmark = pytest.mark.gpu

AMPS = np.array([0.03, -0.02])
PULSE = dict(pulse_len=5.0, pulse_dir=(0.0, 1.0, 0.3))


def _note(*a):
    print("[case]", *a, flush=True)


def _problem(seed=3):
    coords, tets = kuhn_cuboid(6, 4, 2, 2.5)
    mat = material(alpha=0.2)
    g = NativeMagnet(coords, tets, mat.su_coeffs(), nbc_rtol=1e-13, dbc_rtol=1e-13,
                     nbc_maxits=20000, dbc_maxits=20000)
    rng = np.random.default_rng(seed)
    n = g.n
    m0 = random_m(n, seed) * 0.3 + np.array([1.0, 0.2, 0.0])
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    g.set_field("H_ext", np.tile([0.01, 0.0, -0.02], (n, 1)))
    ax = rng.standard_normal(3)
    g.set_field("axis", np.tile(ax / np.linalg.norm(ax), (n, 1)))
    k1 = rng.uniform(0.01, 0.03, n)
    g.set_field("K1", k1)
    V = g.get_field("volumes")
    return g, mat, m0, k1, V


def _loss(g, V):
    return float((g.get_field("m")[:, 0] * V).sum() / V.sum())


def _mbar(V):
    mb = np.zeros((len(V), 3))
    mb[:, 0] = V / V.sum()
    return mb


def _replay_loss(g, V, m0, amps=AMPS):
    g.set_field("m", m0)
    g.bdf_replay(amps, newton_rtol=1e-13)
    return _loss(g, V)


@pytest.mark.parametrize("max_order", [2, 3])
def test_bdf_adjoint_matches_central_differences(max_order):
    _note("BDF adjoint, max_order", max_order)
    g, mat, m0, k1, V = _problem()
    g.set_field("m", m0)
    K = g.bdf_forward(12.0, rtol=1e-7, atol=1e-7, max_order=max_order, max_steps=5000,
                      amps=AMPS, **PULSE)
    t, hs, q = g.bdf_steps()
    assert K == len(t) > 20 and t[-1] == pytest.approx(12.0, abs=1e-12)
    assert q.max() == max_order
    # no step straddles a pulse edge
    edges = np.array([5.0, 10.0])
    t0 = t - hs
    assert not np.any((t0[:, None] < edges - 1e-12) & (t[:, None] > edges + 1e-12))
    _replay_loss(g, V, m0)  # the adjoint runs over the states of the last replay
    grads = g.bdf_adjoint(_mbar(V), AMPS, lin_rtol=1e-12)
    rng = np.random.default_rng(9)
    checks = []
    # initial state
    u = rng.standard_normal((g.n, 3))
    eps = 1e-5
    fd = (_replay_loss(g, V, m0 + eps * u) - _replay_loss(g, V, m0 - eps * u)) / (2 * eps)
    checks.append(("m0", float((grads["m0"] * u).sum()), fd))
    # K1 per node
    dk = rng.standard_normal(g.n) * 0.01
    g.set_field("K1", k1 + eps * dk)
    lp = _replay_loss(g, V, m0)
    g.set_field("K1", k1 - eps * dk)
    lm = _replay_loss(g, V, m0)
    g.set_field("K1", k1)
    checks.append(("K1", float(grads["K1"] @ dk), (lp - lm) / (2 * eps)))
    # Ms at fixed exchange constant
    c = mat.su_coeffs()
    A = -c["negJ"] * c["mu0"] * c["Ms"] / 2

    def with_ms(ms):
        g.set_material(dict(c, Ms=ms, negJ=-2 * A / (c["mu0"] * ms)))
        return _replay_loss(g, V, m0)

    hms = 1e-6 * c["Ms"]
    fd_ms = (with_ms(c["Ms"] + hms) - with_ms(c["Ms"] - hms)) / (2 * hms)
    g.set_material(c)
    checks.append(("Ms", grads["Ms"], fd_ms))
    # pulse amplitudes
    for w in range(len(AMPS)):
        da = np.zeros_like(AMPS)
        da[w] = 1e-5
        fd_a = (_replay_loss(g, V, m0, AMPS + da) - _replay_loss(g, V, m0, AMPS - da)) / 2e-5
        checks.append(("amp%d" % w, grads["amps"][w], fd_a))
    for name, adj, fd in checks:
        assert abs(adj - fd) <= 2e-5 * abs(fd) + 1e-12, (name, adj, fd)


def test_replay_follows_the_recorded_run():
    """The replay solves every recorded step tightly; the recorded run solved
    them to the Newton tolerance of the integrator: the two agree within the
    integration tolerance, and the replay of unchanged parameters is
    reproducible."""
    g, mat, m0, k1, V = _problem(4)
    g.set_field("m", m0)
    g.bdf_forward(8.0, rtol=1e-7, atol=1e-7, max_order=2, max_steps=5000, amps=AMPS, **PULSE)
    recorded = g.get_field("m")
    a = _replay_loss(g, V, m0)
    replayed = g.get_field("m")
    assert np.abs(replayed - recorded).max() < 1e-5
    assert _replay_loss(g, V, m0) == a
    with pytest.raises(Exception):
        g.bdf_adjoint(_mbar(V), AMPS * 2)  # not the amplitudes of the last replay
