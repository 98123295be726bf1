"""Host logic of the nmag front end: simulation units, material coefficients,
nmesh I/O, the convergence countdown (no GPU)."""
import math

import numpy as np
import pytest

from nmag import SI, MagMaterial, NmagUserError, kuhn_cuboid, read_nmesh, to_su, write_nmesh
from nmag.simulation import Convergence, Simulation, every
from nmag.units import mu0


def test_simulation_units_closed_forms():
    # nmag's su: A=1e-3, kg=1e-27, m=1e-9, s=1e-12 (interface/nmag/main.py:188-196)
    assert to_su(SI(8.6e5, "A/m"), "A/m") == pytest.approx(0.86)
    assert to_su(SI(1.3e-11, "J/m"), "J/m") == pytest.approx(13.0)
    assert to_su(SI(2.211e5, "m/A s"), "m/A s") == pytest.approx(0.2211)
    assert to_su(mu0) == pytest.approx(0.4 * math.pi)
    assert to_su(SI(5e4, "J/m^3"), "J/m^3") == pytest.approx(0.05)
    assert to_su(SI(1e-9, "s"), "s") == pytest.approx(1000.0)
    deg_per_ns = SI(math.pi / 180 / 1e-9, "1/s")
    assert to_su(deg_per_ns, "1/s") == pytest.approx(1.745329e-5, rel=1e-6)
    H = SI(25e-3, "T") / mu0
    assert H.is_compatible_with(SI(1, "A/m"))
    assert H.value == pytest.approx(19894.37, rel=1e-6)
    with pytest.raises(ValueError):
        to_su(SI(1.0, "A/m"), "J/m")


def test_magmaterial_coefficients():
    mat = MagMaterial("Py", Ms=SI(8.6e5, "A/m"), exchange_coupling=SI(1.3e-11, "J/m"),
                      llg_gamma_G=SI(2.211e5, "m/A s"), llg_damping=0.5)
    c = mat.su_coeffs()
    assert c["c1"] == pytest.approx(-0.2211 / 1.25)
    assert c["c2"] == pytest.approx(-0.2211 / 1.25 * 0.5)
    assert c["c3"] == pytest.approx(0.1)
    assert c["negJ"] == pytest.approx(-2 * 13.0 / (0.4 * math.pi * 0.86))
    off = MagMaterial("P2", do_precession=False)
    assert off.su_coeffs()["c1"] == 0.0
    with pytest.raises(NmagUserError):
        MagMaterial("bad", exchange_coupling=SI(-1e-11, "J/m"))


def test_nmesh_roundtrip(tmp_path):
    coords, tets = kuhn_cuboid(3, 2, 2, 2.5)
    p = tmp_path / "box.nmesh"
    write_nmesh(str(p), coords, tets)
    c2, t2, reg = read_nmesh(str(p))
    assert np.allclose(c2, coords) and np.array_equal(t2, tets) and (reg == 1).all()


def test_convergence_needs_two_consecutive_hits():
    c = Convergence()
    assert not c.check(0.5, 1.0)
    assert not c.check(2.0, 1.0)   # resets the countdown
    assert not c.check(0.5, 1.0)
    assert c.check(0.5, 1.0)


def test_unsupported_options_fail_loudly():
    with pytest.raises(NmagUserError):
        Simulation(phi_BEM=object())
    with pytest.raises(NmagUserError):
        Simulation(ksp_tolerances={"XYZ.rtol": 1})
    s = Simulation()
    with pytest.raises(NmagUserError):
        s.magnet
    with pytest.raises(NmagUserError):
        Simulation(timestepper_max_order=6)
    with pytest.raises(NmagUserError):
        Simulation(timestepper_krylov_max=600)
    assert every("step", 5).delta == 5
    with pytest.raises(ValueError):
        every("step", 0)


def test_element_to_node_projection_reduces_to_uniform_and_conserves_energy_weights():
    """Per-element K -> nodal K1 (lumped): a uniform K maps to the same nodal
    value, and sum_i V_i K1_i = sum_e vol_e K_e (volume integral kept)."""
    import numpy as np
    from nmag import kuhn_cuboid
    from nmag.adjoint import element_to_node
    coords, tets = kuhn_cuboid(4, 3, 2, 2.5)
    P = element_to_node(coords, tets)
    assert np.allclose(P @ np.full(len(tets), 3.7), 3.7, rtol=0, atol=1e-13)
    x = coords[tets]
    vol = np.abs(np.einsum("ij,ij->i", x[:, 1] - x[:, 0],
                           np.cross(x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]))) / 6
    V = np.bincount(tets.reshape(-1), weights=np.repeat(vol / 4, 4), minlength=len(coords))
    K = np.random.default_rng(1).uniform(0, 1, len(tets))
    assert abs(V @ (P @ K) - vol @ K) <= 1e-12 * (vol @ K)


def test_hmatrix_setup_keeps_the_reference_interface():
    """HMatrixSetup (interface/nmag/hlib.py:28-165): HCA2 only, its defaults,
    user values over set_default_params, and the mapping to the compressed
    route (eps per block)."""
    from nmag import HMatrixSetup, default_hmatrix_setup
    p = default_hmatrix_setup.get_params()
    assert p["eta"] == 2.0 and p["eps"] == 1e-3 and p["eps_aca"] == 1e-7 and p["nmin"] == 30
    assert default_hmatrix_setup.compress_options() == ("hmatrix", 1e-3, 0.0)
    h = HMatrixSetup(eps=1e-7)
    h.set_default_params(eps=1e-2, nmin=10)
    assert h.get_params()["eps"] == 1e-7 and h.get_params()["nmin"] == 10
    assert h.compress_options() == ("hmatrix", 1e-7, 0.0)
    assert HMatrixSetup(eta=2.0).compress_options() == ("hmatrix", 1e-3, 2.0)
    with pytest.raises(NmagUserError):
        HMatrixSetup(algorithm="ACA")
    with pytest.raises(NmagUserError):
        HMatrixSetup(foo=1)
    with pytest.raises(NmagUserError):
        HMatrixSetup(eta=-1.0).compress_options()
    s = Simulation(phi_BEM=HMatrixSetup(eps=1e-7))
    assert (s._potential, s._compress_eps, s._compress_eta) == ("hmatrix", 1e-7, 0.0)
