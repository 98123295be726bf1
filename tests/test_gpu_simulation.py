"""The nmag-compatible Simulation front end driving the HIP path: SI in and out,
set_m normalisation, field probes against the oracle, averages, relax."""
import math

import numpy as np
import pytest

import oracle
from helpers import PERMALLOY, orc_params, rel_err
from nmag import SI, MagMaterial, Simulation, at, every, kuhn_cuboid, uniaxial_anisotropy

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _in_tmp(tmp_path, monkeypatch):
    # relax saves <name>_dat.ndt / _dat.npz in the working directory
    monkeypatch.chdir(tmp_path)


def _sim(nx=6, ny=5, nz=3, h=2.5, alpha=0.5, **mat_kw):
    mat = MagMaterial("Py", llg_damping=alpha, **dict(PERMALLOY, **mat_kw))
    sim = Simulation(ksp_tolerances={"NBC.rtol": 1e-10, "DBC.rtol": 1e-10})
    coords, tets = kuhn_cuboid(nx, ny, nz, h)
    sim.load_mesh((coords, tets), [("Py", mat)], unit_length=SI(1e-9, "m"))
    return sim, mat, coords, tets


def test_fields_in_si_match_oracle():
    sim, mat, coords, tets = _sim()
    sim.set_m([2.0, 1.0, 0.5])  # auto-normalised
    m = sim.get_subfield("m_Py")
    assert np.allclose(np.linalg.norm(m, axis=1), 1.0)
    sim.set_H_ext([0.0, 0.0, 1.0], unit=SI(1e5, "A/m"))
    orc = oracle.OracleMesh(coords, tets)
    H = np.tile([0.0, 0.0, 0.1], (len(coords), 1))
    ref = orc.fields(orc_params(mat), m, H_ext=H, rtol=1e-13)
    assert rel_err(sim.get_subfield("H_demag") * 1e-6, ref["H_demag"]) < 1e-9
    assert rel_err(sim.get_subfield("H_total_Py") * 1e-6, ref["H_total"]) < 1e-9
    assert rel_err(sim.get_subfield("dmdt_Py") * 1e-12, ref["dmdt"]) < 1e-9
    assert np.allclose(sim.get_subfield("M_Py"), 8.6e5 * m)
    avg = sim.get_subfield_average_siv("m", "Py")
    assert np.allclose(avg, m[0], rtol=1e-12)
    assert sim.get_subfield_average_siv("H_ext", None)[2] == pytest.approx(1e5)


def test_advance_time_reaches_target_and_counts_steps():
    sim, _, _, _ = _sim()
    sim.set_m(lambda p: [1.0, math.sin(p[0] * 1e8), 0.2])
    t = sim.advance_time(SI(5e-12, "s"))
    assert t.value == pytest.approx(5e-12, rel=1e-12)
    assert sim.clock.step > 0 and sim.max_dm_dt > 0


def test_relax_elongated_particle_along_long_axis():
    sim, _, _, _ = _sim()
    sim.set_m([1.0, 1.0, 0.3])
    sim.relax(convergence_check=every("step", 5))
    assert sim.max_dm_dt < sim.stopping_dm_dt
    assert sim.get_subfield_average_siv("m", "Py")[0] > 0.98


def test_uniaxial_anisotropy_through_the_front_end():
    anis = uniaxial_anisotropy([0, 0, 1], SI(4e4, "J/m^3"))
    sim, mat, coords, tets = _sim(anisotropy=anis)
    sim.set_m([1.0, 0.0, 1.0])
    Ha = sim.get_subfield("H_anis_Py")
    m = sim.get_subfield("m_Py")
    u = m[:, 2]
    expect = (2 * 4e4 / (4e-7 * math.pi * 8.6e5)) * u[:, None] * (np.array([0, 0, 1.0]) - u[:, None] * m)
    assert np.allclose(Ha, expect, rtol=1e-10, atol=1e-6)


def test_saved_averages_fields_and_restart_on_the_device():
    """relax saving averages every 20 steps and at the end of the stage, and
    the fields at the end, then a restart round trip; the saved averages are
    checked against the device fields."""
    from nmag import datafile, get_subfield_from_h5file
    sim, _, _, tets = _sim()
    sim.set_m([1.0, 1.0, 0.3])
    sim.relax(save=[("averages", every("step", 20) | at("stage_end")),
                    ("fields", at("stage_end"))])
    names, units, rows = datafile.read_ndt(sim.name + "_dat.ndt")
    col = {n: i for i, n in enumerate(names)}
    last = rows[-1]
    assert int(last[col["step"]]) == sim.clock.step
    assert [int(r[col["step"]]) % 20 for r in rows[:-1]] == [0] * (len(rows) - 1)
    m_avg = sim.get_subfield_average_siv("m", "Py")
    Hd_avg = sim.get_subfield_average_siv("H_demag")
    for k in range(3):
        assert last[col["m_Py_%d" % k]] == pytest.approx(m_avg[k], rel=1e-12, abs=1e-14)
        assert last[col["H_demag_%d" % k]] == pytest.approx(Hd_avg[k], rel=1e-9, abs=1e-6)
    m = sim.get_subfield("m_Py")
    edges = datafile.mesh_edges(tets)
    assert last[col["maxangle_m_Py"]] == pytest.approx(datafile.max_angle_deg(m, edges))
    assert np.allclose(get_subfield_from_h5file(sim.name + "_dat.npz", "m_Py"), m)
    sim.save_restart_file()
    sim2 = Simulation(name="second", ksp_tolerances={"NBC.rtol": 1e-10, "DBC.rtol": 1e-10})
    coords, _ = kuhn_cuboid(6, 5, 3, 2.5)
    sim2.load_mesh((coords, tets), [("Py", MagMaterial("Py", llg_damping=0.5, **PERMALLOY))],
                   unit_length=SI(1e-9, "m"))
    sim2.load_restart_file(sim.get_restart_file_name())
    # loading goes through set_m, which renormalises (the integrated m is
    # unit length only to the integrator's tolerance)
    m_unit = m / np.linalg.norm(m, axis=1, keepdims=True)
    assert np.abs(np.linalg.norm(m, axis=1) - 1).max() < 1e-5
    assert np.allclose(sim2.get_subfield("m_Py"), m_unit, rtol=1e-14, atol=1e-15)
    assert sim2.clock.step == sim.clock.step
    sim.set_m(m)
    assert rel_err(sim2.get_subfield("H_demag"), sim.get_subfield("H_demag")) < 1e-9


def test_phi_bem_hmatrix_setup_selects_the_compressed_operator():
    """Simulation(phi_BEM=HMatrixSetup(eps=...)) (interface/nmag/main.py:964-981)
    runs the H-matrix route: H_demag within the requested tolerance of the
    oracle's."""
    from nmag import HMatrixSetup
    mat = MagMaterial("Py", llg_damping=0.5, **PERMALLOY)
    coords, tets = kuhn_cuboid(12, 10, 3, 2.5)
    orc = oracle.OracleMesh(coords, tets)
    for eps, bar in ((1e-10, 1e-9), (1e-4, 1e-3)):
        sim = Simulation(phi_BEM=HMatrixSetup(eps=eps))
        sim.load_mesh((coords, tets), [("Py", mat)], unit_length=SI(1e-9, "m"))
        sim.set_m(lambda p: [1.0, math.sin(p[0] * 2e8), math.cos(p[1] * 3e8)])
        assert sim._magnet.hmatrix_potential
        m = sim.get_subfield("m_Py")
        ref = orc.fields(orc_params(mat), m, rtol=1e-13)
        assert rel_err(sim.get_subfield("H_demag") * 1e-6, ref["H_demag"]) < bar, eps
