"""The nmag-compatible Simulation front end driving the HIP path: SI in and out,
set_m normalisation, field probes against the oracle, averages, relax."""
import math

import numpy as np
import pytest

import oracle
from helpers import PERMALLOY, orc_params, rel_err
from nmag import SI, MagMaterial, Simulation, every, kuhn_cuboid, uniaxial_anisotropy

pytestmark = pytest.mark.gpu


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
