"""Host logic of the nmag front end's data files and stage loop: event
schedules, the ndt table, the field store, restart files, and relax /
hysteresis with save and do actions.  The device is replaced by a stand-in
magnet with simple damped dynamics, so these run without a GPU."""
import math
import os

import numpy as np
import pytest

import nmag
from nmag import SI, MagMaterial, Simulation, at, every, kuhn_cuboid, never
from nmag import datafile
from nmag import simulation as simmod
from nmag.when import next_deltas

CLOCK = dict(step=0, stage_step=0, stage=1, time=0.0, stage_time=0.0, real_time=0.0,
             convergence=False, stage_end=False, zero_stage_step=0, zero_stage_time=0.0)


def _walk(w, ident, start, n=20, tols=None):
    c = dict(CLOCK)
    out, t = [], start
    for _ in range(n):
        c[ident] = t
        nxt = w.next_time(ident, c, tols=tols)
        out.append(nxt)
        if isinstance(nxt, bool):
            break
        t = nxt
    return out


def test_schedule_sequences():
    # the reference's own demonstration cases (interface/nsim/when.py:258-290)
    w = every("step", 2, last=21) & every("step", 4, first=10) | at("step", 15)
    assert _walk(w, "step", 0) == [10, 14, 15, 18, False]
    w = every("time", SI(100e-12, "s")) | every("time", SI(30e-12, "s"))
    got = _walk(w, "time", 0.0, n=8, tols={"time": 1e-20})
    assert np.allclose(got, [30e-12, 60e-12, 90e-12, 100e-12, 120e-12, 150e-12, 180e-12,
                             200e-12], rtol=1e-12)
    # first / last bounds, every(10, 'step') argument order
    assert _walk(every(5, "step", first=3, last=20), "step", 0) == [3, 8, 13, 18, False]
    c = dict(CLOCK, step=8)
    assert every("step", 5, first=3).match_time(c)
    assert not every("step", 5, first=3).match_time(dict(c, step=9))
    # flags: at('convergence') is not a step event
    assert at("convergence").next_time("step", c) is True
    assert at("convergence").match_time(dict(c, convergence=True))
    assert never.next_time("step", c) is False and not never.match_time(c)
    # steps to the next event from step 8: every 5 -> 2, every 3 -> 1
    assert next_deltas(every("step", 5), c)[0] == 2
    assert next_deltas(every("step", 3), c, suggest=next_deltas(every("step", 5), c))[0] == 1
    with pytest.raises(ValueError):
        every("step", 5, first=10, last=10)


def test_dense_unit_strings():
    # the reference's docstring example (interface/nsim/si_units/lib.py:309-317)
    assert datafile.dens_str(SI(1e6, "A/m")) == "<1e+06A/m>"
    assert datafile.dens_str(SI(1, "A/m")) == "<A/m>"
    assert datafile.dens_str(SI(1)) == "<1>"
    assert datafile.dens_str(SI(1, "s")) == "<s>"
    assert datafile.dens_str(SI(1, "1/s")) == "</s>"
    # J/m^3 = kg m^-1 s^-2, base order m, kg, s, A
    assert datafile.dens_str(SI(1, "J/m^3")) == "<kg/ms^2>"
    assert datafile.dens_str("<deg>") == "<deg>"


def test_ndt_layout(tmp_path):
    p = str(tmp_path / "x_dat.ndt")
    names = ["time", "step", "m_Py", "localtime"]
    units = [SI(1, "s"), SI(1), SI(1), "<>"]
    datafile.write_ndt_row(p, names, [1.5e-12, 3, [0.6, 0.8, 0.0], "2026/01/01-00:00:00"],
                           units, header=True)
    datafile.write_ndt_row(p, names, [2.5e-12, 7, [1.0, 0.0, 0.0], "2026/01/01-00:00:01"],
                           units, header=False)
    lines = open(p).read().splitlines()
    cols = ["time", "step", "m_Py_0", "m_Py_1", "m_Py_2", "localtime"]
    assert lines[0] == "#" + "".join("%-25s " % c for c in cols)
    assert lines[1] == "#" + "".join("%-25s " % u for u in
                                     ["<s>", "<1>", "<1>", "<1>", "<1>", "<>"])
    assert lines[2] == ("% 25.13g " % 1.5e-12 + "% 25d " % 3 + "% 25.13g " % 0.6
                        + "% 25.13g " % 0.8 + "% 25.13g " % 0.0 + "% 25s " % "2026/01/01-00:00:00")
    names_r, units_r, rows = datafile.read_ndt(p)
    assert names_r == cols and units_r[0] == "<s>"
    assert rows[1][:5] == [2.5e-12, 7.0, 1.0, 0.0, 0.0]


def test_max_angle_matches_brute_force():
    coords, tets = kuhn_cuboid(2, 2, 1, 1.0)
    rng = np.random.default_rng(3)
    v = rng.normal(size=(len(coords), 3))
    best = 0.0
    for t in tets:
        for a in range(4):
            for b in range(4):
                if a != b:
                    x, y = v[t[a]], v[t[b]]
                    c = x @ y / np.linalg.norm(x) / np.linalg.norm(y)
                    best = max(best, math.degrees(math.acos(max(-1.0, min(1.0, c)))))
    assert datafile.max_angle_deg(v, datafile.mesh_edges(tets)) == pytest.approx(best, 1e-12)
    # uniform field: no angle; Kuhn 2x2x1 cuboid has 18 nodes and 6*4 tets
    assert datafile.max_angle_deg(np.tile([0, 0, 1.0], (len(coords), 1)),
                                  datafile.mesh_edges(tets)) == 0.0


class StandInMagnet:
    """Damped rotation of m towards +x with fixed steps of 0.5 ps: enough of
    NativeMagnet's interface for the front end's stage loop."""

    def __init__(self, coords, tets, *args, **kw):
        self.n = len(coords)
        self.f = {k: np.zeros((self.n, 3)) for k in
                  ("m", "H_exch", "H_demag", "H_anis", "H_total", "H_ext", "dmdt")}
        self.f["m"][:, 0] = 1.0
        self.f.update(phi=np.zeros(self.n), rho=np.zeros(self.n),
                      volumes=np.full(self.n, 1.0 / self.n))
        self.t, self.steps, self.h, self.max_dm = 0.0, 0, 0.5, 0.0
        self.evals = 0

    def set_field(self, name, data):
        self.f[name] = np.array(data, dtype=np.float64).reshape(self.f.get(name, data).shape
                                                               if name in self.f else -1)

    def get_field(self, name):
        return self.f[name].copy()

    def execute(self, program="update_dmdt"):
        self.evals += 1
        self.f["dmdt"] = -0.2 * (self.f["m"] - [1.0, 0.0, 0.0])

    def ts_init(self, t0=0.0, rtol=1e-6, atol=1e-6, max_order=2, krylov_max=300):
        self.t = t0

    def advance(self, t_final, max_steps=-1, exact_tstop=True):
        m0 = self.f["m"].copy()
        k = 0
        while self.t < t_final and (max_steps < 0 or k < max_steps):
            dt = min(self.h, t_final - self.t) if exact_tstop else self.h
            m = self.f["m"]
            m = np.array([1.0, 0, 0]) + (m - [1.0, 0, 0]) * math.exp(-0.2 * dt)
            self.f["m"] = m / np.linalg.norm(m, axis=1, keepdims=True)
            self.t = t_final if exact_tstop and dt < self.h else self.t + dt
            k += 1
        self.steps += k
        self.max_dm = float(np.abs(self.f["m"] - m0).max())
        return self.t

    def stat(self, name):
        return {"ts_steps": self.steps, "ts_h": self.h, "ts_max_dm": self.max_dm}[name]


@pytest.fixture
def standin(monkeypatch, tmp_path):
    monkeypatch.setattr(simmod._native, "NativeMagnet", StandInMagnet)
    monkeypatch.chdir(tmp_path)
    monkeypatch.setattr(simmod.sys, "argv", ["run.py"])
    return tmp_path


def _sim(name="run"):
    sim = Simulation(name=name)
    coords, tets = kuhn_cuboid(3, 2, 1, 2.5)
    sim.load_mesh((coords, tets), [("Py", MagMaterial("Py"))], unit_length=SI(1e-9, "m"))
    sim.set_m([0.0, 1.0, 0.2])
    return sim


def test_relax_default_saves_averages_and_fields_at_stage_end(standin):
    sim = _sim()
    sim.relax()
    names, units, rows = datafile.read_ndt("run_dat.ndt")
    assert len(rows) == 1
    col = {n: i for i, n in enumerate(names)}
    for c in ("time", "id", "step", "last_step_dt", "stage_time", "stage_step", "stage",
              "m_Py_0", "M_Py_0", "H_demag_0", "H_ext_2", "phi", "rho", "E_total_Py",
              "maxangle_m_Py", "localtime", "unixtime"):
        assert c in col, c
    assert units[col["M_Py_0"]] == "<A/m>" and units[col["time"]] == "<s>"
    assert rows[0][col["id"]] == 0 and rows[0][col["stage"]] == 1
    assert rows[0][col["step"]] == sim.clock.step and sim.clock.step > 0
    assert rows[0][col["m_Py_0"]] == pytest.approx(sim.get_subfield_average_siv("m", "Py")[0])
    assert rows[0][col["M_Py_0"]] == pytest.approx(8.6e5 * rows[0][col["m_Py_0"]])
    m = nmag.get_subfield_from_h5file("run_dat.npz", "m_Py")
    assert np.allclose(m, sim.get_subfield("m_Py"))
    assert nmag.get_subfield_positions_from_h5file("run_dat.npz", "m_Py").shape == (24, 3)
    assert datafile.read_meta("run_dat.npz")[0]["step"] == sim.clock.step
    # a new run under the same name refuses to overwrite the data
    with pytest.raises(nmag.NmagUserError):
        Simulation(name="run")


def test_step_and_time_schedules_drive_the_saves(standin):
    sim = _sim()
    sim.relax(save=[("averages", every("step", 7)),
                    ("field_m", every("time", SI(4e-12, "s")) | at("convergence"))])
    names, _, rows = datafile.read_ndt("run_dat.ndt")
    col = {n: i for i, n in enumerate(names)}
    steps = [int(r[col["step"]]) for r in rows]
    t_ps = [r[col["time"]] / 1e-12 for r in rows]
    # saving m writes an averages row too (save_data with fields), so every
    # row is on a multiple of 7 steps or of 4 ps, or is the converged end
    assert steps[0] == 0
    assert all(s % 7 == 0 or abs(t / 4 - round(t / 4)) < 1e-9
               for s, t in zip(steps[:-1], t_ps[:-1]))
    assert [s for s in range(0, steps[-1] + 1, 7)] == [s for s in steps if s % 7 == 0]
    assert [r[col["id"]] for r in rows] == list(range(len(rows)))
    assert len(steps) == len(set(steps))   # no step saved twice
    meta = datafile.read_meta("run_dat.npz")
    times = [m_["time"] for m_ in meta]
    # every 4 ps exactly (exact stops), and a last save at convergence
    assert np.allclose(np.array(times[:-1]) / 4e-12, np.round(np.array(times[:-1]) / 4e-12),
                       atol=1e-9)
    assert meta[-1]["step"] == sim.clock.step
    assert len(times) >= 5


def test_do_next_stage_and_exit(standin):
    sim = _sim()
    H = [SI(1e4, "A/m"), 0, 0]
    avgs = sim.hysteresis([H, H, H], save=[("averages", at("stage_end"))],
                          do=[("next_stage", at("stage_step", 4))])
    names, _, rows = datafile.read_ndt("run_dat.ndt")
    col = {n: i for i, n in enumerate(names)}
    assert [r[col["stage"]] for r in rows] == [1, 2, 3]
    assert [r[col["stage_step"]] for r in rows] == [4, 4, 4]
    assert [r[col["step"]] for r in rows] == [4, 8, 12]
    assert len(avgs) == 3
    assert rows[0][col["H_ext_0"]] == pytest.approx(1e4)
    seen = []
    sim.hysteresis([None, None], save=[(lambda s: seen.append(s.clock.stage), at("stage_end"))],
                   do=[("exit", at("stage_step", 2))])
    assert seen == [4]          # a new call starts a new stage; exit ends both
    with pytest.raises(nmag.NmagUserError):
        sim.relax(save=[("nonsense", at("stage_end"))])


def test_hysteresis_restart_continues_with_the_next_stage(standin, monkeypatch):
    """A loop that saved a restart file at the end of each stage and was cut
    after stage 2 continues at stage 3 under --restart (hysteresis.py:432-451,
    main.py:2876-2885): the stages already done are not run again."""
    H = [[SI(v, "A/m"), 0, 0] for v in (1e4, 2e4, 3e4, 4e4)]
    sim = _sim()
    seen = []

    def cut_after_two(s):
        seen.append(s.clock.stage)
        if s.clock.stage == 2:
            Simulation.hysteresis_exit(s)

    sim.hysteresis(H, save=[("restart", at("stage_end"))],
                   do=[(cut_after_two, at("stage_end"))])
    assert seen == [1, 2]
    steps_done = sim.clock.step
    monkeypatch.setattr(simmod.sys, "argv", ["run.py", "--restart"])
    sim2 = _sim()
    seen2 = []
    avgs = sim2.hysteresis(H, save=[(lambda s: seen2.append((s.clock.stage,
                                                              s.get_subfield_average_siv(
                                                                  "H_ext")[0])),
                                     at("stage_end"))])
    assert [st for st, _ in seen2] == [3, 4]
    assert [h for _, h in seen2] == pytest.approx([3e4, 4e4])
    assert len(avgs) == 2 and sim2.clock.step > steps_done


def test_restart_round_trip(standin, monkeypatch):
    sim = _sim()
    sim.advance_time(SI(3e-12, "s"))
    sim.save_restart_file()
    m_saved = sim.get_subfield("m_Py")
    clock = dict(sim.clock)
    assert os.path.exists("run_restart.npz") and not os.path.exists("run_restart.npz.internal")
    monkeypatch.setattr(simmod.sys, "argv", ["run.py", "--restart"])
    sim2 = _sim()
    sim2.load_restart_file()
    assert np.allclose(sim2.get_subfield("m_Py"), m_saved, rtol=1e-14, atol=1e-15)
    assert sim2.clock.step == clock["step"] and sim2.clock.time == clock["time"]
    sim2.set_m([1.0, 0.0, 0.0])
    sim2.load_m_from_h5file("run_restart.npz")
    assert np.allclose(sim2.get_subfield("m_Py"), m_saved, rtol=1e-14, atol=1e-15)
    with pytest.raises(nmag.NmagUserError):
        sim2.load_m_from_h5file("missing.h5")
    # --clean moves old data files aside
    sim.save_data()
    monkeypatch.setattr(simmod.sys, "argv", ["run.py", "--clean"])
    Simulation(name="run")
    assert os.path.exists("run_dat.ndt.old") and not os.path.exists("run_dat.ndt")
