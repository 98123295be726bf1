"""The Simulation object of the nmag front end, for the LLG right-hand-side path.

Mirrors interface/nmag/main.py (Simulation: load_mesh :1075-1226, advance_time
:1325-1499, set_H_ext :1726-1786, set_m :1787-1874, get_subfield :1960-2052,
averages :2234-2288, reinitialise :2707-2731, is_converged :2737-2756) and
interface/nmag/hysteresis.py (relax :207-242, stage loop :471-549).  Values go
in and come out in SI; internally everything is in nmag's simulation units.
All computation runs in libnsim_hip.so through the C ABI.
"""
import math
import os
import re
import sys
import time
import zipfile

import numpy as np

from . import _native
from . import datafile
from .anisotropy import poly_energy
from .material import MagMaterial, NmagUserError
from .mesh import read_nmesh
from .units import SI, su_factor, to_su
from .when import at, every, next_deltas, next_occurrence

# SI units of each field (the reference's fieldunits_by_fieldname)
FIELD_UNITS = {
    "m": SI(1), "M": SI(1, "A/m"), "H_total": SI(1, "A/m"), "H_exch": SI(1, "A/m"),
    "H_anis": SI(1, "A/m"), "H_demag": SI(1, "A/m"), "H_ext": SI(1, "A/m"),
    "dmdt": SI(1, "1/s"), "phi": SI(1, "A"), "rho": SI(1, "A/m^2"),
    "E_exch": SI(1, "J/m^3"), "E_demag": SI(1, "J/m^3"), "E_ext": SI(1, "J/m^3"),
    "E_anis": SI(1, "J/m^3"), "E_total": SI(1, "J/m^3"),
}
# fields defined per material carry the material name in their subfield name
PER_MATERIAL = ("m", "M", "H_total", "H_exch", "H_anis", "dmdt", "E_exch", "E_demag",
                "E_ext", "E_anis", "E_total")
DEGREES_PER_NS = SI(math.pi / 180.0 / 1e-9, "1/s")

# the clock columns that open every ndt row (interface/nmag/main.py:233-240,
# :2337-2345)
NDT_CLOCK = (("time", SI(1, "s")), ("id", SI(1)), ("step", SI(1)), ("last_step_dt", SI(1, "s")),
             ("stage_time", SI(1, "s")), ("stage_step", SI(1)), ("stage", SI(1)))
# matching tolerance of time events (interface/nmag/hysteresis.py:421-424)
_TIME_TOLS = {"time": 1e-20, "stage_time": 1e-20}


def _runid():
    """Default simulation name: the running script's name without extension."""
    base = os.path.splitext(os.path.basename(sys.argv[0] if sys.argv else ""))[0]
    return base if base and base not in ("-c", "__main__", "-") else "simulation"


class Clock(dict):
    def __getattr__(self, k):
        return self[k]

    def __setattr__(self, k, v):
        self[k] = v


class Convergence:
    """2-hit countdown on max|dm/dt| (interface/nmag/convergence.py:15-97)."""

    def __init__(self, countdown=2):
        self.countdown_initial = countdown
        self.countdown = countdown

    def check(self, max_dm_dt, stopping_dm_dt):
        if max_dm_dt < stopping_dm_dt:
            self.countdown -= 1
            if self.countdown < 1:
                self.countdown = self.countdown_initial
                return True
            return False
        self.countdown = self.countdown_initial
        return False


class Simulation:
    def __init__(self, name=None, phi_BEM=None, periodic_bc=None, do_demag=True,
                 do_anisotropy_jacobian=False, temperature=None, thermal_delta_t=None,
                 timestepper_max_order=2, timestepper_krylov_max=300, ksp_tolerances={},
                 adjust_tolerances=False, precision="fp64"):
        if not 1 <= int(timestepper_max_order) <= 5:
            raise NmagUserError("timestepper_max_order %s: BDF orders 1 to 5 are available"
                                % timestepper_max_order)
        if not 1 <= int(timestepper_krylov_max) <= 500:
            raise NmagUserError("timestepper_krylov_max %s: Krylov bases of 1 to 500 vectors "
                                "are available" % timestepper_krylov_max)
        # phi_BEM=HMatrixSetup(...) (main.py:964-981): the compressed demag
        # operator, to HMatrixSetup's eps per block
        self._potential, self._compress_eps, self._compress_eta = "auto", 0.0, 0.0
        if phi_BEM is not None:
            if not hasattr(phi_BEM, "compress_options"):
                raise NmagUserError("phi_BEM must be an HMatrixSetup")
            self._potential, self._compress_eps, self._compress_eta = phi_BEM.compress_options()
        if periodic_bc is not None:
            raise NmagUserError("periodic boundary conditions are not available on this path")
        if temperature is not None:
            raise NmagUserError("thermal integration is not available on this path")
        for k in ksp_tolerances:
            if k not in ("NBC.rtol", "NBC.atol", "NBC.dtol", "NBC.maxits",
                         "DBC.rtol", "DBC.atol", "DBC.dtol", "DBC.maxits"):
                raise NmagUserError("unknown ksp_tolerances key '%s'" % k)
        self.name = name or _runid()
        self.do_demag = do_demag
        self.precision = precision
        self.ksp_tolerances = dict(ksp_tolerances)
        self.max_order = timestepper_max_order
        self.krylov_max = timestepper_krylov_max
        # LAMTimestepper defaults (interface/nmag/main.py:721-740)
        self.ts_rel_tol = 1e-6
        self.ts_abs_tol = 1e-6
        self.stopping_dm_dt = float(to_su(1.0 * DEGREES_PER_NS, "1/s"))
        self._default_exact_tstop = True
        self.max_dm_dt = None
        self.convergence = Convergence()
        # SimulationClock (interface/nmag/clock.py:21-37); times in SI seconds
        self.clock = Clock(id=-1, time=0.0, step=0, stage=1, stage_step=0, stage_time=0.0,
                           real_time=0.0, stage_end=False, convergence=False,
                           exit_hysteresis=False, zero_stage_step=0, zero_stage_time=0.0,
                           time_reached_si=0.0, time_reached_su=0.0, last_step_dt_si=0.0,
                           last_step_dt_su=0.0)
        self._magnet = None
        self._mat = None
        self._ts_initialised = False
        self._stages_run = 0
        self._saved_steps = set()
        # the stand-in for "no time limit" of relax/hysteresis
        # (simulation_core.py:170-180)
        self.max_time_reached = SI(1.0, "s")
        # relax/hysteresis 'save' and 'do' names (simulation_core.py:185-197)
        self.action_abbreviations = {}
        self.add_save_abbrev("save_averages", lambda s: s.save_data(avoid_same_step=True))
        self.add_save_abbrev("save_fields",
                             lambda s: s.save_data(fields="all", avoid_same_step=True))
        self.add_save_abbrev("save_field_m",
                             lambda s: s.save_data(fields=["m"], avoid_same_step=True))
        self.add_save_abbrev("save_restart", lambda s: s.save_restart_file())
        self.add_do_abbrev("do_next_stage", Simulation.hysteresis_next_stage)
        self.add_do_abbrev("do_exit", Simulation.hysteresis_exit)
        # existing data files: an error unless the script runs with --clean
        # (old files renamed *.old) or --restart (continued); the command-line
        # flags of the reference (simulation_core.py:143-166, snippets.py:151-156)
        self._restarting = False
        argv = sys.argv[1:] if sys.argv else []
        data_files = [self._ndtfilename(), self._datafilename()]
        if "--clean" in argv:
            for f in data_files:
                if os.path.exists(f):
                    os.replace(f, f + ".old")
        elif "--restart" in argv:
            self._restarting = True
        else:
            for f in data_files:
                if os.path.exists(f):
                    raise NmagUserError(
                        "found old file %s -- cannot proceed: run with --clean to move old "
                        "data files aside or with --restart to continue the run" % f)

    # ------------------------------------------------------------- set-up
    def load_mesh(self, filename, region_names_and_mag_mats, unit_length=SI(1e-9, "m"),
                  do_reorder=False, manual_distribution=None):
        """filename: an nmesh ASCII file, or a (coords, tets) pair in units of
        unit_length.  region_names_and_mag_mats: [(region_name, MagMaterial)]."""
        if isinstance(filename, str):
            coords, tets, regions = read_nmesh(filename)
            keep = regions > 0
            tets = tets[keep]
        else:
            coords, tets = filename
        mats = [m for _, m in region_names_and_mag_mats
                if isinstance(m, MagMaterial) or (isinstance(m, list) and m)]
        mats = [m[0] if isinstance(m, list) else m for m in mats]
        if len(set(id(m) for m in mats)) != 1:
            raise NmagUserError("this path simulates one magnetic material per Simulation")
        self._mat = mats[0]
        self._region_name = region_names_and_mag_mats[0][0]
        scale = unit_length.value / 1e-9 if isinstance(unit_length, SI) else float(unit_length)
        self._coords_su = np.ascontiguousarray(np.asarray(coords, dtype=np.float64) * scale)
        self._tets = np.ascontiguousarray(tets, dtype=np.int32)
        self._unit_length = unit_length
        kt = self.ksp_tolerances
        self._magnet = _native.NativeMagnet(
            self._coords_su, self._tets, self._mat.su_coeffs(), precision=self.precision,
            do_demag=self.do_demag, nbc_rtol=kt.get("NBC.rtol", 1e-5),
            dbc_rtol=kt.get("DBC.rtol", 1e-5), nbc_maxits=int(kt.get("NBC.maxits", 10000)),
            dbc_maxits=int(kt.get("DBC.maxits", 10000)), potential=self._potential,
            compress_eps=self._compress_eps, compress_eta=self._compress_eta)
        an = self._mat.su_anisotropy()
        if an is not None:
            k1, k2, axis = an
            n = self._magnet.n
            self._magnet.set_field("axis", np.tile(axis, (n, 1)))
            self._magnet.set_field("K1", np.full(n, k1))
            if k2:
                self._magnet.set_field("K2", np.full(n, k2))
        poly = self._mat.su_anisotropy_poly()
        if poly is not None:
            self._magnet.set_anisotropy(*poly)
        self._volumes = self._magnet.get_field("volumes")
        return self

    @property
    def magnet(self):
        if self._magnet is None:
            raise NmagUserError("load_mesh has not been called")
        return self._magnet

    def _node_values(self, values, comps, unit_su):
        n = self.magnet.n
        if callable(values):
            pos_si = self._coords_su * 1e-9
            arr = np.array([values(p) for p in pos_si], dtype=np.float64)
        else:
            arr = np.asarray(values if not isinstance(values, SI) else values.value,
                             dtype=np.float64)
        arr = np.broadcast_to(arr.reshape(-1, comps) if arr.ndim > 1 or arr.size == n * comps
                              else arr, (n, comps)).astype(np.float64)
        return arr / unit_su

    def set_m(self, values, subfieldname=None):
        """Set the normalised magnetisation: a 3-vector, an (N,3) array or a
        function of the position (m, SI); normalised per node (main.py:1873)."""
        m = self._node_values(values, 3, 1.0).copy()
        nrm = np.linalg.norm(m, axis=1, keepdims=True)
        if np.any(nrm == 0):
            raise NmagUserError("set_m: zero vector given")
        self.magnet.set_field("m", m / nrm)
        self._ts_initialised = False
        self.max_dm_dt = None

    def set_H_ext(self, values, unit=None):
        """Applied field: 3-vector, (N,3) array or function; in A/m unless unit
        is given (e.g. unit=SI(1e6, 'A/m'))."""
        if isinstance(values, (list, tuple)) and any(isinstance(v, SI) for v in values):
            # e.g. relax(H_applied=[SI(1e5, 'A/m'), 0, 0])
            base = float(to_su(unit if unit is not None else SI(1.0, "A/m"), "A/m"))
            vec = [float(to_su(v, "A/m")) if isinstance(v, SI) else float(v) * base
                   for v in values]
            H_su = np.tile(vec, (self.magnet.n, 1))
        elif isinstance(values, SI):
            H_su = self._node_values(values.value, 3, 1.0) * float(
                to_su(SI(1.0, values.dims), "A/m"))
        else:
            factor = unit if unit is not None else SI(1.0, "A/m")
            H_su = self._node_values(values, 3, 1.0) * float(to_su(factor, "A/m"))
        self.magnet.set_field("H_ext", H_su)
        self._ts_initialised = False

    def set_pinning(self, values):
        """Per-node pinning factor (1 free, 0 pinned), the reference's 'pin' field."""
        self.magnet.set_field("pin", self._node_values(values, 1, 1.0).reshape(-1))
        self._ts_initialised = False

    def set_params(self, stopping_dm_dt=None, ts_rel_tol=None, ts_abs_tol=None,
                   exact_tstop=None):
        if stopping_dm_dt is not None:
            self.stopping_dm_dt = float(to_su(stopping_dm_dt, "1/s"))
        if ts_rel_tol is not None:
            self.ts_rel_tol = ts_rel_tol
            self._ts_initialised = False
        if ts_abs_tol is not None:
            self.ts_abs_tol = ts_abs_tol
            self._ts_initialised = False
        if exact_tstop is not None:
            self._default_exact_tstop = exact_tstop

    def reinitialise(self, rel_tolerance=None, abs_tolerance=None, initial_time=None):
        if initial_time is not None:
            t0 = float(to_su(initial_time, "s"))
            self.clock.time_reached_su = t0
            self.clock.time_reached_si = t0 * su_factor(SI(1, "s").dims)
        if rel_tolerance is not None:
            self.ts_rel_tol = rel_tolerance
        if abs_tolerance is not None:
            self.ts_abs_tol = abs_tolerance
        self.magnet.ts_init(self.clock.time_reached_su, self.ts_rel_tol, self.ts_abs_tol,
                            self.max_order, self.krylov_max)
        self._ts_initialised = True

    # ----------------------------------------------------------- dynamics
    def advance_time(self, target_time, max_it=-1, exact_tstop=None):
        """Integrate the LLG to target_time (SI seconds); returns the time reached."""
        if max_it is None or max_it < 0:
            max_it = -1
        if exact_tstop is None:
            exact_tstop = self._default_exact_tstop
        target_su = float(to_su(target_time, "s"))
        if target_su == 0:
            return target_time
        if not self._ts_initialised:
            self.reinitialise()
        g = self.magnet
        steps_before = g.stat("ts_steps")
        wall = time.perf_counter()
        t_reached = g.advance(target_su, max_it, exact_tstop)
        self.clock.real_time += time.perf_counter() - wall
        delta_step = int(g.stat("ts_steps") - steps_before)
        delta_time = t_reached - self.clock.time_reached_su
        time_to_si = su_factor(SI(1, "s").dims)
        self.clock.time_reached_su = t_reached
        self.clock.time_reached_si = t_reached * time_to_si
        self.clock.last_step_dt_su = g.stat("ts_h")
        self.clock.last_step_dt_si = self.clock.last_step_dt_su * time_to_si
        if delta_time > 0.0:
            self.max_dm_dt = g.stat("ts_max_dm") / delta_time
        self.clock.step += delta_step
        self.clock.stage_step += delta_step
        self.clock.time += delta_time * time_to_si
        self.clock.stage_time = self.clock.time_reached_si
        return SI(self.clock.time_reached_si, "s")

    def is_converged(self):
        self.clock.convergence = False
        if self.max_dm_dt is None:
            return False
        c = self.convergence.check(self.max_dm_dt, self.stopping_dm_dt)
        self.clock.convergence = c
        return c

    def relax(self, H_applied=None, save=[("averages", "fields", at("stage_end"))], do=[],
              convergence_check=every("step", 5)):
        """Integrate until max|dm/dt| < stopping_dm_dt on two consecutive checks
        (one stage of hysteresis; interface/nmag/hysteresis.py:207-242).
        Returns the time reached in the stage (SI seconds)."""
        self.hysteresis([H_applied], save=save, do=do, convergence_check=convergence_check)
        return self.clock.time_reached_si

    def hysteresis(self, H_ext_list, save=[("averages", "fields", at("stage_end"))], do=[],
                   convergence_check=every("step", 5)):
        """Relax at every applied field of H_ext_list in turn (None keeps the
        current field), running the `do` and `save` actions whenever their
        schedules fall due (interface/nmag/hysteresis.py:245-549).  Each entry
        of `save`/`do` is (thing, ..., when): a thing is a function of the
        Simulation or one of the names 'averages', 'fields', 'field_m',
        'restart' (save) and 'next_stage', 'exit' (do).  Returns the average
        m (SI) at the end of each stage."""
        actions = self._join_actions(do, "do") + self._join_actions(save, "save")
        planned = {}
        for fn, _ in actions:
            if id(fn) in planned:
                raise NmagUserError("'save'/'do' name the same thing twice: join the "
                                    "schedules with | instead")
            planned[id(fn)] = None
        # the first stage keeps the clock's stage number; a restarted run
        # continues from the stage it was saved in (hysteresis.py:432-451).
        # Later calls on the same object start a new stage (the reference would
        # skip fields here, by indexing the list with the old stage number).
        skip = 0
        if self._restarting:
            self.load_restart_file()
            self._restarting = False
            skip = max(0, self.clock.stage - 1)
            stage = self.clock.stage
        else:
            stage = self.clock.stage if self._stages_run == 0 else self.clock.stage + 1
        out = []
        c = self.clock
        # a clock restored from a file saved by an exiting loop carries the flag
        c.exit_hysteresis = False
        for H in H_ext_list[skip:]:
            self.do_next_stage(stage=stage)
            stage = None
            self._stages_run += 1
            c.stage_end = False
            if H is not None and not (isinstance(H, (list, tuple)) and len(H) == 0):
                self.set_H_ext(H)
            self.reinitialise(initial_time=0)
            for fn, when in actions:
                planned[id(fn)] = next_occurrence(when, c, _TIME_TOLS)
            while True:
                c.stage_end = self.is_converged()
                deltas = next_deltas(convergence_check, c, tols=_TIME_TOLS)
                for fn, when in actions:
                    due = next_occurrence(when, c, _TIME_TOLS)
                    # run when the event is now, or when the planned occurrence
                    # has been passed (the next one moved on)
                    if when.match_time(c) or due != planned[id(fn)]:
                        fn(self)
                    planned[id(fn)] = due
                    deltas = next_deltas(when, c, suggest=deltas, tols=_TIME_TOLS)
                if c.exit_hysteresis:
                    c.exit_hysteresis = False
                    return out
                if c.stage_end:
                    break
                d_step, d_time, _ = deltas
                if d_time is None:
                    target = self.max_time_reached
                else:
                    target = SI(c.stage_time + d_time, "s")
                t = self.advance_time(target, max_it=-1 if d_step is None else int(d_step))
                if t.value > 0.99 * self.max_time_reached.value:
                    raise NmagUserError("Simulation time reached %s s: are you starting from "
                                        "a zero torque configuration?"
                                        % self.max_time_reached.value)
            out.append(self.get_subfield_average_siv("m", self._mat_name()))
        return out

    # ------------------------------------------------- stages and actions
    def do_next_stage(self, stage=None):
        """Start a stage: number `stage`, or the next one (clock.py inc_stage)."""
        self.max_dm_dt = None
        self.convergence = Convergence()
        c = self.clock
        c.stage = c.stage + 1 if stage is None else stage
        c.stage_step = 0
        c.stage_time = 0.0
        c.convergence = False
        c.zero_stage_step = c.step
        c.zero_stage_time = c.time

    @staticmethod
    def hysteresis_next_stage(sim):
        """A 'do' action: end the current stage (simulation_core.py:256-262)."""
        sim.clock.stage_end = True

    @staticmethod
    def hysteresis_exit(sim):
        """A 'do' action: leave relax/hysteresis (simulation_core.py:264-271)."""
        sim.clock.exit_hysteresis = True
        sim.clock.stage_end = True

    def _add_abbrev(self, prefix, abbreviation, function):
        name = abbreviation if abbreviation.startswith(prefix + "_") else prefix + "_" + abbreviation
        self.action_abbreviations[name] = function

    def add_save_abbrev(self, abbreviation, function):
        """Name a function for the 'save' list: add_save_abbrev('funky', f)
        allows save=[('funky', every('step', 10))]."""
        self._add_abbrev("save", abbreviation, function)

    def add_do_abbrev(self, abbreviation, function):
        """Name a function for the 'do' list."""
        self._add_abbrev("do", abbreviation, function)

    def _join_actions(self, items, prefix):
        """[(function, when)] from a save/do list of (thing, ..., when)
        (hysteresis.py:55-126)."""
        out = []
        try:
            for item in items:
                item = list(item)
                things, when = item[:-1], item[-1]
                for thing in things:
                    if callable(thing):
                        out.append((thing, when))
                        continue
                    if isinstance(thing, str):
                        norm = re.sub(r"[ \t_]+", "_", thing.lower())
                        fn = self.action_abbreviations.get(prefix + "_" + norm,
                                                           self.action_abbreviations.get(norm))
                        if fn is not None:
                            out.append((fn, when))
                            continue
                    raise NmagUserError(
                        "'%s' list: cannot %s '%s'; give a function or one of: %s"
                        % (prefix, prefix, thing, ", ".join(sorted(self.action_abbreviations))))
        except TypeError:
            raise NmagUserError("'%s' must be a list of tuples (thing, ..., when)" % prefix)
        return out

    # ------------------------------------------------------- data files
    def _ndtfilename(self):
        return self.name + "_dat.ndt"

    def _datafilename(self):
        """Field store: the counterpart of the reference's _dat.h5."""
        return self.name + "_dat.npz"

    def get_restart_file_name(self):
        return self.name + "_restart.npz"

    def _mat_name(self):
        return self._mat.name if self._mat is not None else self._region_name

    def _subfield_names(self, fields):
        mat = self._mat_name()
        return ["%s_%s" % (f, mat) if f in PER_MATERIAL else f for f in fields]

    def _all_fields_su(self):
        """Every field of FIELD_UNITS at the current m, from one evaluation."""
        g = self.magnet
        g.execute("update_dmdt")
        out = {f: g.get_field(f) for f in ("m", "H_exch", "H_demag", "H_anis", "H_total",
                                           "H_ext", "dmdt", "phi", "rho")}
        out["M"] = self._mat.su_Ms * out["m"]
        out.update(self._energies_su(out))
        return out

    def _energies_su(self, f):
        m = f["m"]
        Ms, mu0 = self._mat.su_Ms, self._mat.su_mu0
        # energy densities (nmag_lam.py template_E)
        e = {"E_demag": -0.5 * mu0 * Ms * (m * f["H_demag"]).sum(1),
             "E_exch": -0.5 * mu0 * Ms * (m * f["H_exch"]).sum(1),
             "E_ext": -1.0 * mu0 * Ms * (m * f["H_ext"]).sum(1)}
        an, poly = self._mat.su_anisotropy(), self._mat.su_anisotropy_poly()
        e["E_anis"] = np.zeros(len(m))
        if an is not None:
            k1, k2, axis = an
            u = m @ axis
            e["E_anis"] = -k1 * u ** 2 - k2 * u ** 4
        if poly is not None:
            e["E_anis"] = e["E_anis"] + poly_energy(poly[0], poly[1], m)
        e["E_total"] = e["E_demag"] + e["E_exch"] + e["E_ext"] + e["E_anis"]
        return e

    def _edges(self):
        if getattr(self, "_edge_list", None) is None:
            self._edge_list = datafile.mesh_edges(self._tets)
        return self._edge_list

    def _compute_averages(self):
        """(names, values, units) of one ndt row: the clock, the volume average
        of every field in SI, the largest angle of m across a mesh edge, local
        and unix time (interface/nmag/main.py:2313-2479)."""
        c = self.clock
        names = [n for n, _ in NDT_CLOCK]
        units = [u for _, u in NDT_CLOCK]
        values = [float(c.time), int(c.id), int(c.step), float(c.last_step_dt_si),
                  float(c.stage_time), int(c.stage_step), int(c.stage)]
        f = self._all_fields_su()
        V = self._volumes
        for field, sub in zip(FIELD_UNITS, self._subfield_names(FIELD_UNITS)):
            si = FIELD_UNITS[field]
            vals = f[field] * su_factor(si.dims)
            avg = (V @ vals) / V.sum()
            names.append(sub)
            units.append(si)
            values.append([float(a) for a in avg] if vals.ndim > 1 else float(avg))
        names.append("maxangle_m_%s" % self._mat_name())
        units.append("<deg>")
        values.append(datafile.max_angle_deg(f["m"], self._edges()))
        now = time.localtime()
        names += ["localtime", "unixtime"]
        units += ["<>", SI(1, "s")]
        values += ["%04d/%02d/%02d-%02d:%02d:%02d" % now[:6], time.time()]
        return names, values, units, f

    def save_data(self, fields=None, avoid_same_step=False):
        """Append the averages to <name>_dat.ndt and, with `fields` ('all', a
        field name or a list of names), the fields to <name>_dat.npz
        (interface/nmag/main.py:2481-2535).  With avoid_same_step, a step
        whose averages were saved already is not saved again."""
        f = None
        if not (avoid_same_step and self.clock.step in self._saved_steps):
            self.clock.id += 1
            names, values, units, f = self._compute_averages()
            path = self._ndtfilename()
            header = not os.path.exists(path) or os.path.getsize(path) == 0
            datafile.write_ndt_row(path, names, values, units, header)
            self._saved_steps.add(self.clock.step)
        if fields:
            if isinstance(fields, str):
                fields = [] if fields.lower() == "all" else [fields]
            self._save_fields(list(fields), f)

    def _save_fields(self, fieldnames, f=None):
        for name in fieldnames:
            if name not in FIELD_UNITS:
                raise NmagUserError("cannot save unknown field '%s'" % name)
        names = fieldnames or list(FIELD_UNITS)
        if f is None:
            f = self._all_fields_su()
        arrays = {sub: f[n] * su_factor(FIELD_UNITS[n].dims)
                  for n, sub in zip(names, self._subfield_names(names))}
        c = self.clock
        meta = {"id": c.id, "step": c.step, "stage": c.stage, "stage_step": c.stage_step,
                "time": c.time, "stage_time": c.stage_time}
        path = self._datafilename()
        row = len(datafile.read_meta(path)) if os.path.exists(path) else 0
        datafile.append_fields(path, row, meta, arrays, self._coords_su * 1e-9, self._tets)

    def save_restart_file(self, filename=None, fieldnames=["m"], all=False):
        """Write the fields (default m) and the clock, from which a run
        continues with --restart (interface/nmag/main.py:2764-2857)."""
        names = list(FIELD_UNITS) if all else list(fieldnames)
        for n in names:
            if n not in FIELD_UNITS:
                raise NmagUserError("cannot save unknown field '%s'" % n)
        f = self._all_fields_su() if any(n != "m" for n in names) else \
            {"m": self.magnet.get_field("m")}
        arrays = {sub: f[n] * su_factor(FIELD_UNITS[n].dims)
                  for n, sub in zip(names, self._subfield_names(names))}
        datafile.write_restart(filename or self.get_restart_file_name(), arrays,
                               dict(self.clock), self._coords_su * 1e-9)

    def load_restart_file(self, file_name=None):
        """Restore m and the clock from a restart file; a file saved at
        convergence continues with the next stage (main.py:2859-2885)."""
        path = file_name or self.get_restart_file_name()
        fields, clock, pos = datafile.read_restart(path)
        if pos.shape != self._coords_su.shape:
            raise NmagUserError("restart file %s holds %d nodes, this mesh %d"
                                % (path, len(pos), len(self._coords_su)))
        self.load_m_from_h5file(path)
        self.clock.update(clock)
        if self.clock.convergence:
            self.clock.stage += 1

    def load_m_from_h5file(self, filename, id=None, row=None):
        """Set m from a field store or restart file written by this package
        (the first saved m unless id/row say otherwise; main.py:1876-1901).
        Files are .npz archives: no HDF5 library is available here."""
        if not (os.path.exists(filename) and zipfile.is_zipfile(filename)):
            raise NmagUserError("%s is not a data file of this package (.npz); HDF5 files "
                                "cannot be read without an HDF5 library" % filename)
        self.set_m(datafile.read_subfield(filename, "m_%s" % self._mat_name(), id=id, row=row))

    # -------------------------------------------------------- observables
    def _field_su(self, field):
        g = self.magnet
        if field in ("m", "H_exch", "H_demag", "H_anis", "H_total", "H_ext", "dmdt", "phi",
                     "rho"):
            if field not in ("m", "H_ext"):
                g.execute("update_dmdt")
            return g.get_field(field)
        if field == "M":
            return self._mat.su_Ms * g.get_field("m")
        if field.startswith("E_") and field in FIELD_UNITS:
            g.execute("update_dmdt")
            f = {k: g.get_field(k) for k in ("m", "H_demag", "H_exch", "H_ext")}
            return self._energies_su(f)[field]
        raise NmagUserError("unknown field '%s'" % field)

    def _split(self, subfieldname):
        for f in sorted(FIELD_UNITS, key=len, reverse=True):
            if subfieldname == f or subfieldname.startswith(f + "_"):
                return f
        raise NmagUserError("unknown subfield '%s'" % subfieldname)

    def get_subfield(self, subfieldname, units=None):
        """(N,3) or (N,) array of a subfield in SI (or in `units`)."""
        field = self._split(subfieldname)
        vals = self._field_su(field)
        si = FIELD_UNITS[field]
        out = vals * su_factor(si.dims)
        if units is not None:
            out = out / units.value
        return out

    def get_subfield_positions(self, subfieldname=None, pos_units=SI(1, "m")):
        return self._coords_su * 1e-9 / pos_units.value

    def get_subfield_average_siv(self, field_name, subfield_name=None):
        """Volume average of a field in SI (interface/nmag/main.py:2234-2288)."""
        vals = self._field_su(field_name) * su_factor(FIELD_UNITS[field_name].dims)
        V = self._volumes
        if vals.ndim == 1:
            return float(V @ vals / V.sum())
        return list((V @ vals) / V.sum())

    def get_timers(self):
        """Per-kernel device time of the last timed evaluations (ms, launches)."""
        g = self.magnet
        return {k: g.timer(k) for k in ("exch_div", "cg_nbc", "bem", "cg_dbc", "demag_llg",
                                        "jtimes")}


def get_subfield_from_h5file(filename, subfieldname, id=None, row=None):
    """SI array of a subfield saved by save_data(fields=...) or
    save_restart_file: the save with clock id `id`, else the row-th save
    holding it (default the first); interface/nmag/main.py:2915-2956."""
    return datafile.read_subfield(filename, subfieldname, id=id, row=row)


def get_subfield_positions_from_h5file(filename, subfieldname=None):
    """Node positions (m) of the mesh the file's fields live on
    (interface/nmag/main.py:2958-2985)."""
    return datafile.read_positions(filename)
