"""The Simulation object of the nmag front end, for the LLG right-hand-side path.

Mirrors interface/nmag/main.py (Simulation: load_mesh :1075-1226, advance_time
:1325-1499, set_H_ext :1726-1786, set_m :1787-1874, get_subfield :1960-2052,
averages :2234-2288, reinitialise :2707-2731, is_converged :2737-2756) and
interface/nmag/hysteresis.py (relax :207-242, stage loop :471-549).  Values go
in and come out in SI; internally everything is in nmag's simulation units.
All computation runs in libnsim_hip.so through the C ABI.
"""
import math

import numpy as np

from . import _native
from .material import MagMaterial, NmagUserError
from .mesh import read_nmesh
from .units import SI, su_factor, to_su

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


class every:
    """Schedule 'every n steps' (interface/nsim/when.py); only ('step', n) is used
    by relax's convergence check."""

    def __init__(self, what, value):
        self.what, self.value = what, value


class at:
    def __init__(self, what, value=None):
        self.what, self.value = what, value


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
        if phi_BEM is not None:
            raise NmagUserError("H-matrix BEM compression (phi_BEM) is not available on this path")
        if periodic_bc is not None:
            raise NmagUserError("periodic boundary conditions are not available on this path")
        if temperature is not None:
            raise NmagUserError("thermal integration is not available on this path")
        for k in ksp_tolerances:
            if k not in ("NBC.rtol", "NBC.atol", "NBC.dtol", "NBC.maxits",
                         "DBC.rtol", "DBC.atol", "DBC.dtol", "DBC.maxits"):
                raise NmagUserError("unknown ksp_tolerances key '%s'" % k)
        self.name = name or "simulation"
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
        self.clock = Clock(time=0.0, step=0, stage=1, stage_step=0, stage_time=0.0,
                           time_reached_si=0.0, time_reached_su=0.0, last_step_dt_si=0.0,
                           last_step_dt_su=0.0, convergence=False)
        self._magnet = None
        self._mat = None
        self._ts_initialised = False
        self.max_time_reached = SI(1e-6, "s")

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
            dbc_maxits=int(kt.get("DBC.maxits", 10000)))
        an = self._mat.su_anisotropy()
        if an is not None:
            k1, k2, axis = an
            n = self._magnet.n
            self._magnet.set_field("axis", np.tile(axis, (n, 1)))
            self._magnet.set_field("K1", np.full(n, k1))
            if k2:
                self._magnet.set_field("K2", np.full(n, k2))
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
        t_reached = g.advance(target_su, max_it, exact_tstop)
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

    def relax(self, H_applied=None, save=None, do=None, convergence_check=every("step", 5)):
        """Integrate until max|dm/dt| < stopping_dm_dt on two consecutive checks,
        checking every convergence_check.value steps (hysteresis.py:207-549)."""
        if H_applied is not None:
            self.set_H_ext(H_applied)
        steps = convergence_check.value if convergence_check.what == "step" else 5
        self.max_dm_dt = None
        self.convergence = Convergence()
        self.clock.stage_step = 0
        max_t = self.max_time_reached
        while True:
            if self.is_converged():
                break
            t = self.advance_time(max_t, max_it=steps)
            if t.value > 0.99 * max_t.value:
                raise NmagUserError("Simulation time reached %s: are you starting from a "
                                    "zero torque configuration?" % max_t)
        self.clock.stage += 1
        return self.clock.time_reached_si

    def hysteresis(self, H_ext_list, save=None, do=None, convergence_check=every("step", 5)):
        """Relax at every applied field of H_ext_list in turn; returns the
        average m (SI list) at the end of each stage."""
        out = []
        for H in H_ext_list:
            self.relax(H, convergence_check=convergence_check)
            out.append(self.get_subfield_average_siv("m", self._region_name))
        return out

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
        if field.startswith("E_"):
            g.execute("update_dmdt")
            m = g.get_field("m")
            Ms, mu0 = self._mat.su_Ms, self._mat.su_mu0
            # energy densities (nmag_lam.py template_E)
            e = {"E_demag": -0.5 * mu0 * Ms * (m * g.get_field("H_demag")).sum(1),
                 "E_exch": -0.5 * mu0 * Ms * (m * g.get_field("H_exch")).sum(1),
                 "E_ext": -1.0 * mu0 * Ms * (m * g.get_field("H_ext")).sum(1)}
            an = self._mat.su_anisotropy()
            if an is not None:
                k1, k2, axis = an
                u = m @ axis
                e["E_anis"] = -k1 * u ** 2 - k2 * u ** 4
            else:
                e["E_anis"] = np.zeros(len(m))
            e["E_total"] = e["E_demag"] + e["E_exch"] + e["E_ext"] + e["E_anis"]
            return e[field]
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
            return float((vals * V).sum() / V.sum())
        return list((vals * V[:, None]).sum(0) / V.sum())

    def get_timers(self):
        """Per-kernel device time of the last timed evaluations (ms, launches)."""
        g = self.magnet
        return {k: g.timer(k) for k in ("exch_div", "cg_nbc", "bem", "cg_dbc", "demag_llg",
                                        "jtimes")}
