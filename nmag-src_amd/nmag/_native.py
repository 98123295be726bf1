"""ctypes binding of the HIP library (C ABI declared in include/nsim.h).

The library is the only compute path: if it is missing, or no HIP device is
visible, creating a magnet raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnsim_hip.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

FP64, MIXED_FP32, MIXED_BF16, MIXED_F24 = 0, 1, 2, 3
PRECISIONS = {"fp64": FP64, "mixed": MIXED_FP32, "fp32": MIXED_FP32, "bf16": MIXED_BF16,
              "mixed24": MIXED_F24}


class NsimMesh(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int), ("n_tets", ctypes.c_int),
                ("coords", _dp), ("tets", _ip)]


class NsimMaterial(ctypes.Structure):
    _fields_ = [("Ms", ctypes.c_double), ("negJ", ctypes.c_double), ("c1", ctypes.c_double),
                ("c2", ctypes.c_double), ("c3", ctypes.c_double), ("mu0", ctypes.c_double)]


class NsimOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("do_demag", ctypes.c_int),
                ("nbc_rtol", ctypes.c_double), ("dbc_rtol", ctypes.c_double),
                ("nbc_maxits", ctypes.c_int), ("dbc_maxits", ctypes.c_int),
                ("potential", ctypes.c_int), ("compress_eps", ctypes.c_double),
                ("compress_eta", ctypes.c_double)]


POTENTIAL = {"auto": 0, "dense": 1, "iterative": -1, "hmatrix": 2}


class NsimTraj(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("n_steps", ctypes.c_int),
                ("checkpoint_every", ctypes.c_int), ("n_pulses", ctypes.c_int),
                ("pulse_len", ctypes.c_double), ("pulse_dir", ctypes.c_double * 3)]


class NsimBdfTraj(ctypes.Structure):
    _fields_ = [("t_final", ctypes.c_double), ("rtol", ctypes.c_double),
                ("atol", ctypes.c_double), ("max_order", ctypes.c_int),
                ("max_steps", ctypes.c_int), ("n_pulses", ctypes.c_int),
                ("pulse_len", ctypes.c_double), ("pulse_dir", ctypes.c_double * 3)]


# name -> (restype, argtypes); every symbol include/nsim.h declares
SIGNATURES = {
    "nsim_create": (ctypes.c_int, [ctypes.POINTER(NsimMesh), ctypes.POINTER(NsimMaterial),
                                   ctypes.POINTER(NsimOptions), ctypes.POINTER(_vp)]),
    "nsim_destroy": (None, [_vp]),
    "nsim_last_error": (ctypes.c_char_p, []),
    "nsim_get_sizes": (ctypes.c_int, [_vp, _ip]),
    "nsim_set_field": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp]),
    "nsim_get_field": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp]),
    "nsim_execute": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "nsim_get_bem": (ctypes.c_int, [_vp, _dp]),
    "nsim_rhs": (ctypes.c_int, [_vp, ctypes.c_double, _vp, _vp, ctypes.c_int]),
    "nsim_jtimes": (ctypes.c_int, [_vp, ctypes.c_double, _vp, _vp, _vp, _vp, ctypes.c_int]),
    "nsim_psolve_noalloc": (ctypes.c_int, [_vp, ctypes.c_double, _vp, _vp, _vp, _vp,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_int]),
    "nsim_ts_set_tols": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "nsim_get_stat": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp]),
    "nsim_timers_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nsim_get_timer": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp]),
    "nsim_set_tuning": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int]),
    "nsim_synchronize": (ctypes.c_int, [_vp]),
    "nsim_psetup": (ctypes.c_int, [_vp, ctypes.c_double, _vp, _vp, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_double, ctypes.c_int]),
    "nsim_psolve": (ctypes.c_int, [_vp, ctypes.c_double, _vp, _vp, _vp, _vp, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "nsim_ts_init": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_int]),
    "nsim_advance": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double)]),
    "nsim_create_batch": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(NsimMesh),
                                         ctypes.POINTER(NsimMaterial),
                                         ctypes.POINTER(NsimOptions), ctypes.POINTER(_vp)]),
    "nsim_get_batch_layout": (ctypes.c_int, [_vp, _ip, _ip]),
    "nsim_autotune": (ctypes.c_int, [_vp, ctypes.c_int]),
    "nsim_get_potential": (ctypes.c_int, [_vp, _dp]),
    "nsim_set_material": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(NsimMaterial)]),
    "nsim_set_anisotropy": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_double)]),
    "nsim_rhs_vjp": (ctypes.c_int, [_vp, ctypes.c_double, _dp, _dp, _dp, _dp, _dp, _dp]),
    "nsim_traj_forward": (ctypes.c_int, [_vp, ctypes.POINTER(NsimTraj), _dp]),
    "nsim_traj_adjoint": (ctypes.c_int, [_vp, ctypes.POINTER(NsimTraj), _dp, _dp, _dp, _dp,
                                         _dp, _dp]),
    "nsim_bdf_forward": (ctypes.c_int, [_vp, ctypes.POINTER(NsimBdfTraj), _dp,
                                        ctypes.POINTER(ctypes.c_int)]),
    "nsim_bdf_steps": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp,
                                      ctypes.POINTER(ctypes.c_int)]),
    "nsim_bdf_replay": (ctypes.c_int, [_vp, _dp, ctypes.c_double]),
    "nsim_bdf_adjoint": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double]),
}

_lib = None


def lib():
    """Load libnsim_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libnsim_hip.so is not built (run __graft_entry__.build() or "
                "make -C nmag-src_amd/csrc); there is no fallback compute path")
        # PyTorch ships its own HIP runtime under the same soname.  Whichever
        # is loaded first serves the whole process; if the library's came
        # first, a later torch finds no GPU.  Load torch's first when present,
        # so device tensors and the library share one runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class NsimError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        msg = lib().nsim_last_error()
        raise NsimError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


def _dptr(a):
    return a.ctypes.data_as(_dp)


class NativeMagnet:
    """One magnet on the GPU: operators, fields and solver state in HBM."""

    VEC3 = ("m", "H_ext", "axis", "H_exch", "H_demag", "H_anis", "H_total", "dmdt")
    SCALAR = ("K1", "K2", "pin", "rho", "phi", "phi1", "phi2", "volumes")

    def __init__(self, coords, tets, material, precision="fp64", do_demag=True,
                 nbc_rtol=1e-5, dbc_rtol=1e-5, nbc_maxits=10000, dbc_maxits=10000,
                 potential="auto", compress_eps=0.0, compress_eta=0.0):
        """potential: "iterative" (the two Poisson solves around the BEM, as
        the reference), "dense" (the precomputed potential operator), "hmatrix"
        (the same operator compressed to compress_eps per block with
        admissibility compress_eta; 0: the defaults) or "auto" (dense when it
        fits)."""
        self._create([(coords, tets)], [material], precision, do_demag, nbc_rtol, dbc_rtol,
                     nbc_maxits, dbc_maxits, potential, compress_eps, compress_eta)

    def _create(self, meshes, materials, precision, do_demag, nbc_rtol, dbc_rtol, nbc_maxits,
                dbc_maxits, potential="auto", compress_eps=0.0, compress_eta=0.0):
        L = lib()
        self._h = None
        if len(meshes) != len(materials) or not meshes:
            raise ValueError("need one material per mesh and at least one mesh")
        self._keep = []
        cmeshes = (NsimMesh * len(meshes))()
        cmats = (NsimMaterial * len(meshes))()
        for k, ((coords, tets), material) in enumerate(zip(meshes, materials)):
            c = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 3)
            t = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
            self._keep += [c, t]
            cmeshes[k] = NsimMesh(len(c), len(t), _dptr(c), t.ctypes.data_as(_ip))
            cmats[k] = NsimMaterial(*[float(material[key])
                                      for key in ("Ms", "negJ", "c1", "c2", "c3", "mu0")])
        self._coords = (self._keep[0] if len(meshes) == 1
                        else np.concatenate(self._keep[0::2]))
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        pot = POTENTIAL[potential] if isinstance(potential, str) else int(potential)
        opt = NsimOptions(prec, 1 if do_demag else 0, nbc_rtol, dbc_rtol, nbc_maxits, dbc_maxits,
                          pot, float(compress_eps), float(compress_eta))
        h = _vp()
        if len(meshes) == 1:
            _check(L.nsim_create(cmeshes, cmats, ctypes.byref(opt), ctypes.byref(h)),
                   "nsim_create")
        else:
            _check(L.nsim_create_batch(len(meshes), cmeshes, cmats, ctypes.byref(opt),
                                       ctypes.byref(h)), "nsim_create_batch")
        self._h = h
        s = np.zeros(6, dtype=np.int32)
        _check(L.nsim_get_sizes(h, s.ctypes.data_as(_ip)), "nsim_get_sizes")
        self.n, self.nb, self.ntri, self.ntet, self.slots = [int(v) for v in s[:5]]
        self.nnz = int(s[5])
        self.precision = precision
        self.n_magnets = len(meshes)
        self.node_off = np.zeros(self.n_magnets + 1, dtype=np.int32)
        self.boundary_off = np.zeros(self.n_magnets + 1, dtype=np.int32)
        _check(L.nsim_get_batch_layout(h, self.node_off.ctypes.data_as(_ip),
                                       self.boundary_off.ctypes.data_as(_ip)),
               "nsim_get_batch_layout")

    def close(self):
        if self._h is not None and _lib is not None:
            _lib.nsim_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    def _size(self, name):
        if name in self.VEC3:
            return (self.n, 3)
        if name == "solid_angles" or name == "boundary_nodes":
            return (self.nb,)
        return (self.n,)

    def set_field(self, name, data):
        shape = self._size(name)
        a = np.asarray(data, dtype=np.float64)
        if a.size == int(np.prod(shape)):
            a = a.reshape(shape)
        a = np.ascontiguousarray(np.broadcast_to(a, shape))
        _check(lib().nsim_set_field(self._h, name.encode(), _dptr(a)), "nsim_set_field(%s)" % name)

    def get_field(self, name):
        out = np.zeros(self._size(name))
        _check(lib().nsim_get_field(self._h, name.encode(), _dptr(out)), "nsim_get_field(%s)" % name)
        return out

    def execute(self, program="update_dmdt"):
        _check(lib().nsim_execute(self._h, program.encode()), "nsim_execute(%s)" % program)

    def rhs(self, y, t=0.0):
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
        yd = np.zeros_like(y)
        _check(lib().nsim_rhs(self._h, t, y.ctypes.data, yd.ctypes.data, 0), "nsim_rhs")
        return yd

    def rhs_device(self, y_ptr, ydot_ptr, t=0.0):
        _check(lib().nsim_rhs(self._h, t, y_ptr, ydot_ptr, 1), "nsim_rhs")

    def jtimes(self, y, v, t=0.0):
        """J v with the Jacobian of the last re-Jacobi (psetup), or, before any
        psetup, at y with H_total of the last rhs (y may then not be None)."""
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        yp = None
        if y is not None:
            y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
            yp = y.ctypes.data
        jv = np.zeros_like(v)
        _check(lib().nsim_jtimes(self._h, t, v.ctypes.data, jv.ctypes.data, yp, None, 0),
               "nsim_jtimes")
        return jv

    def bem(self):
        """Dense BEM matrix; for a batch, the list of the magnets' matrices."""
        sizes = np.diff(self.boundary_off)
        flat = np.zeros(int(np.sum(sizes.astype(np.int64) ** 2)))
        _check(lib().nsim_get_bem(self._h, _dptr(flat)), "nsim_get_bem")
        mats, o = [], 0
        for nb in sizes:
            mats.append(flat[o:o + nb * nb].reshape(nb, nb))
            o += nb * nb
        return mats[0] if self.n_magnets == 1 else mats

    def potential(self):
        """Dense potential operator M (phi = M rho); a list for a batch."""
        sizes = np.diff(self.node_off)
        flat = np.zeros(int(np.sum(sizes.astype(np.int64) ** 2)))
        _check(lib().nsim_get_potential(self._h, _dptr(flat)), "nsim_get_potential")
        mats, o = [], 0
        for n in sizes:
            mats.append(flat[o:o + n * n].reshape(n, n))
            o += n * n
        return mats[0] if self.n_magnets == 1 else mats

    @property
    def dense_potential(self):
        return self.stat("potential") == 1.0

    @property
    def hmatrix_potential(self):
        return self.stat("potential") == 2.0

    def stat(self, name):
        v = ctypes.c_double(0)
        _check(lib().nsim_get_stat(self._h, name.encode(), ctypes.byref(v)), "nsim_get_stat")
        return v.value

    def timers(self, on=True):
        _check(lib().nsim_timers_enable(self._h, 1 if on else 0), "nsim_timers_enable")

    def timer(self, name):
        out = np.zeros(2)
        _check(lib().nsim_get_timer(self._h, name.encode(), _dptr(out)), "nsim_get_timer")
        return float(out[0]), int(out[1])

    def tune(self, key, value):
        _check(lib().nsim_set_tuning(self._h, key.encode(), int(value)), "nsim_set_tuning")

    def autotune(self, bem=True, cg=True, potential=True):
        """Time the tile shapes on the device and keep the fastest; returns the
        choice.  The CG leg solves the current m's load, so set m first."""
        what = (1 if bem else 0) | (2 if cg else 0) | (4 if potential else 0)
        _check(lib().nsim_autotune(self._h, what), "nsim_autotune")
        return {k: self.stat(k) for k in ("bem_rows", "bem_wg_per_cu", "bem_waves",
                                          "cg_rows_per_thread", "pot_rows", "pot_wg_per_cu",
                                          "pot_waves", "tuned_bem_us", "tuned_cg_us",
                                          "tuned_pot_us")}

    def synchronize(self):
        _check(lib().nsim_synchronize(self._h), "nsim_synchronize")

    def psetup(self, y, gamma, t=0.0, jok=False):
        """CVODE's psetup: returns jcur (1 if the Jacobian was rebuilt at y)."""
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
        jcur = ctypes.c_int(0)
        _check(lib().nsim_psetup(self._h, t, y.ctypes.data, None, 1 if jok else 0,
                                 ctypes.byref(jcur), gamma, 0), "nsim_psetup")
        return jcur.value

    def psolve_noalloc(self, r, gamma=0.0, t=0.0, lr=1):
        """psolve with its workspace in the library's static device scratch."""
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1)
        z = np.zeros_like(r)
        _check(lib().nsim_psolve_noalloc(self._h, t, None, None, r.ctypes.data, z.ctypes.data,
                                         gamma, 0.0, lr, 0), "nsim_psolve_noalloc")
        return z

    def set_tols(self, ksp_rtol=-1.0, ksp_atol=-1.0, ksp_dtol=-1.0, ksp_maxit=0,
                 cvode_rtol=-1.0, cvode_atol=-1.0):
        """lam_ts_set_tols: negative values (maxit <= 0) keep the setting."""
        _check(lib().nsim_ts_set_tols(self._h, ksp_rtol, ksp_atol, ksp_dtol, int(ksp_maxit),
                                      cvode_rtol, cvode_atol), "nsim_ts_set_tols")

    def psolve(self, r, gamma=0.0, t=0.0, lr=1):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(-1)
        z = np.zeros_like(r)
        _check(lib().nsim_psolve(self._h, t, None, None, r.ctypes.data, z.ctypes.data, gamma, 0.0,
                                 lr, 0), "nsim_psolve")
        return z

    def set_material(self, material, magnet=0):
        mat = NsimMaterial(*[float(material[k]) for k in ("Ms", "negJ", "c1", "c2", "c3", "mu0")])
        _check(lib().nsim_set_material(self._h, int(magnet), ctypes.byref(mat)),
               "nsim_set_material")

    def set_anisotropy(self, powers, coeffs, magnet=-1):
        """Polynomial anisotropy energy sum_j coeffs[j] m^powers[j] (simulation
        units) of one magnet (-1: all); empty arrays clear it."""
        P = np.ascontiguousarray(powers, dtype=np.int32).reshape(-1, 3)
        c = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1)
        if len(P) != len(c):
            raise ValueError("one exponent triple per coefficient")
        _check(lib().nsim_set_anisotropy(
            self._h, int(magnet), len(c), P.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "nsim_set_anisotropy")

    def rhs_vjp(self, y, lam, t=0.0):
        """J(y)^T lam and the parameter cotangents (K1 per node, Ms, H_ext)."""
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(-1)
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1)
        yb = np.zeros_like(y)
        k1b = np.zeros(self.n)
        hb = np.zeros_like(y)
        ms = ctypes.c_double(0.0)
        _check(lib().nsim_rhs_vjp(self._h, t, _dptr(y), _dptr(lam), _dptr(yb), _dptr(k1b),
                                  ctypes.byref(ms), _dptr(hb)), "nsim_rhs_vjp")
        return {"y": yb, "K1": k1b, "Ms": ms.value, "H_ext": hb}

    @staticmethod
    def _traj(dt, n_steps, checkpoint_every, n_pulses, pulse_len, pulse_dir):
        d = (ctypes.c_double * 3)(*[float(v) for v in pulse_dir])
        return NsimTraj(float(dt), int(n_steps), int(checkpoint_every), int(n_pulses),
                        float(pulse_len), d)

    def traj_forward(self, dt, n_steps, checkpoint_every=50, amps=(), pulse_len=1.0,
                     pulse_dir=(1.0, 0.0, 0.0)):
        """Fixed-step RK4 from the current m with pulses amps[w] * pulse_dir in
        window w; m becomes m(n_steps * dt)."""
        a = np.ascontiguousarray(amps, dtype=np.float64).reshape(-1)
        tr = self._traj(dt, n_steps, checkpoint_every, len(a), pulse_len, pulse_dir)
        _check(lib().nsim_traj_forward(self._h, ctypes.byref(tr),
                                       _dptr(a) if len(a) else None), "nsim_traj_forward")

    def traj_adjoint(self, m_bar_final, dt, n_steps, checkpoint_every=50, amps=(),
                     pulse_len=1.0, pulse_dir=(1.0, 0.0, 0.0)):
        """Discrete adjoint of the last traj_forward: cotangents of m(0), K1 per
        node, Ms and the pulse amplitudes for dLoss/dm(T) = m_bar_final."""
        a = np.ascontiguousarray(amps, dtype=np.float64).reshape(-1)
        tr = self._traj(dt, n_steps, checkpoint_every, len(a), pulse_len, pulse_dir)
        mb = np.ascontiguousarray(m_bar_final, dtype=np.float64).reshape(-1)
        m0b = np.zeros_like(mb)
        k1b = np.zeros(self.n)
        ms = ctypes.c_double(0.0)
        ab = np.zeros(max(len(a), 1))
        _check(lib().nsim_traj_adjoint(self._h, ctypes.byref(tr), _dptr(a) if len(a) else None,
                                       _dptr(mb), _dptr(m0b), _dptr(k1b), ctypes.byref(ms),
                                       _dptr(ab)), "nsim_traj_adjoint")
        return {"m0": m0b.reshape(-1, 3), "K1": k1b, "Ms": ms.value, "amps": ab[:len(a)]}

    def bdf_forward(self, t_final, rtol=1e-6, atol=1e-6, max_order=2, max_steps=100000,
                    amps=(), pulse_len=1.0, pulse_dir=(1.0, 0.0, 0.0)):
        """BDF trajectory from the current m at t = 0 to t_final, every accepted
        step recorded for the adjoint; returns the number of steps."""
        a = np.ascontiguousarray(amps, dtype=np.float64).reshape(-1)
        d = (ctypes.c_double * 3)(*[float(v) for v in pulse_dir])
        tr = NsimBdfTraj(float(t_final), float(rtol), float(atol), int(max_order),
                         int(max_steps), len(a), float(pulse_len), d)
        k = ctypes.c_int(0)
        _check(lib().nsim_bdf_forward(self._h, ctypes.byref(tr), _dptr(a) if len(a) else None,
                                      ctypes.byref(k)), "nsim_bdf_forward")
        self._bdf_pulses = len(a)
        return k.value

    def bdf_steps(self):
        """(end times, step sizes, orders) of the recorded steps."""
        k = lib().nsim_bdf_steps(self._h, 0, None, None, None)
        _check(0 if k >= 0 else k, "nsim_bdf_steps")
        t, hs, q = np.zeros(max(k, 1)), np.zeros(max(k, 1)), np.zeros(max(k, 1), dtype=np.int32)
        lib().nsim_bdf_steps(self._h, k, _dptr(t), _dptr(hs), q.ctypes.data_as(_ip))
        return t[:k], hs[:k], q[:k]

    def bdf_replay(self, amps=(), newton_rtol=1e-12):
        """Re-integrate the recorded step sequence from the current m with the
        current parameters; m becomes the final state."""
        a = np.ascontiguousarray(amps, dtype=np.float64).reshape(-1)
        _check(lib().nsim_bdf_replay(self._h, _dptr(a) if len(a) else None, float(newton_rtol)),
               "nsim_bdf_replay")

    def bdf_adjoint(self, m_bar_final, amps=(), lin_rtol=1e-10):
        """Discrete adjoint of the last bdf_forward / bdf_replay: cotangents of
        m(0), K1 per node, Ms and the pulse amplitudes for dLoss/dm(T)."""
        a = np.ascontiguousarray(amps, dtype=np.float64).reshape(-1)
        mb = np.ascontiguousarray(m_bar_final, dtype=np.float64).reshape(-1)
        m0b = np.zeros_like(mb)
        k1b = np.zeros(self.n)
        ms = ctypes.c_double(0.0)
        ab = np.zeros(max(len(a), 1))
        _check(lib().nsim_bdf_adjoint(self._h, _dptr(a) if len(a) else None, _dptr(mb),
                                      _dptr(m0b), _dptr(k1b), ctypes.byref(ms), _dptr(ab),
                                      float(lin_rtol)), "nsim_bdf_adjoint")
        return {"m0": m0b.reshape(-1, 3), "K1": k1b, "Ms": ms.value, "amps": ab[:len(a)]}

    def ts_init(self, t0=0.0, rtol=1e-6, atol=1e-6, max_order=2, krylov_max=300):
        _check(lib().nsim_ts_init(self._h, t0, rtol, atol, max_order, krylov_max), "nsim_ts_init")

    def advance(self, t_final, max_steps=-1, exact_tstop=True):
        """Integrate towards t_final (simulation time units); returns t reached."""
        tr = ctypes.c_double(0)
        rc = lib().nsim_advance(self._h, t_final, int(max_steps), 1 if exact_tstop else 0,
                                ctypes.byref(tr))
        _check(rc, "nsim_advance")
        return tr.value


class NativeBatch(NativeMagnet):
    """Many independent magnets in one handle (nsim_create_batch).

    Field vectors are the concatenation of the magnets' vectors in the order
    given; `node_off` / `boundary_off` are the instance-offset table, and
    `split` cuts a concatenated array into per-magnet pieces.
    """

    def __init__(self, meshes, materials, precision="fp64", do_demag=True, nbc_rtol=1e-5,
                 dbc_rtol=1e-5, nbc_maxits=10000, dbc_maxits=10000, potential="auto",
                 compress_eps=0.0, compress_eta=0.0):
        self._create(list(meshes), list(materials), precision, do_demag, nbc_rtol, dbc_rtol,
                     nbc_maxits, dbc_maxits, potential, compress_eps, compress_eta)

    def split(self, a):
        a = np.asarray(a)
        return [a[self.node_off[k]:self.node_off[k + 1]] for k in range(self.n_magnets)]

    def magnet_of_node(self):
        return np.repeat(np.arange(self.n_magnets), np.diff(self.node_off))
