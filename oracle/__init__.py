"""CPU oracle for the nsim LLG right-hand-side path -- TEST INFRASTRUCTURE ONLY.

FP64 restatement of the reference algorithm (nsim_oracle.c; every function cites
the nmag-src file:line it follows).  Used only by tests/, the smoke check and the
CPU-baseline leg of bench.py, always as the checker and never as the measured or
shipped path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libnsim_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrcParams(ctypes.Structure):
    _fields_ = [("Ms", ctypes.c_double), ("negJ", ctypes.c_double),
                ("c1", ctypes.c_double), ("c2", ctypes.c_double),
                ("c3", ctypes.c_double), ("mu0", ctypes.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_setup.restype = ctypes.c_void_p
        L.orc_setup.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, ctypes.c_int]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_sizes.argtypes = [ctypes.c_void_p, _ip]
        L.orc_get_volumes.argtypes = [ctypes.c_void_p, _dp]
        L.orc_get_stiffness.argtypes = [ctypes.c_void_p, _ip, _ip, _dp]
        L.orc_get_boundary.argtypes = [ctypes.c_void_p, _ip, _ip, _dp, _dp]
        L.orc_get_bem.argtypes = [ctypes.c_void_p, _dp]
        L.orc_kuhn_cuboid.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, _dp, _ip]
        L.orc_lindholm.argtypes = [_dp] * 5 + [ctypes.c_double, _dp]
        L.orc_lindholm_acos.argtypes = [_dp] * 5 + [ctypes.c_double, _dp]
        L.orc_vertex_solid_angle.restype = ctypes.c_double
        L.orc_vertex_solid_angle.argtypes = [_dp] * 4
        L.orc_fields.argtypes = ([ctypes.c_void_p, ctypes.POINTER(OrcParams)] + [_dp] * 5
                                 + [ctypes.c_double, ctypes.c_int, ctypes.c_int]
                                 + [_dp] * 8 + [_ip])
        L.orc_apply_K.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_div.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, _dp]
        L.orc_grad.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_cg.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_void_p, ctypes.c_double,
                             ctypes.c_int, _dp]
        L.orc_jtimes.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcParams)] + [_dp] * 4
        L.orc_poly_anis.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, ctypes.c_double,
                                    _dp]
        _lib = L
    return _lib


def poly_anis(m, powers, coeffs, s):
    """H = -s (g - m (m.g)), g the gradient of sum_j c_j m^p_j (orc_poly_anis)."""
    m = np.ascontiguousarray(m, dtype=np.float64).reshape(-1, 3)
    P = np.ascontiguousarray(powers, dtype=np.int32).reshape(-1, 3)
    c = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1)
    H = np.zeros_like(m)
    lib().orc_poly_anis(len(m), _p(m), len(c), _p(P, np.int32), _p(c), float(s), _p(H))
    return H


def _p(a, dtype=np.float64):
    if a is None:
        return None
    assert a.dtype == dtype and a.flags.c_contiguous
    return a.ctypes.data_as(_dp if dtype == np.float64 else _ip)


def set_threads(n):
    """OpenMP threads of the oracle from now on (n <= 0: all processors);
    returns the count in effect."""
    return int(lib().orc_set_threads(int(n)))


def kuhn_cuboid(nx, ny, nz, h):
    n = (nx + 1) * (ny + 1) * (nz + 1)
    coords = np.zeros((n, 3))
    tets = np.zeros((6 * nx * ny * nz, 4), dtype=np.int32)
    lib().orc_kuhn_cuboid(nx, ny, nz, h, _p(coords), _p(tets, np.int32))
    return coords, tets


def lindholm(observer, p0, p1, p2, normal, geo_eps=0.0, acos_form=False):
    out = np.zeros(3)
    args = [np.ascontiguousarray(v, dtype=np.float64) for v in (observer, p0, p1, p2, normal)]
    fn = lib().orc_lindholm_acos if acos_form else lib().orc_lindholm
    fn(*[_p(a) for a in args], geo_eps, _p(out))
    return out


def vertex_solid_angle(x, p0, p1, p2):
    args = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, p0, p1, p2)]
    return lib().orc_vertex_solid_angle(*[_p(a) for a in args])


class OracleMesh:
    """Assembled operators of one tetrahedral mesh (FP64, host)."""

    def __init__(self, coords, tets, build_bem=True, acos_form=False):
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.tets = np.ascontiguousarray(tets, dtype=np.int32)
        self.h = lib().orc_setup(len(self.coords), _p(self.coords), len(self.tets),
                                 _p(self.tets, np.int32),
                                 (2 if acos_form else 1) if build_bem else 0)
        s = np.zeros(5, dtype=np.int32)
        lib().orc_sizes(self.h, _p(s, np.int32))
        self.n, self.nnz, self.nb, self.ntri, self.ntet = [int(v) for v in s]
        self.has_bem = build_bem

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_free(self.h)
            self.h = None

    def volumes(self):
        v = np.zeros(self.n)
        lib().orc_get_volumes(self.h, _p(v))
        return v

    def stiffness(self):
        rp = np.zeros(self.n + 1, dtype=np.int32)
        ci = np.zeros(self.nnz, dtype=np.int32)
        val = np.zeros(self.nnz)
        lib().orc_get_stiffness(self.h, _p(rp, np.int32), _p(ci, np.int32), _p(val))
        return rp, ci, val

    def boundary(self):
        bn = np.zeros(self.nb, dtype=np.int32)
        tri = np.zeros((self.ntri, 3), dtype=np.int32)
        nrm = np.zeros((self.ntri, 3))
        sa = np.zeros(self.nb)
        lib().orc_get_boundary(self.h, _p(bn, np.int32), _p(tri, np.int32), _p(nrm), _p(sa))
        return bn, tri, nrm, sa

    def bem(self):
        assert self.has_bem
        B = np.zeros((self.nb, self.nb))
        lib().orc_get_bem(self.h, _p(B))
        return B

    def apply_K(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        lib().orc_apply_K(self.h, _p(x), _p(y))
        return y

    def div(self, Ms, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        rho = np.zeros(self.n)
        lib().orc_div(self.h, Ms, _p(m), _p(rho))
        return rho

    def grad(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        H = np.zeros((self.n, 3))
        lib().orc_grad(self.h, _p(phi), _p(H))
        return H

    def cg(self, b, x0, fixed=None, rtol=1e-12, maxit=100000):
        x = np.array(x0, dtype=np.float64)
        res = ctypes.c_double(0)
        fx = None
        if fixed is not None:
            fixed = np.ascontiguousarray(fixed, dtype=np.uint8)
            fx = fixed.ctypes.data_as(ctypes.c_void_p)
        b = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        it = lib().orc_cg(self.h, _p(b), _p(x), fx, rtol, maxit, ctypes.byref(res))
        return x, it, res.value

    def fields(self, params, m, H_ext=None, k1=None, k2=None, axis=None, rtol=1e-13,
               maxit=100000, do_demag=True, phi1=None, phi2=None, poly=None):
        """All effective-field contributions and dm/dt for magnetisation m (n,3).
        poly = (powers (k,3), coeffs (k,)): a polynomial anisotropy energy
        (SU), its field added to H_anis and H_total."""
        n = self.n
        m = np.ascontiguousarray(m, dtype=np.float64).reshape(n, 3)
        if poly is not None:
            Hp = poly_anis(m, poly[0], poly[1], 1.0 / (params.mu0 * params.Ms))
            He0 = np.zeros((n, 3)) if H_ext is None else np.broadcast_to(
                np.asarray(H_ext, dtype=np.float64), (n, 3))
            out = self.fields(params, m, He0 + Hp, k1, k2, axis, rtol, maxit, do_demag,
                              phi1, phi2)
            out["H_anis"] = out["H_anis"] + Hp
            return out
        out = {k: np.zeros((n, 3)) for k in ("H_exch", "H_demag", "H_anis", "H_total", "dmdt")}
        out["rho"] = np.zeros(n)
        p1 = np.zeros(n) if phi1 is None else np.array(phi1, dtype=np.float64)
        p2 = np.zeros(n) if phi2 is None else np.array(phi2, dtype=np.float64)
        it = np.zeros(2, dtype=np.int32)
        cvt = lambda a, shape: None if a is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(a, dtype=np.float64), shape))
        He, K1, K2, ax = cvt(H_ext, (n, 3)), cvt(k1, (n,)), cvt(k2, (n,)), cvt(axis, (n, 3))
        lib().orc_fields(self.h, ctypes.byref(params), _p(m), _p(He), _p(K1), _p(K2), _p(ax),
                         rtol, maxit, 1 if do_demag else 0, _p(p1), _p(p2), _p(out["rho"]),
                         _p(out["H_exch"]), _p(out["H_demag"]), _p(out["H_anis"]),
                         _p(out["H_total"]), _p(out["dmdt"]), _p(it, np.int32))
        out["phi1"], out["phi2"], out["phi"] = p1, p2, p1 + p2
        out["iters"] = (int(it[0]), int(it[1]))
        return out

    def jtimes(self, params, m, H_total, v):
        m = np.ascontiguousarray(m, dtype=np.float64)
        H = np.ascontiguousarray(H_total, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        Jv = np.zeros_like(v)
        lib().orc_jtimes(self.h, ctypes.byref(params), _p(m), _p(H), _p(v), _p(Jv))
        return Jv


def params_from_si(Ms=8.6e5, A=1.3e-11, alpha=0.5, gamma=2.211e5, normfac=0.1e12):
    """Simulation-unit LLG coefficients (interface/nmag/main.py:423-471)."""
    mu0 = 4e-7 * np.pi / 1e-6
    Ms_su = Ms * 1e-6
    A_su = A * 1e12
    g_su = gamma * 1e-6
    c1 = -g_su / (1.0 + alpha ** 2)
    return OrcParams(Ms=Ms_su, negJ=-2.0 * A_su / (mu0 * Ms_su), c1=c1, c2=c1 * alpha,
                     c3=normfac * 1e-12, mu0=mu0)
