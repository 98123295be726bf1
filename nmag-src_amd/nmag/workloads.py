"""Seeded synthetic workloads of the path's benchmark configurations.

The reference has no generator for these inputs (its only timing harness,
devel/bugs/serializing_bigmesh/bigbar_par.py:73-106, reads a mesh that is not
shipped), so they are defined here once, with numpy's PCG64, and shared by the
benchmark and the tests:

* the 20k-node permalloy slab: 160 x 80 x 10 nm, 2 nm Kuhn mesh, an in-plane
  S-state with 1 degree of seeded noise, 25 mT applied at 170 degrees;
* 64 independent cuboids of uneven sizes (the batching form);
* the 100 x 50 x 5 nm differentiation slab with per-element anisotropy and
  16 field pulses.
"""
import math

import numpy as np

from .material import MagMaterial
from .mesh import kuhn_cuboid
from .units import SI, mu0, to_su


def permalloy(alpha, Ms=8.6e5, name="Py"):
    return MagMaterial(name, Ms=SI(Ms, "A/m"), exchange_coupling=SI(1.3e-11, "J/m"),
                       llg_gamma_G=SI(2.211e5, "m/A s"), llg_damping=alpha)


def s_state(coords, seed, noise_deg=1.0):
    """In-plane S-state phi(u) = (pi/4) u^3, u = (x - xc)/(Lx/2), plus seeded
    isotropic noise of noise_deg degrees, renormalised."""
    x = coords[:, 0]
    u = (x - 0.5 * (x.min() + x.max())) / (0.5 * (x.max() - x.min()))
    phi = 0.25 * math.pi * u ** 3
    m = np.stack([np.cos(phi), np.sin(phi), np.zeros_like(phi)], axis=1)
    rng = np.random.default_rng(seed)
    m += math.radians(noise_deg) * rng.standard_normal(m.shape)
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def applied_field(n, mT=25.0, deg=170.0):
    """A uniform in-plane field of mT millitesla (B = mu0 H) at deg degrees, in
    simulation units, one row per node."""
    H = to_su(SI(mT * 1e-3, "T") / mu0, "A/m")
    return np.tile([H * math.cos(math.radians(deg)), H * math.sin(math.radians(deg)), 0.0],
                   (n, 1))


def slab(dims=(80, 40, 5), h=2.0, seed=2024):
    """The 20k-node slab: (coords, tets, material at alpha 0.02, H_ext, m0)."""
    coords, tets = kuhn_cuboid(*dims, h)
    return coords, tets, permalloy(0.02), applied_field(len(coords)), s_state(coords, seed)


def cuboid_batch(count=64, seed=4, h=2.5):
    """Independent permalloy-like cuboids: sides U[20,120] x U[20,80] x U[5,15] nm
    rounded to h-nm cubes, Ms U[6e5,1e6] A/m, uniaxial K U[0,5e4] J/m^3 along a
    random axis, applied field a random unit vector times U[0,50] mT, alpha
    0.5; m0 normalised i.i.d. Gaussian per node.  Returns one dict per magnet
    (dims, coords, tets, mat, fields in simulation units)."""
    rng = np.random.default_rng(seed)
    mags = []
    for k in range(count):
        L = rng.uniform([20.0, 20.0, 5.0], [120.0, 80.0, 15.0])
        nx, ny, nz = [max(1, int(round(v / h))) for v in L]
        Ms = rng.uniform(6e5, 1e6)
        K1 = rng.uniform(0.0, 5e4)
        ax = rng.standard_normal(3)
        ax /= np.linalg.norm(ax)
        hd = rng.standard_normal(3)
        hd /= np.linalg.norm(hd)
        H = hd * rng.uniform(0.0, 50.0) * 1e-3 / (4e-7 * math.pi)  # A/m
        coords, tets = kuhn_cuboid(nx, ny, nz, h)
        n = len(coords)
        mat = permalloy(0.5, Ms=Ms, name="Py%d" % k)
        m0 = rng.standard_normal((n, 3))
        m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
        fields = {"m": m0,
                  "H_ext": np.tile(float(to_su(SI(1.0, "A/m"), "A/m")) * H, (n, 1)),
                  "K1": np.full(n, float(to_su(SI(K1, "J/m^3"), "J/m^3"))),
                  "axis": np.tile(ax, (n, 1))}
        mags.append(dict(dims=(nx, ny, nz), coords=coords, tets=tets, mat=mat, fields=fields))
    return mags


def differentiation_slab(seed=55, dims=(40, 20, 2), h=2.5, n_pulses=16):
    """The differentiation workload: alpha 0.02, per-element K ~ N(2e4, 5e3)
    J/m^3 along z, n_pulses x-field pulse amplitudes ~ N(0, 5) mT, m0 along
    (1,1,0) with 1 degree of seeded noise.  Returns coords, tets, material,
    K per element, pulse amplitudes and m0 (simulation units)."""
    coords, tets = kuhn_cuboid(*dims, h)
    mat = permalloy(0.02)
    rng = np.random.default_rng(seed)
    n = len(coords)
    K = rng.normal(2e4, 5e3, len(tets)) * float(to_su(SI(1.0, "J/m^3"), "J/m^3"))
    mt = float(to_su(SI(1e-3 / (4e-7 * math.pi), "A/m"), "A/m"))
    amps = rng.normal(0.0, 5.0, n_pulses) * mt
    m0 = np.tile([1.0, 1.0, 0.0], (n, 1)) / math.sqrt(2.0)
    m0 += math.radians(1.0) * rng.standard_normal((n, 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    return coords, tets, mat, K, amps, m0
