"""Shared set-up of the parity tests: meshes, materials and seeded inputs."""
import numpy as np

import oracle
from nmag import MagMaterial, SI, kuhn_cuboid
from nmag._native import NativeMagnet

PERMALLOY = dict(Ms=SI(8.6e5, "A/m"), exchange_coupling=SI(1.3e-11, "J/m"),
                 llg_gamma_G=SI(2.211e5, "m/A s"))


def material(alpha=0.5, **kw):
    args = dict(PERMALLOY)
    args.update(kw)
    return MagMaterial("Py", llg_damping=alpha, **args)


def orc_params(mat):
    c = mat.su_coeffs()
    return oracle.OrcParams(Ms=c["Ms"], negJ=c["negJ"], c1=c["c1"], c2=c["c2"], c3=c["c3"],
                            mu0=c["mu0"])


def random_m(n, seed):
    """Normalised i.i.d. Gaussian unit vectors per node (seeded PCG64)."""
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, 3))
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def rel_err(a, b):
    """max-norm error of a relative to the max-norm of b."""
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (scale if scale > 0 else 1.0))


def make_pair(nx, ny, nz, h, mat=None, precision="fp64", rtol=1e-13, build_bem=True,
              potential="auto"):
    mat = mat or material()
    coords, tets = kuhn_cuboid(nx, ny, nz, h)
    orc = oracle.OracleMesh(coords, tets, build_bem=build_bem)
    gpu = NativeMagnet(coords, tets, mat.su_coeffs(), precision=precision, nbc_rtol=rtol,
                       dbc_rtol=rtol, nbc_maxits=20000, dbc_maxits=20000, potential=potential)
    return coords, tets, orc, gpu, mat
