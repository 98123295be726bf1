"""The 20k-node slab (160 x 80 x 10 nm, 2 nm Kuhn mesh, 19,926 nodes, 7,602
boundary nodes) at full size, through properties that hold at any size:
the two demag routes (dense potential operator; CG solves around the BEM
GEMV) are independent algorithms for the same field, the field is linear in
m, LLG conserves |m|, a uniform state has no exchange field and Aharoni's
demagnetising factor, and the mixed modes stay within their per-field
bounds."""
import numpy as np
import pytest

from helpers import material, rel_err
from nmag import kuhn_cuboid
from nmag._native import NativeMagnet
from test_oracle import aharoni_nz

pytestmark = pytest.mark.gpu

DIMS = (80, 40, 5, 2.0)


@pytest.fixture(scope="module")
def slab():
    coords, tets = kuhn_cuboid(*DIMS)
    mat = material(alpha=0.02)
    rng = np.random.default_rng(2024)
    # a smooth S-like state with seeded noise, as in the dynamics workload
    x = coords[:, 0] / coords[:, 0].max()
    m = np.stack([np.cos(np.pi * (x - 0.5)), np.sin(np.pi * (x - 0.5)) * 0.8,
                  0.1 * np.ones_like(x)], axis=1) + 0.02 * rng.standard_normal((len(x), 3))
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    return coords, tets, mat, m


def _make(slab, **kw):
    coords, tets, mat, _ = slab
    return NativeMagnet(coords, tets, mat.su_coeffs(), **kw)


def _fields(g, m):
    g.set_field("m", m)
    g.execute()
    return {k: g.get_field(k) for k in ("H_exch", "H_demag", "dmdt", "phi")}


def test_full_size_routes_agree_and_invariants_hold(slab):
    coords, tets, mat, m = slab
    dense = _make(slab, potential="dense")
    assert dense.n == 19926 and dense.stat("potential") == 1
    fd = _fields(dense, m)
    it = _make(slab, potential="iterative", nbc_rtol=1e-12, dbc_rtol=1e-12,
               nbc_maxits=50000, dbc_maxits=50000)
    fi = _fields(it, m)
    it.close()
    # two independent algorithms: agreement at the CG tolerance
    assert rel_err(fi["H_demag"], fd["H_demag"]) < 1e-8
    assert rel_err(fi["dmdt"], fd["dmdt"]) < 1e-8
    # linear in m (exact up to rounding on the dense route)
    assert rel_err(-_fields(dense, -m)["H_demag"], fd["H_demag"]) < 1e-12
    # LLG: dm/dt is orthogonal to m at every node
    dots = np.abs(np.einsum("ij,ij->i", m, fd["dmdt"]))
    assert dots.max() <= 1e-12 * np.abs(fd["dmdt"]).max()
    # uniform state: no exchange field, Aharoni's factor along z
    mz = np.tile([0.0, 0.0, 1.0], (dense.n, 1))
    fu = _fields(dense, mz)
    assert np.abs(fu["H_exch"]).max() <= 1e-12 * np.abs(fd["H_exch"]).max()
    V = dense.get_field("volumes")
    Nz = -(fu["H_demag"][:, 2] * V).sum() / V.sum() / mat.su_Ms
    assert abs(Nz - aharoni_nz(160.0, 80.0, 10.0)) < 0.02
    dense.close()


@pytest.mark.parametrize("precision,tol_demag,tol_dmdt", [("mixed", 1e-6, 1e-6),
                                                          ("mixed24", 5e-5, 1e-5)])
def test_full_size_mixed_modes_within_bounds(slab, precision, tol_demag, tol_dmdt):
    """Bounds from the storage rounding: FP32 entries carry 2^-24, 24-bit
    entries 2^-16 (about 1.5e-5) of relative error."""
    _, _, _, m = slab
    ref = _make(slab, potential="dense")
    f64 = _fields(ref, m)
    ref.close()
    g = _make(slab, potential="dense", precision=precision)
    f = _fields(g, m)
    g.close()
    assert rel_err(f["H_exch"], f64["H_exch"]) < 1e-7
    assert rel_err(f["H_demag"], f64["H_demag"]) < tol_demag
    assert rel_err(f["dmdt"], f64["dmdt"]) < tol_dmdt
