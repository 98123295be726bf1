"""The H-matrix route of the demag potential (nsim_options.potential = 2):
M = (I + E B R) K^+ with dense near-field blocks and low-rank far-field
blocks, applied in two launches.

* FP64 storage at the default tolerance (1e-12 per block) against the oracle
  at the FP64 bar (1e-10), on one magnet and on a batch;
* FP32 storage at its default tolerance (1e-7 per block: the order of FP32
  rounding) against the FP64 dense route, per field, at the 20k-node slab;
* a magnet above the dense route's 20,480-node limit against the iterative
  route (the reference's structure) solved tightly;
* graph replay bit-identical to eager launches, and <m>(t) of a short
  integration on the compressed operator against the dense FP64 one.
"""
import numpy as np
import pytest

import oracle
from helpers import material, orc_params, random_m, rel_err
from nmag import kuhn_cuboid
from nmag._native import NativeBatch, NativeMagnet
from nmag.workloads import slab

pytestmark = pytest.mark.gpu


def _note(*a):
    print("[case]", *a, flush=True)


def test_fp64_hmatrix_matches_the_oracle():
    _note("H-matrix fp64, 12x12x4 cuboid")
    coords, tets = kuhn_cuboid(12, 12, 4, 2.5)
    mat = material()
    g = NativeMagnet(coords, tets, mat.su_coeffs(), potential="hmatrix")
    assert g.hmatrix_potential and g.stat("hm_eps") == 1e-12
    m = random_m(g.n, 61)
    H = np.tile([0.02, -0.01, 0.03], (g.n, 1))
    g.set_field("H_ext", H)
    g.set_field("m", m)
    g.execute()
    ref = oracle.OracleMesh(coords, tets).fields(orc_params(mat), m, H_ext=H, rtol=1e-13)
    for key in ("H_demag", "H_total", "dmdt"):
        assert rel_err(g.get_field(key), ref[key]) < 1e-10, key
    with pytest.raises(Exception):
        g.potential()  # the compressed operator has no dense copy


def test_fp64_hmatrix_batch_matches_the_oracle_per_magnet():
    _note("H-matrix fp64 batch of 3 magnets")
    dims = [(10, 6, 3, 2.5), (6, 6, 2, 2.5), (14, 5, 2, 2.0)]
    meshes = [kuhn_cuboid(*d) for d in dims]
    mats = [material(alpha=a) for a in (0.5, 0.3, 0.1)]
    b = NativeBatch(meshes, [m.su_coeffs() for m in mats], potential="hmatrix")
    assert b.hmatrix_potential
    ms = [random_m(len(c), 70 + k) for k, (c, _) in enumerate(meshes)]
    b.set_field("m", np.concatenate(ms))
    b.execute()
    got = b.split(b.get_field("H_demag"))
    for k, (c, t) in enumerate(meshes):
        ref = oracle.OracleMesh(c, t).fields(orc_params(mats[k]), ms[k], rtol=1e-13)
        assert rel_err(got[k], ref["H_demag"]) < 1e-10, k


@pytest.fixture(scope="module")
def slab_fields():
    coords, tets, mat, Hext, m0 = slab()
    out = {}
    for prec, route in (("fp64", "dense"), ("mixed", "hmatrix")):
        _note("slab %s/%s" % (prec, route))
        g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=prec, potential=route)
        g.set_field("H_ext", Hext)
        g.set_field("m", m0)
        g.execute()
        out[(prec, route)] = {k: g.get_field(k) for k in ("H_exch", "H_demag", "dmdt")}
        if route == "hmatrix":
            out["stats"] = {k: g.stat(k) for k in ("potential_bytes", "hm_eps", "hm_eta", "hm_max_rank",
                                                    "hm_lowrank_blocks", "hm_p1_panels",
                                                    "hm_p2_panels")}
        g.close()
    return out


def test_mixed_hmatrix_error_at_the_slab(slab_fields):
    """FP32 entries of the operator compressed at 1e-7 per block: per-field
    errors of the order of FP32 storage of the dense operator (4.8e-8
    H_demag, 1.8e-8 dm/dt measured)."""
    ref, got = slab_fields[("fp64", "dense")], slab_fields[("mixed", "hmatrix")]
    st = slab_fields["stats"]
    assert st["hm_eps"] == 1e-7 and st["hm_eta"] == 3.0
    assert st["potential_bytes"] < 0.3 * 4 * 19926 ** 2  # under 30 % of the FP32 dense bytes
    assert rel_err(got["H_exch"], ref["H_exch"]) < 1e-7
    assert rel_err(got["H_demag"], ref["H_demag"]) < 3e-7
    assert rel_err(got["dmdt"], ref["dmdt"]) < 1e-7


def test_graph_replay_is_bit_identical_on_the_hmatrix_route():
    _note("H-matrix graph replay, 20x10x3 slab")
    coords, tets = kuhn_cuboid(20, 10, 3, 2.0)
    g = NativeMagnet(coords, tets, material().su_coeffs(), precision="mixed",
                     potential="hmatrix")
    import torch
    rng = np.random.default_rng(5)
    Y = torch.from_numpy(np.stack([random_m(g.n, 80 + k).reshape(-1) for k in range(3)])).cuda()
    eager = []
    g.tune("rhs_graph", 0)
    for k in range(3):
        out = torch.zeros_like(Y[0])
        g.rhs_device(Y[k].data_ptr(), out.data_ptr())
        g.synchronize()
        eager.append(out.cpu().numpy())
    g.tune("rhs_graph", 1)
    out = torch.zeros_like(Y[0])
    for rep in range(4):  # eager, capture, then replays
        for k in range(3):
            g.rhs_device(Y[k].data_ptr(), out.data_ptr())
            g.synchronize()
            assert np.array_equal(out.cpu().numpy(), eager[k]), (rep, k)
    assert g.stat("graph_launches") > 0
    del rng


def test_hmatrix_above_the_dense_limit_matches_the_iterative_route():
    """25,116 nodes: beyond the dense operator's 20,480-node staging limit the
    compressed operator still applies; against the reference's structure
    (two Poisson solves around the BEM) solved to 1e-12."""
    coords, tets = kuhn_cuboid(90, 45, 5, 2.0)
    assert len(coords) > 20480
    mat = material(alpha=0.02)
    m = random_m(len(coords), 90)
    H = {}
    for route, prec in (("iterative", "fp64"), ("hmatrix", "mixed")):
        _note("%d nodes, %s/%s" % (len(coords), prec, route))
        g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=prec, potential=route,
                         nbc_rtol=1e-12, dbc_rtol=1e-12, nbc_maxits=20000, dbc_maxits=20000)
        g.set_field("m", m)
        g.execute()
        H[route] = g.get_field("H_demag")
        g.close()
    assert rel_err(H["hmatrix"], H["iterative"]) < 3e-7
    with pytest.raises(Exception):
        NativeMagnet(coords, tets, mat.su_coeffs(), precision="mixed", potential="dense")


def test_hmatrix_integration_follows_the_dense_fp64_route():
    """20 ps of the BDF integrator at rtol 1e-6 on a 40x20x3 slab from the
    S-state: <m>(t) on the mixed compressed operator within 1e-5 of FP64
    dense at every 2 ps."""
    coords, tets, mat, Hext, m0 = slab(dims=(40, 20, 3))
    avgs = {}
    for prec, route in (("fp64", "dense"), ("mixed", "hmatrix")):
        _note("integration %s/%s" % (prec, route))
        g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=prec, potential=route)
        g.set_field("H_ext", Hext)
        g.set_field("m", m0)
        g.ts_init(0.0, 1e-6, 1e-6)
        V = g.get_field("volumes")
        out, t = [], 0.0
        while t < 20.0 - 1e-9:
            t = g.advance(t + 2.0)
            out.append((g.get_field("m") * V[:, None]).sum(0) / V.sum())
        g.close()
        avgs[route] = np.array(out)
    assert np.max(np.abs(avgs["hmatrix"] - avgs["dense"])) < 1e-5


def test_options_are_checked():
    coords, tets = kuhn_cuboid(2, 2, 1, 2.5)
    for kw in ({"potential": 3}, {"potential": "hmatrix", "compress_eps": 1.5},
               {"potential": "hmatrix", "compress_eps": -1.0},
               {"potential": "hmatrix", "compress_eta": -1.0},
               {"potential": "hmatrix", "compress_eta": 1e3}):
        with pytest.raises(Exception):
            NativeMagnet(coords, tets, material().su_coeffs(), **kw)
