"""Batching form: many independent magnets of uneven sizes in one handle.

Every magnet of a batch must give the fields that the oracle gives for that
magnet alone (FP64, Poisson solves to 1e-13), and the per-magnet BEM segments
must equal the matrices of single-magnet handles bit for bit.
"""
import numpy as np
import pytest

import oracle
from helpers import material, orc_params, random_m, rel_err
from nmag import SI, kuhn_cuboid
from nmag._native import NativeBatch, NativeMagnet

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10

# uneven magnets: (nx, ny, nz, h), Ms in A/m, damping
MAGNETS = [((6, 4, 2, 2.5), 8.6e5, 0.5), ((3, 3, 3, 2.5), 6.0e5, 0.3),
           ((10, 5, 2, 2.0), 1.0e6, 0.5), ((4, 2, 1, 2.5), 7.5e5, 0.02)]


def _batch_inputs(seed=3):
    rng = np.random.default_rng(seed)
    meshes, mats, fields = [], [], []
    for k, (dims, Ms, alpha) in enumerate(MAGNETS):
        coords, tets = kuhn_cuboid(*dims)
        mat = material(alpha=alpha, Ms=SI(Ms, "A/m"))
        n = len(coords)
        ax = rng.standard_normal(3)
        ax /= np.linalg.norm(ax)
        fields.append(dict(m=random_m(n, 100 + k),
                           H_ext=np.tile(rng.uniform(-0.03, 0.03, 3), (n, 1)),
                           K1=np.full(n, rng.uniform(0.0, 0.05)),
                           axis=np.tile(ax, (n, 1))))
        meshes.append((coords, tets))
        mats.append(mat)
    return meshes, mats, fields


def _cat(fields, key):
    return np.concatenate([f[key] for f in fields])


def _set_all(mag, fields):
    for key in ("m", "H_ext", "axis", "K1"):
        mag.set_field(key, _cat(fields, key) if isinstance(fields, list) else fields[key])


def _batch(meshes, mats, precision="fp64", rtol=1e-13):
    return NativeBatch(meshes, [m.su_coeffs() for m in mats], precision=precision,
                       nbc_rtol=rtol, dbc_rtol=rtol, nbc_maxits=20000, dbc_maxits=20000)


def test_layout_of_single_and_batch():
    meshes, mats, _ = _batch_inputs()
    b = _batch(meshes, mats)
    sizes = [len(c) for c, _ in meshes]
    assert b.n_magnets == len(MAGNETS)
    assert list(b.node_off) == list(np.concatenate([[0], np.cumsum(sizes)]))
    assert b.boundary_off[-1] == b.nb
    s = NativeMagnet(*meshes[0], mats[0].su_coeffs())
    assert list(s.node_off) == [0, s.n] and list(s.boundary_off) == [0, s.nb]


@pytest.mark.parametrize("precision", ["fp64", "mixed", "bf16"])
def test_bem_segments_equal_single_magnet_matrices(precision):
    meshes, mats, _ = _batch_inputs()
    b = _batch(meshes, mats, precision)
    segs = b.bem()
    assert len(segs) == len(MAGNETS)
    for k, (coords, tets) in enumerate(meshes):
        single = NativeMagnet(coords, tets, mats[k].su_coeffs(), precision=precision).bem()
        assert segs[k].shape == single.shape
        assert np.array_equal(segs[k], single)
        assert int(b.boundary_off[k + 1] - b.boundary_off[k]) == single.shape[0]


def test_batch_fields_match_oracle_per_magnet():
    meshes, mats, fields = _batch_inputs()
    b = _batch(meshes, mats)
    _set_all(b, fields)
    b.execute("update_dmdt")
    got = {k: b.split(b.get_field(k)) for k in ("H_exch", "H_demag", "H_anis", "H_total",
                                                 "dmdt", "rho")}
    for k, (coords, tets) in enumerate(meshes):
        f = fields[k]
        ref = oracle.OracleMesh(coords, tets).fields(orc_params(mats[k]), f["m"],
                                                     H_ext=f["H_ext"], k1=f["K1"],
                                                     axis=f["axis"], rtol=1e-13)
        assert rel_err(got["H_exch"][k], ref["H_exch"]) < 1e-12, k
        assert rel_err(got["rho"][k], ref["rho"]) < 1e-12, k
        assert rel_err(got["H_anis"][k], ref["H_anis"]) < 1e-13, k
        assert rel_err(got["H_demag"][k], ref["H_demag"]) < FP64_TOL, k
        assert rel_err(got["H_total"][k], ref["H_total"]) < FP64_TOL, k
        assert rel_err(got["dmdt"][k], ref["dmdt"]) < FP64_TOL, k
    assert b.stat("cg_unconverged") == 0


def test_batch_jtimes_and_preconditioner_match_single_magnets():
    meshes, mats, fields = _batch_inputs()
    b = _batch(meshes, mats)
    _set_all(b, fields)
    rng = np.random.default_rng(8)
    v = rng.standard_normal((b.n, 3))
    y = _cat(fields, "m")
    b.rhs(y.reshape(-1))
    jv = b.split(b.jtimes(y.reshape(-1), v.reshape(-1)).reshape(-1, 3))
    b.psetup(y.reshape(-1), gamma=0.05)
    # solved to near rounding, so that the batch's one solve and the
    # per-magnet solves agree (to pc_rtol 1e-2 they would stop differently)
    b.set_tols(ksp_rtol=1e-11, ksp_atol=0.0)
    z = b.split(b.psolve(v.reshape(-1)).reshape(-1, 3))
    vs = b.split(v)
    for k, (coords, tets) in enumerate(meshes):
        s = NativeMagnet(coords, tets, mats[k].su_coeffs(), nbc_rtol=1e-13, dbc_rtol=1e-13,
                         nbc_maxits=20000, dbc_maxits=20000)
        _set_all(s, fields[k])
        yk = fields[k]["m"].reshape(-1)
        s.rhs(yk)
        ref = s.jtimes(yk, vs[k].reshape(-1)).reshape(-1, 3)
        assert rel_err(jv[k], ref) < 1e-9, k
        s.psetup(yk, gamma=0.05)
        s.set_tols(ksp_rtol=1e-11, ksp_atol=0.0)
        zr = s.psolve(vs[k].reshape(-1)).reshape(-1, 3)
        assert rel_err(z[k], zr) < 1e-9, k
        assert s.stat("pc_unconverged") == 0
    assert b.stat("pc_unconverged") == 0


@pytest.mark.parametrize("prec,tol_demag,tol_dmdt", [("mixed", 1e-6, 1e-6),
                                                     ("mixed24", 5e-5, 1e-5)])
def test_batch_mixed_precision_error_bounds(prec, tol_demag, tol_dmdt):
    """FP32 (24-bit potential operator) storage in a batch of uneven magnets:
    per-magnet bounds from the storage rounding, 2^-24 (2^-16) per entry."""
    meshes, mats, fields = _batch_inputs()
    out = {}
    for p in ("fp64", prec):
        b = _batch(meshes, mats, p, rtol=1e-12)
        _set_all(b, fields)
        b.execute("update_dmdt")
        out[p] = {k: b.split(b.get_field(k)) for k in ("H_exch", "H_demag", "dmdt")}
    for k in range(len(MAGNETS)):
        assert rel_err(out[prec]["H_exch"][k], out["fp64"]["H_exch"][k]) < 1e-6, k
        assert rel_err(out[prec]["H_demag"][k], out["fp64"]["H_demag"][k]) < tol_demag, k
        assert rel_err(out[prec]["dmdt"][k], out["fp64"]["dmdt"][k]) < tol_dmdt, k


def test_batch_integration_matches_single_magnets():
    """One integrator over the batch (shared steps) vs each magnet alone."""
    meshes, mats, fields = _batch_inputs()
    b = _batch(meshes, mats, rtol=1e-10)
    _set_all(b, fields)
    b.ts_init(0.0, rtol=1e-8, atol=1e-8)
    t_end = 2.0  # ps
    assert abs(b.advance(t_end) - t_end) < 1e-12
    mb = b.split(b.get_field("m"))
    for k, (coords, tets) in enumerate(meshes):
        s = NativeMagnet(coords, tets, mats[k].su_coeffs(), nbc_rtol=1e-10, dbc_rtol=1e-10,
                         nbc_maxits=20000, dbc_maxits=20000)
        _set_all(s, fields[k])
        s.ts_init(0.0, rtol=1e-8, atol=1e-8)
        s.advance(t_end)
        ms = s.get_field("m")
        V = s.get_field("volumes")
        avg_b = (V[:, None] * mb[k]).sum(0) / V.sum()
        avg_s = (V[:, None] * ms).sum(0) / V.sum()
        assert np.max(np.abs(avg_b - avg_s)) < 1e-5, k
        assert np.max(np.abs(mb[k] - ms)) < 1e-4, k


def test_iterative_route_batch_matches_oracle_with_either_coarse_space():
    """The batch on the iterative route: the CG's sub-part aggregates (a few
    of them straddle two magnets, whose constants the coarse pseudo-inverse
    then shares) leave every magnet's demag field that of the oracle for the
    magnet alone, as the aggregates of whole parts do."""
    meshes, mats, fields = _batch_inputs()
    b = NativeBatch(meshes, [m.su_coeffs() for m in mats], precision="fp64", nbc_rtol=1e-13,
                    dbc_rtol=1e-13, nbc_maxits=20000, dbc_maxits=20000, potential="iterative")
    assert not b.dense_potential
    grid = b.stat("cg_grid")
    assert b.stat("cg_aggregates") in (2 * grid, 4 * grid)
    _set_all(b, fields)
    got = {}
    for fine in (1, 0):
        b.tune("cg_fine", fine)
        b.set_field("phi1", np.zeros(b.n))
        b.set_field("phi2", np.zeros(b.n))
        b.execute("update_dmdt")
        got[fine] = b.split(b.get_field("H_demag"))
    assert b.stat("cg_aggregates") <= grid
    for k, (coords, tets) in enumerate(meshes):
        f = fields[k]
        ref = oracle.OracleMesh(coords, tets).fields(orc_params(mats[k]), f["m"],
                                                     H_ext=f["H_ext"], k1=f["K1"],
                                                     axis=f["axis"], rtol=1e-13)
        for fine in (1, 0):
            assert rel_err(got[fine][k], ref["H_demag"]) < FP64_TOL, (k, fine)
    assert b.stat("cg_unconverged") == 0 and b.stat("cg_error") == 0


@pytest.mark.parametrize("route", ["iterative", "dense"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotuner_keeps_the_fields(precision, route):
    """Every tile shape the autotuner may pick computes the same fields."""
    meshes, mats, fields = _batch_inputs()
    b = NativeBatch(meshes, [m.su_coeffs() for m in mats], precision=precision,
                    nbc_rtol=1e-12, dbc_rtol=1e-12, nbc_maxits=20000, dbc_maxits=20000,
                    potential=route)
    assert b.dense_potential == (route == "dense")
    _set_all(b, fields)
    b.execute("update_dmdt")
    ref = {k: b.get_field(k) for k in ("H_demag", "dmdt")}
    choice = b.autotune()
    assert choice["bem_rows"] in (1, 2, 4) and 1 <= choice["bem_wg_per_cu"] <= 8
    assert choice["cg_rows_per_thread"] in (1, 2, 4)
    assert choice["tuned_bem_us"] > 0
    if route == "dense":
        assert choice["tuned_pot_us"] > 0 and choice["pot_waves"] in (4, 16)
    else:
        assert choice["tuned_cg_us"] > 0
    op = "pot" if route == "dense" else "bem"
    for r in (1, 4):
        for w in (1, 8):
            for v in (4, 16):
                b.tune(op + "_rows", r)
                b.tune(op + "_wg_per_cu", w)
                b.tune(op + "_waves", v)
                if route == "iterative":
                    b.set_field("phi1", np.zeros(b.n))
                    b.set_field("phi2", np.zeros(b.n))
                b.execute("update_dmdt")
                assert rel_err(b.get_field("H_demag"), ref["H_demag"]) < 1e-9, (r, w, v)
                assert rel_err(b.get_field("dmdt"), ref["dmdt"]) < 1e-9, (r, w, v)
    assert b.stat("cg_error") == 0 and b.stat("cg_unconverged") == 0
