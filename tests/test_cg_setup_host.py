"""The persistent CG's operator set-up on the host (no GPU):
nsim_cg_setup.cpp compiled with a small harness, checked against numpy.

* The coarse inverses: the Neumann Galerkin matrix P^T K P is singular along
  the constants of every connected set of aggregates; the set-up forms its
  pseudo-inverse by deflating them and must agree with numpy's pseudo-inverse
  of the same matrix (assembled here from the oracle's stiffness and the
  aggregate map read off the row order).  The Dirichlet one is the inverse
  over the free rows (zero where an aggregate has no free node).
* The fine coarse space: every part's aggregate list starts with its own
  sub-parts and holds every aggregate its rows' columns touch, so the kernel,
  which evaluates only those rows of the inverse, never reads one it did not
  form.
* The number of sub-parts follows the rule (at most 4, at least 64 rows
  each, at most 320 aggregates), else runs of whole parts (at most 64)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sps

import oracle
from nmag import kuhn_cuboid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nmag-src_amd", "csrc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("cgop")
    so = d / "cgop.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", CSRC,
                    os.path.join(ROOT, "tests", "cgop", "cgop_harness.cpp"),
                    os.path.join(CSRC, "nsim_cg_setup.cpp"), os.path.join(CSRC, "nsim_setup.cpp"),
                    "-o", str(so)], check=True, capture_output=True, text=True)
    lib = ctypes.CDLL(str(so))
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.cg_build.argtypes = [ctypes.c_int, dp, ctypes.c_int, ip, ctypes.c_int, ctypes.c_int, ip,
                             ctypes.c_char_p, ctypes.c_int]
    lib.cg_fetch.argtypes = [ip, dp, dp, ip, ip]
    return lib


def build(lib, coords, tets, cap, fine):
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    xyz = np.ascontiguousarray(coords, dtype=np.float64)
    tt = np.ascontiguousarray(tets, dtype=np.int32)
    info = np.zeros(7, dtype=np.int32)
    msg = ctypes.create_string_buffer(256)
    rc = lib.cg_build(len(xyz), xyz.ctypes.data_as(dp), len(tt), tt.ctypes.data_as(ip), cap,
                      1 if fine else 0, info.ctypes.data_as(ip), msg, 256)
    assert rc == 0, msg.value
    parts, npad, na, sub_shift, agg_shift, need_max, need_lg = (int(v) for v in info)
    p2o = np.zeros(npad, dtype=np.int32)
    An = np.zeros(na * na)
    Ad = np.zeros(na * na)
    fine_space = sub_shift > 0
    kneed = 64
    need = np.zeros(parts * kneed if fine_space else 1, dtype=np.int32)
    need_n = np.zeros(parts if fine_space else 1, dtype=np.int32)
    lib.cg_fetch(p2o.ctypes.data_as(ip), An.ctypes.data_as(dp), Ad.ctypes.data_as(dp),
                 need.ctypes.data_as(ip) if fine_space else None,
                 need_n.ctypes.data_as(ip) if fine_space else None)
    return dict(parts=parts, npad=npad, na=na, sub_shift=sub_shift, agg_shift=agg_shift,
                need_max=need_max, need_lg=need_lg, p2o=p2o, An=An.reshape(na, na),
                Ad=Ad.reshape(na, na), need=need.reshape(-1, kneed) if fine_space else None,
                need_n=need_n, cap=npad // parts)


def aggregate_map(op, n):
    agg = np.full(n, -1)
    rows = np.nonzero(op["p2o"] >= 0)[0]
    if op["sub_shift"] > 0:
        agg[op["p2o"][rows]] = rows // (op["cap"] >> op["sub_shift"])
    else:
        agg[op["p2o"][rows]] = (rows // op["cap"]) >> op["agg_shift"]
    assert (agg >= 0).all()
    return agg


def two_magnets(nx, ny, nz, h):
    c, t = kuhn_cuboid(nx, ny, nz, h)
    c2 = c + np.array([10 * nx * h, 0.0, 0.0])
    return np.vstack([c, c2]), np.vstack([t, t + len(c)])


@pytest.mark.parametrize("mesh,cap,fine,subs", [
    ((80, 40, 5, 2.0, False), 256, True, 4),    # the 20k-node slab: 78 parts x 4
    ((80, 40, 5, 2.0, False), 512, True, 4),    # 39 parts x 4
    ((80, 40, 5, 2.0, False), 256, False, 1),   # 39 aggregates of two parts
    ((30, 30, 30, 2.0, False), 256, True, 2),   # 117 parts x 2
    ((40, 20, 2, 2.5, True), 256, True, 4),     # two disjoint magnets
    ((12, 12, 4, 2.5, True), 256, False, 1),
    ((1, 1, 1, 2.5, False), 256, True, 4),      # one part of 8 nodes
])
def test_cg_coarse_space_and_inverses(harness, mesh, cap, fine, subs):
    nx, ny, nz, h, two = mesh
    coords, tets = two_magnets(nx, ny, nz, h) if two else kuhn_cuboid(nx, ny, nz, h)
    n = len(coords)
    op = build(harness, coords, tets, cap, fine)
    assert op["parts"] == (n + cap - 1) // cap
    assert 1 << op["sub_shift"] == (subs if subs > 1 else 1)
    if subs > 1:
        assert op["na"] == op["parts"] * subs <= 320
        assert (cap >> op["sub_shift"]) >= 64
    else:
        assert op["na"] <= 64
    # every node appears once in the row order
    rows = op["p2o"][op["p2o"] >= 0]
    assert np.array_equal(np.sort(rows), np.arange(n))

    o = oracle.OracleMesh(coords, tets, build_bem=False)
    rp, ci, val = o.stiffness()
    K = sps.csr_matrix((val, ci, rp), shape=(n, n))
    agg = aggregate_map(op, n)
    P = sps.csr_matrix((np.ones(n), (np.arange(n), agg)), shape=(n, op["na"]))
    Ac = (P.T @ K @ P).toarray()
    ref = np.linalg.pinv(Ac, rcond=1e-9, hermitian=True)
    scale = np.abs(ref).max()
    assert np.abs(op["An"] - ref).max() <= 1e-10 * scale
    assert np.abs(op["An"] - op["An"].T).max() <= 1e-12 * scale
    # the null directions get no correction: A^+ applied to the constant of
    # each connected set of aggregates vanishes (a magnet's aggregates, or all
    # of them when an aggregate straddles the two magnets)
    for nodes in ((np.arange(n // 2), np.arange(n // 2, n)) if two else (np.arange(n),)):
        sets = np.unique(agg[nodes])
        if two and len(np.intersect1d(np.unique(agg[: n // 2]), np.unique(agg[n // 2:]))):
            sets = np.arange(op["na"])
        comp = np.zeros(op["na"])
        comp[sets] = 1.0
        assert np.abs(op["An"] @ comp).max() <= 1e-10 * scale * np.sqrt(op["na"])

    # Dirichlet: free rows (interior nodes) only
    bn = np.asarray(o.boundary()[0], dtype=int)
    free = np.ones(n, dtype=bool)
    free[bn] = False
    Kf = K[free][:, free]
    Pf = P[free]
    Ad = (Pf.T @ Kf @ Pf).toarray()
    has = np.asarray(Pf.sum(axis=0)).ravel() > 0
    refd = np.zeros_like(Ad)
    refd[np.ix_(has, has)] = np.linalg.inv(Ad[np.ix_(has, has)])
    assert np.abs(op["Ad"] - refd).max() <= 1e-10 * np.abs(refd).max()

    if subs > 1:
        S = subs
        assert op["need_max"] <= 64 and (op["need_max"] << op["need_lg"]) <= 256
        assert op["need_lg"] <= 6
        for g in range(op["parts"]):
            nd = op["need"][g, : op["need_n"][g]]
            assert list(nd[:S]) == list(range(g * S, g * S + S))
            assert len(set(nd)) == len(nd) and (nd >= 0).all() and (nd < op["na"]).all()
            prow = op["p2o"][g * op["cap"]:(g + 1) * op["cap"]]
            prow = prow[prow >= 0]
            touched = set(agg[K[prow].indices]) if len(prow) else set()
            assert touched <= set(nd.tolist())
            assert op["need_n"][g] <= op["need_max"]
