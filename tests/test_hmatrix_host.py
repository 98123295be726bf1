"""The H-matrix compression of the demag potential operator, on the host
(no GPU): the builder (nsim_hmatrix_build.cpp) compiled with a small harness,
applied from its own panel format, against the dense operator.

The dense operator is M = (I + E B R) K^+ assembled in numpy from the
oracle's stiffness and BEM matrices (the construction the device set-up uses),
so the test pins both the compression error and the panel layout: every
entry of M must be reached exactly once, through a dense block or through the
U V^T of a low-rank block.  The tolerance is a relative Frobenius bound per
block, so the product's relative error stays near it."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sps

import oracle
from nmag import kuhn_cuboid
from nmag.workloads import s_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nmag-src_amd", "csrc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("hmat")
    so = d / "hmat.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread",
                    "-I", CSRC, os.path.join(ROOT, "tests", "hmat", "hmat_harness.cpp"),
                    os.path.join(CSRC, "nsim_hmatrix_build.cpp"), "-o", str(so)],
                   check=True, capture_output=True, text=True)
    lib = ctypes.CDLL(str(so))
    dp = ctypes.POINTER(ctypes.c_double)
    lib.hm_apply.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double, ctypes.c_double,
                             ctypes.c_int, ctypes.c_int, dp, dp,
                             ctypes.POINTER(ctypes.c_longlong), ctypes.c_char_p, ctypes.c_int]
    lib.hm_apply.restype = ctypes.c_int
    return lib


def potential_operator(coords, tets):
    o = oracle.OracleMesh(coords, tets)
    rp, ci, val = o.stiffness()
    n = len(coords)
    K = sps.csr_matrix((val, ci, rp), shape=(n, n)).toarray()
    bn = np.asarray(o.boundary()[0], dtype=int)
    B = o.bem()
    tau = np.trace(K) / n
    Kp = np.linalg.inv(K + tau / n) - 1.0 / (tau * n)
    inter = np.setdiff1d(np.arange(n), bn)
    W = B @ Kp[bn, :]
    M = Kp.copy()
    M[bn, :] += W
    M[inter, :] -= np.linalg.solve(K[np.ix_(inter, inter)], K[np.ix_(inter, bn)] @ W)
    return M, o


def apply(lib, M, coords, X, eps, eta=2.0, group=4):
    n = len(coords)
    A = np.asfortranarray(M)
    xyz = np.ascontiguousarray(coords, dtype=np.float64)
    X = np.ascontiguousarray(X.T)  # one vector after another
    Y = np.full_like(X, np.nan)
    stats = (ctypes.c_longlong * 12)()
    msg = ctypes.create_string_buffer(256)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.hm_apply(n, A.ctypes.data_as(dp), xyz.ctypes.data_as(dp), eps, eta, group,
                      X.shape[0], X.ctypes.data_as(dp), Y.ctypes.data_as(dp), stats, msg, 256)
    assert rc == 0, msg.value
    keys = ("dense", "lowrank", "full", "lr_blocks", "dense_blocks", "max_rank", "fallback",
            "nt", "p1", "p2", "p1_entries", "p2_entries")
    return Y.T, dict(zip(keys, list(stats)))


@pytest.fixture(scope="module")
def slab():
    coords, tets = kuhn_cuboid(30, 15, 3, 2.0)  # 1,984 nodes
    M, o = potential_operator(coords, tets)
    return coords, M, o


def test_every_entry_is_reached_once(harness, slab):
    """At a tolerance far below rounding, H x equals M x to rounding for unit
    vectors, so every column of M is covered by exactly one block."""
    coords, M, _ = slab
    n = len(coords)
    cols = np.random.default_rng(1).choice(n, 40, replace=False)
    X = np.zeros((n, len(cols)))
    X[cols, np.arange(len(cols))] = 1.0
    Y, st = apply(harness, M, coords, X, 1e-14)
    assert st["full"] == n * n
    err = np.abs(Y - M[:, cols]).max() / np.abs(M).max()
    assert err < 1e-12, err
    assert not np.isnan(Y).any()


@pytest.mark.parametrize("eps", [1e-12, 1e-9, 1e-7, 1e-6])
def test_product_error_follows_the_tolerance(harness, slab, eps):
    coords, M, o = slab
    m = s_state(coords, 2024)
    rho = o.div(0.86, m)
    X = np.stack([rho, np.random.default_rng(2).standard_normal(len(coords))], axis=1)
    Y, st = apply(harness, M, coords, X, eps)
    ref = M @ X
    err = np.abs(Y - ref).max(axis=0) / np.abs(ref).max(axis=0)
    print(eps, err, st)
    assert np.all(err < 20 * eps), (err, st)
    # storage of the compressed form below the dense one, and shrinking with eps
    assert st["dense"] + st["lowrank"] < st["full"]
    assert st["p2_entries"] >= st["dense"]
    assert st["lr_blocks"] > 0 and st["max_rank"] > 0


def test_group_width_does_not_change_the_product(harness, slab):
    coords, M, _ = slab
    x = np.random.default_rng(3).standard_normal((len(coords), 1))
    y4, _ = apply(harness, M, coords, x, 1e-9, group=4)
    y2, _ = apply(harness, M, coords, x, 1e-9, group=2)
    assert np.abs(y4 - y2).max() <= 1e-13 * np.abs(y4).max()
