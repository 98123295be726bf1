"""HIP path vs the CPU oracle on the same meshes and seeded inputs.

FP64 bar: H_eff within 1e-10 relative (max-norm) of the oracle, both solving
the Poisson problems to 1e-13.  Mixed precision is held to the per-field bounds
stated in each test.
"""
import numpy as np
import pytest

from helpers import make_pair, material, orc_params, random_m, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def _fields_gpu(gpu, m, H_ext=None, k1=None, axis=None):
    gpu.set_field("m", m)
    if H_ext is not None:
        gpu.set_field("H_ext", H_ext)
    if k1 is not None:
        gpu.set_field("axis", axis)
        gpu.set_field("K1", k1)
    gpu.execute("update_dmdt")
    return {k: gpu.get_field(k) for k in ("H_exch", "H_demag", "H_anis", "H_total", "dmdt",
                                          "rho", "phi")}


@pytest.mark.parametrize("dims", [(1, 1, 1, 2.5), (3, 2, 2, 2.0), (12, 12, 4, 2.5)])
def test_bem_matrix_matches_oracle(dims):
    nx, ny, nz, h = dims
    _, _, orc, gpu, _ = make_pair(nx, ny, nz, h)
    Bo, Bg = orc.bem(), gpu.bem()
    assert Bg.shape == Bo.shape
    assert np.max(np.abs(Bg - Bo)) <= 1e-12 * max(1.0, np.max(np.abs(Bo)))
    # Fredkin-Koehler invariant: B 1 = -1
    assert np.allclose(Bg.sum(axis=1), -1.0, atol=1e-10)


@pytest.mark.parametrize("route", ["iterative", "dense"])
@pytest.mark.parametrize("dims,seed", [((12, 12, 4, 2.5), 1), ((5, 3, 2, 2.0), 7),
                                       ((2, 2, 1, 2.5), 3), ((1, 1, 1, 2.5), 4)])
def test_fields_fp64_match_oracle(dims, seed, route):
    nx, ny, nz, h = dims
    _, _, orc, gpu, mat = make_pair(nx, ny, nz, h, potential=route)
    assert gpu.dense_potential == (route == "dense")
    m = random_m(orc.n, seed)
    P = orc_params(mat)
    ref = orc.fields(P, m, rtol=1e-13)
    got = _fields_gpu(gpu, m)
    assert rel_err(got["H_exch"], ref["H_exch"]) < 1e-12
    assert rel_err(got["rho"], ref["rho"]) < 1e-12
    assert rel_err(got["H_demag"], ref["H_demag"]) < FP64_TOL
    assert rel_err(got["H_total"], ref["H_total"]) < FP64_TOL
    assert rel_err(got["dmdt"], ref["dmdt"]) < FP64_TOL
    # phi is defined up to a constant (Neumann gauge): compare mean-free parts
    pg, po = got["phi"] - got["phi"].mean(), ref["phi"] - ref["phi"].mean()
    assert rel_err(pg, po) < FP64_TOL


def test_fields_with_anisotropy_and_applied_field():
    _, _, orc, gpu, mat = make_pair(8, 4, 2, 2.5)
    rng = np.random.default_rng(11)
    m = random_m(orc.n, 12)
    k1 = rng.normal(0.02, 0.005, orc.n)
    ax = np.tile([0.0, 0.0, 1.0], (orc.n, 1))
    H = np.tile([0.01, -0.02, 0.005], (orc.n, 1))
    P = orc_params(mat)
    ref = orc.fields(P, m, H_ext=H, k1=k1, axis=ax, rtol=1e-13)
    got = _fields_gpu(gpu, m, H_ext=H, k1=k1, axis=ax)
    assert rel_err(got["H_anis"], ref["H_anis"]) < 1e-13
    assert rel_err(got["H_total"], ref["H_total"]) < FP64_TOL
    assert rel_err(got["dmdt"], ref["dmdt"]) < FP64_TOL


def test_rhs_host_pointers_and_jtimes():
    _, _, orc, gpu, mat = make_pair(6, 5, 3, 2.5)
    P = orc_params(mat)
    m = random_m(orc.n, 21)
    ydot = gpu.rhs(m.reshape(-1)).reshape(-1, 3)
    ref = orc.fields(P, m, rtol=1e-13)
    assert rel_err(ydot, ref["dmdt"]) < FP64_TOL
    v = random_m(orc.n, 22) * 1e-3
    jv = gpu.jtimes(m.reshape(-1), v.reshape(-1)).reshape(-1, 3)
    jref = orc.jtimes(P, m, ref["H_total"], v)
    assert rel_err(jv, jref) < 1e-10


@pytest.mark.parametrize("route", ["iterative", "dense"])
@pytest.mark.parametrize("precision,tol_exch,tol_demag", [("fp32", 1e-7, 1e-6),
                                                           ("bf16", 1e-7, 2e-2),
                                                           ("mixed24", 1e-7, 5e-5)])
def test_mixed_precision_field_error(precision, tol_exch, tol_demag, route):
    """FP32 (BF16, 24-bit) operator storage: the per-field error against the
    FP64 oracle is that of rounding the stored operators; solves run to 1e-10.
    (The 24-bit mode stores only the dense potential operator in 24 bits; on
    the iterative route it is the FP32 mode.)"""
    _, _, orc, gpu, mat = make_pair(12, 12, 4, 2.5, precision=precision, rtol=1e-10,
                                    potential=route)
    m = random_m(orc.n, 1)
    ref = orc.fields(orc_params(mat), m, rtol=1e-13)
    got = _fields_gpu(gpu, m)
    assert rel_err(got["H_exch"], ref["H_exch"]) < tol_exch
    assert rel_err(got["H_demag"], ref["H_demag"]) < tol_demag


def test_warm_start_reduces_iterations():
    _, _, orc, gpu, mat = make_pair(12, 12, 4, 2.5, rtol=1e-8, potential="iterative")
    m = random_m(orc.n, 5)
    gpu.set_field("m", m)
    gpu.execute()
    cold = gpu.stat("cg_iters_nbc")
    gpu.execute()
    warm = gpu.stat("cg_iters_nbc")
    assert cold > 5 and warm <= 1
    assert gpu.stat("cg_error") == 0 and gpu.stat("cg_unconverged") == 0


@pytest.mark.parametrize("dims", [(1, 1, 1, 2.5), (6, 5, 3, 2.5), (40, 20, 2, 2.5)])
def test_two_level_preconditioner_matches_jacobi_solution(dims):
    """The coarse correction changes the iteration, not the solution: the
    two-level solve (sub-part aggregates, one part holds two or four) gives
    the field of the Jacobi-only solve, in fewer iterations on three parts
    or more."""
    nx, ny, nz, h = dims
    _, _, orc, gpu, mat = make_pair(nx, ny, nz, h, rtol=1e-12, potential="iterative")
    m = random_m(orc.n, 13)
    a = _fields_gpu(gpu, m)
    its2 = gpu.stat("cg_iters_nbc")
    gpu.tune("cg_coarse", 0)
    gpu.set_field("phi1", np.zeros(orc.n))
    gpu.set_field("phi2", np.zeros(orc.n))
    b = _fields_gpu(gpu, m)
    its1 = gpu.stat("cg_iters_nbc")
    assert rel_err(a["H_demag"], b["H_demag"]) < 1e-10
    assert gpu.stat("cg_unconverged") == 0
    if gpu.stat("cg_grid") > 2:
        assert its2 < its1


@pytest.mark.parametrize("dims", [(6, 5, 3, 2.5), (40, 20, 2, 2.5), (30, 24, 6, 2.5)])
def test_fine_coarse_space_matches_part_aggregates(dims):
    """Sub-part aggregates (two or four per part, their coarse inverse in FP32)
    change the iteration, not the solution: both Poisson solves at 1e-12 give
    the field of the coarse space of whole parts, in no more iterations."""
    nx, ny, nz, h = dims
    _, _, orc, gpu, mat = make_pair(nx, ny, nz, h, rtol=1e-12, potential="iterative")
    parts = gpu.stat("cg_grid")
    assert gpu.stat("cg_aggregates") in (2 * parts, 4 * parts)
    m = random_m(orc.n, 17)
    a = _fields_gpu(gpu, m)
    its_fine = (gpu.stat("cg_iters_nbc"), gpu.stat("cg_iters_dbc"))
    gpu.tune("cg_fine", 0)
    assert gpu.stat("cg_aggregates") == parts
    gpu.set_field("phi1", np.zeros(orc.n))
    gpu.set_field("phi2", np.zeros(orc.n))
    b = _fields_gpu(gpu, m)
    its_part = (gpu.stat("cg_iters_nbc"), gpu.stat("cg_iters_dbc"))
    assert rel_err(a["H_demag"], b["H_demag"]) < 1e-10
    assert gpu.stat("cg_unconverged") == 0 and gpu.stat("cg_error") == 0
    assert its_fine[0] <= its_part[0] and its_fine[1] <= its_part[1]
    ref = orc.fields(orc_params(mat), m, rtol=1e-13)
    assert rel_err(a["H_demag"], ref["H_demag"]) < FP64_TOL


def test_fine_coarse_space_at_30k_nodes():
    """A cube of 30^3 cells (29,791 nodes): 117 parts of 256 rows with two
    sub-part aggregates each, and 59 parts of 512 rows with four; the field
    matches the coarse space of whole parts (59 aggregates of two parts) with
    both solves at 1e-11, in fewer iterations."""
    _, _, orc, gpu, mat = make_pair(30, 30, 30, 2.0, rtol=1e-11, build_bem=False,
                                    potential="iterative")
    m = random_m(orc.n, 23)
    got, its = {}, {}
    for rpt, subs in ((1, 2), (2, 4)):
        gpu.tune("cg_rows_per_thread", rpt)
        assert gpu.stat("cg_aggregates") == subs * gpu.stat("cg_grid")
        gpu.set_field("phi1", np.zeros(orc.n))
        gpu.set_field("phi2", np.zeros(orc.n))
        got[rpt] = _fields_gpu(gpu, m)["H_demag"]
        its[rpt] = gpu.stat("cg_iters_nbc")
    gpu.tune("cg_rows_per_thread", 1)
    gpu.tune("cg_fine", 0)
    assert gpu.stat("cg_aggregates") == (gpu.stat("cg_grid") + 1) // 2
    gpu.set_field("phi1", np.zeros(orc.n))
    gpu.set_field("phi2", np.zeros(orc.n))
    ref = _fields_gpu(gpu, m)["H_demag"]
    assert gpu.stat("cg_unconverged") == 0 and gpu.stat("cg_error") == 0
    for rpt in (1, 2):
        assert rel_err(got[rpt], ref) < 1e-9
    assert its[1] < gpu.stat("cg_iters_nbc")


def test_demag_linear_and_uniform_cuboid_factors():
    """H_demag(-m) = -H_demag(m); uniform m gives Aharoni's factors (coarse mesh)."""
    from test_oracle import aharoni_nz
    _, _, orc, gpu, mat = make_pair(12, 12, 4, 2.5)
    m = random_m(orc.n, 9)
    a = _fields_gpu(gpu, m)["H_demag"]
    b = _fields_gpu(gpu, -m)["H_demag"]
    assert rel_err(-b, a) < 1e-10
    V = orc.volumes()
    mz = np.tile([0.0, 0.0, 1.0], (orc.n, 1))
    hz = _fields_gpu(gpu, mz)["H_demag"]
    Nz = -(hz[:, 2] * V).sum() / V.sum() / mat.su_Ms
    assert abs(Nz - aharoni_nz(30.0, 30.0, 10.0)) < 0.02


def test_dense_potential_operator_properties():
    """M = (I + E B R) K^+: M 1 = 0 (K^+ 1 = 0), M rho matches the two solves
    of the iterative route, and on the boundary M = K^+ + B K^+ restricted."""
    coords, tets, orc, gpu, mat = make_pair(6, 4, 3, 2.5, potential="dense")
    M = gpu.potential()
    assert M.shape == (gpu.n, gpu.n)
    assert np.max(np.abs(M.sum(axis=1))) <= 1e-10 * np.max(np.abs(M))
    it = make_pair(6, 4, 3, 2.5, potential="iterative")[3]
    m = random_m(gpu.n, 31)
    for g in (gpu, it):
        g.set_field("m", m)
        g.execute()
    rho = it.get_field("rho")
    phi_it = it.get_field("phi")
    phi = M @ rho
    # the Neumann gauge constant of phi1 cancels in phi1 + phi2 (B 1 = -1), so
    # the two routes give the same potential, not just the same field
    assert np.max(np.abs(phi - phi_it)) <= 1e-10 * np.max(np.abs(phi_it))
    assert np.max(np.abs(gpu.get_field("phi") - phi_it)) <= 1e-10 * np.max(np.abs(phi_it))
    with pytest.raises(Exception):
        gpu.get_field("phi1")


def test_24bit_potential_operator_is_rounded_fp64_operator():
    """The 24-bit potential operator holds, entry by entry, the FP64 operator
    rounded to FP32 and then to nearest (ties to even) at 15 stored
    significand bits.  585 nodes: rows of 640 entries span a full 512-entry
    block and a partial one, so both block layouts are read back.  (The bit
    formula below equals frexp-based rounding of the significand to 16 bits
    with ties to even.)"""
    M64 = make_pair(12, 8, 4, 2.5, potential="dense", build_bem=False)[3].potential()
    M24 = make_pair(12, 8, 4, 2.5, precision="mixed24", potential="dense",
                    build_bem=False)[3].potential()
    assert M24.shape == M64.shape == (585, 585)
    u = M64.astype(np.float32).view(np.uint32).astype(np.uint64)
    u24 = (u + 0x7F + ((u >> 8) & 1)) >> 8
    expect = (u24 << 8).astype(np.uint32).view(np.float32).astype(np.float64)
    # the two handles build M with the same FP64 kernels
    assert np.mean(M24 != expect) < 1e-3
    nz = M64 != 0
    rel = np.abs(M24[nz] - M64[nz]) / np.abs(M64[nz])
    assert rel.max() <= 2.0 ** -16 * (1 + 2.0 ** -6)
    # 15 significand bits kept, not 23: the rounding error reaches 2^-17
    assert rel.max() > 2.0 ** -17


def test_route_selection_at_the_dense_size_limit():
    """The dense potential route stages rho in 160 KiB of LDS: at most 20,480
    FP64 entries.  A 20,000-node magnet takes it by default; a 20,500-node one
    falls back to the iterative route, and asking for the dense route there is
    an error that names the limit.  The fallback's field of a uniform state
    has Aharoni's demagnetising factor."""
    from nmag import kuhn_cuboid
    from nmag._native import NativeMagnet
    from test_oracle import aharoni_nz
    mat = material()
    c0, t0 = kuhn_cuboid(39, 24, 19, 2.5)     # 40 x 25 x 20 = 20,000 nodes
    c1, t1 = kuhn_cuboid(40, 24, 19, 2.5)     # 41 x 25 x 20 = 20,500 nodes
    g0 = NativeMagnet(c0, t0, mat.su_coeffs(), potential="auto")
    assert g0.n == 20000 and g0.stat("potential") == 1
    g0.close()
    with pytest.raises(Exception) as err:
        NativeMagnet(c1, t1, mat.su_coeffs(), potential="dense")
    assert "20480" in str(err.value)
    g1 = NativeMagnet(c1, t1, mat.su_coeffs(), potential="auto", nbc_rtol=1e-10,
                      dbc_rtol=1e-10)
    assert g1.n == 20500 and g1.stat("potential") == 0
    g1.set_field("m", np.tile([0.0, 0.0, 1.0], (g1.n, 1)))
    g1.execute()
    hz = g1.get_field("H_demag")[:, 2]
    V = g1.get_field("volumes")
    g1.close()
    # 40 x 24 x 19 cubes of 2.5 nm: a 100 x 60 x 47.5 nm cuboid
    Nz = -(hz * V).sum() / V.sum() / mat.su_Ms
    assert abs(Nz - aharoni_nz(100.0, 60.0, 47.5)) < 0.02


@pytest.mark.parametrize("route,precision", [("dense", "fp64"), ("dense", "mixed"),
                                             ("dense", "mixed24"),
                                             ("iterative", "fp64"), ("iterative", "bf16")])
def test_dynamic_row_gemv_matches_static_plan(route, precision):
    """Rows claimed from a device counter (rows = 0) go through the same
    per-row dot product as the static one-row plan, so every shape gives
    bit-identical fields; the counter resets itself, so repeated launches and
    graph replays keep doing so."""
    coords, tets, orc, gpu, mat = make_pair(12, 8, 3, 2.5, precision=precision,
                                            potential=route, build_bem=False)
    op = "pot" if route == "dense" else "bem"
    m = random_m(gpu.n, 17)
    gpu.set_field("m", m)

    def fields():
        if route == "iterative":
            gpu.set_field("phi1", np.zeros(gpu.n))
            gpu.set_field("phi2", np.zeros(gpu.n))
        gpu.execute()
        return gpu.get_field("H_demag")

    gpu.tune(op + "_rows", 1)
    ref = fields()
    for w in (1, 2, 8):
        for v in (4, 16):
            gpu.tune(op + "_rows", 0)
            gpu.tune(op + "_wg_per_cu", w)
            gpu.tune(op + "_waves", v)
            assert gpu.stat(op + "_rows") == 0
            for _ in range(3):
                assert np.array_equal(fields(), ref), (w, v)
    if route == "dense":  # graph replays of the dynamic variant
        import torch
        n3 = 3 * gpu.n
        y = torch.tensor(m.reshape(-1), device="cuda:0")
        o = torch.zeros(n3, dtype=torch.float64, device="cuda:0")
        outs = []
        for _ in range(4):  # seen, captured, replayed, replayed
            o.zero_()
            gpu.rhs_device(y.data_ptr(), o.data_ptr())
            gpu.synchronize()
            outs.append(o.cpu().numpy())
        assert gpu.stat("graph_launches") >= 2
        gpu.tune("rhs_graph", 0)
        o.zero_()
        gpu.rhs_device(y.data_ptr(), o.data_ptr())
        gpu.synchronize()
        for a in outs:
            assert np.array_equal(a, o.cpu().numpy())
