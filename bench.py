"""Benchmark: LLG field evaluations per second on the 20k-node slab.

Workload (the metric's configuration): permalloy slab 160 x 80 x 10 nm, 2 nm
Kuhn-split cubes (80 x 40 x 5 cubes, 96,000 tets, 19,926 nodes, 7,602 boundary
nodes), alpha = 0.02, 25 mT applied at 170 deg in-plane, seeded S-state.  One
step = one full right-hand-side evaluation y -> H_exch + H_demag + H_ext ->
dm/dt on a sequence of magnetisation states along a trajectory.  The demag is
timed on three routes: the potential operator M compressed as an H-matrix
(phi = M rho in two launches; the headline, FP32-mixed), the same operator
dense (one GEMV), and the reference's structure (two warm-started PCG solves
around the BEM GEMV).  Inputs are resident in HBM before timing starts.

Further legs: sim-ns/hour of the BDF integrator, the 64-magnet batch against
64 sequential handles, the differentiation form (adjoint of a 2 ns RK4
trajectory of the 100 x 50 x 5 nm slab), and the CPU oracle as the baseline.

Prints ONE JSON line.  Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "nmag-src_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_material(alpha):
    from nmag.workloads import permalloy
    return permalloy(alpha)


def trajectory_states(magnet, m0, count, dt):
    """count states m_{k+1} = normalize(m_k + dt f(m_k)) (explicit Euler, dt in ps)."""
    states = [m0]
    m = m0
    for _ in range(count - 1):
        f = magnet.rhs(m.reshape(-1)).reshape(-1, 3)
        m = m + dt * f
        m = m / np.linalg.norm(m, axis=1, keepdims=True)
        states.append(m)
    return np.stack(states)


def gemv_bytes(nb, esz):
    return esz * nb * nb + 16 * nb


AUTOTUNE = True
# --no-graph: eager launches only (rocprofv3's kernel trace cannot follow
# HIP-graph capture in this image)
GRAPHS = True


def graphs(g):
    if not GRAPHS:
        g.tune("rhs_graph", 0)
    return g


# HBM bytes per launch of the potential GEMV at the 20k-node slab, measured
# with rocprofv3 (FETCH_SIZE x 2.000, the factor a 1 GiB device copy gives,
# plus WRITE_SIZE; profiles/potential_gemv_c2_slab_hbm_counters.md).  Counters
# cannot be read inside a timed run, so the figure is reported for this
# workload only.
MEASURED_TRAFFIC = {(4, 19926): 777846 * 2048 + 179 * 1024,
                    (8, 19926): 1554955 * 2048 + 185 * 1024,
                    # the 24-bit operator, one handle per profiled process
                    # (profiles/potential_gemv_f24_c2_slab_hbm_counters.md)
                    (3, 19926): int(583578.2 * 2048 + 194.9 * 1024),
                    # both k_panel_gemv<float> launches of the compressed
                    # operator (eps 1e-7, eta 3: 292,995,808 stored bytes;
                    # profiles/potential_hmatrix_c2_slab.md); any other
                    # operator gets null
                    ("hmatrix", 19926, 292995808): int((48021.4 + 99293.5) * 2048
                                                       + (777.8 + 843.7) * 1024)}
KERNELS = {"dense": ("exch_div", "potential", "demag_llg"),
           "hmatrix": ("exch_div", "potential", "demag_llg"),
           "iterative": ("exch_div", "cg_nbc", "bem", "cg_dbc", "demag_llg")}


def run_gpu(prec, route, coords, tets, mat, Hext, states, warmup, steps, rtol, torch):
    from nmag._native import NativeMagnet
    t0 = time.perf_counter()
    g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=prec, nbc_rtol=rtol,
                     dbc_rtol=rtol, potential=route)
    t_setup = time.perf_counter() - t0
    g.set_field("H_ext", Hext)
    g.set_field("m", states[0])
    tuned = g.autotune() if AUTOTUNE else {}
    log("%s/%s set-up %.2f s, autotuned: %s" % (prec, route, t_setup, tuned))
    n3 = 3 * g.n
    Y = torch.from_numpy(states.reshape(len(states), n3)).to("cuda")
    YD = torch.zeros_like(Y)
    torch.cuda.synchronize()
    for k in range(warmup):
        g.rhs_device(Y[k].data_ptr(), YD[k].data_ptr())
    g.synchronize()
    if route == "iterative":
        p1, p2 = g.get_field("phi1"), g.get_field("phi2")
    it0 = (g.stat("cg_iters_nbc_total"), g.stat("cg_iters_dbc_total"))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        g.rhs_device(Y[warmup + k].data_ptr(), YD[warmup + k].data_ptr())
    g.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    it1 = (g.stat("cg_iters_nbc_total"), g.stat("cg_iters_dbc_total"))
    assert g.stat("cg_error") == 0, "CG reported an error"
    # second pass over the same states with per-kernel HIP-event timers (on
    # the handle's stream, around each launch)
    if route == "iterative":
        g.set_field("phi1", p1)
        g.set_field("phi2", p2)
    g.timers(True)
    for k in range(steps):
        g.rhs_device(Y[warmup + k].data_ptr(), YD[warmup + k].data_ptr())
    g.synchronize()
    kern = {}
    for name in KERNELS[route]:
        ms, cnt = g.timer(name)
        kern[name] = ms / max(cnt, 1)
    g.timers(False)
    # the CVODE-callback form: host arrays in and out (nsim_rhs with
    # device_ptrs = 0), so each evaluation carries its PCIe copies
    host_states = states[warmup:warmup + steps].reshape(steps, n3)
    if route == "iterative":
        g.set_field("phi1", p1)
        g.set_field("phi2", p2)
    t0 = time.perf_counter()
    for k in range(steps):
        g.rhs(host_states[k])
    dt_host = time.perf_counter() - t0
    out = YD[warmup:warmup + steps].cpu().numpy()
    res = {
        "evals_per_s": steps / dt,
        "ms_per_eval": 1e3 * dt / steps,
        "host_pointer_evals_per_s": steps / dt_host,
        "cg_iters_nbc": (it1[0] - it0[0]) / steps,
        "cg_iters_dbc": (it1[1] - it0[1]) / steps,
        "kernel_ms": kern,
        "nb": g.nb, "n": g.n, "nnz": g.nnz,
        "cg_grid": g.stat("cg_grid"),
        "autotuned": tuned,
        "setup_s": round(t_setup, 3),
        "potential_bytes": g.stat("potential_bytes"),
        "bem_bytes": g.stat("bem_bytes"),
    }
    if route == "hmatrix":
        res["hmatrix"] = {k[3:]: g.stat(k) for k in (
            "hm_eps", "hm_eta", "hm_rank_entries", "hm_full_entries", "hm_dense_entries",
            "hm_lowrank_entries", "hm_lowrank_blocks", "hm_dense_blocks", "hm_max_rank",
            "hm_fallback_blocks", "hm_build_s", "hm_p1_panels", "hm_p2_panels",
            "hm_p1_entries", "hm_p2_entries")}
    g.close()
    return res, out


def run_dynamics(prec, route, coords, tets, mat, Hext, m0, window_ps, checkpoint_ps, ksp_rtol):
    """advance_time over window_ps from the S-state (BDF, rtol = atol = 1e-6);
    returns wall time, integrator statistics and <m>(t) at the checkpoints."""
    from nmag._native import NativeMagnet
    t_set = time.perf_counter()
    g = graphs(NativeMagnet(coords, tets, mat.su_coeffs(), precision=prec, nbc_rtol=ksp_rtol,
                            dbc_rtol=ksp_rtol, potential=route))
    g.set_field("H_ext", Hext)
    g.set_field("m", m0)
    g.ts_init(0.0, 1e-6, 1e-6)
    g.synchronize()
    setup_s = time.perf_counter() - t_set
    V = g.get_field("volumes")
    wv = V / V.sum()  # <m> = wv . m, one BLAS product per checkpoint
    avgs, t = [], 0.0
    t0 = time.perf_counter()
    while t < window_ps - 1e-9:
        t = g.advance(min(window_ps, t + checkpoint_ps))
        m = g.get_field("m")
        avgs.append(wv @ m)
        if len(avgs) % 10 == 0 or t >= window_ps - 1e-9:
            log("%s dynamics t=%.1f ps wall=%.1f s steps=%d" % (
                prec, t, time.perf_counter() - t0, g.stat("ts_steps")))
    wall = time.perf_counter() - t0
    stats = {k: g.stat("ts_" + k) for k in ("steps", "rhs", "newton", "linear", "err_fails",
                                            "conv_fails")}
    stats["setup_s"] = setup_s
    g.close()
    return wall, stats, np.array(avgs)


def run_cpu(coords, tets, mat, Hext, states, rtol, budget_s):
    import oracle
    c = mat.su_coeffs()
    P = oracle.OrcParams(Ms=c["Ms"], negJ=c["negJ"], c1=c["c1"], c2=c["c2"], c3=c["c3"],
                         mu0=c["mu0"])
    t0 = time.perf_counter()
    M = oracle.OracleMesh(coords, tets)
    setup = time.perf_counter() - t0
    log("cpu oracle setup %.1f s" % setup)
    phi1 = phi2 = None
    done, t_used = 0, 0.0
    outs = []
    while done < len(states) and t_used < budget_s:
        t1 = time.perf_counter()
        f = M.fields(P, states[done], H_ext=Hext, rtol=rtol, phi1=phi1, phi2=phi2)
        t_used += time.perf_counter() - t1
        phi1, phi2 = f["phi1"], f["phi2"]
        outs.append(f["dmdt"])
        done += 1
    return done / t_used, done, setup, outs


def run_cpu_leg(coords, tets, mat, Hext, states, rtol, budget_s):
    """The oracle on the host cores (all threads, then one core) on a bounded
    sample of the same trajectory states."""
    import oracle
    log("cpu baseline (oracle, %.0f s budget) ..." % budget_s)
    threads = oracle.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    rate, done, setup, _ = run_cpu(coords, tets, mat, Hext, states, rtol, budget_s)
    # the same on one core, a smaller bounded sample
    oracle.set_threads(1)
    rate1, done1, _, _ = run_cpu(coords, tets, mat, Hext, states, rtol, budget_s / 2)
    oracle.set_threads(threads)
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f
                          if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"value": round(rate, 3), "unit": "field evals/s", "cores": threads,
            "kind": "port",
            "sample": "%d FP64 field evals of the same trajectory states, warm-started "
                      "Jacobi-PCG rtol %g (setup %.1f s excluded)" % (done, rtol, setup),
            "one_core": {"value": round(rate1, 3), "unit": "field evals/s",
                         "sample": "%d evals of the same states" % done1},
            "cpu_model": model}


def run_batch(steps, warmup, rtol, window_ps, torch):
    """Batching form: one handle for all magnets (grouped SpMV/BEM, one
    persistent CG) against one FP64 handle per magnet run one after another."""
    from nmag._native import NativeBatch, NativeMagnet
    from nmag.workloads import cuboid_batch
    mags = cuboid_batch()
    meshes = [(g["coords"], g["tets"]) for g in mags]
    mats = [g["mat"].su_coeffs() for g in mags]
    sizes = np.array([len(g["coords"]) for g in mags])
    log("batch: %d magnets, %d nodes total (%d..%d)" % (len(mags), sizes.sum(), sizes.min(),
                                                         sizes.max()))
    # a sequence of states: each a small random rotation of the previous one
    rng = np.random.default_rng(5)
    m = np.concatenate([g["fields"]["m"] for g in mags])
    states = []
    for _ in range(warmup + steps):
        m = m + 0.002 * rng.standard_normal(m.shape)
        m /= np.linalg.norm(m, axis=1, keepdims=True)
        states.append(m.reshape(-1).copy())
    states = np.array(states)
    cat = {key: np.concatenate([g["fields"][key] for g in mags]) for key in
           ("H_ext", "K1", "axis")}
    out = {"magnets": len(mags), "nodes_total": int(sizes.sum()),
           "nodes_min": int(sizes.min()), "nodes_max": int(sizes.max())}

    def setup(mag, f):
        for key in ("H_ext", "axis", "K1"):
            mag.set_field(key, f[key])

    # batched field evaluations
    for prec in ("fp64", "mixed", "mixed24"):
        t0 = time.perf_counter()
        b = graphs(NativeBatch(meshes, mats, precision=prec, nbc_rtol=rtol, dbc_rtol=rtol))
        t_setup = time.perf_counter() - t0
        setup(b, cat)
        b.set_field("m", states[0])
        tuned = b.autotune()
        Y = torch.from_numpy(states).to("cuda")
        YD = torch.zeros_like(Y)
        torch.cuda.synchronize()
        for k in range(warmup):
            b.rhs_device(Y[k].data_ptr(), YD[k].data_ptr())
        b.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            b.rhs_device(Y[warmup + k].data_ptr(), YD[warmup + k].data_ptr())
        b.synchronize()
        dt = time.perf_counter() - t0
        assert b.stat("cg_error") == 0
        out["batched_" + prec] = {"batch_evals_per_s": round(steps / dt, 2),
                                  "magnet_evals_per_s": round(len(mags) * steps / dt, 1),
                                  "ms_per_batch_eval": round(1e3 * dt / steps, 4),
                                  "setup_s": round(t_setup, 2), "cg_grid": b.stat("cg_grid"),
                                  "autotuned": tuned}
        log("batched %s: %.1f batch evals/s (setup %.1f s)" % (prec, steps / dt, t_setup))
        if prec == "fp64":
            ref_out = YD[warmup:].cpu().numpy()
        else:
            got = YD[warmup:].cpu().numpy()
            out["batched_" + prec]["dmdt_max_rel_err_vs_fp64"] = float(
                np.max(np.abs(got - ref_out)) / np.max(np.abs(ref_out)))
        del Y, YD
        b.close()

    # 64 sequential FP64 solves: one handle per magnet, one magnet after another
    t0 = time.perf_counter()
    singles = [NativeMagnet(c, t, mt, precision="fp64", nbc_rtol=rtol, dbc_rtol=rtol)
               for (c, t), mt in zip(meshes, mats)]
    t_setup = time.perf_counter() - t0
    off = np.concatenate([[0], np.cumsum(sizes)])
    Ys, YDs = [], []
    for j, s in enumerate(singles):
        setup(s, mags[j]["fields"])
        Ys.append(torch.from_numpy(np.ascontiguousarray(
            states.reshape(len(states), -1, 3)[:, off[j]:off[j + 1]].reshape(len(states), -1))
        ).to("cuda"))
        YDs.append(torch.zeros_like(Ys[-1]))
    torch.cuda.synchronize()
    for j, s in enumerate(singles):
        for k in range(warmup):
            s.rhs_device(Ys[j][k].data_ptr(), YDs[j][k].data_ptr())
        s.synchronize()
    t0 = time.perf_counter()
    for j, s in enumerate(singles):
        for k in range(steps):
            s.rhs_device(Ys[j][warmup + k].data_ptr(), YDs[j][warmup + k].data_ptr())
        s.synchronize()
    dt = time.perf_counter() - t0
    seq = np.concatenate([YDs[j][warmup:].cpu().numpy().reshape(steps, -1, 3)
                          for j in range(len(singles))], axis=1).reshape(steps, -1)
    out["sequential_fp64"] = {"batch_evals_per_s": round(steps / dt, 2),
                              "magnet_evals_per_s": round(len(mags) * steps / dt, 1),
                              "setup_s": round(t_setup, 2)}
    out["batched_vs_sequential_fp64"] = round(
        out["batched_fp64"]["batch_evals_per_s"] / out["sequential_fp64"]["batch_evals_per_s"], 3)
    for prec in ("mixed", "mixed24"):
        out["batched_%s_vs_sequential_fp64" % prec] = round(
            out["batched_" + prec]["batch_evals_per_s"] /
            out["sequential_fp64"]["batch_evals_per_s"], 3)
    out["batched_vs_sequential_max_rel_diff"] = float(
        np.max(np.abs(ref_out - seq)) / np.max(np.abs(seq)))
    log("sequential fp64: %.1f batch evals/s" % (steps / dt))

    # relaxation window at alpha 0.5: one integrator over the batch against one
    # integrator per magnet, FP64, rtol = atol = 1e-6
    if window_ps > 0:
        t0 = time.perf_counter()
        b = graphs(NativeBatch(meshes, mats, precision="fp64", nbc_rtol=rtol, dbc_rtol=rtol))
        setup(b, cat)
        b.set_field("m", np.concatenate([g["fields"]["m"] for g in mags]))
        b.ts_init(0.0, 1e-6, 1e-6)
        t1 = time.perf_counter()
        b.advance(window_ps)
        wall_b = time.perf_counter() - t1
        steps_b = b.stat("ts_steps")
        b.close()
        wall_s, steps_s = 0.0, 0
        for j, s in enumerate(singles):
            s.set_field("m", mags[j]["fields"]["m"])
            s.ts_init(0.0, 1e-6, 1e-6)
            t1 = time.perf_counter()
            s.advance(window_ps)
            wall_s += time.perf_counter() - t1
            steps_s += int(s.stat("ts_steps"))
        out["relax_window"] = {"window_ps": window_ps, "batched_wall_s": round(wall_b, 3),
                               "batched_steps": int(steps_b),
                               "sequential_wall_s": round(wall_s, 3),
                               "sequential_steps_total": steps_s,
                               "speedup": round(wall_s / wall_b, 3)}
        log("relax window %.1f ps: batched %.2f s, sequential %.2f s" % (window_ps, wall_b,
                                                                       wall_s))
    for s in singles:
        s.close()
    return out


def run_precision(slab, window_ps=1000.0, checkpoint_ps=10.0, seed=1, rtol=1e-10):
    """Precision form on the 12x12x4-cube cuboid (2.5 nm, 845 nodes, alpha 0.5,
    seeded Gaussian m, no applied field): per-field max relative error of
    H_exch, H_demag and dm/dt against FP64 at the seeded state, and the
    <m>(t) deviation from FP64 over window_ps of relaxation; the same
    per-field errors at one state of the 20k-node slab (slab = coords, tets,
    mat, Hext, m)."""
    from nmag import kuhn_cuboid
    from nmag._native import NativeMagnet
    coords, tets = kuhn_cuboid(12, 12, 4, 2.5)
    mat = make_material(alpha=0.5)
    rng = np.random.default_rng(seed)
    m0 = rng.standard_normal((len(coords), 3))
    m0 /= np.linalg.norm(m0, axis=1, keepdims=True)
    fields = ("H_exch", "H_demag", "dmdt")

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    out = {"workload": "Py 30x30x10 nm cuboid, 2.5 nm Kuhn mesh (845 nodes), alpha 0.5, "
                       "seeded Gaussian m (seed %d), relaxed %.0f ps" % (seed, window_ps)}
    ref, avg = {}, {}
    # "<prec>_tight": integrated at rtol = atol = 1e-8.  FP64 tight against
    # FP64 at the working tolerance 1e-6 is the size of the integration error
    # itself; mixed / 24-bit tight against FP64 tight is the storage error's
    # effect once the integration error no longer dominates
    for prec in ("fp64", "mixed", "mixed24", "bf16", "fp64_tight", "mixed_tight",
                 "mixed24_tight"):
        base = prec.replace("_tight", "")
        g = graphs(NativeMagnet(coords, tets, mat.su_coeffs(), precision=base,
                                nbc_rtol=rtol, dbc_rtol=rtol))
        g.set_field("m", m0)
        g.execute()
        f = {k: g.get_field(k) for k in fields}
        if prec == "fp64":
            ref = f
        elif prec == base:
            out[prec] = {k + "_max_rel_err": rel(f[k], ref[k]) for k in fields}
        tol = 1e-8 if prec.endswith("_tight") else 1e-6
        g.ts_init(0.0, tol, tol)
        V = g.get_field("volumes")
        wv = V / V.sum()
        t, series = 0.0, []
        while t < window_ps - 1e-9:
            t = g.advance(min(window_ps, t + checkpoint_ps))
            m = g.get_field("m")
            series.append(wv @ m)
        avg[prec] = np.array(series)
        g.close()
    for prec in ("mixed", "mixed24", "bf16"):
        out[prec]["mean_m_max_dev_vs_fp64"] = float(np.abs(avg[prec] - avg["fp64"]).max())
    out["fp64_rtol_1e-8_mean_m_max_dev_vs_fp64"] = float(
        np.abs(avg["fp64_tight"] - avg["fp64"]).max())
    for prec in ("mixed", "mixed24"):
        out[prec]["mean_m_max_dev_vs_fp64_both_rtol_1e-8"] = float(
            np.abs(avg[prec + "_tight"] - avg["fp64_tight"]).max())
    # the slab at one state
    coords, tets, mat, Hext, m = slab
    sref = {}
    out["slab"] = {}
    for prec in ("fp64", "mixed", "mixed24", "bf16"):
        g = graphs(NativeMagnet(coords, tets, mat.su_coeffs(), precision=prec))
        g.set_field("H_ext", Hext)
        g.set_field("m", m)
        g.execute()
        f = {k: g.get_field(k) for k in fields}
        g.close()
        if prec == "fp64":
            sref = f
        else:
            out["slab"][prec] = {k + "_max_rel_err": rel(f[k], sref[k]) for k in fields}
    log("precision form: %s" % json.dumps(out))
    return out


def run_adjoint(t_ns, rtol, fd_checks=True):
    """Differentiation form: permalloy slab 100 x 50 x 5 nm, 2.5 nm Kuhn mesh
    (40 x 20 x 2 cubes, 2583 nodes), alpha 0.02, per-element K ~ N(2e4, 5e3)
    J/m^3 along z, 16 piecewise-constant x-field pulses over t_ns ~ N(0, 5) mT,
    loss <m_x>(t_ns); fixed-step RK4 at 0.2 ps, checkpoints every 50 steps."""
    from nmag._native import NativeMagnet
    from nmag.adjoint import TrajectoryProblem
    from nmag.workloads import differentiation_slab
    coords, tets, mat, K, amps, m0 = differentiation_slab()
    n = len(coords)
    dt = 0.2
    n_steps = int(round(t_ns * 1000.0 / dt))
    out = {"workload": "Py 100x50x5 nm, 2.5 nm Kuhn mesh, %d nodes, %d tets, %d pulses, "
                       "RK4 dt=%.1f ps x %d steps, checkpoint every 50" % (
                           n, len(tets), len(amps), dt, n_steps),
           "ksp_rtol": rtol}
    grads = {}
    for prec in ("mixed", "fp64"):
        g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=prec, nbc_rtol=rtol,
                         dbc_rtol=rtol)
        prob = TrajectoryProblem(g, mat.su_coeffs(), coords, tets, m0,
                                 np.tile([0.0, 0.0, 1.0], (n, 1)), dt=dt, n_steps=n_steps,
                                 n_pulses=len(amps), pulse_len=dt * n_steps / len(amps),
                                 checkpoint_every=50)
        Ms = mat.su_coeffs()["Ms"]
        t0 = time.perf_counter()
        loss = prob.run(K, Ms, amps)
        t_fwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        gK, gMs, gA, gm0 = prob.gradient(amps)
        t_adj = time.perf_counter() - t0
        grads[prec] = (gK, gMs, gA)
        rhs = g.stat("rhs_evals")
        out[prec] = {"loss": loss, "forward_s": round(t_fwd, 2), "adjoint_s": round(t_adj, 2),
                     "adjoint_over_forward": round(t_adj / t_fwd, 2), "rhs_evals": int(rhs),
                     "dLoss_dMs": gMs, "dLoss_damps_first4": [float(v) for v in gA[:4]],
                     "cg_unconverged": int(g.stat("cg_unconverged"))}
        log("adjoint %s: loss %.6f forward %.1f s adjoint %.1f s" % (prec, loss, t_fwd, t_adj))
        if prec == "fp64" and fd_checks:
            # central differences at relative step 1e-3, and Richardson's
            # extrapolation from steps 1e-3 and 5e-4 (error O(eps^4)): over
            # 2 ns the precession phase makes the loss strongly nonlinear in Ms,
            # so the plain 1e-3 difference carries its own ~1e-3 error there
            k = int(np.argmax(np.abs(gA)))

            def fd(fun, x0, h):
                return (fun(x0 + h) - fun(x0 - h)) / (2 * h)

            def amp_fun(v):
                a = amps.copy()
                a[k] = v
                return prob.run(K, Ms, a)

            checks = {}
            for name, fun, x0, scale, adj in (
                    ("Ms", lambda v: prob.run(K, v, amps), Ms, Ms, gMs),
                    ("amp_%d" % k, amp_fun, amps[k], np.max(np.abs(amps)), float(gA[k]))):
                d1 = fd(fun, x0, 1e-3 * scale)
                d2 = fd(fun, x0, 5e-4 * scale)
                rich = (4.0 * d2 - d1) / 3.0
                checks[name] = {"adjoint": adj, "fd_1e-3": float(d1), "fd_richardson": float(rich),
                                "rel_err_fd_1e-3": float(abs(adj - d1) / abs(d1)),
                                "rel_err_richardson": float(abs(adj - rich) / abs(rich))}
            out["fd_check_fp64"] = checks
            log("fd check: %s" % checks)
        g.close()
    rel = lambda a, b: float(np.max(np.abs(np.atleast_1d(a) - np.atleast_1d(b))) /
                             np.max(np.abs(np.atleast_1d(b))))
    out["mixed_vs_fp64_gradient_rel_err"] = {
        "K_elem": rel(grads["mixed"][0], grads["fp64"][0]),
        "Ms": rel(grads["mixed"][1], grads["fp64"][1]),
        "amps": rel(grads["mixed"][2], grads["fp64"][2])}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--rtol", type=float, default=1e-6)
    ap.add_argument("--dt", type=float, default=0.1, help="ps between trajectory states")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--dims", default="80,40,5,2.0")
    ap.add_argument("--sim-window", type=float, default=1000.0,
                    help="ps of LLG dynamics timed for sim-ns/hour (1000 = the full 1 ns)")
    ap.add_argument("--no-dynamics", action="store_true")
    ap.add_argument("--dynamics-routes", default="dense,hmatrix",
                    help="demag routes timed in the dynamics leg")
    ap.add_argument("--dense-precs", default="fp64,mixed,mixed24",
                    help="precisions timed on the dense route (fp64 first when given; "
                         "one alone builds one operator handle, as the counter passes need)")
    ap.add_argument("--routes", default="dense,iterative,hmatrix",
                    help="demag routes timed in the field-evaluation leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile shapes (short profiler runs)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 64-magnet batching leg")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-window", type=float, default=2.0,
                    help="ps of relaxation timed for the batch (0 skips)")
    ap.add_argument("--adjoint-ns", type=float, default=2.0,
                    help="ns of the differentiation-form trajectory (0 skips)")
    ap.add_argument("--adjoint-rtol", type=float, default=1e-8)
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the precision-form leg (per-field errors, <m> deviation)")
    ap.add_argument("--no-graph", action="store_true",
                    help="no HIP-graph replay of field evaluations (for kernel traces)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run a plain device copy of this many MiB (HBM counter "
                         "calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    global AUTOTUNE, GRAPHS
    AUTOTUNE = not args.no_autotune
    GRAPHS = not args.no_graph
    dp = args.dense_precs.split(",")
    if not dp or not set(dp) <= {"fp64", "mixed", "mixed24"} or ("fp64" in dp and dp[0] != "fp64"):
        print("--dense-precs: any of fp64, mixed, mixed24 (fp64 first when given)",
              file=sys.stderr)
        sys.exit(2)
    if not args.routes or not set(args.routes.split(",")) <= {"dense", "iterative", "hmatrix"}:
        print("--routes: any of dense, iterative, hmatrix", file=sys.stderr)
        sys.exit(2)
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(each magnet runs on a single MI355X); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from nmag import kuhn_cuboid
    from nmag._native import NativeMagnet

    nx, ny, nz, h = [float(v) for v in args.dims.split(",")]
    coords, tets = kuhn_cuboid(int(nx), int(ny), int(nz), h)
    from nmag.workloads import applied_field, s_state
    mat = make_material(alpha=0.02)
    Hext = applied_field(len(coords))
    m0 = s_state(coords, seed=2024)
    total = args.warmup + args.steps
    log("mesh n=%d tets=%d; generating %d trajectory states" % (len(coords), len(tets), total))
    gen = NativeMagnet(coords, tets, mat.su_coeffs(), precision="fp64", nbc_rtol=1e-8,
                       dbc_rtol=1e-8)
    gen.set_field("H_ext", Hext)
    states = trajectory_states(gen, m0, total, args.dt)
    gen.close()

    # Every leg runs guarded: a failing leg is recorded in the line's
    # "errors" and the run exits non-zero, but what was measured is printed.
    errors = {}

    def leg(name, fn, *a, **kw):
        try:
            return fn(*a, **kw)
        except Exception as e:  # noqa: BLE001 (recorded, then the run fails)
            import traceback
            errors[name] = "%s: %s" % (type(e).__name__, e)
            log("leg %s FAILED: %s" % (name, errors[name]))
            traceback.print_exc(file=sys.stderr)
            return None

    results, outs = {}, {}
    # the precisions timed per route: the 24-bit mode differs from FP32-mixed
    # only in the dense operator; the compressed operator is timed with FP32
    # entries (in FP64 at 1e-12 per block it compresses too little to pay)
    route_precs = {"dense": tuple(args.dense_precs.split(",")), "iterative": ("fp64", "mixed"),
                   "hmatrix": ("mixed",)}
    for route in ("dense", "iterative", "hmatrix"):
        if route not in args.routes.split(","):
            continue
        for prec in route_precs[route]:
            key = prec + "/" + route
            log("timing %s ..." % key)
            r = leg("field_evals:" + key, run_gpu, prec, route, coords, tets, mat, Hext, states,
                    args.warmup, args.steps, args.rtol, torch)
            if r is None:
                continue
            results[key], outs[key] = r
            log("%s: %.1f evals/s, kernels(ms) %s, cg iters %.1f/%.1f" % (
                key, results[key]["evals_per_s"], results[key]["kernel_ms"],
                results[key]["cg_iters_nbc"], results[key]["cg_iters_dbc"]))

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    def roofline(kernel, alg_bytes, ms, traffic=None):
        ach = alg_bytes / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "kernel": kernel, "achieved": round(ach, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": traffic, "alg_bytes_per_launch": int(alg_bytes)}

    # headline: FP32-mixed on the compressed (H-matrix) potential operator when
    # it was timed, else on the dense one; the FP64 path is the dense operator
    # (exact up to rounding)
    head_key = "mixed/hmatrix" if "mixed/hmatrix" in results else "mixed/dense"
    mixed, fp64 = results.get(head_key), results.get("fp64/dense")
    mixed_dense = results.get("mixed/dense")
    m24 = results.get("mixed24/dense")
    head = mixed or fp64 or m24
    nb = head["nb"] if head else None
    n = head["n"] if head else len(coords)
    nnz = head["nnz"] if head else None
    kern = mixed["kernel_ms"] if mixed else None
    dominant = max(kern, key=kern.get) if kern else None
    err = rel(outs[head_key], outs["fp64/dense"]) if mixed and fp64 else None

    # dense potential GEMV: the N x N operator once, rho in, phi out
    pot_bytes = lambda esz: esz * n * n + 16 * n
    rooflines = {}
    if mixed and head_key == "mixed/hmatrix":
        # the compressed operator: its FP32 entries and index arrays once, rho
        # in and phi out (two launches, timed together)
        rooflines["potential"] = roofline(
            "k_panel_gemv<float> x2 H-matrix potential operator (N=%d, %.0f MB)" % (
                n, mixed["potential_bytes"] / 1e6),
            mixed["potential_bytes"] + 16 * n, kern["potential"],
            MEASURED_TRAFFIC.get(("hmatrix", n, int(mixed["potential_bytes"]))))
    if mixed_dense:
        rooflines["potential_dense" if head_key == "mixed/hmatrix" else "potential"] = roofline(
            "k_dense_gemv<float> potential operator (N=%d)" % n, pot_bytes(4),
            mixed_dense["kernel_ms"]["potential"], MEASURED_TRAFFIC.get((4, n)))
    if fp64:
        rooflines["potential_fp64"] = roofline(
            "k_dense_gemv<double> potential operator (N=%d)" % n, pot_bytes(8),
            fp64["kernel_ms"]["potential"], MEASURED_TRAFFIC.get((8, n)))
    if m24:
        rooflines["potential_f24"] = roofline(
            "k_dense_gemv<F24> potential operator (N=%d)" % n, pot_bytes(3),
            m24["kernel_ms"]["potential"], MEASURED_TRAFFIC.get((3, n)))
    roof = rooflines.get(dominant) if dominant else None
    pot64_gbs = (pot_bytes(8) / (fp64["kernel_ms"]["potential"] * 1e-3) / 1e9) if fp64 else None
    # iterative route (the reference's structure): per-iteration algorithmic
    # bytes of one PCG iteration (SpMV over the stiffness pattern with FP64
    # values + 4-byte columns, row pointers, about eight FP64 vector passes)
    # times the measured iterations per solve; the BEM GEMV streams B once
    rooflines_iterative = None
    if "mixed/iterative" in results:
        it_m = results["mixed/iterative"]
        kit = it_m["kernel_ms"]
        cg_iter_bytes = 12 * it_m["nnz"] + 4 * (n + 1) + 64 * n
        rooflines_iterative = {
            "cg_nbc": roofline("k_cg<double> Neumann solve",
                               cg_iter_bytes * max(it_m["cg_iters_nbc"], 1), kit["cg_nbc"]),
            "cg_dbc": roofline("k_cg<double> Dirichlet solve",
                               cg_iter_bytes * max(it_m["cg_iters_dbc"], 1), kit["cg_dbc"]),
            "bem": roofline("k_dense_gemv<float> BEM", gemv_bytes(it_m["nb"], 4), kit["bem"]),
        }
        # the CG's bytes are priced against HBM for comparison only: an
        # iteration is one grid-wide exchange and a gather, so latency bounds it
        for key, its in (("cg_nbc", it_m["cg_iters_nbc"]), ("cg_dbc", it_m["cg_iters_dbc"])):
            rooflines_iterative[key]["limited_by"] = "latency (one grid exchange per iteration)"
            rooflines_iterative[key]["us_per_iteration"] = round(
                1e3 * kit[key] / max(its, 1), 2)

    dyn = None
    if not args.no_dynamics:
        dyn = {}
        all_avgs = {}
        for route in args.dynamics_routes.split(","):
            dyn[route] = {}
            avgs = {}
            precs = {"dense": ("fp64", "mixed", "mixed24"), "hmatrix": ("mixed",)}.get(
                route, ("fp64", "mixed"))
            for prec in precs:
                r = leg("dynamics:%s/%s" % (prec, route), run_dynamics, prec, route, coords,
                        tets, mat, Hext, m0, args.sim_window, 5.0, args.rtol)
                if r is None:
                    continue
                wall, stats, avgs[prec] = r
                all_avgs[(prec, route)] = avgs[prec]
                setup_s = stats.pop("setup_s")
                hours = args.sim_window * 1e-3
                dyn[route][prec] = {
                    "sim_ns_per_hour": round(hours / (wall / 3600.0), 3),
                    "wall_s": round(wall, 3),
                    # end to end: handle creation (operator set-up) included
                    "setup_s": round(setup_s, 3),
                    "sim_ns_per_hour_incl_setup": round(hours / ((wall + setup_s) / 3600.0), 3),
                    **{k: int(v) for k, v in stats.items()}}
            if "mixed" in avgs and "fp64" in avgs:
                dyn[route]["mean_m_max_dev_mixed_vs_fp64"] = float(
                    np.abs(avgs["mixed"] - avgs["fp64"]).max())
            if "mixed24" in avgs and "fp64" in avgs:
                dyn[route]["mean_m_max_dev_mixed24_vs_fp64"] = float(
                    np.abs(avgs["mixed24"] - avgs["fp64"]).max())
            if "mixed" in dyn[route] and "fp64" in dyn[route]:
                dyn[route]["mixed_vs_fp64_sim_ns_per_hour"] = round(
                    dyn[route]["mixed"]["sim_ns_per_hour"] /
                    dyn[route]["fp64"]["sim_ns_per_hour"], 3)
            # against the dense FP64 run (the FP64 path) when this route has none
            ref64 = all_avgs.get(("fp64", "dense"))
            if route != "dense" and "mixed" in avgs and ref64 is not None:
                dyn[route]["mean_m_max_dev_mixed_vs_fp64_dense"] = float(
                    np.abs(avgs["mixed"] - ref64).max())
                if "fp64" in dyn.get("dense", {}):
                    dyn[route]["mixed_vs_fp64_dense_sim_ns_per_hour"] = round(
                        dyn[route]["mixed"]["sim_ns_per_hour"] /
                        dyn["dense"]["fp64"]["sim_ns_per_hour"], 3)
        dyn["window_ps"] = args.sim_window
        dyn["integrator"] = "BDF order<=2, Newton-GMRES, rtol=atol=1e-6, from the seeded S-state"

    if args.copy_mb > 0:
        # a copy of known size: FETCH_SIZE / WRITE_SIZE of its kernel calibrate
        # the counters read for the potential GEMV
        src = torch.ones(args.copy_mb * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy of %d MiB done (3 launches)" % args.copy_mb)

    batch = None
    if not args.no_batch:
        batch = leg("batch", run_batch, args.batch_steps, 3, args.rtol, args.batch_window, torch)

    precision = None
    if not args.no_precision:
        precision = leg("precision", run_precision, (coords, tets, mat, Hext, states[args.warmup]))

    adjoint = None
    if args.adjoint_ns > 0:
        adjoint = leg("differentiation", run_adjoint, args.adjoint_ns, args.adjoint_rtol)

    cpu = None
    if not args.no_cpu:
        cpu = leg("cpu_baseline", run_cpu_leg, coords, tets, mat, Hext, states[args.warmup:],
                  args.rtol, args.cpu_budget)

    line = {
        "metric": "LLG field evals/s & sim-ns/hour, 20k-node mesh, mixed vs FP64 at "
                  "⟨m⟩ err≤1e-5",
        "value": round(mixed["evals_per_s"], 2) if mixed else None,
        "unit": "field evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_eval"], 4) if mixed else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (FP32 potential operator and exchange/gradient SpMV, FP64 "
                 "accumulation, divergence, dots and state)" + (
                     "; potential operator compressed to 1e-7 per block, the order of FP32 "
                     "rounding" if head_key == "mixed/hmatrix" else ""),
        "data": "synthetic (seeded S-state trajectory, explicit-Euler states 0.1 ps apart)",
        "config": {"workload": "Py slab 160x80x10 nm, 2 nm Kuhn mesh, %d nodes, "
                               "Nb=%s, alpha=0.02, 25 mT @170deg" % (n, nb),
                   "demag": None if not mixed else
                   ("H-matrix potential operator M=(I+EBR)K^+ (%.3f GB: FP32 entries and "
                    "indices), built in %.1f s" if head_key == "mixed/hmatrix" else
                    "dense potential operator M=(I+EBR)K^+ (%.2f GB FP32), built in %.1f s") % (
                        mixed["potential_bytes"] / 1e9, mixed["setup_s"]),
                   "ksp_rtol_iterative_route": args.rtol},
        "fp64": None if not fp64 else {
            "value": round(fp64["evals_per_s"], 2),
            "ms_per_step": round(fp64["ms_per_eval"], 4),
            "kernel_ms": fp64["kernel_ms"],
            "potential_gemv_GBs": round(pot64_gbs, 1),
            "setup_s": fp64["setup_s"],
            "autotuned": fp64["autotuned"]},
        "mixed_autotuned": mixed["autotuned"] if mixed else None,
        "mixed_kernel_ms": kern,
        "mixed_vs_fp64_speedup": (round(mixed["evals_per_s"] / fp64["evals_per_s"], 3)
                                  if mixed and fp64 else None),
        "mixed_dmdt_max_rel_err": err,
        # FP32-mixed on the dense operator (the same arithmetic without the
        # compression)
        "mixed_dense": None if not mixed_dense else {
            "value": round(mixed_dense["evals_per_s"], 2),
            "ms_per_step": round(mixed_dense["ms_per_eval"], 4),
            "vs_fp64_speedup": (round(mixed_dense["evals_per_s"] / fp64["evals_per_s"], 3)
                                if fp64 else None),
            "dmdt_max_rel_err_vs_fp64": (rel(outs["mixed/dense"], outs["fp64/dense"])
                                         if fp64 else None),
            "kernel_ms": mixed_dense["kernel_ms"],
            "setup_s": mixed_dense["setup_s"]},
        "headline_route": head_key,
        # the compressed operator: blocks, ranks, entries per launch
        "hmatrix": results["mixed/hmatrix"].get("hmatrix") if "mixed/hmatrix" in results
        else None,
        # host arrays in and out per evaluation (the CVODE-callback form),
        # PCIe copies included; never the headline value
        "host_pointer_evals_per_s": {
            key: round(results[key]["host_pointer_evals_per_s"], 2)
            for key in ("mixed/hmatrix", "mixed/dense", "fp64/dense", "mixed24/dense")
            if key in results},
        # the potential operator in 24 bits (FP32 rounded to 15 significand
        # bits, 3 bytes per entry), everything else as FP32-mixed
        "mixed24": None if m24 is None else {
            "value": round(m24["evals_per_s"], 2),
            "ms_per_step": round(m24["ms_per_eval"], 4),
            "vs_fp64_speedup": (round(m24["evals_per_s"] / fp64["evals_per_s"], 3)
                                if fp64 else None),
            "dmdt_max_rel_err_vs_fp64": (rel(outs["mixed24/dense"], outs["fp64/dense"])
                                         if fp64 else None),
            "kernel_ms": m24["kernel_ms"],
            "potential_bytes": m24["potential_bytes"],
            "autotuned": m24["autotuned"]},
        "dense_vs_iterative_fp64_dmdt_rel_diff": (
            rel(outs["fp64/dense"], outs["fp64/iterative"])
            if "fp64/iterative" in outs and "fp64/dense" in outs else None),
        "iterative_route": {
            key.split("/")[0]: {"value": round(results[key]["evals_per_s"], 2),
                                "kernel_ms": results[key]["kernel_ms"],
                                "cg_iters": [results[key]["cg_iters_nbc"],
                                             results[key]["cg_iters_dbc"]],
                                "autotuned": results[key]["autotuned"]}
            for key in ("fp64/iterative", "mixed/iterative") if key in results},
        "dominant_kernel": dominant,
        "roofline": roof,
        "roofline_all": rooflines,
        "roofline_iterative_route": rooflines_iterative,
        "cpu_baseline": cpu,
        "dynamics": dyn,
        "precision": precision,
        "batch": batch,
        "differentiation": adjoint,
        "errors": errors or None,
    }
    print(json.dumps(line), flush=True)
    if errors:
        sys.exit(1)


if __name__ == "__main__":
    main()
