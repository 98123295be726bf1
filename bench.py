"""Benchmark: LLG field evaluations per second on the 20k-node slab.

Workload (the metric's configuration): permalloy slab 160 x 80 x 10 nm, 2 nm
Kuhn-split cubes (80 x 40 x 5 cubes, 96,000 tets, 19,926 nodes, 7,602 boundary
nodes, dense 7,602^2 BEM), alpha = 0.02, 25 mT applied at 170 deg in-plane,
seeded S-state.  One step = one full right-hand-side evaluation
y -> H_exch + H_demag (2 PCG solves + BEM GEMV) + H_ext -> dm/dt, on a
sequence of magnetisation states along a trajectory (warm-started solves, as
the reference's KSPs are).  Inputs are resident in HBM before timing starts.

Prints ONE JSON line.  Progress goes to stderr.
"""
import argparse
import json
import math
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


def make_material(alpha):
    from nmag import MagMaterial, SI
    return MagMaterial("Py", Ms=SI(8.6e5, "A/m"), exchange_coupling=SI(1.3e-11, "J/m"),
                       llg_gamma_G=SI(2.211e5, "m/A s"), llg_damping=alpha)


def applied_field(n, mT=25.0, deg=170.0):
    from nmag.units import to_su, SI, mu0
    H = to_su(SI(mT * 1e-3, "T") / mu0, "A/m")
    return np.tile([H * math.cos(math.radians(deg)), H * math.sin(math.radians(deg)), 0.0],
                   (n, 1))


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


def run_gpu(prec, coords, tets, mat, Hext, states, warmup, steps, rtol, torch):
    from nmag._native import NativeMagnet
    g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=prec, nbc_rtol=rtol,
                     dbc_rtol=rtol)
    g.set_field("H_ext", Hext)
    n3 = 3 * g.n
    Y = torch.from_numpy(states.reshape(len(states), n3)).to("cuda")
    YD = torch.zeros_like(Y)
    torch.cuda.synchronize()
    for k in range(warmup):
        g.rhs_device(Y[k].data_ptr(), YD[k].data_ptr())
    g.synchronize()
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
    # second pass over the same states with per-kernel HIP-event timers
    g.set_field("phi1", p1)
    g.set_field("phi2", p2)
    g.timers(True)
    for k in range(steps):
        g.rhs_device(Y[warmup + k].data_ptr(), YD[warmup + k].data_ptr())
    g.synchronize()
    kern = {}
    for name in ("exch_div", "cg_nbc", "bem", "cg_dbc", "demag_llg"):
        ms, cnt = g.timer(name)
        kern[name] = ms / max(cnt, 1)
    g.timers(False)
    out = YD[warmup:warmup + steps].cpu().numpy()
    res = {
        "evals_per_s": steps / dt,
        "ms_per_eval": 1e3 * dt / steps,
        "cg_iters_nbc": (it1[0] - it0[0]) / steps,
        "cg_iters_dbc": (it1[1] - it0[1]) / steps,
        "kernel_ms": kern,
        "nb": g.nb, "n": g.n, "nnz": g.nnz,
        "cg_grid": g.stat("cg_grid"),
    }
    g.close()
    return res, out


def run_dynamics(prec, coords, tets, mat, Hext, m0, window_ps, checkpoint_ps, ksp_rtol):
    """advance_time over window_ps from the S-state (BDF, rtol = atol = 1e-6);
    returns wall time, integrator statistics and <m>(t) at the checkpoints."""
    from nmag._native import NativeMagnet
    g = NativeMagnet(coords, tets, mat.su_coeffs(), precision=prec, nbc_rtol=ksp_rtol,
                     dbc_rtol=ksp_rtol)
    g.set_field("H_ext", Hext)
    g.set_field("m", m0)
    g.ts_init(0.0, 1e-6, 1e-6)
    V = g.get_field("volumes")
    avgs, t = [], 0.0
    t0 = time.perf_counter()
    while t < window_ps - 1e-9:
        t = g.advance(min(window_ps, t + checkpoint_ps))
        m = g.get_field("m")
        avgs.append((m * V[:, None]).sum(0) / V.sum())
        log("%s dynamics t=%.1f ps wall=%.1f s steps=%d" % (prec, t, time.perf_counter() - t0,
                                                          g.stat("ts_steps")))
    wall = time.perf_counter() - t0
    stats = {k: g.stat("ts_" + k) for k in ("steps", "rhs", "newton", "linear", "err_fails",
                                            "conv_fails")}
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
    ap.add_argument("--sim-window", type=float, default=20.0,
                    help="ps of LLG dynamics timed for sim-ns/hour (1000 = the full 1 ns)")
    ap.add_argument("--no-dynamics", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(each magnet runs on a single MI355X); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from nmag import kuhn_cuboid
    from nmag._native import NativeMagnet

    nx, ny, nz, h = [float(v) for v in args.dims.split(",")]
    coords, tets = kuhn_cuboid(int(nx), int(ny), int(nz), h)
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

    results = {}
    outs = {}
    for prec in ("fp64", "mixed"):
        log("timing %s ..." % prec)
        results[prec], outs[prec] = run_gpu(prec, coords, tets, mat, Hext, states, args.warmup,
                                            args.steps, args.rtol, torch)
        log("%s: %.1f evals/s, kernels(ms) %s, cg iters %.1f/%.1f" % (
            prec, results[prec]["evals_per_s"], results[prec]["kernel_ms"],
            results[prec]["cg_iters_nbc"], results[prec]["cg_iters_dbc"]))
    err = float(np.max(np.abs(outs["mixed"] - outs["fp64"])) / np.max(np.abs(outs["fp64"])))

    mixed = results["mixed"]
    nb, n, nnz = mixed["nb"], mixed["n"], mixed["nnz"]
    kern = mixed["kernel_ms"]
    dominant = max(kern, key=kern.get)

    def roofline(kernel, alg_bytes, ms):
        ach = alg_bytes / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "kernel": kernel, "achieved": round(ach, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": None, "alg_bytes_per_launch": int(alg_bytes)}

    # per-iteration algorithmic bytes of one PCG iteration (SpMV over the
    # stiffness pattern with FP32 values + 4-byte columns, row pointers, and
    # about eight FP64 vector passes), times the measured iterations per solve
    cg_iter_bytes = 8 * nnz + 4 * (n + 1) + 64 * n
    rooflines = {
        "cg_nbc": roofline("k_cg<float> Neumann solve",
                           cg_iter_bytes * max(mixed["cg_iters_nbc"], 1), kern["cg_nbc"]),
        "cg_dbc": roofline("k_cg<float> Dirichlet solve",
                           cg_iter_bytes * max(mixed["cg_iters_dbc"], 1), kern["cg_dbc"]),
        "bem": roofline("k_bem_gemv<float>", gemv_bytes(nb, 4), kern["bem"]),
    }
    roof = rooflines.get(dominant)
    bem64 = results["fp64"]["kernel_ms"]["bem"]
    bem_fp64_gbs = gemv_bytes(nb, 8) / (bem64 * 1e-3) / 1e9

    dyn = None
    if not args.no_dynamics:
        dyn = {}
        avgs = {}
        for prec in ("fp64", "mixed"):
            wall, stats, avgs[prec] = run_dynamics(prec, coords, tets, mat, Hext, m0,
                                                   args.sim_window, 5.0, args.rtol)
            dyn[prec] = {"sim_ns_per_hour": round(args.sim_window * 1e-3 / (wall / 3600.0), 3),
                         "wall_s": round(wall, 3), **{k: int(v) for k, v in stats.items()}}
        dyn["window_ps"] = args.sim_window
        dyn["mean_m_max_dev_mixed_vs_fp64"] = float(np.abs(avgs["mixed"] - avgs["fp64"]).max())
        dyn["integrator"] = "BDF order<=2, Newton-GMRES, rtol=atol=1e-6, from the seeded S-state"

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle, %.0f s budget) ..." % args.cpu_budget)
        rate, done, setup, _ = run_cpu(coords, tets, mat, Hext, states[args.warmup:],
                                       args.rtol, args.cpu_budget)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": round(rate, 3), "unit": "field evals/s", "cores": threads,
               "kind": "port",
               "sample": "%d FP64 field evals of the same trajectory states, warm-started "
                         "Jacobi-PCG rtol %g (setup %.1f s excluded)" % (done, args.rtol, setup)}

    line = {
        "metric": "LLG field evals/s & sim-ns/hour, 20k-node mesh, mixed vs FP64 at "
                  "⟨m⟩ err≤1e-5",
        "value": round(mixed["evals_per_s"], 2),
        "unit": "field evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_eval"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (FP32 SpMV/BEM, FP64 dots/state)",
        "data": "synthetic (seeded S-state trajectory, explicit-Euler states 0.1 ps apart)",
        "config": {"workload": "Py slab 160x80x10 nm, 2 nm Kuhn mesh, 19926 nodes, "
                               "Nb=%d dense BEM, alpha=0.02, 25 mT @170deg" % nb,
                   "ksp_rtol": args.rtol},
        "fp64": {"value": round(results["fp64"]["evals_per_s"], 2),
                 "ms_per_step": round(results["fp64"]["ms_per_eval"], 4),
                 "kernel_ms": results["fp64"]["kernel_ms"],
                 "cg_iters": [results["fp64"]["cg_iters_nbc"], results["fp64"]["cg_iters_dbc"]],
                 "bem_gemv_GBs": round(bem_fp64_gbs, 1)},
        "mixed_kernel_ms": kern,
        "mixed_cg_iters": [mixed["cg_iters_nbc"], mixed["cg_iters_dbc"]],
        "mixed_vs_fp64_speedup": round(mixed["evals_per_s"] / results["fp64"]["evals_per_s"], 3),
        "mixed_dmdt_max_rel_err": err,
        "dominant_kernel": dominant,
        "roofline": roof,
        "roofline_all": rooflines,
        "cpu_baseline": cpu,
        "dynamics": dyn,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
