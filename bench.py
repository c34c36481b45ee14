"""Benchmark of the TwoScaleRich time-stepping hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--plain-tend T] [--full-plain]
                    [--cpu-sample-s S] [--no-cpu] [--basis f32|f64]

Metric (the reference's headline for this path): wall-time to t_end of the exponential
integrator, learned + batched GPU EXPREM against plain GPU EXPREM, on the plain GPU-scale
workload (macro 101x101 nodes, 10,201 micro cells of 21x21 nodes, N = 4,508,842 FP64
unknowns, t in [0, 1] day, rtol = atol = 1e-6), with the RHS throughput beside it.

A "step" is one pass of the hot path over the workload: the integration from t = 0 to
t_end with the learned controller on the batched solver path (every accepted EXPREM
step, all on the device; inputs resident in HBM before the timer starts; DCGS2 reads
FP32 copies of the Krylov basis unless --basis f64, the arithmetic is FP64 either way).  `value` is
its wall-time in seconds (lower is better).  W warm-up steps integrate a short horizon
first (kernel loading, launch-shape tuning); K timed steps each integrate the full
horizon from the same initial state.

Extra fields on the JSON line:
  plain         the plain GPU path (the reference algorithm: classical controller, CGS2,
                FD products of the full RHS, a host check per Arnoldi iteration) timed to
                --plain-tend in this run, the learned+batched run's time to the same point
                and the solutions' difference there; with --full-plain the plain path is
                also timed to t_end (about 5 minutes).  `full_committed` repeats the
                plain path's full-horizon time from profiles/ for reference.
  gfunc         G evaluation times per kernel for the three micro-kernel modes.
  roofline      the dominant kernel of the timed run (k_micro in its linearised mode, the
                FD products Jv = (G(u + eps v) - G(u)) / eps): algorithmic bytes of one
                launch = those of an FD product with the perturbation fused into the
                state load, 32 B per unknown (read u, v and G(u), write the product)
                plus the distinct per-cell / per-scenario data, over its average
                duration from HIP events on the solver stream; peak 8 TB/s.
                traffic = FETCH+WRITE bytes per launch from rocprofv3 counters
                (profiles/), scaled by the copy-kernel calibration, or null; it exceeds
                the algorithmic bytes because the linearised form also streams the
                recorded expansion (three FP64 values and three FP32 derivatives per
                interior micro node).
  cpu_baseline  the CPU oracle (FP64 C++/OpenMP restatement of the plain algorithm,
                "port") timed on this host: EXPREM steps of the same workload for a
                bounded sample, scaled to t_end by the plain GPU path's time profile.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_micro_lin_plain_traffic.json")
PLAIN_FULL_FILE = os.path.join(ROOT, "profiles", "exprem_plain_full.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def integrate(T, s, u0, tend, chunk, label, stops=()):
    """Run solver s from (u0, 0) to tend; returns (wall, totals, {stop: wall at stop})."""
    L = T.lib()
    s.reset()
    s.set_state(u0, 0.0)
    L.tsr_device_sync()
    t0 = time.perf_counter()
    tot = {"steps": 0, "rejects": 0, "fevals": 0, "host_dense_s": 0.0, "sync_wait_s": 0.0, "t": 0.0}
    at = {}
    for target in [x for x in stops if x < tend] + [tend]:
        while True:
            rc, st, _ = s.run(target, chunk)
            for key in ("steps", "rejects", "fevals", "host_dense_s", "sync_wait_s"):
                tot[key] += getattr(st, key)
            tot["t"] = st.t
            log(f"{label}: t={st.t:.4e}/{tend} exprem_steps={tot['steps']} wall={time.perf_counter() - t0:.1f}s")
            if rc:
                raise RuntimeError(f"{label}: integration stopped at t={st.t} rc={rc}: " + T.last_error())
            if st.t >= target:
                break
        if target != tend:
            L.tsr_device_sync()
            at[target] = (time.perf_counter() - t0, s.get_state()[0])
    L.tsr_device_sync()
    return time.perf_counter() - t0, tot, at


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--warmup-tend", type=float, default=1e-6)
    ap.add_argument("--plain-tend", type=float, default=0.25, help="horizon of the plain-path comparison run")
    ap.add_argument("--tend", type=float, default=0.0, help="override t_end (reported in config; default the workload's)")
    ap.add_argument("--full-plain", action="store_true", help="also time the plain path to t_end")
    ap.add_argument("--cpu-sample-s", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--chunk", type=int, default=1000, help="EXPREM steps between progress lines")
    ap.add_argument("--basis", choices=("f32", "f64"), default="f32",
                    help="storage of the Krylov basis columns on the batched path (arithmetic is FP64)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only time the G kernels in their three modes on a perturbed initial state (for profilers)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU (--gpus must be 1); the path's sharding across GPUs is "
              "described in the design notes, not built", file=sys.stderr)
        sys.exit(2)

    import twoscalerich_amd as T
    from twoscalerich_amd import _ffi as F
    from twoscalerich_amd import controller as CT
    L = T.lib()

    p = T.Problem.config(F.CFG_PLAIN)
    opt = p.options()
    tend = args.tend if args.tend > 0 else p.tend()
    args.plain_tend = min(args.plain_tend, tend)
    u0 = p.initial()

    fast = T.Solver(p, opt)
    fast.set_path(F.PATH_BATCHED | (F.PATH_F32BASIS if args.basis == "f32" else 0))
    CT.attach(fast)
    if args.kernels_only:
        rng = np.random.default_rng(1)
        fast.set_state(u0 + 0.5 * rng.standard_normal(p.N), 0.0)
        copy_ms = C.c_double()
        L.tsr_solver_time_copy(fast.h, args.kernel_reps, C.byref(copy_ms))
        print(json.dumps({"gfunc": gfunc_modes(T, fast, args.kernel_reps),
                          "copy_16N_ms": round(copy_ms.value, 4), "copy_bytes": 16 * p.N}), flush=True)
        fast.close()
        return
    plain = T.Solver(p, opt)     # plain path, classical controller (the defaults)

    for _ in range(args.warmup):
        for s, lab in ((fast, "warm-up learned+batched"), (plain, "warm-up plain")):
            integrate(T, s, u0, args.warmup_tend, args.chunk, lab)

    # timed: learned + batched to t_end (the metric), noting the wall time at plain_tend
    walls, runs, at_prefix = [], [], None
    for k in range(args.steps):
        w, tot, at = integrate(T, fast, u0, tend, args.chunk, f"learned+batched step {k}", stops=(args.plain_tend,))
        walls.append(w)
        runs.append(tot)
        at_prefix = at.get(args.plain_tend, (w, fast.get_state()[0]))
    wall = max(walls)
    st = runs[-1]
    u_end, _ = fast.get_state()

    # plain path to plain_tend (and optionally to t_end)
    w_plain, tot_plain, _ = integrate(T, plain, u0, args.plain_tend, args.chunk, "plain")
    u_plain, _ = plain.get_state()
    d2 = float(np.linalg.norm(at_prefix[1] - u_plain) / np.linalg.norm(u_plain))
    dinf = float(np.max(np.abs(at_prefix[1] - u_plain)) / np.max(np.abs(u_plain)))
    plain_info = {"t_end_days": args.plain_tend, "wall_s": round(w_plain, 3), "steps": tot_plain["steps"],
                  "rhs_evals": tot_plain["fevals"], "learned_batched_wall_s": round(at_prefix[0], 3),
                  "speedup": round(w_plain / at_prefix[0], 3),
                  "rel_diff_2norm": d2, "rel_diff_max": dinf}
    if args.full_plain:
        w_full, tot_full, _ = integrate(T, plain, u0, tend, args.chunk, "plain full")
        u_pf, _ = plain.get_state()
        plain_info["full"] = {"wall_s": round(w_full, 3), "steps": tot_full["steps"], "rhs_evals": tot_full["fevals"],
                              "speedup": round(w_full / wall, 3),
                              "rel_diff_2norm_t_end": float(np.linalg.norm(u_end - u_pf) / np.linalg.norm(u_pf)),
                              "rel_diff_max_t_end": float(np.max(np.abs(u_end - u_pf)) / np.max(np.abs(u_pf)))}
    if os.path.exists(PLAIN_FULL_FILE):
        with open(PLAIN_FULL_FILE) as f:
            ref = json.load(f)
        plain_info["full_committed"] = {"wall_s": ref["wall_s"], "speedup": round(ref["wall_s"] / wall, 3),
                                        "source": os.path.relpath(PLAIN_FULL_FILE, ROOT)}

    # G kernels per mode and the dominant kernel's roofline (HIP events on the solver stream)
    gf = gfunc_modes(T, fast, args.kernel_reps)
    v = p.view.contents
    M = v.scen[0].ud.macro_mesh.no_nodes
    n_micro = p.N - M
    r = v.tmpl[0].no_nodes
    n_cells = v.n_cells
    n_dir = n_cells * (4 * (int(round(r ** 0.5)) - 1))
    n_int = n_micro - n_dir
    lin_bytes = 32.0 * p.N + p.distinct_bytes()
    # what the linearised form reads and writes: interior node u, v, g0, 3 FP64 + 3 FP32 recorded
    # components, the product; Dirichlet node g0, the recorded head, 2 FP64 + 2 FP32 conductivity
    # components, the product
    lin_impl_bytes = 68.0 * n_int + 48.0 * n_dir + p.distinct_bytes()
    lin_ms = gf["linearised_fd"]["micro_ms"]
    achieved = lin_bytes / (lin_ms * 1e-3) / 1e9
    full_bytes = 16.0 * p.N + p.distinct_bytes()
    gf["full"]["GB_per_s_16N"] = round(full_bytes / (gf["full"]["micro_ms"] * 1e-3) / 1e9, 1)
    copy_ms = C.c_double()
    L.tsr_solver_time_copy(fast.h, args.kernel_reps, C.byref(copy_ms))
    copy_gbs = 16.0 * p.N / (copy_ms.value * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(T, p, opt, tend, args.cpu_sample_s)
        # the plain GPU path's time for the same first n steps scales the sample to t_end
        plain.reset()
        plain.set_state(u0, 0.0)
        L.tsr_device_sync()
        t0 = time.perf_counter()
        plain.run(tend, cpu["_n"])
        L.tsr_device_sync()
        gpu_n = time.perf_counter() - t0
        full_plain = plain_info.get("full", plain_info.get("full_committed", {})).get("wall_s")
        if full_plain:
            cpu["value"] = round(cpu["_cpu_s"] * full_plain / gpu_n, 1)
            cpu["sample"] += (f"; the plain GPU path took {gpu_n:.3f} s for the same {cpu['_n']} steps and "
                              f"{full_plain:.1f} s to t_end, so the CPU time to t_end is estimated as "
                              f"{cpu['_cpu_s']:.1f} s x {full_plain / gpu_n:.1f}")
        cpu["speedup_plain_gpu_same_steps"] = round(cpu["_cpu_s"] / gpu_n, 1)
        del cpu["_n"], cpu["_cpu_s"]

    line = {
        "metric": "wall-time to t_end, learned+batched vs plain GPU EXPREM at <=1e-6 rel err; G GB/s",
        "value": round(wall, 3),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 1),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain GPU scale: macro 101x101, micro 21x21 per macro node, loam matrix + "
                               "centred sand disc r=0.25, Dirichlet -0.1 m top, h0=-5 m, t in [0, 1] d",
                   "N": p.N, "t_end_days": tend, "rtol": opt.RelTol, "atol": opt.AbsTol,
                   "max_krylov_plain": opt.MaxKrylov, "controller": "learned (Krylov range [5, 60])",
                   "path": "batched (DCGS2" + (" on FP32 copies of the basis" if args.basis == "f32" else "")
                           + ", deferred Krylov checks, linearised FD products)"},
        "exprem": {"accepted_steps": st["steps"], "rejected": st["rejects"], "rhs_evals": st["fevals"],
                   "host_dense_s": round(st["host_dense_s"], 3), "sync_wait_s": round(st["sync_wait_s"], 3),
                   "u_end_min": float(u_end.min()), "u_end_max": float(u_end.max())},
        "plain": plain_info,
        "gfunc": gf,
        "roofline": {"kernel": "k_micro (linearised FD product)", "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes": lin_bytes, "implementation_bytes": lin_impl_bytes,
                     "avg_launch_ms": lin_ms,
                     "copy_kernel_GB_per_s": round(copy_gbs, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    fast.close()
    plain.close()


def gfunc_modes(T, s, reps):
    """Per-kernel device times of G about the solver's state in its three k_micro modes
    (0 full, 1 recording the expansion point, 2 linearised FD product), HIP events on the
    solver stream; the first launches tune the micro kernel's launch shape per mode."""
    L = T.lib()
    du = L.tsr_solver_dev_u(s.h)
    g = L.tsr_solver_dev_scratch(s.h, 1)
    gf = {}
    for mode, name in ((0, "full"), (1, "record"), (2, "linearised_fd")):
        ms3 = (C.c_double * 3)()
        L.tsr_solver_time_gfunc_mode(s.h, g, du, mode, 3, ms3)
        L.tsr_solver_time_gfunc_mode(s.h, g, du, mode, reps, ms3)
        # micro_ms: the k_micro launch (micro cells, and the macro elements in its leading
        # workgroups); macro_node_ms: k_macro_node
        gf[name] = {"micro_ms": round(ms3[1], 4), "macro_node_ms": round(ms3[2], 4)}
    return gf


def cpu_baseline(T, p, opt, tend, budget_s):
    """The CPU oracle (FP64 C++/OpenMP restatement of the plain algorithm) on the same
    workload from t = 0, bounded: EXPREM steps until the time budget is spent (at least
    one step)."""
    from twoscalerich_amd import _ffi as F
    from oracle.oracle import Oracle
    o = Oracle(p.view, p.ud)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ud = p.ud.contents
    ud.NUM_THREADS = cores          # Gfunc_parallel's thread count (user_data.h:11)
    u = p.initial()
    st = F.exprem_stats_struct()
    t = 0.0
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and t < tend:
        rc, t, _ = o.exprem_step(ud, u, t, tend, opt, st)
        if rc:
            break
        n += 1
        log(f"cpu oracle step {n}: t={t:.3e} ({time.perf_counter() - t0:.1f} s)")
    el = time.perf_counter() - t0
    o.close()
    return {"value": None, "unit": "s", "cores": cores, "kind": "port", "_n": n, "_cpu_s": el,
            "sample": f"CPU oracle: the first {n} EXPREM steps of the plain run from t=0 took {el:.1f} s "
                      f"(FuncEvals {st.FuncEvals}, t reached {t:.3e} d, OpenMP threads {cores})"}


if __name__ == "__main__":
    main()
