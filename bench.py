"""Benchmark of the TwoScaleRich time-stepping hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--controller classical|learned]
                    [--workload plain] [--tend T] [--cpu-sample-s S] [--no-cpu]

Metric (the reference's headline for this path): wall-time to t_end of the exponential
integrator on the plain GPU-scale workload (macro 101x101 nodes, 10,201 micro cells of
21x21 nodes, N = 4,508,842 FP64 unknowns, t in [0, 1] day, rtol = atol = 1e-6, Krylov
dimension <= 30), with the RHS throughput in GB/s beside it.

A "step" is one pass of the hot path over the workload: the integration from t = 0 to
t_end (every accepted EXPREM step, all on the device; inputs resident in HBM before the
timer starts).  `value` is its wall-time in seconds (lower is better).  W warm-up steps
integrate a short horizon first (kernel loading, allocation); K timed steps each
integrate the full horizon from the same initial state.

Extra fields on the JSON line:
  roofline      the dominant kernel (k_micro, one workgroup per micro cell):
                achieved = algorithmic bytes of one launch (16 B per micro unknown: read
                u, write g, plus per-cell distinct data) / its average duration from HIP
                events on the solver stream; peak 8 TB/s; traffic = FETCH+WRITE bytes per
                launch from rocprofv3 counters (profiles/), or null.
  cpu_baseline  the CPU oracle (FP64 C++/OpenMP restatement of the same method, "port")
                timed on this host: EXPREM steps of the same workload for a bounded
                sample, scaled to the GPU run's step count (labelled in "sample").
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_micro_plain_traffic.json")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-kernels", action="store_true", help="skip the per-kernel timing section")
    ap.add_argument("--controller", default="classical", choices=["classical", "learned"])
    ap.add_argument("--workload", default="plain", choices=["plain"])
    ap.add_argument("--tend", type=float, default=0.0, help="override t_end (reported in config)")
    ap.add_argument("--warmup-tend", type=float, default=1e-6)
    ap.add_argument("--cpu-sample-s", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--max-steps", type=int, default=10 ** 7)
    ap.add_argument("--chunk", type=int, default=500, help="EXPREM steps between progress lines")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the path's sharding is described in DESIGN.md "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import twoscalerich_amd as T
    from twoscalerich_amd import _ffi as F
    L = T.lib()

    p = T.Problem.config(F.CFG_PLAIN)
    opt = p.options()
    tend = args.tend if args.tend > 0 else p.tend()
    s = T.Solver(p, opt)
    if args.controller == "learned":
        from twoscalerich_amd import controller as CT
        CT.attach(s)
    u0 = p.initial()

    # warm-up: short horizon from the same initial state
    for _ in range(args.warmup):
        s.reset()
        s.set_state(u0, 0.0)
        rc, st, _ = s.run(args.warmup_tend, args.max_steps)
        if rc:
            raise RuntimeError("warm-up failed: " + T.last_error())
    L.tsr_device_sync()

    walls, runs = [], []
    for k in range(args.steps):
        s.reset()
        s.set_state(u0, 0.0)
        L.tsr_device_sync()
        t0 = time.perf_counter()
        tot = {"steps": 0, "rejects": 0, "fevals": 0, "host_dense_s": 0.0, "sync_wait_s": 0.0, "t": 0.0}
        while True:
            rc, st, _ = s.run(tend, min(args.chunk, args.max_steps - tot["steps"]))
            for key in ("steps", "rejects", "fevals", "host_dense_s", "sync_wait_s"):
                tot[key] += getattr(st, key)
            tot["t"] = st.t
            print(f"[bench] step {k}: t={st.t:.4e}/{tend} exprem_steps={tot['steps']} "
                  f"wall={time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
            if rc or st.t >= tend or tot["steps"] >= args.max_steps:
                break
        L.tsr_device_sync()
        walls.append(time.perf_counter() - t0)
        runs.append(tot)
        if rc or tot["t"] < tend:
            raise RuntimeError(f"integration stopped at t={tot['t']} rc={rc}: " + T.last_error())
    wall = max(walls)
    st = runs[-1]
    u_end, _ = s.get_state()

    # RHS throughput and the dominant kernel's roofline (HIP events on the solver stream)
    du = L.tsr_solver_dev_u(s.h)
    g = L.tsr_solver_dev_scratch(s.h, 1)
    L.tsr_solver_gfunc_dev(s.h, g, du, 3)
    L.tsr_solver_gfunc_dev(s.h, g, du, args.kernel_reps)
    g_ms = L.tsr_solver_last_gfunc_ms(s.h)
    ms3 = (C.c_double * 3)()
    L.tsr_solver_time_gfunc_parts(s.h, g, du, args.kernel_reps, ms3)
    v = p.view.contents
    M = v.scen[0].ud.macro_mesh.no_nodes
    n_micro = p.N - M
    micro_bytes = 16.0 * n_micro + p.distinct_bytes()
    g_bytes = 16.0 * p.N + p.distinct_bytes()
    micro_gbs = micro_bytes / (ms3[1] * 1e-3) / 1e9
    shape = (C.c_int(), C.c_int())
    L.tsr_solver_micro_shape(s.h, C.byref(shape[0]), C.byref(shape[1]))
    copy_ms = C.c_double()
    L.tsr_solver_time_copy(s.h, args.kernel_reps, C.byref(copy_ms))
    copy_gbs = 16.0 * p.N / (copy_ms.value * 1e-3) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p, opt, tend, args.cpu_sample_s)
        # the GPU's time for the same first n steps, to scale the sample to t_end
        s.reset()
        s.set_state(u0, 0.0)
        L.tsr_device_sync()
        t0 = time.perf_counter()
        s.run(tend, cpu["_n"])
        L.tsr_device_sync()
        gpu_n = time.perf_counter() - t0
        cpu["value"] = round(cpu["_cpu_s"] * wall / gpu_n, 1)
        cpu["sample"] += (f"; the GPU took {gpu_n:.3f} s for the same {cpu['_n']} steps and {wall:.1f} s to t_end, "
                          f"so the CPU time to t_end is estimated as {cpu['_cpu_s']:.1f} s x {wall / gpu_n:.1f}")
        cpu["speedup_same_steps"] = round(cpu["_cpu_s"] / gpu_n, 1)
        del cpu["_n"], cpu["_cpu_s"]

    line = {
        "metric": "wall-time to t_end, learned+batched vs plain GPU EXPREM at <=1e-6 rel err; G GB/s",
        "value": round(wall, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain GPU scale: macro 101x101, micro 21x21 per macro node, loam matrix + "
                               "centred sand disc r=0.25, Dirichlet -0.1 m top, h0=-5 m",
                   "N": p.N, "t_end_days": tend, "rtol": opt.RelTol, "atol": opt.AbsTol,
                   "max_krylov": opt.MaxKrylov, "controller": args.controller, "mode": "plain"},
        "exprem": {"accepted_steps": st["steps"], "rejected": st["rejects"], "rhs_evals": st["fevals"],
                   "host_dense_s": round(st["host_dense_s"], 3), "sync_wait_s": round(st["sync_wait_s"], 3),
                   "u_end_min": float(u_end.min()), "u_end_max": float(u_end.max())},
        "gfunc": {"ms": round(g_ms, 4), "GB_per_s": round(g_bytes / (g_ms * 1e-3) / 1e9, 1),
                  "kernel_ms": {"macro_elem": round(ms3[0], 4), "micro": round(ms3[1], 4),
                                "macro_node": round(ms3[2], 4)}},
        "roofline": {"kernel": "k_micro", "bound": "hbm", "achieved": round(micro_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(micro_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": micro_bytes, "avg_launch_ms": round(ms3[1], 4),
                     "copy_kernel_GB_per_s": round(copy_gbs, 1),
                     "launch_shape": {"workgroup": shape[0].value, "waves_per_eu": shape[1].value}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    s.close()


def cpu_baseline(p, opt, tend, budget_s):
    """The CPU oracle (FP64 C++/OpenMP restatement of the same method) on the same
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
        print(f"[bench] cpu oracle step {n}: t={t:.3e} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    o.close()
    return {"value": None, "unit": "s", "cores": cores, "kind": "port", "_n": n, "_cpu_s": el,
            "sample": f"CPU oracle: the first {n} EXPREM steps of the same run from t=0 took {el:.1f} s "
                      f"(FuncEvals {st.FuncEvals}, t reached {t:.3e} d, OpenMP threads {cores})"}


if __name__ == "__main__":
    main()
