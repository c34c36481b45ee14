"""Benchmark of the TwoScaleRich time-stepping hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--budget-s S] [--basis f64|f32]
                    [--no-cpu] [--ktimer-every E] [--kernels-only]

Metric (the reference's headline for this path): wall-time to t_end of the exponential
integrator, learned + batched GPU EXPREM against plain GPU EXPREM, on the plain GPU-scale
workload (macro 101x101 nodes, 10,201 micro cells of 21x21 nodes, N = 4,508,842 FP64
unknowns, t in [0, 1] day, rtol = atol = 1e-6), with the RHS throughput beside it.

A "step" is one pass of the hot path over the workload: the integration from t = 0 to
t_end with the learned controller on the batched solver path (every accepted EXPREM
step, all on the device, FP64 arithmetic and an FP64 Krylov basis; inputs resident in
HBM before the timer starts).  `value` is its wall-time in seconds (lower is better),
the mean over the timed steps.  W warm-up steps integrate a short horizon first (kernel
loading, launch-shape tuning).  Then the plain GPU path (the reference algorithm:
classical controller, CGS2, FD products of the full RHS, a host check per Arnoldi
iteration) integrates the same workload to t_end once, in this process, and the timed
steps follow; fewer than K are run when the next one would overrun the wall-clock
budget (--budget-s, default 540 s for the whole process), and `steps` says how many ran.

Extra fields on the JSON line:
  plain         the plain path's wall-time to t_end, its step and RHS counts, the speed-up
                and the two solutions' difference at t_end (2-norm and max-norm, relative)
  kernels       per kernel class, from HIP events on the solver stream around a sample of
                the launches of the timed steps (every E-th launch per class): share of the
                timed device time, average launch duration, algorithmic bytes per launch
                and the rate they give
  roofline      the dominant kernel of the timed steps (largest summed device time):
                algorithmic bytes per launch / its average launch duration, against
                8 TB/s; traffic = this run's algorithmic bytes per launch x the HBM bytes per
                algorithmic byte that rocprofv3 FETCH_SIZE / WRITE_SIZE runs measured for
                the kernel class (profiles/hbm_traffic_plain_workload.json, copy-kernel
                calibrated), or null
  gfunc         G alone: device time per launch of the micro-cell kernel in its three
                modes (HIP events) and the HBM rates of its algorithmic bytes (16 B per
                unknown for G, 32 B for an FD product), roofline_gfunc for the FD product
  cpu_baseline  the CPU oracle (FP64 C++/OpenMP restatement of the plain algorithm,
                "port") on this host's cores: the plain workload sampled from four states
                of the plain GPU run (t = 0 and about 0.25, 0.5, 0.75 d), a few seconds of
                EXPREM steps each, with the plain GPU path timed over the same steps from
                the same states; the CPU time to t_end is estimated as the sum over the four
                segments of the plain GPU path's wall time times the CPU/GPU ratio sampled
                at the segment's start (an extrapolation, labelled as such), and the plain
                CPU-runnable workload (macro 21x21, micro 11x11, t in [0, 0.1] d) integrated
                to t_end on both, measured whole
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# name -> (problem configuration, description); every one integrates from h0 = -5 m
WORKLOADS = {
    "small": (0, "plain CPU-runnable: macro 21x21, micro 11x11 per macro node, centred sand disc r=0.25, "
                 "Dirichlet -0.1 m top, h0=-5 m, t in [0, 0.1] d"),
    "plain": (1, "plain GPU scale: macro 101x101, micro 21x21 per macro node, loam matrix + centred sand disc "
                 "r=0.25, Dirichlet -0.1 m top, h0=-5 m, t in [0, 1] d"),
    "ragged": (2, "ragged batch: macro 101x101, micro r in {11,16,21,31} per cell (seeded), 1-3 sand discs per "
                  "cell, Dirichlet -0.1 m top, h0=-5 m"),
    "scenarios": (3, "scenario batch: 64 scenarios of macro 41x41 / micro 16x16 with sampled van Genuchten "
                     "parameters, one launch per kernel for all, h0=-5 m"),
    "large": (4, "large: macro 201x201, micro 31x31 per macro node, one random sand disc per cell, Krylov "
                 "dimension up to 40 on the plain path, h0=-5 m"),
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_plain_workload.json")
T_START = time.perf_counter()


def log(msg):
    print(f"[bench {time.perf_counter() - T_START:6.1f}s] {msg}", file=sys.stderr, flush=True)


def integrate(T, s, u0, t0_, tend, chunk, label, snap_at=(), tau0=None, max_steps=None):
    """Run solver s from (u0, t0_) to tend (or for max_steps steps).  Returns (wall,
    totals, snapshots): at the first chunk boundary past each time in snap_at, a snapshot
    (t, u, proposal, steps and wall so far); the snapshot copies are not in the wall time."""
    L = T.lib()
    s.reset()
    s.set_state(u0, t0_)
    if tau0 is not None:
        s.set_step(tau0)
    L.tsr_device_sync()
    tot = {"steps": 0, "rejects": 0, "fevals": 0, "host_dense_s": 0.0, "sync_wait_s": 0.0, "t": t0_}
    snaps, pending = [], sorted(x for x in snap_at if x < tend)
    paused = 0.0
    t0 = time.perf_counter()
    last_log = t0
    while True:
        n = chunk if max_steps is None else min(chunk, max_steps - tot["steps"])
        rc, st, _ = s.run(tend, n)
        for key in ("steps", "rejects", "fevals", "host_dense_s", "sync_wait_s"):
            tot[key] += getattr(st, key)
        tot["t"] = st.t
        now = time.perf_counter()
        if now - last_log > 10.0 or st.t >= tend or rc:
            log(f"{label}: t={st.t:.4e}/{tend} exprem_steps={tot['steps']} wall={now - t0 - paused:.1f}s")
            last_log = now
        if rc:
            raise RuntimeError(f"{label}: integration stopped at t={st.t} rc={rc}: " + T.last_error())
        if st.t >= tend or (max_steps is not None and tot["steps"] >= max_steps):
            break
        if pending and st.t >= pending[0]:
            while pending and st.t >= pending[0]:
                pending.pop(0)
            p0 = time.perf_counter()
            u, t = s.get_state()
            snaps.append({"t": t, "u": u, "tau": s.stats().CurrentStep, "steps": tot["steps"],
                          "wall": p0 - t0 - paused})
            paused += time.perf_counter() - p0
    L.tsr_device_sync()
    return time.perf_counter() - t0 - paused, tot, snaps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--warmup-tend", type=float, default=1e-6)
    ap.add_argument("--budget-s", type=float, default=540.0,
                    help="wall-clock budget of the whole process; timed steps stop before overrunning it")
    ap.add_argument("--tend", type=float, default=0.0, help="override t_end (reported in config; default the workload's)")
    ap.add_argument("--cpu-sample-s", type=float, default=6.0, help="CPU oracle seconds per sampled state")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain path (profiling runs)")
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--ktimer-every", type=int, default=4, help="time every E-th launch of each kernel class")
    ap.add_argument("--chunk", type=int, default=250, help="EXPREM steps between progress checks")
    ap.add_argument("--rtol", type=float, default=0.0,
                    help="run both paths at rtol = atol = this value instead of the workload's 1e-6 (matched-accuracy "
                         "comparisons)")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched path without HIP graphs (by default each Arnoldi window shape is captured into a "
                         "HIP graph once and replayed: TSR_PATH_GRAPH)")
    ap.add_argument("--basis", choices=("f64", "f32"), default="f64",
                    help="storage of the Krylov basis columns on the batched path (arithmetic is FP64 either way)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only time the G kernels in their three modes on a perturbed initial state (for profilers)")
    ap.add_argument("--workload", choices=tuple(WORKLOADS), default="plain",
                    help="the timed workload (the headline is 'plain'; the others are measured with a bounded --tend)")
    ap.add_argument("--acc-tols", default="1e-6,1e-7,1e-8",
                    help="accuracy study: the rtol = atol values of the runs compared with the reference")
    ap.add_argument("--accuracy", type=float, default=0.0, metavar="T",
                    help="accuracy study instead of the benchmark: to t = T d, a plain run at rtol = atol = 1e-10 "
                         "as the reference, plain and learned+batched at 1e-6, 1e-7, 1e-8 against it")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU (--gpus must be 1); the path's sharding across GPUs is "
              "described in the design notes, not built", file=sys.stderr)
        sys.exit(2)

    import twoscalerich_amd as T
    from twoscalerich_amd import _ffi as F
    from twoscalerich_amd import controller as CT
    L = T.lib()

    p = T.Problem.config(WORKLOADS[args.workload][0])
    opt = p.options()
    if args.rtol > 0:
        opt.RelTol = opt.AbsTol = args.rtol
    tend = args.tend if args.tend > 0 else p.tend()
    u0 = p.initial()
    path_fast = F.PATH_BATCHED | (F.PATH_F32BASIS if args.basis == "f32" else 0) | (0 if args.no_graph else F.PATH_GRAPH)

    fast = T.Solver(p, opt)
    fast.set_path(path_fast)
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
    if args.accuracy > 0:
        fast.close()
        accuracy_study(T, F, CT, p, u0, args.accuracy, path_fast, [float(x) for x in args.acc_tols.split(",")])
        return
    plain = T.Solver(p, opt)     # plain path, classical controller (the defaults)

    for w in range(args.warmup):
        for s, lab in ((fast, "warm-up learned+batched"), (plain, "warm-up plain")):
            integrate(T, s, u0, 0.0, args.warmup_tend, args.chunk, f"{lab} {w}")

    # the plain path to t_end (the reference algorithm), with snapshots for the CPU samples
    plain_info, snaps, u_plain = None, [], None
    if not args.no_plain:
        w_plain, tot_plain, snaps = integrate(T, plain, u0, 0.0, tend, args.chunk, "plain",
                                              snap_at=(0.25 * tend, 0.5 * tend, 0.75 * tend))
        u_plain, _ = plain.get_state()
        plain_info = {"wall_s": round(w_plain, 3), "steps": tot_plain["steps"], "rhs_evals": tot_plain["fevals"],
                      "rejected": tot_plain["rejects"], "host_dense_s": round(tot_plain["host_dense_s"], 3),
                      "sync_wait_s": round(tot_plain["sync_wait_s"], 3)}
        snaps = [{"t": 0.0, "u": u0, "tau": opt.InitStep, "steps": 0, "wall": 0.0}] + snaps
        for k, sn in enumerate(snaps):
            nxt = snaps[k + 1] if k + 1 < len(snaps) else {"steps": tot_plain["steps"], "wall": w_plain}
            sn["seg_steps"] = nxt["steps"] - sn["steps"]
            sn["seg_wall"] = nxt["wall"] - sn["wall"]

    # timed: learned + batched to t_end, as many times as the budget allows (at least once)
    reserve = 12.0 + (0.0 if args.no_cpu else 4 * args.cpu_sample_s * 2.5 + 40.0)
    walls, runs = [], []
    fast.ktimer(args.ktimer_every)
    for k in range(max(args.steps, 1)):
        elapsed = time.perf_counter() - T_START
        if walls and elapsed + 1.05 * walls[-1] + reserve > args.budget_s:
            log(f"timed step {k}: {elapsed:.0f} s elapsed, the next step (~{walls[-1]:.0f} s) would overrun the "
                f"{args.budget_s:.0f} s budget: {len(walls)} timed step(s)")
            break
        w, tot, _ = integrate(T, fast, u0, 0.0, tend, 1000, f"learned+batched step {k}")
        walls.append(w)
        runs.append(tot)
    kt = fast.ktimer_read()
    fast.ktimer(0)
    wall = float(np.mean(walls))
    st = runs[-1]
    u_end, _ = fast.get_state()

    acc = None
    if plain_info is not None:
        plain_info["speedup"] = round(plain_info["wall_s"] / wall, 3)
        acc = {"rel_diff_2norm_t_end": float(np.linalg.norm(u_end - u_plain) / np.linalg.norm(u_plain)),
               "rel_diff_max_t_end": float(np.max(np.abs(u_end - u_plain)) / np.max(np.abs(u_plain))),
               "note": f"learned+batched vs plain GPU path at t_end, both at rtol = atol = {opt.RelTol:g} (different step "
                       "sequences); each path's own global error against a tight-tolerance solution is of the "
                       "same size (profiles/accuracy_plain_workload.txt)"}

    # per-kernel classes of the timed steps, and the dominant one's roofline
    kernels, total_ms = {}, sum(v["ms"] * v["seen"] / max(v["timed"], 1) for v in kt.values())
    for name, v in kt.items():
        if not v["timed"]:
            continue
        avg_ms = v["ms"] / v["timed"]
        kernels[name] = {"launches": v["seen"], "timed": v["timed"], "avg_launch_ms": round(avg_ms, 5),
                         "bytes_per_launch": round(v["bytes"] / v["timed"]),
                         "GB_per_s": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1),
                         "share_among_timed_classes": round(avg_ms * v["seen"] / total_ms, 4)}
    dom = max(kernels, key=lambda n: kernels[n]["share_among_timed_classes"])
    dk = kernels[dom]
    # HBM traffic per launch: the counter runs' (FETCH_SIZE + WRITE_SIZE, copy-calibrated)
    # bytes per algorithmic byte of this kernel class, times this run's algorithmic bytes
    # per launch (the DCGS2 passes' bytes grow with the Krylov index, so the ratio, not a
    # byte count, carries over from the counter run's shorter horizon)
    traffic, traffic_ratio = None, None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tj = json.load(f)
        traffic_ratio = tj.get("per_class", {}).get(dom, {}).get("traffic_over_algorithmic")
        if traffic_ratio is not None:
            traffic = round(traffic_ratio * dk["bytes_per_launch"])

    # G kernels per mode (HIP events on the solver stream), on the final state
    gf = gfunc_modes(T, fast, args.kernel_reps)
    n_micro = p.N - sum(p.view.contents.scen[q].ud.macro_mesh.no_nodes for q in range(p.view.contents.n_scen))
    for name, bpu in (("full", 16.0), ("record", 16.0), ("linearised_fd", 32.0)):
        b = bpu * n_micro + p.distinct_bytes()
        gf[name]["algorithmic_bytes"] = b
        gf[name]["GB_per_s"] = round(b / (gf[name]["micro_ms"] * 1e-3) / 1e9, 1)
    copy_ms = C.c_double()
    L.tsr_solver_time_copy(fast.h, args.kernel_reps, C.byref(copy_ms))
    copy_gbs = 16.0 * p.N / (copy_ms.value * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(T, F, p, opt, plain, snaps, tend, args.cpu_sample_s, plain_info)

    line = {
        "metric": "wall-time to t_end, learned+batched vs plain GPU EXPREM at ≤1e-6 rel err; G GB/s",
        "value": round(wall, 3),
        "unit": "s",
        "n_gpus": 1,
        "steps": len(walls),
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 1),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload][1],
                   "N": p.N, "t_end_days": tend, "rtol": opt.RelTol, "atol": opt.AbsTol,
                   "max_krylov_plain": opt.MaxKrylov, "controller": "learned (Krylov range [5, 60])",
                   "path": "batched (DCGS2" + (" on FP32 copies of the basis" if args.basis == "f32" else
                                               ", FP64 basis") + ", deferred Krylov checks, linearised FD products" +
                           ("" if args.no_graph else ", Arnoldi windows replayed from HIP graphs") + ")",
                   "steps_requested": args.steps, "budget_s": args.budget_s},
        "walls_s": [round(x, 3) for x in walls],
        "exprem": {"accepted_steps": st["steps"], "rejected": st["rejects"], "rhs_evals": st["fevals"],
                   "host_dense_s": round(st["host_dense_s"], 3), "sync_wait_s": round(st["sync_wait_s"], 3),
                   "u_end_min": float(u_end.min()), "u_end_max": float(u_end.max())},
        "plain": plain_info,
        "accuracy": acc,
        "kernels": kernels,
        "kernel_timing": f"HIP events on the solver stream around every {args.ktimer_every}th launch of each "
                         "class during the timed steps (inside the replayed HIP graphs: event-record nodes of a second "
                         "capture of each window, replayed every 8th time); bytes are algorithmic (reads and writes "
                         "the kernel must do)",
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": dk["GB_per_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dk["GB_per_s"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_over_algorithmic": traffic_ratio,
                     "algorithmic_bytes_per_launch": dk["bytes_per_launch"], "avg_launch_ms": dk["avg_launch_ms"],
                     "copy_kernel_GB_per_s": round(copy_gbs, 1)},
        "roofline_gfunc": {"kernel": "k_micro (linearised FD product)", "bound": "hbm",
                           "achieved": gf["linearised_fd"]["GB_per_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(gf["linearised_fd"]["GB_per_s"] / HBM_PEAK_GBS, 4),
                           "full_G_frac": round(gf["full"]["GB_per_s"] / HBM_PEAK_GBS, 4)},
        "gfunc": gf,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    fast.close()
    plain.close()


def accuracy_study(T, F, CT, p, u0, horizon, path_fast, tols):
    """Global error of each path at t = horizon against a plain run at rtol = atol = 1e-10,
    at rtol = atol in tols; one JSON line per run.  Every run starts from InitStep 1e-9
    with MinStep 1e-16 (the tightest tolerance's first steps need that)."""
    def run(tol, learned):
        opt = p.options()
        opt.RelTol = opt.AbsTol = tol
        opt.InitStep, opt.MinStep = 1e-9, 1e-16
        s = T.Solver(p, opt)
        if learned:
            s.set_path(path_fast)
            CT.attach(s)
        integrate(T, s, u0, 0.0, 1e-9, 10, "accuracy warm-up")
        label = f"accuracy {'learned+batched' if learned else 'plain'} tol {tol:g}"
        w, tot, _ = integrate(T, s, u0, 0.0, horizon, 1000, label)
        u, _ = s.get_state()
        s.close()
        return u, w, tot
    u_ref, w_ref, t_ref = run(1e-10, False)
    print(json.dumps({"accuracy_reference": {"t": horizon, "rtol_atol": 1e-10, "path": "plain", "wall_s": round(w_ref, 2),
                                             "steps": t_ref["steps"], "rhs_evals": t_ref["fevals"]}}), flush=True)
    for tol in tols:
        for learned in (False, True):
            u, w, tot = run(tol, learned)
            print(json.dumps({"accuracy": {"t": horizon, "rtol_atol": tol,
                                           "path": "learned+batched" if learned else "plain", "wall_s": round(w, 3),
                                           "steps": tot["steps"], "rhs_evals": tot["fevals"],
                                           "rel_err_max": float(np.max(np.abs(u - u_ref)) / np.max(np.abs(u_ref))),
                                           "rel_err_2norm": float(np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref))}}),
                  flush=True)


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _oracle_steps(o, ud, u, t, tau, tend, opt, F, budget_s, max_steps=10 ** 9):
    st = F.exprem_stats_struct()
    st.CurrentStep, st.CurrentTime, st.NumSteps, st.FuncEvals = tau, t, 0, 0
    n = 0
    t0 = time.perf_counter()
    while n < max_steps and t < tend and time.perf_counter() - t0 < budget_s:
        rc, t, _ = o.exprem_step(ud, u, t, tend, opt, st)
        if rc:
            raise RuntimeError(f"CPU oracle step failed at t={t}")
        n += 1
    return n, time.perf_counter() - t0, st.FuncEvals, t


def cpu_baseline(T, F, p, opt, plain, snaps, tend, sample_s, plain_info):
    """The CPU oracle (FP64 C++/OpenMP restatement of the plain algorithm) on this host."""
    from oracle.oracle import Oracle
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"value": None, "unit": "s", "cores": cores, "kind": "port", "cpu_model": cpu_model()}
    # 1. the plain CPU-runnable workload to t_end, whole, on the CPU and on the GPU
    ps = T.Problem.config(F.CFG_PLAIN_SMALL)
    os_ = ps.options()
    o = Oracle(ps.view, ps.ud)
    ud = ps.ud.contents
    ud.NUM_THREADS = cores
    u = ps.initial()
    n, cpu_s, fe, _ = _oracle_steps(o, ud, u, 0.0, os_.InitStep, ps.tend(), os_, F, 1e9)
    o.close()
    sg = T.Solver(ps, os_)
    w_gpu, tot, _ = integrate(T, sg, ps.initial(), 0.0, ps.tend(), 10 ** 6, "small workload, plain GPU path")
    ug, _ = sg.get_state()
    sg.close()
    out["small_workload_measured"] = {
        "workload": "plain CPU-runnable: macro 21x21, micro 11x11, t in [0, 0.1] d, N = %d" % ps.N,
        "cpu_s": round(cpu_s, 3), "cpu_steps": n, "cpu_rhs_evals": fe,
        "gpu_plain_s": round(w_gpu, 3), "gpu_steps": tot["steps"], "gpu_rhs_evals": tot["fevals"],
        "cpu_over_gpu": round(cpu_s / w_gpu, 2),
        "rel_diff_max_t_end": float(np.max(np.abs(ug - u)) / np.max(np.abs(u)))}
    log(f"small workload: CPU {cpu_s:.1f} s ({n} steps), GPU plain {w_gpu:.2f} s ({tot['steps']} steps)")
    ps.close()
    if not snaps:
        out["sample"] = "small workload only (the plain path was not run)"
        return out
    # 2. the plain workload, sampled from the plain GPU run's states
    o = Oracle(p.view, p.ud)
    ud = p.ud.contents
    ud.NUM_THREADS = cores
    est, parts = 0.0, []
    for sn in snaps:
        u = sn["u"].copy()
        n, c_s, fe, t_c = _oracle_steps(o, ud, u, sn["t"], sn["tau"], tend, opt, F, sample_s)
        g_s, g_tot, _ = integrate(T, plain, sn["u"], sn["t"], tend, n, f"plain GPU path, {n} steps from t={sn['t']:.3g}",
                                  tau0=sn["tau"], max_steps=n)
        ratio = c_s / g_s
        est += ratio * sn["seg_wall"]
        parts.append({"t_start": round(sn["t"], 5), "cpu_steps": n, "cpu_s": round(c_s, 3), "cpu_rhs_evals": fe,
                      "gpu_steps": g_tot["steps"], "gpu_rhs_evals": g_tot["fevals"], "gpu_s": round(g_s, 4),
                      "cpu_over_gpu": round(ratio, 1), "segment_gpu_steps": sn["seg_steps"],
                      "segment_gpu_wall_s": round(sn["seg_wall"], 3)})
        log(f"CPU sample from t={sn['t']:.4g}: {n} steps in {c_s:.1f} s, GPU {g_s:.3f} s, ratio {ratio:.0f}")
    o.close()
    out["value"] = round(est, 1)
    out["samples"] = parts
    out["sample"] = ("plain workload: a few seconds of CPU oracle EXPREM steps from each of four states of the plain "
                     "GPU run (t = 0 and the first step ends past 0.25, 0.5, 0.75 d); the plain GPU path timed over "
                     "the same steps from the same states; value = sum over the four segments of the plain GPU "
                     "path's wall time x the CPU/GPU ratio sampled at the segment's start (an extrapolation: "
                     "%.0f s of CPU steps stand for the whole horizon)" % sum(x["cpu_s"] for x in parts))
    return out


if __name__ == "__main__":
    main()
