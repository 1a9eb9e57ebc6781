"""Benchmark: rod-steps/s of the full collision-constraint timestep on one MI355X.

A step is one pass of the hot path over the device-resident rod system: U_nc (mobility + Brownian),
cell list + pair sweep + rows, persistent BBPGD to residual tol, write-back, Euler move.
Default workload: the 1,000,000-rod lognormal 3-D box (lengths lognormal(ln 1 um, 0.3) resampled
above 3 um, periodic 25^3 um^3, isotropic), FP64 plain BBPGD, each step one replay of a captured HIP
graph.  Beside it: the chosen form (FP32 + FP64 refinement, learned warm start) with its accuracy and
same-state iteration counts, and the FP64 oracle on the host cores for the same step, with gamma and
U compared bit for bit.  Prints one JSON line.  The library is driven through its C ABI only (no
torch in the process, so profiled runs load one HIP runtime).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lognormal_3d|monolayer_slab|small_3d|replicas]
                  [--precision 0|1] [--graph 1|0] [--no-cpu] [--no-form] [--no-autotune]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
SPEEDUP_TARGET = 20.0  # one 1M-rod GPU step against the host-CPU restatement of the reference path
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "k_bbpgd_traffic.json")


def bytes_per_matvec(m, n, w=8):
    """Algorithmic HBM bytes of one BBPGD matvec + vector update (compulsory traffic):
    per constraint 2 x (9w geometry + 8 B rod ids) for D and D^T, plus 5w for x, g (read),
    g_prev, b and the new x/g (write); per rod 6w mobility, 6w U written, 6w U read."""
    return m * (2 * (9 * w + 8) + 5 * w) + n * 18 * w


def cpu_model():
    """CPU model string plus family/model numbers from /proc/cpuinfo."""
    info = {}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, v = line.partition(":")
                k = k.strip()
                if k in ("model name", "cpu family", "model", "stepping") and k not in info:
                    info[k] = v.strip()
    except OSError:
        pass
    name = info.get("model name") or platform.processor() or "unknown"
    return f"{name} (family {info.get('cpu family', '?')}, model {info.get('model', '?')})"


def cpu_topology():
    """Logical CPUs this process may run on, and the physical cores behind them (distinct (package,
    core) pairs from sysfs; equal to the logical count when SMT siblings are not both visible)."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "physical_package_id") as f1, open(base + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    return len(cpus), len(cores)


def setup_cpu_threads(requested=0):
    """OpenMP placement for the CPU leg, set before the oracle's OpenMP runtime loads: one thread per
    physical core, spread over the cores (OMP_PROC_BIND=spread, OMP_PLACES=cores)."""
    logical, physical = cpu_topology()
    threads = requested or physical
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["OMP_PROC_BIND"] = "spread"
    os.environ["OMP_PLACES"] = "cores"
    return threads, logical, physical


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64).reshape(-1)
    b = np.ascontiguousarray(b, np.float64).reshape(-1)
    return a.size == b.size and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def load_traffic(workload):
    """Measured solver traffic of a workload (rocprofv3 FETCH_SIZE + WRITE_SIZE passes), or None."""
    if not os.path.exists(TRAFFIC_JSON):
        return None
    with open(TRAFFIC_JSON) as fh:
        tr = json.load(fh)
    return tr.get(workload)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="lognormal_3d",
                    choices=["lognormal_3d", "monolayer_slab", "small_3d", "replicas"])
    ap.add_argument("--precision", type=int, default=0, help="1: FP32 BBPGD + FP64 refinement for the timed pass")
    ap.add_argument("--n", type=int, default=0, help="rod count (default: the workload's own)")
    ap.add_argument("--cell-size", type=float, default=0.0,
                    help="cell-list edge in um (default: the library's choice, a quarter of the mean rod length)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-form", action="store_true", help="skip the chosen-form (FP32 + warm start) pass")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default solver schedule (otherwise the autotuner picks it after warm-up)")
    ap.add_argument("--graph", type=int, default=1, choices=[0, 1],
                    help="1: each step after the first is one replay of a captured HIP graph (alens_set_graph)")
    ap.add_argument("--progress", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="before the run, one device-to-device copy of this many MiB (a kernel of known "
                         "traffic for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--rocprof-summary", nargs=3, metavar=("TRACE_DIR", "FETCH_DIR", "WRITE_DIR"),
                    help="summarise rocprofv3 output of a bench run made with the same --warmup / --steps "
                         "(--no-form --no-cpu --calib-copy-mib): prints markdown, writes --traffic-out")
    ap.add_argument("--traffic-out", default=TRAFFIC_JSON,
                    help="JSON of measured solver traffic, one record per workload (merged)")
    args = ap.parse_args()
    if args.rocprof_summary:
        return rocprof_summary(args)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    from alens_amd import capi, systems

    if args.calib_copy_mib:
        # one k_copy_probe launch of known traffic (reads and writes calib_copy_mib MiB) for scaling
        # the profiler's FETCH_SIZE / WRITE_SIZE counters
        capi.copy_probe(args.calib_copy_mib << 20, reps=1)
    # (skipped in profiled runs: its copy kernels would sit beside the calibration copy)
    copy_gbs = None if args.calib_copy_mib else capi.copy_probe(1 << 30, reps=10)
    if args.workload == "replicas":
        return bench_replicas(args, capi, systems, copy_gbs)
    gen = {"monolayer_slab": systems.monolayer_slab, "lognormal_3d": systems.lognormal_3d,
           "small_3d": systems.small_3d}[args.workload]
    rods, cfg = gen(n=args.n) if args.n else gen()
    if args.cell_size > 0:
        cfg = dict(cfg, cell_size=args.cell_size)
    n = len(rods["length"])
    ctx = capi.Context(**dict(cfg, precision=args.precision))
    ctx.set_rods(rods)
    ctx.set_graph(bool(args.graph))
    # spatial (Morton cell) order, as the reference's FDPS containers keep it; later arrays follow it
    perm = ctx.sort_rods()
    rods = {k: (v[perm] if isinstance(v, np.ndarray) else v) for k, v in rods.items()}

    def one_step():
        rep = ctx.step()
        ctx.euler()
        return rep

    for w in range(args.warmup):
        rep = one_step()
        if args.progress:
            print(f"warmup {w}: M={rep['n_con']} iters={rep['iters']} step={rep['t_step_ms']:.2f} ms",
                  file=sys.stderr, flush=True)
    tune = None
    if not args.no_autotune and args.warmup > 0:
        # schedule picked on the last warm-up step's rows (outputs of that step are unchanged)
        ctx.step()
        tune = ctx.autotune()
        ctx.euler()
        one_step()  # untimed: in graph mode the autotuner dropped the graph, and this records it again
    reps = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        reps.append(one_step())
        if args.progress and (s % 20 == 0):
            r = reps[-1]
            print(f"step {s}: M={r['n_con']} iters={r['iters']} step={r['t_step_ms']:.2f} ms",
                  file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    K = args.steps
    iters = np.array([r["iters"] for r in reps])
    it32 = np.array([r["iters_fp32"] for r in reps])
    mv = np.array([r["matvecs"] for r in reps])
    ms = np.array([r["n_con"] for r in reps])
    tsolve = np.array([r["t_solve_ms"] for r in reps]) * 1e-3
    # FP64 matvecs (plain solve, or the FP32 form's residual check and refinement) at 8-byte words,
    # the FP32 form's iterations at 4-byte words
    alg = sum(bytes_per_matvec(int(m), n) * int(v) + bytes_per_matvec(int(m), n, w=4) * int(f)
              for m, v, f in zip(ms, mv, it32))
    achieved = alg / tsolve.sum() / 1e9 if tsolve.sum() > 0 else 0.0
    value = n * K / elapsed

    out = {
        "metric": "rod-steps/sec at BBPGD residual 1e-5; BBPGD iters/step, plain vs chosen form",
        "value": value,
        "unit": "rod-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {n} rods, "
                        f"{'FP32 BBPGD + FP64 refinement' if args.precision else 'FP64 plain BBPGD'}, "
                        f"tol {cfg['tol']}, dt {cfg['dt']} s",
            "rods": n,
            "box": list(cfg["box_hi"]),
            "pbc": list(cfg["pbc"]),
        },
        "solver_schedule": tune,
        "bbpgd_iters_per_step": float((iters + it32).mean()),
        "bbpgd_iters_per_step_plain_fp64": None if args.precision else float(iters.mean()),
        "fp32_iters_per_step": float(it32.mean()) if args.precision else None,
        "matvecs_per_step": [int(v) for v in mv],
        "constraints_per_step": float(ms.mean()),
        "constraints_per_timed_step": [int(v) for v in ms],
        "t_collect_ms_mean": float(np.mean([r["t_collect_ms"] for r in reps])),
        "t_solve_ms_mean": float(1e3 * tsolve.mean()),
        "roofline": {
            "bound": "hbm",
            "kernel": ("k_bbpgd32 + k_bbpgd64 (FP32 BBPGD, FP64 residual check and refinement)"
                       if args.precision else "k_bbpgd64 (persistent BBPGD, FP64)"),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": float(alg / K),
            "stream_copy_gbs": copy_gbs,
            "frac_of_stream_copy": achieved / copy_gbs if copy_gbs else None,
        },
        "graph": ctx.graph_info() if args.graph else None,
    }
    if args.graph and out["graph"]["replays"] == 0:
        print(f"graph mode did not replay: {out['graph']}", file=sys.stderr, flush=True)
    # measured HBM traffic per launch: the ratio of measured to algorithmic bytes from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE passes on this workload, times this run's algorithmic bytes
    tr = load_traffic(args.workload) if (not args.n and args.precision == 0) else None
    if tr:
        out["roofline"]["traffic"] = tr["traffic_over_algorithmic"] * float(alg / K)
        out["roofline"]["traffic_source"] = (
            f"profiles/k_bbpgd_traffic.json[{args.workload}]: rocprofv3 FETCH_SIZE + WRITE_SIZE of the solver, "
            f"{tr['traffic_over_algorithmic']:.2f} x its algorithmic bytes, times this run's algorithmic bytes")

    # chosen form on the same workload: FP32 iterations + FP64 residual/refinement, started from the
    # learned network (the shipped weights); plain BBPGD finishes every solve
    if not args.no_form:
        from alens_amd import warmstart as W

        P, norm = W.load()
        fctx = capi.Context(**dict(cfg, precision=1, warm_start=2))
        fctx.mlp_set(P, norm)
        fctx.set_rods(rods)
        fctx.set_graph(bool(args.graph))
        fctx.sort_rods()
        for w in range(args.warmup):
            fctx.step()
            fctx.euler()
        if not args.no_autotune and args.warmup > 0:
            fctx.step()
            fctx.autotune()
            fctx.euler()
            fctx.step()
            fctx.euler()
        freps = []
        f0 = time.perf_counter()
        for s in range(args.steps):
            freps.append(fctx.step())
            fctx.euler()
        f1 = time.perf_counter()
        fit32 = np.array([r["iters_fp32"] for r in freps])
        fit64 = np.array([r["iters"] for r in freps])
        out["chosen_form"] = {
            "form": "FP32 BBPGD + FP64 residual/refinement, warm start from the learned network",
            "value": n * K / (f1 - f0),
            "unit": "rod-steps/s",
            "ms_per_step": 1e3 * (f1 - f0) / K,
            "bbpgd_iters_per_step": float((fit32 + fit64).mean()),
            "fp32_iters_per_step": float(fit32.mean()),
            "fp64_refine_iters_per_step": float(fit64.mean()),
            "fallbacks": int(sum(r["fallback_used"] for r in freps)),
            "iteration_ratio_plain_over_form": float(iters.mean() / max((fit32 + fit64).mean(), 1e-9)),
        }
        # iterations on identical states (a plain FP64 driver moves the rods; each form keeps its own
        # previous-step impulses): plain FP64 against the FP32 form with each warm start
        ev = W.evaluate(args.workload, n, 4, P, norm, seed=cfg["seed"], warmup=40,
                        forms={"plain_fp64": (0, 0), "fp32_cold": (0, 1), "fp32_prev_gamma": (1, 1),
                               "fp32_network": (2, 1), "fp64_network": (2, 0)}, lead=2)
        pm = float(ev["plain_fp64"].mean())
        out["chosen_form"]["same_state_iterations"] = {
            "states": "4 steps after 40 warm-up steps of a plain FP64 driver from the workload's seeded start "
                      "(each form also steps the 2 states before them, for its previous-step keys)",
            **{k: float(ev[k].mean()) for k in ("plain_fp64", "fp32_cold", "fp32_prev_gamma", "fp32_network",
                                                "fp64_network")},
            "fallbacks_fp32_network": int(ev["fallback_fp32_network"]),
            "ratio_plain_over_fp32_network": pm / max(float(ev["fp32_network"].mean()), 1e-9),
            "target_ratio": 2.0,
        }
        # accuracy of the form, stated per result: from the post-benchmark state, the FP32 form (cold) and
        # the plain FP64 solve (both to residual tol) against an FP64 solve to 1e-11
        c, t = fctx.get_rods()
        st = fctx.step_count
        fctx.close()
        del fctx
        sols = {}
        for name, kw in (("fp32_form", dict(precision=1)), ("plain_fp64", {}),
                         ("tight_fp64", dict(tol=1e-11, max_ite=2_000_000))):
            x = capi.Context(**dict(cfg, **kw))
            x.set_rods(dict(rods, center=c, dir=t))
            x.set_step(st)
            r = x.step()
            sols[name] = dict(gamma=x.gamma(), u=x.velocity()["u"], res=r["res"], iters=r["iters"] + r["iters_fp32"])
            x.close()
        gt, ut = sols["tight_fp64"]["gamma"], sols["tight_fp64"]["u"]
        gs, us = max(np.max(np.abs(gt)), 1e-300), max(np.max(np.abs(ut)), 1e-300)
        out["chosen_form"]["accuracy"] = {
            "reference": "FP64 BBPGD to residual 1e-11 on the same state "
                         f"({sols['tight_fp64']['iters']} iterations)",
            **{name: {"gamma_max_rel_err": float(np.max(np.abs(sols[name]["gamma"] - gt)) / gs),
                      "u_max_rel_err": float(np.max(np.abs(sols[name]["u"] - ut)) / us),
                      "fp64_residual": float(sols[name]["res"]), "iters": int(sols[name]["iters"])}
               for name in ("fp32_form", "plain_fp64")},
        }

    if not args.no_cpu:
        # the same step on both sides: from the post-benchmark state, one more GPU step (timed alone)
        # and the FP64 oracle step on the host cores; gamma and U are compared bit for bit
        threads, logical, physical = setup_cpu_threads(args.cpu_threads)
        from oracle import oracle as O
        O.set_threads(threads)
        c, t = ctx.get_rods()
        state = dict(rods, center=c, dir=t)
        oc = O.cfg_from_dict(cfg)
        oc.step = ctx.step_count
        g0 = time.perf_counter()
        grep = ctx.step()
        tg = time.perf_counter() - g0
        g_gpu, u_gpu = ctx.gamma(), ctx.velocity()["u"]
        ts = time.perf_counter()
        res = O.step(state, oc)
        tc = time.perf_counter() - ts
        same = (res["report"]["iters"] == grep["iters"] and res["I"].size == grep["n_con"]
                and same_bits(g_gpu, res["gamma"]) and same_bits(u_gpu, res["u"]))
        out["cpu_baseline"] = {
            "value": n / tc,
            "unit": "rod-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"1 full step (FP64 oracle, OpenMP, {threads} threads, OMP_PROC_BIND=spread OMP_PLACES=cores; "
                      f"{logical} logical CPUs = {physical} physical cores visible) from the post-benchmark state: "
                      f"{res['I'].size} constraints, {res['report']['iters']} BBPGD iterations, {tc:.3f} s; the GPU "
                      f"took {tg * 1e3:.2f} ms for the same step ({grep['iters']} iterations; gamma and U "
                      f"{'bit-identical' if same else 'MISMATCH'}); CPU {cpu_model()}",
            "same_step_bit_identical": bool(same),
            "logical_cpus": logical,
            "physical_cores": physical,
        }
        out["gpu_over_cpu"] = value / (n / tc)
        out["gpu_over_cpu_same_step"] = tc / tg
        out["gpu_over_cpu_target"] = SPEEDUP_TARGET
    ctx.close()
    print(json.dumps(out), flush=True)


def _csv(path):
    import csv
    with open(path) as fh:
        return list(csv.DictReader(fh))


def _timed_window(rows, steps):
    """Dispatches of the timed steps: from the k_mob_unc that opens the step of the K-th last solve
    (warm-up and tuning solves come before it)."""
    rows = sorted(rows, key=lambda r: int(r["Start_Timestamp"]))
    solves = [i for i, r in enumerate(rows) if "k_bbpgd" in r["Kernel_Name"]]
    start = solves[-steps]
    while start > 0 and "k_mob_unc" not in rows[start]["Kernel_Name"]:
        start -= 1
    return rows[start:]


def rocprof_summary(args):
    """Markdown summary of a profiled bench run (kernel trace plus FETCH_SIZE / WRITE_SIZE passes), and
    the solver's measured traffic per matvec as JSON for the roofline object."""
    tdir, fdir, wdir = args.rocprof_summary
    trace = _timed_window(_csv(os.path.join(tdir, "run_kernel_trace.csv")), args.steps)
    stats = {}
    for r in trace:
        nm = r["Kernel_Name"].split("(")[0].replace("void ", "")
        d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        e = stats.setdefault(nm, [0, 0])
        e[0] += 1
        e[1] += d
    tot = sum(v[1] for v in stats.values())
    print("| kernel | calls | total ms | average us | share |")
    print("|---|---:|---:|---:|---:|")
    for nm, (c, d) in sorted(stats.items(), key=lambda kv: -kv[1][1])[:15]:
        print(f"| `{nm}` | {c} | {d / 1e6:.3f} | {d / c / 1e3:.1f} | {100.0 * d / tot:.1f}% |")
    out = {}
    for key, d, cname in (("fetch", fdir, "FETCH_SIZE"), ("write", wdir, "WRITE_SIZE")):
        rows = [r for r in _csv(os.path.join(d, "run_counter_collection.csv")) if r["Counter_Name"] == cname]
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        # calibration: the --calib-copy-mib launch of k_copy_probe (reads N, writes N bytes)
        calib = [r for r in rows if "k_copy_probe" in r["Kernel_Name"]][0]
        true_bytes = float(args.calib_copy_mib) * (1 << 20)
        factor = true_bytes / (float(calib["Counter_Value"]) * 1024.0)
        solves = [r for r in rows if "k_bbpgd" in r["Kernel_Name"]][-args.steps:]
        out[key] = dict(factor=factor, counter_kib=[float(r["Counter_Value"]) for r in solves],
                        bytes=[float(r["Counter_Value"]) * 1024.0 * factor for r in solves])
    # matvecs per timed solve from the bench line of the FETCH pass (same seed and steps)
    with open(os.path.join(fdir, "..", os.path.basename(fdir) + ".log")) as fh:
        js = [json.loads(l) for l in fh if l.startswith("{")][-1]
    mv = js["matvecs_per_step"]
    alg = [bytes_per_matvec(int(m), int(js["config"]["rods"])) for m in js["constraints_per_timed_step"]]
    per_mv = [(f + w) / v for f, w, v in zip(out["fetch"]["bytes"], out["write"]["bytes"], mv)]
    res = dict(kernel="k_bbpgd<double>", workload=js["config"]["workload"], rods=int(js["config"]["rods"]),
               constraints_per_timed_step=js["constraints_per_timed_step"], matvecs_per_timed_step=mv,
               fetch_factor=out["fetch"]["factor"], write_factor=out["write"]["factor"],
               traffic_bytes_per_matvec=float(np.mean(per_mv)),
               algorithmic_bytes_per_matvec=float(np.mean(alg)),
               traffic_over_algorithmic=float(np.mean(per_mv) / np.mean(alg)))
    print()
    print("| solver traffic (timed steps) | value |")
    print("|---|---:|")
    print(f"| FETCH_SIZE scale (calibration copy of {args.calib_copy_mib} MiB) | {res['fetch_factor']:.3f} |")
    print(f"| WRITE_SIZE scale | {res['write_factor']:.3f} |")
    print(f"| measured bytes per matvec (fetch + write, scaled) | {res['traffic_bytes_per_matvec'] / 1e6:.1f} MB |")
    print(f"| algorithmic bytes per matvec (200 M + 144 N) | {res['algorithmic_bytes_per_matvec'] / 1e6:.1f} MB |")
    print(f"| measured / algorithmic | {res['traffic_over_algorithmic']:.2f} |")
    allw = {}
    if os.path.exists(args.traffic_out):
        with open(args.traffic_out) as fh:
            allw = json.load(fh)
    allw[res["workload"].split(":")[0]] = res
    with open(args.traffic_out, "w") as fh:
        json.dump(allw, fh, indent=1)


def _run_replicas(args, capi, parts, cfg, precision, warm_start, P=None, norm=None, keep=False):
    ctx = capi.Context(**dict(cfg, precision=precision, warm_start=warm_start))
    if warm_start == 2:
        ctx.mlp_set(P, norm)
    offs = ctx.set_systems(parts)
    ctx.set_graph(bool(args.graph))
    perm = ctx.sort_rods()
    n = int(offs[-1])
    for w in range(args.warmup):
        ctx.step()
        ctx.euler()
    tune = None
    if not args.no_autotune and args.warmup > 0:
        ctx.step()
        tune = ctx.autotune()
        ctx.euler()
        ctx.step()
        ctx.euler()
    reps, sreps = [], []
    t0 = time.perf_counter()
    for s in range(args.steps):
        reps.append(ctx.step())
        sreps.append(ctx.system_reports())
        ctx.euler()
    el = time.perf_counter() - t0
    it = np.array([[r["iters"] + r["iters_fp32"] for r in sr] for sr in sreps])  # steps x replicas
    hist, edges = np.histogram(it.mean(0), bins=[0, 25, 50, 100, 200, 400, 800, 1600, 1e9])
    # algorithmic solver bytes: every system moves its own rows and rods once per matvec
    nsys_rods = np.diff(offs)
    w = 4 if precision else 8
    alg = 0.0
    for sr in sreps:
        for q, r in enumerate(sr):
            alg += r["matvecs"] * bytes_per_matvec(r["n_con"], int(nsys_rods[q]))
            if precision:
                alg += r["iters_fp32"] * bytes_per_matvec(r["n_con"], int(nsys_rods[q]), w=4)
    res = dict(n=n, elapsed=el, reps=reps, sreps=sreps, tune=tune, graph=ctx.graph_info() if args.graph else None,
               iters_mean_per_replica=float(it.mean()),
               hist={"edges": [float(e) for e in edges[:-1]], "counts": hist.tolist()},
               fallbacks=int(sum(r["fallback_used"] for sr in sreps for r in sr)),
               alg_bytes=alg, t_solve_s=float(sum(r["t_solve_ms"] for r in reps)) * 1e-3, offs=offs, perm=perm)
    if keep:
        res["ctx"] = ctx
    else:
        ctx.close()
    return res


def bench_replicas(args, capi, systems, copy_gbs=None):
    """64 independent replicas (rod counts log-uniform in [2k, 50k], 3-D boxes and slabs) stepped
    together: one collection pass and one persistent solver launch per step for all of them (one
    graph replay per step in graph mode).  Plain FP64 from a cold start gives `value`; the chosen
    form (FP32 + FP64 refinement, learned warm start) runs the same workload beside it, and the FP64
    oracle takes the same step of every replica on the host cores."""
    parts, cfg = systems.replicas()
    K = args.steps
    a = _run_replicas(args, capi, parts, cfg, args.precision, 0, keep=True)
    n, reps = a["n"], a["reps"]
    achieved = a["alg_bytes"] / a["t_solve_s"] / 1e9 if a["t_solve_s"] > 0 else 0.0
    out = {
        "metric": "rod-steps/sec at BBPGD residual 1e-5 (64 replicas in one launch per step)",
        "value": n * K / a["elapsed"], "unit": "rod-steps/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": 1e3 * a["elapsed"] / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.precision else "f64", "data": "synthetic",
        "config": {"workload": f"replicas: {len(parts)} systems, {n} rods, cold start", "rods": n,
                   "replicas": len(parts)},
        "constraints_per_step": float(np.mean([r["n_con"] for r in reps])),
        "solver_iters_max_per_step": float(np.mean([r["iters"] + r["iters_fp32"] for r in reps])),
        "replica_iters_mean": a["iters_mean_per_replica"],
        "replica_iters_histogram": a["hist"],
        "t_solve_ms_mean": float(np.mean([r["t_solve_ms"] for r in reps])),
        "t_collect_ms_mean": float(np.mean([r["t_collect_ms"] for r in reps])),
        "solver_schedule": a["tune"],
        "graph": a["graph"],
        "roofline": {
            "bound": "hbm", "kernel": "k_bbpgd64 (persistent BBPGD over all replicas)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None, "algorithmic_bytes_per_launch": a["alg_bytes"] / K,
            "note": "algorithmic bytes summed per system: matvecs_s x (200 M_s + 144 N_s)",
            "stream_copy_gbs": copy_gbs,
            "frac_of_stream_copy": achieved / copy_gbs if copy_gbs else None,
        },
    }
    tr = load_traffic("replicas")
    if tr and args.precision == 0:
        out["roofline"]["traffic"] = tr["traffic_over_algorithmic"] * a["alg_bytes"] / K
        out["roofline"]["traffic_source"] = "profiles/k_bbpgd_traffic.json[replicas], scaled by this run's algorithmic bytes"
    if not args.no_cpu:
        ctx = a["ctx"]
        threads, logical, physical = setup_cpu_threads(args.cpu_threads)
        from oracle import oracle as O
        O.set_threads(threads)
        offs, perm = a["offs"], a["perm"]
        cat = {k: np.concatenate([np.asarray(r[k]).reshape(len(r["length"]), -1) for r, _ in parts])[perm]
               for k in ("gid", "length", "radius")}
        c, t = ctx.get_rods()
        st = ctx.step_count
        g0 = time.perf_counter()
        grep = ctx.step()
        tg = time.perf_counter() - g0
        g = ctx.constraint_arrays()
        gam, u = ctx.gamma(), ctx.velocity()["u"]
        tc, same, ncon = 0.0, True, 0
        for s_, (_, box) in enumerate(parts):
            lo, hi = int(offs[s_]), int(offs[s_ + 1])
            rods = dict(gid=cat["gid"][lo:hi, 0], length=cat["length"][lo:hi, 0], radius=cat["radius"][lo:hi, 0],
                        center=c[lo:hi], dir=t[lo:hi], immovable=None)
            oc = O.cfg_from_dict(dict(cfg, **box))
            oc.step = st
            ts = time.perf_counter()
            res = O.step(rods, oc)
            tc += time.perf_counter() - ts
            sel = (g["I"] >= lo) & (g["I"] < hi)
            same &= bool(same_bits(gam[sel], res["gamma"]) and same_bits(u[6 * lo:6 * hi], res["u"]))
            ncon += res["I"].size
        out["cpu_baseline"] = {
            "value": n / tc, "unit": "rod-steps/s", "cores": threads, "kind": "port",
            "sample": f"1 full step of all {len(parts)} replicas one after another (FP64 oracle, OpenMP, {threads} "
                      f"threads, OMP_PROC_BIND=spread OMP_PLACES=cores; {logical} logical CPUs = {physical} physical "
                      f"cores) from the post-benchmark state: {ncon} constraints, {tc:.3f} s; the GPU took "
                      f"{tg * 1e3:.2f} ms for the same step (gamma and U {'bit-identical' if same else 'MISMATCH'} "
                      f"per replica); CPU {cpu_model()}",
            "same_step_bit_identical": bool(same),
        }
        out["gpu_over_cpu"] = out["value"] / (n / tc)
    a["ctx"].close()
    if not args.no_form:
        from alens_amd import warmstart as W

        P, norm = W.load()
        b = _run_replicas(args, capi, parts, cfg, 1, 2, P, norm, keep=True)
        out["chosen_form"] = {
            "form": "FP32 BBPGD + FP64 residual/refinement, warm start from the learned network",
            "value": n * K / b["elapsed"], "unit": "rod-steps/s", "ms_per_step": 1e3 * b["elapsed"] / K,
            "replica_iters_mean": b["iters_mean_per_replica"], "replica_iters_histogram": b["hist"],
            "fallbacks": b["fallbacks"], "solver_schedule": b["tune"], "graph": b["graph"],
            "accuracy": replica_accuracy(capi, parts, cfg, b),
        }
        b["ctx"].close()
    print(json.dumps(out), flush=True)


def replica_accuracy(capi, parts, cfg, run):
    """Error of the FP32 form per replica, stated per result: from the form run's final state, the FP32
    form (cold, to tol) and the plain FP64 solve (to tol) against an FP64 solve to 1e-11, each as the
    largest |gamma - gamma_ref| of the replica over the replica's largest |gamma_ref| (and the same for
    U).  All replicas are solved together in one context per solve, as in the benchmark."""
    ctx, offs, perm = run["ctx"], run["offs"], run["perm"]
    c, t = ctx.get_rods()
    st = ctx.step_count
    cat = {k: np.concatenate([np.asarray(r[k]).reshape(len(r["length"]), -1) for r, _ in parts])[perm]
           for k in ("gid", "length", "radius")}
    sparts = []
    for s_, (_, box) in enumerate(parts):
        lo, hi = int(offs[s_]), int(offs[s_ + 1])
        sparts.append((dict(gid=cat["gid"][lo:hi, 0], length=cat["length"][lo:hi, 0], radius=cat["radius"][lo:hi, 0],
                            center=c[lo:hi], dir=t[lo:hi], immovable=None), box))
    sols = {}
    for name, kw in (("fp32_form", dict(precision=1)), ("plain_fp64", {}),
                     ("tight_fp64", dict(tol=1e-11, max_ite=2_000_000))):
        x = capi.Context(**dict(cfg, **kw))
        x.set_systems(sparts)
        x.set_step(st)
        x.step()
        g = x.constraint_arrays()
        sols[name] = dict(I=g["I"], gamma=x.gamma(), u=x.velocity()["u"])
        x.close()
    ref = sols["tight_fp64"]
    res = {}
    for name in ("fp32_form", "plain_fp64"):
        eg, eu = [], []
        for s_ in range(len(parts)):
            lo, hi = int(offs[s_]), int(offs[s_ + 1])
            sel = (ref["I"] >= lo) & (ref["I"] < hi)
            gr, ur = ref["gamma"][sel], ref["u"][6 * lo:6 * hi]
            eg.append(float(np.max(np.abs(sols[name]["gamma"][sel] - gr)) / max(np.max(np.abs(gr)), 1e-300))
                      if gr.size else 0.0)
            eu.append(float(np.max(np.abs(sols[name]["u"][6 * lo:6 * hi] - ur)) / max(np.max(np.abs(ur)), 1e-300)))
        res[name] = {"gamma_max_rel_err": {"median": float(np.median(eg)), "max": float(np.max(eg))},
                     "u_max_rel_err": {"median": float(np.median(eu)), "max": float(np.max(eu))}}
    res["reference"] = "per replica: FP64 BBPGD to residual 1e-11 on the same state"
    return res


if __name__ == "__main__":
    main()
