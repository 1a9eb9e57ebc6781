"""Benchmark: rod-steps/s of the full collision-constraint timestep on one MI355X.

A step is one pass of the hot path over the device-resident rod system: U_nc (mobility + Brownian),
cell list + pair sweep + rows, persistent BBPGD to residual tol, write-back, Euler move.
Default workload: the dense quasi-2D monolayer (200,000 rods, L = 1 um, 100 x 100 um slab, nematic
von Mises kappa = 4), FP64 plain BBPGD, each step one replay of a captured HIP graph.  Beside it: the
chosen form (FP32 + FP64 refinement, learned warm start) with its accuracy and same-state iteration
counts, and the FP64 oracle on the host cores for the same step.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload monolayer_slab|lognormal_3d|small_3d|replicas]
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


def bytes_per_matvec(m, n, w=8):
    """Algorithmic HBM bytes of one BBPGD matvec + vector update (compulsory traffic):
    per constraint 2 x (9w geometry + 8 B rod ids) for D and D^T, plus 5w for x, g (read),
    g_prev, b and the new x/g (write); per rod 6w mobility, 6w U written, 6w U read."""
    return m * (2 * (9 * w + 8) + 5 * w) + n * 18 * w


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="monolayer_slab",
                    choices=["monolayer_slab", "lognormal_3d", "small_3d", "replicas"])
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
    ap.add_argument("--traffic-out", default=os.path.join(ROOT, "profiles", "k_bbpgd_traffic.json"))
    args = ap.parse_args()
    if args.rocprof_summary:
        return rocprof_summary(args)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    from alens_amd import capi, systems

    if args.calib_copy_mib:
        import torch

        nb = args.calib_copy_mib << 20
        src = torch.empty(nb, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        torch.add(src, 1, out=dst)  # elementwise: reads nb bytes, writes nb bytes
        torch.cuda.synchronize()
        del src, dst
    # (skipped in profiled runs: its copy kernels would sit beside the calibration copy)
    copy_gbs = None if args.calib_copy_mib else stream_copy_gbs()
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
    mv = np.array([r["matvecs"] for r in reps])
    ms = np.array([r["n_con"] for r in reps])
    tsolve = np.array([r["t_solve_ms"] for r in reps]) * 1e-3
    alg = sum(bytes_per_matvec(int(m), n) * int(v) for m, v in zip(ms, mv))
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
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {n} rods, FP64 plain BBPGD, tol {cfg['tol']}, dt {cfg['dt']} s",
            "rods": n,
            "box": list(cfg["box_hi"]),
            "pbc": list(cfg["pbc"]),
        },
        "solver_schedule": tune,
        "bbpgd_iters_per_step": float(iters.mean()),
        "bbpgd_iters_per_step_plain_fp64": float(iters.mean()),
        "matvecs_per_step": [int(v) for v in mv],
        "constraints_per_step": float(ms.mean()),
        "constraints_per_timed_step": [int(v) for v in ms],
        "t_collect_ms_mean": float(np.mean([r["t_collect_ms"] for r in reps])),
        "t_solve_ms_mean": float(1e3 * tsolve.mean()),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_bbpgd64 (persistent BBPGD, FP64)",
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
    # measured HBM traffic per launch: bytes per matvec from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE passes on this workload, times this run's matvecs per launch
    tp = os.path.join(ROOT, "profiles", "k_bbpgd_traffic.json")
    if os.path.exists(tp) and not args.n and args.precision == 0:
        with open(tp) as fh:
            tr = json.load(fh)
        if tr.get("workload", "").split(":")[0] == args.workload:
            out["roofline"]["traffic"] = tr["traffic_bytes_per_matvec"] * float(mv.mean())
            out["roofline"]["traffic_source"] = ("profiles/k_bbpgd_traffic.json: rocprofv3 FETCH_SIZE + WRITE_SIZE "
                                                 f"per matvec ({tr['traffic_over_algorithmic']:.2f} x algorithmic)")

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
                               "fp32_network": (2, 1), "fp64_network": (2, 0)})
        pm = float(ev["plain_fp64"].mean())
        out["chosen_form"]["same_state_iterations"] = {
            "states": "4 steps after 40 warm-up from the workload's seeded start",
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
        # and the FP64 oracle step on the host cores; the two are bit-identical, iterations included
        from oracle import oracle as O
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        O.set_threads(threads)
        c, t = ctx.get_rods()
        state = dict(rods, center=c, dir=t)
        oc = O.cfg_from_dict(cfg)
        oc.step = ctx.step_count
        g0 = time.perf_counter()
        grep = ctx.step()
        tg = time.perf_counter() - g0
        ts = time.perf_counter()
        res = O.step(state, oc)
        tc = time.perf_counter() - ts
        same = res["report"]["iters"] == grep["iters"] and res["I"].size == grep["n_con"]
        out["cpu_baseline"] = {
            "value": n / tc,
            "unit": "rod-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"1 full step (FP64 oracle, OpenMP) from the post-benchmark state: {res['I'].size} constraints, "
                      f"{res['report']['iters']} BBPGD iterations, {tc:.3f} s; the GPU took {tg * 1e3:.2f} ms for the "
                      f"same step ({grep['iters']} iterations{', bit-identical' if same else ', MISMATCH'}); "
                      f"CPU {cpu_model()}",
        }
        out["gpu_over_cpu"] = value / (n / tc)
        out["gpu_over_cpu_same_step"] = tc / tg
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
        # calibration: the second elementwise kernel is the --calib-copy-mib add (reads N, writes N bytes)
        calib = [r for r in rows if "elementwise" in r["Kernel_Name"]][1]
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
    res = dict(kernel="k_bbpgd<double>", workload=js["config"]["workload"],
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
    with open(args.traffic_out, "w") as fh:
        json.dump(res, fh, indent=1)


def stream_copy_gbs(nbytes=1 << 30, reps=10):
    """Device-to-device copy bandwidth on this GPU (read + write bytes per second, GB/s): the
    practical HBM ceiling beside the published 8 TB/s."""
    try:
        import torch
    except ImportError:
        return None
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2.0 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def _run_replicas(args, capi, parts, cfg, precision, warm_start, P=None, norm=None):
    ctx = capi.Context(**dict(cfg, precision=precision, warm_start=warm_start))
    if warm_start == 2:
        ctx.mlp_set(P, norm)
    offs = ctx.set_systems(parts)
    ctx.set_graph(bool(args.graph))
    ctx.sort_rods()
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
    res = dict(n=n, elapsed=el, reps=reps, tune=tune, graph=ctx.graph_info() if args.graph else None,
               iters_mean_per_replica=float(it.mean()),
               hist={"edges": [float(e) for e in edges[:-1]], "counts": hist.tolist()},
               fallbacks=int(sum(r["fallback_used"] for sr in sreps for r in sr)))
    ctx.close()
    return res


def bench_replicas(args, capi, systems, copy_gbs=None):
    """64 independent replicas (rod counts log-uniform in [2k, 50k], 3-D boxes and slabs) stepped
    together: one collection pass and one persistent solver launch per step for all of them (one
    graph replay per step in graph mode).  Plain FP64 from a cold start gives `value`; the chosen
    form (FP32 + FP64 refinement, learned warm start) runs the same workload beside it."""
    parts, cfg = systems.replicas()
    K = args.steps
    a = _run_replicas(args, capi, parts, cfg, args.precision, 0)
    n, reps = a["n"], a["reps"]
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
        "stream_copy_gbs": copy_gbs,
    }
    if not args.no_form:
        from alens_amd import warmstart as W

        P, norm = W.load()
        b = _run_replicas(args, capi, parts, cfg, 1, 2, P, norm)
        out["chosen_form"] = {
            "form": "FP32 BBPGD + FP64 residual/refinement, warm start from the learned network",
            "value": n * K / b["elapsed"], "unit": "rod-steps/s", "ms_per_step": 1e3 * b["elapsed"] / K,
            "replica_iters_mean": b["iters_mean_per_replica"], "replica_iters_histogram": b["hist"],
            "fallbacks": b["fallbacks"], "solver_schedule": b["tune"], "graph": b["graph"],
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
