"""Learned warm start: data collection from solver runs, GPU training, evaluation.

The network maps 12 per-constraint inputs to an initial impulse gamma0; BBPGD always finishes the
solve from it (and restarts cold if it raises the initial residual).  Everything here drives
libalens.so through the C ABI: features and gamma come from real steps, training runs the library's
MFMA forward/backward kernels and Adam on the GPU.

  python -m alens_amd.warmstart --out alens_amd/data/warmstart_mlp.npz   (on a GPU)
"""
import argparse
import os
import time

import numpy as np

from . import capi, systems

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "data", "warmstart_mlp.npz")


def collect(kind, n, steps, seed=1234, skip=2, sort=True):
    """Run `steps` constraint steps (warm start from the previous gamma) and return the features and
    converged gamma of every step after the first `skip` (initial overlap resolution)."""
    rods, cfg = systems.scaled(kind, n, seed=seed) if kind != "small_3d" else systems.small_3d(seed=seed)
    ctx = capi.Context(**dict(cfg, warm_start=1))
    ctx.set_rods(rods)
    if sort:
        ctx.sort_rods()
    Xs, gs = [], []
    for s in range(steps):
        ctx.step()
        if s >= skip:
            Xs.append(ctx.features())
            gs.append(ctx.gamma().astype(np.float32))
        ctx.euler()
    ctx.close()
    return np.concatenate(Xs), np.concatenate(gs)


def train(X, gamma, epochs=12, batch=4096, lr=2e-3, seed=1):
    ctx = capi.Context()
    losses = ctx.mlp_train(X, gamma, epochs=epochs, batch=batch, lr=lr, seed=seed)
    P, norm = ctx.mlp_get()
    ctx.close()
    return P, norm, losses


def evaluate(kind, n, steps, P, norm, seed=99, warmup=2, precision=0, forms=None, lead=None):
    """Same rod states fed to several contexts at every step (a plain FP64 driver moves the rods);
    returns each context's per-step BBPGD iterations (FP32 + FP64) and its warm-start fallbacks.
    forms: name -> (warm_start, precision); default cold / previous gamma / network at `precision`.
    lead: the forms start stepping this many steps before the measured ones (they need a previous
    step for their warm-start keys); None = from the first step."""
    if forms is None:
        forms = {"cold": (0, precision), "prev": (1, precision), "mlp": (2, precision)}
    rods, cfg = systems.scaled(kind, n, seed=seed) if kind != "small_3d" else systems.small_3d(seed=seed)
    driver = capi.Context(**cfg)
    driver.set_rods(rods)
    perm = driver.sort_rods()
    rods = {k: (v[perm] if isinstance(v, np.ndarray) else v) for k, v in rods.items()}
    ctxs = {}
    for name, (ws, prec) in forms.items():
        c = capi.Context(**dict(cfg, warm_start=ws, precision=prec))
        c.set_rods(rods)
        if ws == 2:
            c.mlp_set(P, norm)
        ctxs[name] = c
    out = {k: [] for k in ctxs}
    for name in ctxs:
        out["fallback_" + name] = 0
    first = 0 if lead is None else max(0, warmup - lead)
    for s in range(warmup + steps):
        c, t = driver.get_rods()
        state = dict(rods, center=c, dir=t)
        for name, ctx in (ctxs.items() if s >= first else ()):
            ctx.set_rods(state)
            ctx.set_step(s)
            r = ctx.step()
            if s >= warmup:
                out[name].append(r["iters"] + r["iters_fp32"])
                out["fallback_" + name] += r["fallback_used"]
        driver.set_step(s)
        driver.step()
        driver.euler()
    for ctx in list(ctxs.values()) + [driver]:
        ctx.close()
    return {k: (np.array(v) if isinstance(v, list) else v) for k, v in out.items()}


def train_energy(P, norm, kind, n, steps, iters=8, lr=3e-4, seed=1234, warmup=10):
    """Fine-tune the network on the BCQP energy of its own start point (alens_mlp_train_energy): a
    plain FP64 driver moves the rods from the seeded start; from step `warmup` on, a context running
    the network takes each state's step (so its features carry its own previous impulses) and then
    `iters` Adam steps on that step's rows.  Returns the parameters and the energies per state."""
    rods, cfg = systems.scaled(kind, n, seed=seed) if kind != "small_3d" else systems.small_3d(seed=seed)
    driver = capi.Context(**cfg)
    driver.set_rods(rods)
    perm = driver.sort_rods()
    rods = {k: (v[perm] if isinstance(v, np.ndarray) else v) for k, v in rods.items()}
    ctx = capi.Context(**dict(cfg, warm_start=2))
    ctx.mlp_set(P, norm)
    ctx.set_rods(rods)
    hist = []
    for s in range(warmup + steps):
        if s >= warmup - 1:
            c, t = driver.get_rods()
            ctx.set_rods(dict(rods, center=c, dir=t))
            ctx.set_step(s)
            ctx.step()
            if s >= warmup:
                hist.append(ctx.mlp_train_energy(iters, lr))
        driver.set_step(s)
        driver.step()
        driver.euler()
    P2, norm2 = ctx.mlp_get()
    for c in (ctx, driver):
        c.close()
    return P2, norm2, hist


def save(path, P, norm, meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=np.asarray(P, np.float32), norm=np.asarray(norm, np.float32),
             meta=np.array(repr(meta)))


def load(path=DEFAULT_MODEL):
    d = np.load(path, allow_pickle=False)
    return d["params"], d["norm"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--pairs", type=int, default=6_000_000, help="training pairs (a random subset)")
    ap.add_argument("--epochs", type=int, default=12)
    args = ap.parse_args()
    t0 = time.time()
    # early and relaxed states: the small 3-D system, the dense slab right after its random start
    # (steps 2-3) and near steady state (steps 40-47), a 20k-rod slab over steps 10-39, and
    # lognormal 3-D boxes of 200k rods (steps 40-47) and 1M rods (steps 40-43)
    runs = (("small_3d", 1000, 80, 11, 2), ("monolayer_slab", 200_000, 4, 1002, 2),
            ("monolayer_slab", 200_000, 48, 1003, 40), ("monolayer_slab", 20_000, 40, 1004, 10),
            ("lognormal_3d", 200_000, 48, 1005, 40), ("lognormal_3d", 1_000_000, 44, 1006, 40))
    Xs, gs = [], []
    for kind, n, steps, seed, skip in runs:
        X, g = collect(kind, n, steps, seed=seed, skip=skip)
        Xs.append(X)
        gs.append(g)
    X = np.concatenate(Xs)
    g = np.concatenate(gs)
    sel = np.random.default_rng(5).permutation(len(g))[: args.pairs]
    X, g = X[sel], g[sel]
    print(f"collected {len(g)} pairs in {time.time() - t0:.1f} s", flush=True)
    P, norm, losses = train(X, g, epochs=args.epochs)
    print("loss per epoch", np.round(losses, 5).tolist(), flush=True)
    save(args.out, P, norm, dict(pairs=len(g), epochs=args.epochs, losses=[float(v) for v in losses]))
    ev = evaluate("monolayer_slab", 20000, 8, P, norm)
    print({k: (v.mean() if hasattr(v, "mean") else v) for k, v in ev.items()}, flush=True)


if __name__ == "__main__":
    main()
