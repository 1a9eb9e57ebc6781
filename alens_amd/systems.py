"""Seeded synthetic rod systems for the benchmark workloads, and a reader/writer for the
reference's rod file format (TubuleInitial.dat / SylinderAscii lines
`C|S gid radius minus_x minus_y minus_z plus_x plus_y plus_z [group]`, docs/source/quickstart.md:215-221).

Each generator returns (rods, cfg): rods = dict(gid, center (n,3), dir (n,3), length, radius,
immovable); cfg = keyword arguments for alens_amd.capi.Context / make_config.
"""
import numpy as np

RADIUS = 0.0125  # um (diameter 0.025)
COMMON = dict(col_buf=0.025, viscosity=0.01, kT=0.00411, dt=1e-4, tol=1e-5, max_ite=100000, seed=1234)


def _iso_dirs(rng, n):
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    s = np.sqrt(1.0 - z * z)
    d = np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def _rods(center, dirs, length, radius=RADIUS, gid=None):
    n = len(length)
    return dict(gid=np.arange(n, dtype=np.int32) if gid is None else np.asarray(gid, np.int32),
                center=np.ascontiguousarray(center, np.float64), dir=np.ascontiguousarray(dirs, np.float64),
                length=np.ascontiguousarray(length, np.float64), radius=np.full(n, float(radius)), immovable=None)


def small_3d(n=1000, box=4.0, seed=1234):
    """1,000 rods, L ~ U[0.5, 1.5] um, isotropic, periodic 4^3 um^3."""
    rng = np.random.default_rng(seed)
    L = rng.uniform(0.5, 1.5, n)
    c = rng.uniform(0.0, box, (n, 3))
    cfg = dict(COMMON, box_lo=(0, 0, 0), box_hi=(box,) * 3, pbc=(1, 1, 1), monolayer=0, seed=seed)
    return _rods(c, _iso_dirs(rng, n), L), cfg


def monolayer_slab(n=200_000, side=100.0, thick=0.05, kappa=4.0, seed=1234):
    """Dense quasi-2D layer: L = 1 um, centres in a side x side x thick slab (periodic in x, y),
    in-plane directions from a von Mises(0, kappa) distribution about x.  Rods move in 3-D (stacked
    crossings inside the slab are resolved by out-of-plane motion)."""
    rng = np.random.default_rng(seed)
    c = np.stack([rng.uniform(0, side, n), rng.uniform(0, side, n), rng.uniform(0, thick, n)], axis=1)
    th = rng.vonmises(0.0, kappa, n)
    d = np.stack([np.cos(th), np.sin(th), np.zeros(n)], axis=1)
    cfg = dict(COMMON, box_lo=(0, 0, 0), box_hi=(side, side, thick), pbc=(1, 1, 0), monolayer=0, seed=seed)
    return _rods(c, d, np.ones(n)), cfg


def lognormal_3d(n=1_000_000, box=25.0, median=1.0, sigma=0.3, lmax=3.0, seed=1234):
    """1M rods, L ~ lognormal(ln median, sigma) resampled above lmax, periodic box^3, isotropic."""
    rng = np.random.default_rng(seed)
    L = rng.lognormal(np.log(median), sigma, n)
    bad = L > lmax
    while bad.any():
        L[bad] = rng.lognormal(np.log(median), sigma, int(bad.sum()))
        bad = L > lmax
    c = rng.uniform(0.0, box, (n, 3))
    cfg = dict(COMMON, box_lo=(0, 0, 0), box_hi=(box,) * 3, pbc=(1, 1, 1), monolayer=0, seed=seed)
    return _rods(c, _iso_dirs(rng, n), L), cfg


def scaled(kind, n, seed=1234):
    """Same density and shape law as a named workload at a smaller rod count."""
    if kind == "small_3d":
        return small_3d(n=n, box=4.0 * (n / 1000.0) ** (1 / 3), seed=seed)
    if kind == "monolayer_slab":
        return monolayer_slab(n=n, side=100.0 * (n / 200_000.0) ** 0.5, seed=seed)
    if kind == "lognormal_3d":
        return lognormal_3d(n=n, box=25.0 * (n / 1_000_000.0) ** (1 / 3), seed=seed)
    raise ValueError(kind)


def read_rod_file(path):
    """Parse the reference's rod text format (TubuleInitial.dat, SylinderAscii_*.dat).  'S' rods are
    stationary (immovable); the optional last column is the group id (-1 when absent)."""
    gid, rad, mi, pl, imm, grp = [], [], [], [], [], []
    with open(path) as fh:
        for line in fh:
            tok = line.split()
            if len(tok) < 9 or tok[0] not in ("C", "S"):
                continue
            gid.append(int(tok[1]))
            rad.append(float(tok[2]))
            mi.append([float(x) for x in tok[3:6]])
            pl.append([float(x) for x in tok[6:9]])
            imm.append(1 if tok[0] == "S" else 0)
            grp.append(int(tok[9]) if len(tok) > 9 else -1)
    mi = np.array(mi, dtype=np.float64).reshape(-1, 3)
    pl = np.array(pl, dtype=np.float64).reshape(-1, 3)
    v = pl - mi
    L = np.linalg.norm(v, axis=1)
    d = v / np.where(L > 0, L, 1.0)[:, None]
    d[L == 0] = (0.0, 0.0, 1.0)
    return dict(gid=np.array(gid, np.int32), center=0.5 * (mi + pl), dir=d, length=L,
                radius=np.array(rad, dtype=np.float64), immovable=np.array(imm, np.uint8),
                group=np.array(grp, np.int32))


def write_rod_file(path, rods, header_time=0.0):
    """Write rods in the reference's text format: two header lines (rod count, time), then one line
    per rod with type, gid, radius, minus end, plus end and group (17 significant digits, so the
    endpoints read back exactly)."""
    c, d, L = rods["center"], rods["dir"], rods["length"]
    imm = rods.get("immovable")
    grp = rods.get("group")
    mi = c - 0.5 * L[:, None] * d
    pl = c + 0.5 * L[:, None] * d
    with open(path, "w") as fh:
        fh.write(f"{len(L)}\n{header_time:.17g}\n")
        for k in range(len(L)):
            t = "S" if (imm is not None and imm[k]) else "C"
            g = int(grp[k]) if grp is not None else -1
            fh.write(f"{t} {int(rods['gid'][k])} {rods['radius'][k]:.17g} " +
                     " ".join(f"{x:.17g}" for x in (*mi[k], *pl[k])) + f" {g}\n")


# Snapshots and restart.  The reference writes SylinderAscii_<snap>.dat into its result folders
# (README.md:128-135; read back by Utils/read_write_utils.py:185-197, 210-213) and restarts when
# TimeStepInfo.txt exists (SRC/main.cpp:28-33), resuming after the last snapshot with the snapshot
# id and step count advanced (SRC/TubuleSystem.cpp:36-38, 48-49).  The content of TimeStepInfo.txt
# is written by SimToolbox (absent); here it is "key value" lines naming the snapshot file.
RESTART_FILE = "TimeStepInfo.txt"


def write_snapshot(folder, snap_id, rods, step, time=0.0):
    """Write SylinderAscii_<snap_id>.dat and the restart record into folder; returns the rod file path."""
    import os

    os.makedirs(folder, exist_ok=True)
    name = f"SylinderAscii_{int(snap_id)}.dat"
    write_rod_file(os.path.join(folder, name), rods, header_time=time)
    with open(os.path.join(folder, RESTART_FILE), "w") as fh:
        fh.write(f"snapID {int(snap_id)}\nstepCount {int(step)}\ntime {float(time):.17g}\nsnapFile {name}\n")
    return os.path.join(folder, name)


def read_restart(folder):
    """Rods of the snapshot named by folder/TimeStepInfo.txt, and where to resume: the step count and
    snapshot id after it (both advanced by one, as the reference's reinitialize does), and the time."""
    import os

    info = {}
    with open(os.path.join(folder, RESTART_FILE)) as fh:
        for line in fh:
            tok = line.split(None, 1)
            if len(tok) == 2:
                info[tok[0]] = tok[1].strip()
    for key in ("snapID", "stepCount", "snapFile"):
        if key not in info:
            raise ValueError(f"{RESTART_FILE}: missing '{key}'")
    rods = read_rod_file(os.path.join(folder, info["snapFile"]))
    return rods, dict(step=int(info["stepCount"]) + 1, snap_id=int(info["snapID"]) + 1,
                      time=float(info.get("time", "0")))


def replicas(nsys=64, nmin=2000, nmax=50000, seed=1234):
    """Independent replica systems with rod counts log-uniform in [nmin, nmax]; even-numbered replicas
    are periodic 3-D boxes at 64 rods/um^3 (lognormal lengths, median 1 um, sigma 0.3, resampled above
    1.5 um so that the smallest boxes satisfy the periodic-image condition), odd-numbered ones are
    quasi-2D slabs at 20 rods/um^2 (L = 1 um, nematic kappa = 4).  gids are unique across replicas so
    every replica draws its own Brownian noise.  Returns ([(rods, box_cfg)], common cfg)."""
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsys))).astype(int)
    out = []
    off = 0
    for s, n in enumerate(ns):
        if s % 2 == 0:
            rods, cfg = lognormal_3d(n=int(n), box=(n / 64.0) ** (1 / 3), lmax=1.5, seed=seed + 7 * s + 1)
        else:
            rods, cfg = monolayer_slab(n=int(n), side=float(np.sqrt(n / 20.0)), seed=seed + 7 * s + 2)
        rods = dict(rods, gid=rods["gid"] + off)
        off += int(n)
        out.append((rods, dict(box_lo=cfg["box_lo"], box_hi=cfg["box_hi"], pbc=cfg["pbc"],
                               monolayer=cfg["monolayer"])))
    return out, dict(COMMON, box_lo=(0, 0, 0), box_hi=(1, 1, 1), pbc=(0, 0, 0), seed=seed)
