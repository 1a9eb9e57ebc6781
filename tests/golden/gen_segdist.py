"""Writes tests/golden/segdist.json: segment-segment closest distances solved exactly in 50-digit
mpmath by enumerating the KKT candidates of min |r + s tI - u tJ|^2 over the box
[-hI, hI] x [-hJ, hJ] (interior stationary point, the four edges' clamped 1-D minimisers).

Inputs are binary64 values stored with repr (exact round trip).  The expected distance is stored as
a 30-digit decimal string; s and u are stored when the minimiser is unique (det > 0).
Run:  python tests/golden/gen_segdist.py
"""
import json
import math
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50


def exact(cI, tI, hI, cJ, tJ, hJ):
    cI = [mp.mpf(x) for x in cI]
    tI = [mp.mpf(x) for x in tI]
    cJ = [mp.mpf(x) for x in cJ]
    tJ = [mp.mpf(x) for x in tJ]
    hI = mp.mpf(hI)
    hJ = mp.mpf(hJ)
    r = [cI[k] - cJ[k] for k in range(3)]
    dot = lambda x, y: sum(x[k] * y[k] for k in range(3))
    a, b, c, d, e = dot(tI, tI), dot(tI, tJ), dot(tJ, tJ), dot(tI, r), dot(tJ, r)

    def f(s, u):
        return sum((r[k] + s * tI[k] - u * tJ[k]) ** 2 for k in range(3))

    def clamp(x, lo, hi):
        return lo if x < lo else (hi if x > hi else x)

    cand = []
    det = a * c - b * b
    if det > 0:
        s = (b * e - c * d) / det
        u = (a * e - b * d) / det
        if abs(s) <= hI and abs(u) <= hJ:
            cand.append((s, u))
    for s in (-hI, hI):
        cand.append((s, clamp((b * s + e) / c, -hJ, hJ)))
    for u in (-hJ, hJ):
        cand.append((clamp((b * u - d) / a, -hI, hI), u))
    best = min(cand, key=lambda su: f(*su))
    unique = det > mp.mpf(10) ** -30 * a * c
    return mp.sqrt(f(*best)), best, unique


def unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def main():
    rng = np.random.default_rng(20240611)
    cases = []

    def add(kind, cI, tI, hI, cJ, tJ, hJ):
        cases.append(dict(kind=kind, cI=[float(x) for x in cI], tI=[float(x) for x in tI], hI=float(hI),
                          cJ=[float(x) for x in cJ], tJ=[float(x) for x in tJ], hJ=float(hJ)))

    for _ in range(300):  # random skew pairs at contact-like distances
        tI, tJ = unit(rng.normal(size=3)), unit(rng.normal(size=3))
        cI = rng.uniform(-1, 1, 3) * 2
        cJ = cI + rng.normal(size=3) * rng.choice([0.05, 0.3, 1.0])
        add("random", cI, tI, rng.uniform(0, 0.75), cJ, tJ, rng.uniform(0, 0.75))
    for _ in range(40):  # parallel / anti-parallel offsets (continuum of minimisers)
        t = unit(rng.normal(size=3))
        sgn = rng.choice([-1.0, 1.0])
        cI = rng.uniform(-1, 1, 3)
        off = unit(np.cross(t, rng.normal(size=3))) * rng.uniform(0.0, 0.1)
        cJ = cI + off + t * rng.uniform(-1.5, 1.5)
        add("parallel", cI, t, rng.uniform(0.1, 0.75), cJ, sgn * t, rng.uniform(0.1, 0.75))
    for _ in range(20):  # collinear overlapping
        t = unit(rng.normal(size=3))
        cI = rng.uniform(-1, 1, 3)
        add("collinear", cI, t, 0.5, cI + t * rng.uniform(-0.8, 0.8), t, 0.5)
    for _ in range(40):  # crossing / touching (distance ~ 0 or tiny)
        tI, tJ = unit(rng.normal(size=3)), unit(rng.normal(size=3))
        cI = rng.uniform(-1, 1, 3)
        sI, sJ = rng.uniform(-0.4, 0.4), rng.uniform(-0.4, 0.4)
        gap = rng.choice([0.0, 1e-9, 1e-6, 0.025])
        nrm = unit(np.cross(tI, tJ))
        cJ = cI + sI * tI - sJ * tJ + gap * nrm
        add("crossing", cI, tI, 0.5, cJ, tJ, 0.5)
    for _ in range(40):  # nearly parallel skew
        t = unit(rng.normal(size=3))
        eps = 10.0 ** rng.uniform(-9, -3)
        tJ = unit(t + eps * unit(np.cross(t, rng.normal(size=3))))
        cI = rng.uniform(-1, 1, 3)
        cJ = cI + unit(np.cross(t, rng.normal(size=3))) * 0.03 + t * rng.uniform(-1, 1)
        add("near_parallel", cI, t, 0.5, cJ, tJ, 0.5)
    for _ in range(30):  # point-segment and point-point (sphere) cases
        tI, tJ = unit(rng.normal(size=3)), unit(rng.normal(size=3))
        cI = rng.uniform(-1, 1, 3)
        cJ = cI + rng.normal(size=3) * 0.3
        kind = rng.integers(3)
        add("sphere", cI, tI, 0.0 if kind != 1 else 0.5, cJ, tJ, 0.0 if kind != 2 else 0.5)
    for _ in range(30):  # large coordinates (100 um box)
        tI, tJ = unit(rng.normal(size=3)), unit(rng.normal(size=3))
        cI = rng.uniform(0, 100, 3)
        cJ = cI + rng.normal(size=3) * 0.4
        add("far_origin", cI, tI, 0.5, cJ, tJ, 0.5)

    for cs in cases:
        dist, (s, u), unique = exact(cs["cI"], cs["tI"], cs["hI"], cs["cJ"], cs["tJ"], cs["hJ"])
        cs["dist"] = mp.nstr(dist, 30)
        cs["unique"] = bool(unique)
        if unique:
            cs["s"] = mp.nstr(s, 30)
            cs["u"] = mp.nstr(u, 30)
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "segdist.json"), "w") as fh:
        json.dump(dict(generator="tests/golden/gen_segdist.py", dps=50, cases=cases), fh, indent=0)
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
