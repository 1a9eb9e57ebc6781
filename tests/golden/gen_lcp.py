"""Writes tests/golden/lcp_small.json: small bounded-constraint QPs (LCPs)
    find x >= 0 with g = A x + b >= 0 and x_k g_k = 0,   A = dt D^T M D,
solved exactly by enumerating active sets in 50-digit mpmath with A assembled densely from the
stored constraint data (normals n, arms rI = posI x n, rJ = posJ x n, rod indices I, J) and the
per-rod mobility blocks {t, 1/zperp, 1/zpar - 1/zperp, 1/zrot}.  The dense assembly here is an
independent path from the matrix-free rod-major operator the solvers use.
Run:  python tests/golden/gen_lcp.py
"""
import itertools
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50


def unit(v):
    return v / np.linalg.norm(v)


def dense_A(n, I, J, nrm, rI, rJ, mob, dt):
    m = len(I)
    D = mp.zeros(6 * n, m)
    for k in range(m):
        for a in range(3):
            D[6 * I[k] + a, k] += mp.mpf(nrm[3 * k + a])
            D[6 * I[k] + 3 + a, k] += mp.mpf(rI[3 * k + a])
            D[6 * J[k] + a, k] -= mp.mpf(nrm[3 * k + a])
            D[6 * J[k] + 3 + a, k] -= mp.mpf(rJ[3 * k + a])
    Mm = mp.zeros(6 * n, 6 * n)
    for i in range(n):
        t = [mp.mpf(mob[6 * i + a]) for a in range(3)]
        iperp, diff, irot = (mp.mpf(mob[6 * i + 3]), mp.mpf(mob[6 * i + 4]), mp.mpf(mob[6 * i + 5]))
        for a in range(3):
            for c in range(3):
                Mm[6 * i + a, 6 * i + c] = diff * t[a] * t[c] + (iperp if a == c else 0)
            Mm[6 * i + 3 + a, 6 * i + 3 + a] = irot
    return mp.mpf(dt) * D.T * Mm * D


def solve_lcp(A, b):
    m = len(b)
    bm = mp.matrix([mp.mpf(x) for x in b])
    best = None
    for r in range(m + 1):
        for S in itertools.combinations(range(m), r):
            x = mp.zeros(m, 1)
            if S:
                As = mp.matrix([[A[i, j] for j in S] for i in S])
                try:
                    xs = mp.lu_solve(As, mp.matrix([-bm[i] for i in S]))
                except ZeroDivisionError:
                    continue
                if any(v < 0 for v in xs):
                    continue
                for q, i in enumerate(S):
                    x[i] = xs[q]
            g = A * x + bm
            if all(g[i] >= -mp.mpf(10) ** -40 for i in range(m) if i not in S):
                best = x
                return best
    return best


def main():
    rng = np.random.default_rng(7)
    cases = []
    for case in range(24):
        n = int(rng.integers(2, 7))
        m = min(int(rng.integers(1, 9)), n * (n - 1) // 2)
        I, J = [], []
        seen = set()
        while len(I) < m:
            i, j = sorted(rng.choice(n, 2, replace=False).tolist())
            if (i, j) in seen:
                continue
            seen.add((i, j))
            I.append(i)
            J.append(j)
        order = sorted(range(m), key=lambda k: (I[k], J[k]))
        I = [I[k] for k in order]
        J = [J[k] for k in order]
        nrm, rI, rJ = [], [], []
        for k in range(m):
            nk = unit(rng.normal(size=3))
            pI = rng.normal(size=3) * 0.4
            pJ = rng.normal(size=3) * 0.4
            nrm += nk.tolist()
            rI += np.cross(pI, nk).tolist()
            rJ += np.cross(pJ, nk).tolist()
        mob = []
        for i in range(n):
            t = unit(rng.normal(size=3))
            L = rng.uniform(0.5, 1.5)
            ipar, iperp, irot = 60.0 / L, 38.0 / L, 450.0 / L ** 3
            mob += t.tolist() + [iperp, ipar - iperp, irot]
        dt = 1e-4
        b = (rng.normal(size=m) * 0.01).tolist()
        A = dense_A(n, I, J, nrm, rI, rJ, mob, dt)
        x = solve_lcp(A, b)
        if x is None:
            continue
        lam = min(mp.eig(A)[0], key=lambda z: mp.re(z))
        cases.append(dict(n=n, I=I, J=J, nrm=nrm, rI=rI, rJ=rJ, mob=mob, dt=dt, b=b,
                          gamma=[mp.nstr(x[k], 30) for k in range(m)],
                          lambda_min=mp.nstr(mp.re(lam), 10)))
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "lcp_small.json"), "w") as fh:
        json.dump(dict(generator="tests/golden/gen_lcp.py", dps=50, cases=cases), fh, indent=0)
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
