"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
cpu_baseline leg, always as the checker or the timed CPU restatement, never as the product path.

Every function mirrors one stage of the reference path (SimToolbox's SylinderSystem::runStep,
called at SRC/TubuleSystem.cpp:223); see oracle/alens_oracle.c for the per-stage anchors.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class OrcCfg(ctypes.Structure):
    _fields_ = [
        ("box_lo", ctypes.c_double * 3),
        ("box_hi", ctypes.c_double * 3),
        ("pbc", ctypes.c_int32 * 3),
        ("monolayer", ctypes.c_int32),
        ("col_buf", ctypes.c_double),
        ("viscosity", ctypes.c_double),
        ("kT", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("diam_col_ratio", ctypes.c_double),
        ("len_col_ratio", ctypes.c_double),
        ("cell_size", ctypes.c_double),
        ("max_ite", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("step", ctypes.c_int64),
    ]


class OrcReport(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int32),
        ("matvecs", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("fallback_used", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("res0", ctypes.c_double),
        ("res", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad0"}


class _OptDouble:
    """ndpointer that also accepts None."""

    @classmethod
    def from_param(cls, obj):
        if obj is None:
            return None
        return _dp.from_param(obj)


class _OptU8:
    @classmethod
    def from_param(cls, obj):
        if obj is None:
            return None
        return _u8p.from_param(obj)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError("oracle/liboracle.so is not built; run oracle/build_oracle.sh")
    L = ctypes.CDLL(_LIB_PATH)
    c_cfg = ctypes.POINTER(OrcCfg)
    L.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
    L.orc_log.argtypes = [ctypes.c_double]
    L.orc_log.restype = ctypes.c_double
    L.orc_sqrt.argtypes = [ctypes.c_double]
    L.orc_sqrt.restype = ctypes.c_double
    L.orc_sincos2pi.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.orc_segseg.argtypes = [_dp, _dp, ctypes.c_double, _dp, _dp, ctypes.c_double, _dp]
    L.orc_block_sum.argtypes = [_dp, ctypes.c_int64]
    L.orc_block_sum.restype = ctypes.c_double
    L.orc_chunked_sum.argtypes = [_dp, ctypes.c_int64]
    L.orc_chunked_sum.restype = ctypes.c_double
    for fn in (L.orc_pairs_brute, L.orc_pairs_cells):
        fn.argtypes = [ctypes.c_int32, _dp, _dp, _dp, _dp, c_cfg, ctypes.c_int64, _ip, _ip, _ip]
        fn.restype = ctypes.c_int64
    L.orc_mobility.argtypes = [ctypes.c_int32, _dp, _dp, _dp, _OptU8, c_cfg, _dp]
    L.orc_unc.argtypes = [ctypes.c_int32, _ip, _dp, _OptDouble, c_cfg, _dp]
    L.orc_rows.argtypes = [ctypes.c_int32, _dp, _dp, _dp, _dp, c_cfg, ctypes.c_int64, _ip, _ip, _ip, _dp,
                           _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_apply_D_M.argtypes = [ctypes.c_int32, ctypes.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, c_cfg, _dp,
                                _OptDouble, _dp]
    L.orc_apply_DT.argtypes = [ctypes.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, _dp]
    L.orc_bbpgd.argtypes = [ctypes.c_int32, ctypes.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, _OptDouble, _OptDouble,
                            _dp, c_cfg, _OptDouble, _dp, ctypes.POINTER(OrcReport)]
    L.orc_rows_ext.argtypes = [ctypes.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, _dp, c_cfg, _dp, _dp, _dp]
    L.orc_stress_sum.argtypes = [ctypes.c_int64, _dp, _dp, _dp, _dp, _OptDouble, _OptU8, _dp]
    L.orc_bbpgd.restype = ctypes.c_int
    L.orc_writeback.argtypes = [ctypes.c_int32, ctypes.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, c_cfg,
                                _dp, _dp]
    L.orc_euler.argtypes = [ctypes.c_int32, _dp, _dp, _dp, c_cfg]
    L.orc_set_threads.argtypes = [ctypes.c_int32]
    L.orc_get_threads.restype = ctypes.c_int32
    _lib = L
    return L


def make_cfg(box_lo=(0, 0, 0), box_hi=(1, 1, 1), pbc=(1, 1, 1), monolayer=0, col_buf=0.025,
             viscosity=0.01, kT=0.00411, dt=1e-4, tol=1e-5, diam_col_ratio=1.0, len_col_ratio=1.0,
             cell_size=0.0, max_ite=100000, seed=1234, step=0):
    c = OrcCfg()
    c.box_lo[:] = [float(v) for v in box_lo]
    c.box_hi[:] = [float(v) for v in box_hi]
    c.pbc[:] = [int(v) for v in pbc]
    c.monolayer = int(monolayer)
    c.col_buf = col_buf
    c.viscosity = viscosity
    c.kT = kT
    c.dt = dt
    c.tol = tol
    c.diam_col_ratio = diam_col_ratio
    c.len_col_ratio = len_col_ratio
    c.cell_size = cell_size
    c.max_ite = int(max_ite)
    c.seed = int(seed)
    c.step = int(step)
    return c


def cfg_from_dict(d):
    keys = {f for f, _ in OrcCfg._fields_}
    return make_cfg(**{k: v for k, v in d.items() if k in keys})


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def philox(ctr, key):
    a = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(a, k, o)
    return list(o)


def log(x):
    return lib().orc_log(float(x))


def sincos2pi(u):
    s = ctypes.c_double()
    c = ctypes.c_double()
    lib().orc_sincos2pi(float(u), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def segseg(cI, tI, hI, cJ, tJ, hJ):
    out = np.zeros(6)
    lib().orc_segseg(np.ascontiguousarray(cI, np.float64), np.ascontiguousarray(tI, np.float64), float(hI),
                     np.ascontiguousarray(cJ, np.float64), np.ascontiguousarray(tJ, np.float64), float(hJ), out)
    return out


def block_sum(v):
    v = np.ascontiguousarray(v, np.float64)
    return lib().orc_block_sum(v, v.size)


def chunked_sum(v):
    v = np.ascontiguousarray(v, np.float64)
    return lib().orc_chunked_sum(v, v.size)


def _rods(rods):
    c = np.ascontiguousarray(rods["center"], np.float64).reshape(-1)
    t = np.ascontiguousarray(rods["dir"], np.float64).reshape(-1)
    L = np.ascontiguousarray(rods["length"], np.float64)
    r = np.ascontiguousarray(rods["radius"], np.float64)
    return c, t, L, r


def pairs(rods, cfg, method="cells"):
    """Canonical (I, J ascending) pair list: (I, J, img[m,3]).  Raises if the box is too small."""
    c, t, L, r = _rods(rods)
    n = L.size
    fn = lib().orc_pairs_cells if method == "cells" else lib().orc_pairs_brute
    cap = max(1024, 16 * n)
    while True:
        I = np.zeros(cap, np.int32)
        J = np.zeros(cap, np.int32)
        img = np.zeros(3 * cap, np.int32)
        m = fn(n, c, t, L, r, ctypes.byref(cfg), cap, I, J, img)
        if m < 0:
            raise ValueError("periodic box shorter than twice the interaction range")
        if m <= cap:
            return I[:m].copy(), J[:m].copy(), img[: 3 * m].reshape(m, 3).copy()
        cap = int(m)


def mobility(rods, cfg):
    c, t, L, r = _rods(rods)
    imm = rods.get("immovable")
    imm = None if imm is None else np.ascontiguousarray(imm, np.uint8)
    mob = np.zeros(6 * L.size)
    lib().orc_mobility(L.size, t, L, r, imm, ctypes.byref(cfg), mob)
    return mob


def unc(rods, mob, cfg, f_ext=None):
    gid = np.ascontiguousarray(rods["gid"], np.int32)
    out = np.zeros(6 * gid.size)
    fe = None if f_ext is None else np.ascontiguousarray(f_ext, np.float64).reshape(-1)
    lib().orc_unc(gid.size, gid, mob, fe, ctypes.byref(cfg), out)
    return out


def rows(rods, cfg, I, J, img, unc_v):
    c, t, L, r = _rods(rods)
    m = I.size
    out = {k: np.zeros(3 * m) for k in ("nrm", "posI", "posJ", "rI", "rJ")}
    out["delta0"] = np.zeros(m)
    out["b"] = np.zeros(m)
    lib().orc_rows(L.size, c, t, L, r, ctypes.byref(cfg), m, np.ascontiguousarray(I, np.int32),
                   np.ascontiguousarray(J, np.int32), np.ascontiguousarray(img, np.int32).reshape(-1), unc_v,
                   out["delta0"], out["nrm"], out["posI"], out["posJ"], out["rI"], out["rJ"], out["b"])
    return out


def apply_D_M(n, I, J, rw, mob, cfg, gamma):
    F = np.zeros(6 * n)
    U = np.zeros(6 * n)
    lib().orc_apply_D_M(n, I.size, I, J, rw["nrm"], rw["rI"], rw["rJ"], mob, ctypes.byref(cfg),
                        np.ascontiguousarray(gamma, np.float64), F, U)
    return F, U


def apply_DT(I, J, rw, U):
    du = np.zeros(I.size)
    lib().orc_apply_DT(I.size, I, J, rw["nrm"], rw["rI"], rw["rJ"], np.ascontiguousarray(U, np.float64), du)
    return du


def bbpgd(n, I, J, rw, mob, cfg, x0=None, lb=None, invk=None):
    """BBPGD on A = dt D^T M D (+ diag(invk)), bounds x >= lb (lb None = 0)."""
    m = I.size
    gamma = np.zeros(m)
    rep = OrcReport()
    opt = lambda v: None if v is None else np.ascontiguousarray(v, np.float64)
    lib().orc_bbpgd(n, m, np.ascontiguousarray(I, np.int32), np.ascontiguousarray(J, np.int32), rw["nrm"],
                    rw["rI"], rw["rJ"], rw["b"], opt(lb), opt(invk), mob, ctypes.byref(cfg), opt(x0), gamma,
                    ctypes.byref(rep))
    return gamma, rep.as_dict()


def rows_ext(ext, unc_v, cfg):
    """Rows of caller-supplied constraint blocks (dict of arrays: I, J, delta0, nrm, posI, posJ)."""
    m = len(ext["I"])
    out = {k: np.zeros(3 * m) for k in ("rI", "rJ")}
    out["b"] = np.zeros(m)
    f = lambda v: np.ascontiguousarray(v, np.float64).reshape(-1)
    lib().orc_rows_ext(m, np.ascontiguousarray(ext["I"], np.int32), np.ascontiguousarray(ext["J"], np.int32),
                       f(ext["delta0"]), f(ext["nrm"]), f(ext["posI"]), f(ext["posJ"]), unc_v, ctypes.byref(cfg),
                       out["rI"], out["rJ"], out["b"])
    return out


def ext_bounds(ext):
    """Lower bound and compliance per caller row: bilateral rows are unbounded and add 1/kappa
    (kappa > 0) to the diagonal; unilateral rows are bounded below by gamma_lb."""
    bil = np.asarray(ext["bilateral"]).astype(bool)
    kap = np.asarray(ext["kappa"], np.float64)
    lb = np.where(bil, -np.inf, np.asarray(ext["gamma_lb"], np.float64))
    invk = np.zeros(len(kap))
    sel = bil & (kap > 0)
    invk[sel] = 1.0 / kap[sel]
    return lb, invk


def stress_sum(nrm, labI, labJ, gamma, stress=None, use_given=None):
    m = np.asarray(gamma).size
    out = np.zeros(9)
    f = lambda v: np.ascontiguousarray(v, np.float64).reshape(-1)
    lib().orc_stress_sum(m, f(nrm), f(labI), f(labJ), f(gamma), None if stress is None else f(stress),
                         None if use_given is None else np.ascontiguousarray(use_given, np.uint8), out)
    return out.reshape(3, 3)


def writeback(n, I, J, rw, gamma, mob, unc_v, cfg):
    fcon = np.zeros(6 * n)
    u = np.zeros(6 * n)
    lib().orc_writeback(n, I.size, I, J, rw["nrm"], rw["rI"], rw["rJ"], np.ascontiguousarray(gamma, np.float64),
                        mob, unc_v, ctypes.byref(cfg), fcon, u)
    return fcon, u


def euler(rods, u, cfg):
    """Returns new (center, dir) arrays; the inputs are not modified."""
    c = np.ascontiguousarray(rods["center"], np.float64).reshape(-1).copy()
    t = np.ascontiguousarray(rods["dir"], np.float64).reshape(-1).copy()
    lib().orc_euler(c.size // 3, c, t, np.ascontiguousarray(u, np.float64), ctypes.byref(cfg))
    return c.reshape(-1, 3), t.reshape(-1, 3)


def step(rods, cfg, f_ext=None, x0=None, method="cells", ext=None):
    """One full constraint step (U_nc, pairs, rows, BBPGD, write-back). Returns a dict.

    ext: caller-supplied constraint blocks (dict of arrays I, J, delta0, nrm, posI, posJ, gamma0,
    bilateral, kappa, gamma_lb), appended after the collision rows in the given order; the solve
    starts from gamma0 on them (and from x0, or 0, on the collision rows)."""
    n = len(rods["length"])
    mob = mobility(rods, cfg)
    un = unc(rods, mob, cfg, f_ext)
    I, J, img = pairs(rods, cfg, method)
    rw = rows(rods, cfg, I, J, img, un)
    lb = invk = None
    mc = I.size
    if ext is not None and len(ext["I"]) > 0:
        re = rows_ext(ext, un, cfg)
        lbe, ike = ext_bounds(ext)
        I = np.concatenate([I, np.asarray(ext["I"], np.int32)])
        J = np.concatenate([J, np.asarray(ext["J"], np.int32)])
        rw = dict(rw)
        for key in ("rI", "rJ", "b"):
            rw[key] = np.concatenate([rw[key], re[key]])
        for key in ("nrm", "posI", "posJ", "delta0"):
            rw[key] = np.concatenate([rw[key], np.asarray(ext[key], np.float64).reshape(-1)])
        lb = np.concatenate([np.zeros(mc), lbe])
        invk = np.concatenate([np.zeros(mc), ike])
        x0 = np.concatenate([np.zeros(mc) if x0 is None else np.asarray(x0, np.float64),
                             np.asarray(ext["gamma0"], np.float64)])
    gamma, rep = bbpgd(n, I, J, rw, mob, cfg, x0=x0, lb=lb, invk=invk)
    fcon, u = writeback(n, I, J, rw, gamma, mob, un, cfg)
    return dict(mob=mob, unc=un, I=I, J=J, img=img, rows=rw, gamma=gamma, report=rep, fcon=fcon, u=u,
                n_collision=mc)
