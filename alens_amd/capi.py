"""ctypes binding of libalens.so (include/alens_capi.h).

This is the Python face of the C ABI: it loads the in-tree library built by alens_amd.build and
raises if it is missing -- there is no CPU fallback.  Array conventions follow the header: rods are
given in caller order (center/dir as (n, 3) float64), constraints come back in canonical order
(ascending rod index I, then J).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libalens.so")

ALENS_OK, ALENS_E_ARG, ALENS_E_HIP, ALENS_E_OOM, ALENS_E_NONFINITE, ALENS_E_NOCONV, ALENS_E_BOX, ALENS_E_STATE = range(8)
STATUS_NAMES = {0: "OK", 1: "E_ARG", 2: "E_HIP", 3: "E_OOM", 4: "E_NONFINITE", 5: "E_NOCONV", 6: "E_BOX",
                7: "E_STATE"}

# every entry point declared in include/alens_capi.h
EXPORTS = [
    "alens_create", "alens_destroy", "alens_last_error", "alens_version", "alens_set_rods", "alens_get_rods",
    "alens_sort_rods", "alens_set_systems", "alens_get_num_systems", "alens_get_system_reports",
    "alens_set_step", "alens_get_step", "alens_step_constraints", "alens_step_euler",
    "alens_get_num_constraints", "alens_get_constraints", "alens_get_constraint_arrays", "alens_get_gamma",
    "alens_get_velocity", "alens_bcqp_solve", "alens_add_constraints", "alens_clear_constraints",
    "alens_get_con_stress", "alens_autotune", "alens_set_graph", "alens_get_graph_info",
    "alens_mlp_set", "alens_mlp_get", "alens_get_features", "alens_mlp_predict", "alens_mlp_gradient",
    "alens_mlp_train", "alens_copy_probe", "alens_bcqp_solve_static", "alens_mlp_energy", "alens_mlp_train_energy",
]
MLP_PARAMS = 9473
MLP_NORM = 25


class Config(ctypes.Structure):
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
        ("solver", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("warm_start", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("device", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class Box(ctypes.Structure):
    _fields_ = [("box_lo", ctypes.c_double * 3), ("box_hi", ctypes.c_double * 3), ("pbc", ctypes.c_int32 * 3),
                ("monolayer", ctypes.c_int32)]


class ConBlock(ctypes.Structure):
    _fields_ = [
        ("delta0", ctypes.c_double), ("gamma", ctypes.c_double), ("kappa", ctypes.c_double),
        ("gamma_lb", ctypes.c_double),
        ("gid_i", ctypes.c_int32), ("gid_j", ctypes.c_int32), ("gidx_i", ctypes.c_int32), ("gidx_j", ctypes.c_int32),
        ("norm_i", ctypes.c_double * 3), ("norm_j", ctypes.c_double * 3), ("pos_i", ctypes.c_double * 3),
        ("pos_j", ctypes.c_double * 3), ("lab_i", ctypes.c_double * 3), ("lab_j", ctypes.c_double * 3),
        ("stress", ctypes.c_double * 9),
        ("one_side", ctypes.c_int32), ("bilateral", ctypes.c_int32),
    ]


class TuneReport(ctypes.Structure):
    _fields_ = [
        ("entries", ctypes.c_int32), ("bar_mode", ctypes.c_int32), ("xcd_map", ctypes.c_int32),
        ("max_blocks", ctypes.c_int32), ("candidates", ctypes.c_int32),
        ("default_ms", ctypes.c_float), ("best_ms", ctypes.c_float), ("group_width", ctypes.c_int32),
        ("stream_hint", ctypes.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class GraphInfo(ctypes.Structure):
    _fields_ = [
        ("captured", ctypes.c_int32), ("reserved", ctypes.c_int32), ("replays", ctypes.c_int64),
        ("captures", ctypes.c_int64), ("capture_failures", ctypes.c_int64), ("overflows", ctypes.c_int64),
        ("last_failure", ctypes.c_char * 192),
    ]


class Report(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int32), ("matvecs", ctypes.c_int32), ("converged", ctypes.c_int32),
        ("fallback_used", ctypes.c_int32), ("status", ctypes.c_int32), ("iters_fp32", ctypes.c_int32),
        ("n_con", ctypes.c_int64), ("n_entries", ctypes.c_int64),
        ("res0", ctypes.c_double), ("res", ctypes.c_double),
        ("t_collect_ms", ctypes.c_double), ("t_solve_ms", ctypes.c_double), ("t_step_ms", ctypes.c_double),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class AlensError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def lib():
    """Load libalens.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m alens_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.alens_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
    L.alens_destroy.argtypes = [vp]
    L.alens_destroy.restype = None
    L.alens_last_error.argtypes = [vp]
    L.alens_last_error.restype = ctypes.c_char_p
    L.alens_version.restype = ctypes.c_char_p
    L.alens_set_rods.argtypes = [vp, ctypes.c_int32, _ip, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_uint8)]
    L.alens_get_rods.argtypes = [vp, _dp, _dp]
    L.alens_sort_rods.argtypes = [vp, _ip]
    L.alens_set_systems.argtypes = [vp, ctypes.c_int32, _ip, ctypes.POINTER(Box), _ip, _dp, _dp, _dp, _dp,
                                    ctypes.POINTER(ctypes.c_uint8)]
    L.alens_get_num_systems.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
    L.alens_get_system_reports.argtypes = [vp, ctypes.POINTER(Report)]
    L.alens_set_step.argtypes = [vp, ctypes.c_int64]
    L.alens_get_step.argtypes = [vp]
    L.alens_get_step.restype = ctypes.c_int64
    L.alens_step_constraints.argtypes = [vp, _dp, ctypes.POINTER(Report)]
    L.alens_step_euler.argtypes = [vp]
    L.alens_get_num_constraints.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.alens_get_constraints.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(ConBlock), ctypes.POINTER(ctypes.c_int64)]
    L.alens_get_constraint_arrays.argtypes = [vp, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.alens_get_gamma.argtypes = [vp, _dp]
    L.alens_get_velocity.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.alens_bcqp_solve.argtypes = [vp, ctypes.c_int32, ctypes.c_int64, _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                   ctypes.POINTER(Report)]
    _fp = ctypes.POINTER(ctypes.c_float)
    L.alens_mlp_set.argtypes = [vp, _fp, _fp]
    L.alens_mlp_get.argtypes = [vp, _fp, _fp]
    L.alens_get_features.argtypes = [vp, _fp]
    L.alens_mlp_predict.argtypes = [vp, ctypes.c_int64, _fp, _fp]
    L.alens_mlp_gradient.argtypes = [vp, ctypes.c_int64, _fp, _fp, _fp, _fp]
    L.alens_mlp_train.argtypes = [vp, ctypes.c_int64, _fp, _fp, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                  ctypes.c_uint64, _fp]
    L.alens_copy_probe.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, _dp]
    L.alens_mlp_energy.argtypes = [vp, _dp, _fp]
    L.alens_mlp_train_energy.argtypes = [vp, ctypes.c_int32, ctypes.c_float, _dp]
    L.alens_bcqp_solve_static.argtypes = L.alens_bcqp_solve.argtypes
    for name in EXPORTS:
        fn = getattr(L, name)
        if name not in ("alens_destroy", "alens_last_error", "alens_version", "alens_get_step"):
            fn.restype = ctypes.c_int
    _lib = L
    return L


def copy_probe(nbytes=1 << 30, reps=10, device=0):
    """Read + write GB/s of a device-to-device copy kernel (alens_copy_probe): the practical HBM
    ceiling beside the published peak.  One repetition of a known size also serves as the profiler
    calibration kernel (k_copy_probe)."""
    g = ctypes.c_double()
    st = lib().alens_copy_probe(int(device), int(nbytes), int(reps), ctypes.byref(g))
    if st != ALENS_OK:
        raise AlensError(st, "alens_copy_probe failed")
    return g.value


def _p(a, ct):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ct))


def make_config(box_lo=(0, 0, 0), box_hi=(1, 1, 1), pbc=(1, 1, 1), monolayer=0, col_buf=0.025, viscosity=0.01,
                kT=0.00411, dt=1e-4, tol=1e-5, diam_col_ratio=1.0, len_col_ratio=1.0, cell_size=0.0,
                max_ite=100000, solver=0, precision=0, warm_start=0, seed=1234, device=0, **_ignored):
    c = Config()
    c.box_lo[:] = [float(v) for v in box_lo]
    c.box_hi[:] = [float(v) for v in box_hi]
    c.pbc[:] = [int(v) for v in pbc]
    c.monolayer = int(monolayer)
    c.col_buf = float(col_buf)
    c.viscosity = float(viscosity)
    c.kT = float(kT)
    c.dt = float(dt)
    c.tol = float(tol)
    c.diam_col_ratio = float(diam_col_ratio)
    c.len_col_ratio = float(len_col_ratio)
    c.cell_size = float(cell_size)
    c.max_ite = int(max_ite)
    c.solver = int(solver)
    c.precision = int(precision)
    c.warm_start = int(warm_start)
    c.seed = int(seed)
    c.device = int(device)
    return c


class Context:
    """One rod system on one GPU (mirrors the lifetime of SimToolbox's SylinderSystem)."""

    def __init__(self, **cfg):
        self.cfg = make_config(**cfg)
        self._h = ctypes.c_void_p()
        st = lib().alens_create(ctypes.byref(self.cfg), ctypes.byref(self._h))
        if st != ALENS_OK:
            raise AlensError(st, "alens_create failed")
        self.n = 0

    def close(self):
        if self._h:
            lib().alens_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st, allow=()):
        if st != ALENS_OK and st not in allow:
            raise AlensError(st, lib().alens_last_error(self._h).decode())
        return st

    def set_rods(self, rods):
        gid = np.ascontiguousarray(rods["gid"], np.int32)
        c = np.ascontiguousarray(rods["center"], np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(rods["dir"], np.float64).reshape(-1, 3)
        L = np.ascontiguousarray(rods["length"], np.float64)
        r = np.ascontiguousarray(rods["radius"], np.float64)
        imm = rods.get("immovable")
        imm = None if imm is None else np.ascontiguousarray(imm, np.uint8)
        n = gid.size
        self._check(lib().alens_set_rods(self._h, n, _p(gid, ctypes.c_int32), _p(c, ctypes.c_double),
                                         _p(t, ctypes.c_double), _p(L, ctypes.c_double), _p(r, ctypes.c_double),
                                         _p(imm, ctypes.c_uint8)))
        self.n = n

    def set_systems(self, systems_list):
        """systems_list: [(rods, box_cfg)] with box_cfg holding box_lo, box_hi, pbc, monolayer."""
        parts = [r for r, _ in systems_list]
        off = np.zeros(len(parts) + 1, np.int32)
        off[1:] = np.cumsum([len(r["length"]) for r in parts])
        cat = lambda k, dt: np.ascontiguousarray(np.concatenate([np.asarray(r[k], dt) for r in parts]), dt)
        gid, L, rad = cat("gid", np.int32), cat("length", np.float64), cat("radius", np.float64)
        c = np.ascontiguousarray(np.concatenate([np.asarray(r["center"]).reshape(-1, 3) for r in parts]))
        t = np.ascontiguousarray(np.concatenate([np.asarray(r["dir"]).reshape(-1, 3) for r in parts]))
        boxes = (Box * len(parts))()
        for b, (_, bc) in zip(boxes, systems_list):
            b.box_lo[:] = [float(v) for v in bc["box_lo"]]
            b.box_hi[:] = [float(v) for v in bc["box_hi"]]
            b.pbc[:] = [int(v) for v in bc["pbc"]]
            b.monolayer = int(bc.get("monolayer", 0))
        self._check(lib().alens_set_systems(self._h, len(parts), _p(off, ctypes.c_int32), boxes,
                                            _p(gid, ctypes.c_int32), _p(c, ctypes.c_double), _p(t, ctypes.c_double),
                                            _p(L, ctypes.c_double), _p(rad, ctypes.c_double), None))
        self.n = int(off[-1])
        self.rod_offset = off
        return off

    def set_graph(self, enable=True):
        """Run later steps as one captured HIP graph each (alens_set_graph)."""
        self._check(lib().alens_set_graph(self._h, 1 if enable else 0))

    def graph_info(self):
        g = GraphInfo()
        self._check(lib().alens_get_graph_info(self._h, ctypes.byref(g)))
        return dict(captured=g.captured, replays=g.replays, captures=g.captures,
                    capture_failures=g.capture_failures, overflows=g.overflows,
                    last_failure=g.last_failure.decode(errors="replace"))

    def autotune(self):
        """Pick the solver schedule by timing the last step's solve under candidates (alens_autotune)."""
        r = TuneReport()
        self._check(lib().alens_autotune(self._h, ctypes.byref(r)))
        return r.as_dict()

    def system_reports(self):
        ns = ctypes.c_int32()
        self._check(lib().alens_get_num_systems(self._h, ctypes.byref(ns)))
        arr = (Report * max(ns.value, 1))()
        self._check(lib().alens_get_system_reports(self._h, arr))
        return [arr[i].as_dict() for i in range(ns.value)]

    def get_rods(self):
        c = np.zeros((self.n, 3))
        t = np.zeros((self.n, 3))
        self._check(lib().alens_get_rods(self._h, _p(c, ctypes.c_double), _p(t, ctypes.c_double)))
        return c, t

    def sort_rods(self):
        """Reorder the rods spatially on the device; returns perm (new index -> previous index)."""
        perm = np.zeros(max(self.n, 1), np.int32)
        self._check(lib().alens_sort_rods(self._h, _p(perm, ctypes.c_int32)))
        return perm[: self.n]

    def set_step(self, step):
        self._check(lib().alens_set_step(self._h, int(step)))

    @property
    def step_count(self):
        return lib().alens_get_step(self._h)

    def step(self, f_ext=None, allow=(ALENS_E_NOCONV,)):
        rep = Report()
        fe = None if f_ext is None else np.ascontiguousarray(f_ext, np.float64).reshape(-1)
        st = lib().alens_step_constraints(self._h, _p(fe, ctypes.c_double), ctypes.byref(rep))
        self._check(st, allow)
        d = rep.as_dict()
        d["status_code"] = st
        return d

    def euler(self):
        self._check(lib().alens_step_euler(self._h))

    def num_constraints(self):
        m = ctypes.c_int64()
        self._check(lib().alens_get_num_constraints(self._h, ctypes.byref(m)))
        return m.value

    def constraint_arrays(self):
        m = self.num_constraints()
        out = dict(I=np.zeros(m, np.int32), J=np.zeros(m, np.int32), img=np.zeros((m, 3), np.int32),
                   delta0=np.zeros(m), nrm=np.zeros((m, 3)), posI=np.zeros((m, 3)), posJ=np.zeros((m, 3)),
                   rI=np.zeros((m, 3)), rJ=np.zeros((m, 3)), b=np.zeros(m))
        if m:
            self._check(lib().alens_get_constraint_arrays(
                self._h, _p(out["I"], ctypes.c_int32), _p(out["J"], ctypes.c_int32), _p(out["img"], ctypes.c_int32),
                _p(out["delta0"], ctypes.c_double), _p(out["nrm"], ctypes.c_double), _p(out["posI"], ctypes.c_double),
                _p(out["posJ"], ctypes.c_double), _p(out["rI"], ctypes.c_double), _p(out["rJ"], ctypes.c_double),
                _p(out["b"], ctypes.c_double)))
        return out

    def constraint_blocks(self):
        m = self.num_constraints()
        arr = (ConBlock * max(m, 1))()
        mm = ctypes.c_int64()
        self._check(lib().alens_get_constraints(self._h, m, arr, ctypes.byref(mm)))
        return arr[:m]

    def add_constraints(self, ext):
        """Queue caller constraint blocks for the next step (the reference's constraint pool).  ext is a
        dict of arrays: gid_i, gid_j, delta0, gamma0, nrm (m x 3, = norm_i; norm_j = -nrm), posI, posJ
        (m x 3), bilateral, kappa, gamma_lb, and optionally labI, labJ (m x 3), stress (m x 9)."""
        m = len(ext["gid_i"])
        arr = (ConBlock * max(m, 1))()
        f3 = lambda key: np.asarray(ext[key], np.float64).reshape(m, 3) if key in ext else np.zeros((m, 3))
        nrm, pI, pJ, lI, lJ = f3("nrm"), f3("posI"), f3("posJ"), f3("labI"), f3("labJ")
        st = np.asarray(ext["stress"], np.float64).reshape(m, 9) if "stress" in ext else np.zeros((m, 9))
        for e in range(m):
            b = arr[e]
            b.delta0 = float(ext["delta0"][e])
            b.gamma = float(ext["gamma0"][e])
            b.kappa = float(ext["kappa"][e])
            b.gamma_lb = float(ext["gamma_lb"][e])
            b.gid_i = int(ext["gid_i"][e])
            b.gid_j = int(ext["gid_j"][e])
            b.bilateral = int(bool(ext["bilateral"][e]))
            for a in range(3):
                b.norm_i[a] = nrm[e, a]
                b.norm_j[a] = -nrm[e, a]
                b.pos_i[a] = pI[e, a]
                b.pos_j[a] = pJ[e, a]
                b.lab_i[a] = lI[e, a]
                b.lab_j[a] = lJ[e, a]
            for q in range(9):
                b.stress[q] = st[e, q]
        self._check(lib().alens_add_constraints(self._h, m, arr))

    def clear_constraints(self):
        self._check(lib().alens_clear_constraints(self._h))

    def con_stress(self):
        """(collision, caller) sums of gamma * per-unit stress over the last step's rows, 3 x 3 each."""
        c, e = np.zeros(9), np.zeros(9)
        self._check(lib().alens_get_con_stress(self._h, _p(c, ctypes.c_double), _p(e, ctypes.c_double)))
        return c.reshape(3, 3), e.reshape(3, 3)

    def gamma(self):
        m = self.num_constraints()
        g = np.zeros(m)
        self._check(lib().alens_get_gamma(self._h, _p(g, ctypes.c_double)))
        return g

    def velocity(self):
        n6 = 6 * self.n
        u, unc, f, mob = np.zeros(n6), np.zeros(n6), np.zeros(n6), np.zeros(n6)
        self._check(lib().alens_get_velocity(self._h, _p(u, ctypes.c_double), _p(unc, ctypes.c_double),
                                             _p(f, ctypes.c_double), _p(mob, ctypes.c_double)))
        return dict(u=u, unc=unc, fcon=f, mob=mob)

    def bcqp_solve(self, n, I, J, nrm, rI, rJ, b, mob, x0=None, allow=(ALENS_E_NOCONV,), static=False):
        """alens_bcqp_solve, or with static=True its allocation-free variant alens_bcqp_solve_static."""
        I = np.ascontiguousarray(I, np.int32)
        J = np.ascontiguousarray(J, np.int32)
        nrm = np.ascontiguousarray(nrm, np.float64).reshape(-1)
        rI = np.ascontiguousarray(rI, np.float64).reshape(-1)
        rJ = np.ascontiguousarray(rJ, np.float64).reshape(-1)
        b = np.ascontiguousarray(b, np.float64)
        mob = np.ascontiguousarray(mob, np.float64).reshape(-1)
        x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        m = I.size
        g = np.zeros(max(m, 1))
        rep = Report()
        fn = lib().alens_bcqp_solve_static if static else lib().alens_bcqp_solve
        st = fn(self._h, int(n), m, _p(I, ctypes.c_int32), _p(J, ctypes.c_int32),
                                    _p(nrm, ctypes.c_double), _p(rI, ctypes.c_double), _p(rJ, ctypes.c_double),
                                    _p(b, ctypes.c_double), _p(mob, ctypes.c_double), _p(x0, ctypes.c_double),
                                    _p(g, ctypes.c_double), ctypes.byref(rep))
        self._check(st, allow)
        d = rep.as_dict()
        d["status_code"] = st
        return g[:m], d

    # ---- learned warm start ----
    def mlp_set(self, params, norm):
        P = np.ascontiguousarray(params, np.float32)
        N = np.ascontiguousarray(norm, np.float32)
        assert P.size == MLP_PARAMS and N.size == MLP_NORM
        self._check(lib().alens_mlp_set(self._h, _p(P, ctypes.c_float), _p(N, ctypes.c_float)))

    def mlp_get(self):
        P = np.zeros(MLP_PARAMS, np.float32)
        N = np.zeros(MLP_NORM, np.float32)
        self._check(lib().alens_mlp_get(self._h, _p(P, ctypes.c_float), _p(N, ctypes.c_float)))
        return P, N

    def features(self):
        m = self.num_constraints()
        X = np.zeros((max(m, 1), 16), np.float32)
        self._check(lib().alens_get_features(self._h, _p(X, ctypes.c_float)))
        return X[:m]

    def mlp_predict(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 16)
        y = np.zeros(max(len(X), 1), np.float32)
        self._check(lib().alens_mlp_predict(self._h, len(X), _p(X, ctypes.c_float), _p(y, ctypes.c_float)))
        return y[: len(X)]

    def mlp_gradient(self, X, target):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 16)
        t = np.ascontiguousarray(target, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = ctypes.c_float()
        self._check(lib().alens_mlp_gradient(self._h, len(X), _p(X, ctypes.c_float), _p(t, ctypes.c_float),
                                             _p(g, ctypes.c_float), ctypes.byref(loss)))
        return g, loss.value

    def mlp_energy(self, grad=False):
        """BCQP energy of the network's start point on the last step's rows (alens_mlp_energy), and with
        grad=True its gradient with respect to the parameters."""
        e = ctypes.c_double()
        g = np.zeros(MLP_PARAMS, np.float32) if grad else None
        self._check(lib().alens_mlp_energy(self._h, ctypes.byref(e), _p(g, ctypes.c_float)))
        return (e.value, g) if grad else e.value

    def mlp_train_energy(self, iters=10, lr=1e-3):
        """Adam on the start point's BCQP energy for the last step's rows (alens_mlp_train_energy);
        returns the energy before each step and after the last."""
        e = np.zeros(iters + 1)
        self._check(lib().alens_mlp_train_energy(self._h, int(iters), float(lr), _p(e, ctypes.c_double)))
        return e

    def mlp_train(self, X, gamma, epochs=10, batch=4096, lr=2e-3, seed=1):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 16)
        gm = np.ascontiguousarray(gamma, np.float32)
        losses = np.zeros(max(epochs, 1), np.float32)
        self._check(lib().alens_mlp_train(self._h, len(X), _p(X, ctypes.c_float), _p(gm, ctypes.c_float),
                                          int(epochs), int(batch), float(lr), int(seed),
                                          _p(losses, ctypes.c_float)))
        return losses[:epochs]
