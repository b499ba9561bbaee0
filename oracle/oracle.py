"""ctypes wrapper of the CPU oracle (oracle/mpm_oracle.c) plus the reference's loss.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.

Conventions (shared with the HIP library): forces has max_steps rows; row s is
m_stick/dt * sum_p dv[s, p] (sf:330-336, fcam:16-19) with row 0 = 0.  The losses follow
sf:330-339 (L1 of the x force over steps 0..T-2, divided by T-1) and the MSE over x and z.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_BUILD selects another build directory, e.g. oracle/_build/asan from `make asan`
_BUILD = os.environ.get("ORACLE_BUILD") or os.path.join(_HERE, "_build")


class OrScene(C.Structure):
    _fields_ = [("n_grid", C.c_int), ("bound", C.c_int), ("model", C.c_int), ("k_from_param", C.c_int),
                ("dx", C.c_double), ("inv_dx", C.c_double), ("dt", C.c_double), ("gravity", C.c_double),
                ("k_bulk", C.c_double), ("wall_r", C.c_double), ("wall_cx", C.c_double),
                ("wall_cz", C.c_double), ("m_liq", C.c_double), ("vol_liq", C.c_double),
                ("m_stick", C.c_double), ("m_cont", C.c_double)]


def _state_struct(real):
    P = C.POINTER(real)

    class OrState(C.Structure):
        _fields_ = [("nl", C.c_int), ("ns", C.c_int), ("nc", C.c_int),
                    ("lx", P), ("lv", P), ("lC", P), ("lF", P),
                    ("sx", P), ("sv", P), ("sC", P), ("cx", P), ("cv", P), ("cC", P)]

    class OrAdj(C.Structure):
        _fields_ = [("lx", P), ("lv", P), ("lC", P), ("lF", P), ("sC", P), ("cC", P)]

    return OrState, OrAdj


ERRORS = {1: "particle outside the grid", 2: "non-finite particle position", 3: "out of memory", 4: "bad argument"}


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)


class Oracle:
    """One build of the oracle: f64 (the oracle), f32 (FP32 drift sizing) or f32at (FP32 with the
    reference's execution structure: per-particle parallel loops with atomic grid adds; the
    benchmark's CPU baseline)."""

    PREFIX = {"f64": "or_", "f32": "or32_", "f32at": "ora32_"}

    def __init__(self, precision: str = "f64"):
        if precision not in self.PREFIX:
            raise ValueError(f"precision must be one of {list(self.PREFIX)}")
        path = os.path.join(_BUILD, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        self.precision = precision
        self.real = C.c_double if precision == "f64" else C.c_float
        self.np_real = np.float64 if precision == "f64" else np.float32
        self.State, self.Adj = _state_struct(self.real)
        pre = self.PREFIX[precision]
        self._set_threads = getattr(self.lib, pre + "set_threads")
        self._set_threads.restype = C.c_int
        self._set_threads.argtypes = [C.c_int]
        self._step = getattr(self.lib, pre + "step")
        self._step_adj = getattr(self.lib, pre + "step_adjoint")
        self._rollout = getattr(self.lib, pre + "rollout")
        self._rollout_grad = getattr(self.lib, pre + "rollout_grad")
        self._gridop_one = getattr(self.lib, pre + "gridop_one")
        self._eta = getattr(self.lib, pre + "eta")
        self._p2g = getattr(self.lib, pre + "p2g")
        self._gather = getattr(self.lib, pre + "gather")
        self._eta.restype = C.c_double
        self._eta.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]

    def set_threads(self, n: int = 0) -> int:
        """OpenMP threads of the following calls (0: keep); returns the count in use."""
        return int(self._set_threads(int(n)))

    # ------------------------------------------------------------------ marshalling
    @staticmethod
    def scene_struct(cfg) -> OrScene:
        return OrScene(int(cfg.n_grid), int(cfg.bound), int(cfg.model), int(bool(cfg.k_from_param)),
                       1.0 / cfg.n_grid, float(cfg.n_grid), float(cfg.dt), float(cfg.gravity),
                       float(cfg.k_bulk), float(cfg.wall_r), float(cfg.wall_cx), float(cfg.wall_cz),
                       float(cfg.m_liq), float(cfg.vol_liq), float(cfg.m_stick), float(cfg.m_cont))

    def _arr(self, a, shape):
        a = np.ascontiguousarray(np.asarray(a, dtype=self.np_real).reshape(shape))
        return a

    def _ptr(self, a):
        return a.ctypes.data_as(C.POINTER(self.real))

    def make_state(self, st: dict):
        """st: dict with lx, lv, lC, lF, sx, sv, sC, cx, cv, cC (any float arrays)."""
        nl = int(np.asarray(st["lx"]).size // 3)
        ns = int(np.asarray(st["sx"]).size // 3)
        nc = int(np.asarray(st["cx"]).size // 3)
        keep = {}
        shapes = {"lx": (nl, 3), "lv": (nl, 3), "lC": (nl, 9), "lF": (nl, 9), "sx": (ns, 3), "sv": (ns, 3),
                  "sC": (ns, 9), "cx": (nc, 3), "cv": (nc, 3), "cC": (nc, 9)}
        for k, shp in shapes.items():
            keep[k] = self._arr(st[k], shp)
        s = self.State(nl, ns, nc, *[self._ptr(keep[k]) for k in shapes])
        return s, keep

    def empty_state(self, nl, ns, nc):
        z = lambda n, c: np.zeros((n, c), self.np_real)
        return self.make_state({"lx": z(nl, 3), "lv": z(nl, 3), "lC": z(nl, 9), "lF": z(nl, 9),
                                "sx": z(ns, 3), "sv": z(ns, 3), "sC": z(ns, 9),
                                "cx": z(nc, 3), "cv": z(nc, 3), "cC": z(nc, 9)})

    def make_adj(self, ad: dict):
        keep = {k: self._arr(ad[k], (-1,)) for k in ("lx", "lv", "lC", "lF", "sC", "cC")}
        return self.Adj(*[self._ptr(keep[k]) for k in ("lx", "lv", "lC", "lF", "sC", "cC")]), keep

    @staticmethod
    def initial_state(scene) -> dict:
        """Step-0 state (sf:301-319): v = 0, F = I, C = 0 for every species."""
        nl, ns, nc = scene.n_liquid, scene.n_stick, scene.n_container
        eye = np.tile(np.eye(3).ravel(), (nl, 1))
        return {"lx": scene.x_liquid, "lv": np.zeros((nl, 3)), "lC": np.zeros((nl, 9)), "lF": eye,
                "sx": scene.x_stick, "sv": np.zeros((ns, 3)), "sC": np.zeros((ns, 9)),
                "cx": scene.x_container, "cv": np.zeros((nc, 3)), "cC": np.zeros((nc, 9))}

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}: {ERRORS.get(rc, 'unknown')}")

    # ------------------------------------------------------------------ API
    def step(self, cfg, params, st: dict, traj_f, traj_f1):
        sc = self.scene_struct(cfg)
        s_in, keep_in = self.make_state(st)
        s_out, keep_out = self.empty_state(s_in.nl, s_in.ns, s_in.nc)
        th = np.ascontiguousarray(params, dtype=np.float64)
        tf = np.ascontiguousarray(traj_f, dtype=np.float64)
        tf1 = np.ascontiguousarray(traj_f1, dtype=np.float64)
        dv = np.zeros(3)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self._check(self._step(C.byref(sc), dp(th), C.byref(s_in), C.byref(s_out), dp(tf), dp(tf1), dp(dv)))
        return keep_out, dv

    def step_adjoint(self, cfg, params, st: dict, traj_f, traj_f1, out_bar: dict, dvbar):
        sc = self.scene_struct(cfg)
        s_in, keep_in = self.make_state(st)
        ob, keep_ob = self.make_adj(out_bar)
        nl, ns, nc = s_in.nl, s_in.ns, s_in.nc
        ib_keep = {"lx": np.zeros(nl * 3, self.np_real), "lv": np.zeros(nl * 3, self.np_real),
                   "lC": np.zeros(nl * 9, self.np_real), "lF": np.zeros(nl * 9, self.np_real),
                   "sC": np.zeros(ns * 9, self.np_real), "cC": np.zeros(nc * 9, self.np_real)}
        ib = self.Adj(*[self._ptr(ib_keep[k]) for k in ("lx", "lv", "lC", "lF", "sC", "cC")])
        th = np.ascontiguousarray(params, dtype=np.float64)
        thb = np.zeros(4)
        dvb = np.ascontiguousarray(dvbar, dtype=np.float64)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        tf = np.ascontiguousarray(traj_f, dtype=np.float64)
        tf1 = np.ascontiguousarray(traj_f1, dtype=np.float64)
        self._check(self._step_adj(C.byref(sc), dp(th), C.byref(s_in), dp(tf), dp(tf1), C.byref(ob), dp(dvb),
                                   C.byref(ib), dp(thb)))
        shp = {"lx": (nl, 3), "lv": (nl, 3), "lC": (nl, 9), "lF": (nl, 9), "sC": (ns, 9), "cC": (nc, 9)}
        return {k: v.reshape(shp[k]) for k, v in ib_keep.items()}, thb

    def rollout(self, scene, params=None, nsteps=None, return_state=False):
        cfg = scene.cfg
        params = scene.params if params is None else params
        nsteps = scene.max_steps - 1 if nsteps is None else nsteps
        sc = self.scene_struct(cfg)
        s0, keep0 = self.make_state(self.initial_state(scene))
        traj = np.ascontiguousarray(scene.traj[:nsteps + 1], dtype=np.float64)
        forces = np.zeros((nsteps + 1, 3))
        th = np.ascontiguousarray(params, dtype=np.float64)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        fin, keep_fin = self.empty_state(s0.nl, s0.ns, s0.nc)
        self._check(self._rollout(C.byref(sc), dp(th), C.byref(s0), dp(traj), int(nsteps), dp(forces),
                                  C.byref(fin)))
        if return_state:
            return forces, keep_fin
        return forces

    def rollout_grad(self, scene, dLdF, params=None, nsteps=None, ckpt=10):
        cfg = scene.cfg
        params = scene.params if params is None else params
        nsteps = scene.max_steps - 1 if nsteps is None else nsteps
        sc = self.scene_struct(cfg)
        s0, keep0 = self.make_state(self.initial_state(scene))
        traj = np.ascontiguousarray(scene.traj[:nsteps + 1], dtype=np.float64)
        forces = np.zeros((nsteps + 1, 3))
        th = np.ascontiguousarray(params, dtype=np.float64)
        g = np.ascontiguousarray(dLdF, dtype=np.float64).reshape(nsteps + 1, 3)
        thb = np.zeros(4)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self._check(self._rollout_grad(C.byref(sc), dp(th), C.byref(s0), dp(traj), int(nsteps), int(ckpt),
                                       dp(g), dp(forces), dp(thb)))
        return forces, thb

    def eta(self, model, params, gamma):
        th = np.ascontiguousarray(params, dtype=np.float64)
        dg = C.c_double(0)
        dth = np.zeros(4)
        e = self._eta(int(model), th.ctypes.data_as(C.POINTER(C.c_double)), float(gamma), C.byref(dg),
                      dth.ctypes.data_as(C.POINTER(C.c_double)))
        return e, dg.value, dth

    def gridop_one(self, cfg, Pn, i, j, k):
        sc = self.scene_struct(cfg)
        pn = self._arr(Pn, (4,))
        out = np.zeros(3, self.np_real)
        self._gridop_one(C.byref(sc), self._ptr(pn), int(i), int(j), int(k), self._ptr(out))
        return out

    def p2g(self, cfg, params, st: dict):
        sc = self.scene_struct(cfg)
        s_in, keep = self.make_state(st)
        n = cfg.n_grid
        P = np.zeros((n, n, n, 4), self.np_real)
        th = np.ascontiguousarray(params, dtype=np.float64)
        self._check(self._p2g(C.byref(sc), th.ctypes.data_as(C.POINTER(C.c_double)), C.byref(s_in), self._ptr(P)))
        return P

    def gather(self, cfg, Vout, x):
        sc = self.scene_struct(cfg)
        V = self._arr(Vout, (-1,))
        X = self._arr(x, (-1, 3))
        n = X.shape[0]
        nv = np.zeros((n, 3), self.np_real)
        nC = np.zeros((n, 9), self.np_real)
        self._check(self._gather(C.byref(sc), self._ptr(V), int(n), self._ptr(X), self._ptr(nv), self._ptr(nC)))
        return nv, nC


# ---------------------------------------------------------------------------- losses

def loss_and_grad(forces: np.ndarray, target: np.ndarray, kind: int):
    """Loss over steps 0..T-2 and its gradient with respect to the forces (T rows).

    kind 0: reference L1 of the x force (sf:330-339); d|.|/dx = sign(x), 0 at 0.
    kind 1: MSE over x and z, sum_s [(Fx-tx)^2 + (Fz-tz)^2] / (2 (T-1))."""
    F = np.asarray(forces, dtype=np.float64)
    T = F.shape[0]
    tgt = np.asarray(target, dtype=np.float64).reshape(T, 3)
    g = np.zeros_like(F)
    d = F[:T - 1] - tgt[:T - 1]
    if kind == 0:
        L = np.abs(d[:, 0]).sum() / (T - 1)
        g[:T - 1, 0] = np.sign(d[:, 0]) / (T - 1)
    else:
        L = ((d[:, 0] ** 2).sum() + (d[:, 2] ** 2).sum()) / (2 * (T - 1))
        g[:T - 1, 0] = d[:, 0] / (T - 1)
        g[:T - 1, 2] = d[:, 2] / (T - 1)
    return float(L), g
