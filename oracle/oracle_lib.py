"""ctypes wrapper of the C oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``; never by the product package.

``OracleScene`` holds one scene's state as NumPy arrays of the chosen
precision (float32 = the reference's Warp arithmetic, float64 = the FP64
oracle) and advances it with ``oracle_step_{f32,f64}`` (a restatement of
``MPM_Simulator_WARP.p2g2p``, engine_stress/mpm_solver_warp.py:156-203).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


class OracleParams(C.Structure):
    _fields_ = [("n_grid", C.c_int32), ("material", C.c_int32),
                ("dx", C.c_double), ("inv_dx", C.c_double),
                ("mu", C.c_double), ("lam", C.c_double), ("alpha", C.c_double),
                ("gravity", C.c_double * 3), ("hardening", C.c_int32), ("xi", C.c_double)]


class OracleCollider(C.Structure):
    _fields_ = [("kind", C.c_int32), ("surface_type", C.c_int32),
                ("point", C.c_double * 3), ("normal", C.c_double * 3),
                ("start_time", C.c_double), ("end_time", C.c_double),
                ("friction", C.c_double), ("velocity", C.c_double * 3),
                ("threshold", C.c_double), ("radius", C.c_double), ("height", C.c_double), ("width", C.c_double),
                ("v_scale", C.c_double), ("x_unit", C.c_double * 3), ("y_unit", C.c_double * 3),
                ("reset", C.c_int32)]


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


def _sfx(dtype) -> str:
    return "_f64" if np.dtype(dtype) == np.float64 else "_f32"


def make_params(n_grid, dx, inv_dx, mu, lam, alpha, gravity, material=2, hardening=0, xi=0.0) -> OracleParams:
    p = OracleParams()
    p.n_grid = int(n_grid)
    p.material = int(material)
    p.hardening = int(hardening)
    p.xi = float(xi)
    p.dx, p.inv_dx, p.mu, p.lam, p.alpha = float(dx), float(inv_dx), float(mu), float(lam), float(alpha)
    for i in range(3):
        p.gravity[i] = float(gravity[i])
    return p


def grid_update(params: dict, grid_m: np.ndarray, grid_mv: np.ndarray, colliders, time: float, dt: float):
    """oracle_grid_update on (n^3,) masses and (n^3, 3) momenta; returns v_out (n^3, 3)."""
    dtype = grid_m.dtype
    P = make_params(params["n_grid"], params["dx"], params["inv_dx"], params["mu"], params["lam"],
                    params["alpha"], params["gravity"], params.get("material", 2))
    gm = np.ascontiguousarray(grid_m)
    gmv = np.ascontiguousarray(grid_mv)
    gv = np.zeros_like(gmv)
    cc = make_colliders(colliders)
    getattr(lib(), "oracle_grid_update" + _sfx(dtype))(C.byref(P), _p(gm), _p(gmv), _p(gv), cc,
                                                         C.c_int(len(colliders)), C.c_double(time), C.c_double(dt))
    return gv


def make_colliders(colliders):
    arr = (OracleCollider * max(1, len(colliders)))()
    for i, c in enumerate(colliders):
        arr[i].kind = int(c["kind"])
        arr[i].surface_type = int(c.get("surface_type", 0))
        for d in range(3):
            arr[i].point[d] = float(c.get("point", (0, 0, 0))[d])
            arr[i].normal[d] = float(c.get("normal", (0, 0, 0))[d])
        arr[i].start_time = float(c.get("start_time", 0.0))
        arr[i].end_time = float(c.get("end_time", 999.0))
        arr[i].friction = float(c.get("friction", 0.0))
        for d in range(3):
            arr[i].velocity[d] = float(c.get("velocity", (0, 0, 0))[d])
            arr[i].x_unit[d] = float(c.get("x_unit", (0, 0, 0))[d])
            arr[i].y_unit[d] = float(c.get("y_unit", (0, 0, 0))[d])
        for k in ("threshold", "radius", "height", "width", "v_scale"):
            setattr(arr[i], k, float(c.get(k, 0.0)))
        arr[i].reset = int(c.get("reset", 0))
    return arr


def move_collider(c: dict, time: float, dt: float):
    """The host `modify` callbacks of the moving colliders, after the substep
    at `time`: kinds 2, 5, 6 move the point by dt * velocity
    (mpm_solver_warp.py:560-563, 688-691, 715-717), kind 3 along its circle
    (:594-603); Python arithmetic stored back into a float32 vec3, as Warp's
    struct fields are."""
    k = c["kind"]
    if k not in (2, 3, 5, 6) or not (time >= c["start_time"] and time < c["end_time"]):
        return
    p = c["point"]
    if k == 3:
        z = [np.float32(p[0] - 0.5), np.float32(p[1]), np.float32(p[2] - 0.5)]
        ln = np.sqrt(np.float32(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]))
        th = float(np.arccos(np.float32(z[0] / ln)))
        if z[2] < 0:
            th = -th
        vel = (-c["v_scale"] * np.sin(th), 0.0, c["v_scale"] * np.cos(th))
    else:
        vel = c["velocity"]
    c["point"] = tuple(float(np.float32(p[d] + dt * vel[d])) for d in range(3))


# ---- single-function entry points ---------------------------------------

def svd3(A: np.ndarray):
    """oracle_svd3 on one 3x3 matrix; returns U, sig, V."""
    A = np.ascontiguousarray(A)
    U = np.zeros((3, 3), A.dtype); V = np.zeros((3, 3), A.dtype); s = np.zeros(3, A.dtype)
    getattr(lib(), "oracle_svd3" + _sfx(A.dtype))(_p(A), _p(U), _p(s), _p(V))
    return U, s, V


def weights(x: np.ndarray, inv_dx: float):
    x = np.ascontiguousarray(x)
    base = np.zeros(3, np.int32); fx = np.zeros(3, x.dtype)
    w = np.zeros((3, 3), x.dtype); dw = np.zeros((3, 3), x.dtype)
    ct = C.c_double if x.dtype == np.float64 else C.c_float
    getattr(lib(), "oracle_weights" + _sfx(x.dtype))(_p(x), ct(inv_dx), _p(base), _p(fx), _p(w), _p(dw))
    return base, fx, w, dw


def sand_return_map(Ftr: np.ndarray, mu, lam, alpha):
    Ftr = np.ascontiguousarray(Ftr)
    Fe = np.zeros((3, 3), Ftr.dtype)
    ct = C.c_double if Ftr.dtype == np.float64 else C.c_float
    getattr(lib(), "oracle_sand_return_map" + _sfx(Ftr.dtype))(_p(Ftr), ct(mu), ct(lam), ct(alpha), _p(Fe))
    return Fe


def dp_stress(F: np.ndarray, mu, lam):
    F = np.ascontiguousarray(F)
    tau = np.zeros((3, 3), F.dtype)
    ct = C.c_double if F.dtype == np.float64 else C.c_float
    getattr(lib(), "oracle_dp_stress" + _sfx(F.dtype))(_p(F), ct(mu), ct(lam), _p(tau))
    return tau


def von_mises(Ftr: np.ndarray, mu, lam, yield_stress, hardening=0, xi=0.0):
    """oracle_von_mises on one 3x3 matrix; returns (F_e, updated yield stress)."""
    Ftr = np.ascontiguousarray(Ftr)
    Fe = np.zeros((3, 3), Ftr.dtype)
    ys = np.array([yield_stress], Ftr.dtype)
    ct = C.c_double if Ftr.dtype == np.float64 else C.c_float
    getattr(lib(), "oracle_von_mises" + _sfx(Ftr.dtype))(_p(Ftr), ct(mu), ct(lam), _p(ys), C.c_int(hardening),
                                                           ct(xi), _p(Fe))
    return Fe, float(ys[0])


def stvk_stress(F: np.ndarray, mu, lam):
    F = np.ascontiguousarray(F)
    tau = np.zeros((3, 3), F.dtype)
    ct = C.c_double if F.dtype == np.float64 else C.c_float
    getattr(lib(), "oracle_stvk_stress" + _sfx(F.dtype))(_p(F), ct(mu), ct(lam), _p(tau))
    return tau


def fcr_stress(F: np.ndarray, mu, lam):
    F = np.ascontiguousarray(F)
    tau = np.zeros((3, 3), F.dtype)
    ct = C.c_double if F.dtype == np.float64 else C.c_float
    getattr(lib(), "oracle_fcr_stress" + _sfx(F.dtype))(_p(F), ct(mu), ct(lam), _p(tau))
    return tau


# ---- scene state ----------------------------------------------------------

class OracleScene:
    """Particle state + grid scratch of one scene in float32 or float64."""

    FIELDS = ("x0", "u", "v", "C", "F", "tau", "mass", "vol", "yield_stress")

    def __init__(self, params: dict, state: dict, colliders=(), dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.params = dict(params)
        self.n = int(state["x0"].shape[0])
        shapes = dict(x0=(self.n, 3), u=(self.n, 3), v=(self.n, 3), C=(self.n, 3, 3),
                      F=(self.n, 3, 3), tau=(self.n, 3, 3), mass=(self.n,), vol=(self.n,), yield_stress=(self.n,))
        state = dict(state)
        state.setdefault("yield_stress", np.zeros(self.n))
        self.s = {k: np.array(state[k], dtype=self.dtype, copy=True, order="C").reshape(shapes[k])
                  for k in self.FIELDS}
        ng = int(params["n_grid"])
        self.x = np.zeros((self.n, 3), self.dtype)
        self.grid_m = np.zeros(ng ** 3, self.dtype)
        self.grid_mv = np.zeros((ng ** 3, 3), self.dtype)
        self.grid_v = np.zeros((ng ** 3, 3), self.dtype)
        self.colliders = [dict(c) for c in colliders]
        self._cp = make_params(ng, params["dx"], params["inv_dx"], params["mu"], params["lam"],
                               params["alpha"], params["gravity"], params.get("material", 2),
                               params.get("hardening", 0), params.get("xi", 0.0))
        self._cc = make_colliders(self.colliders)
        self.time = float(params.get("time", 0.0))

    def step(self, dt: float, nthreads: int = 1):
        s = self.s
        if any(c["kind"] in (2, 3, 5, 6) for c in self.colliders):
            self._cc = make_colliders(self.colliders)
        fn = getattr(lib(), "oracle_step" + _sfx(self.dtype))
        fn.restype = C.c_int
        fn(C.byref(self._cp), C.c_int64(self.n), _p(s["x0"]), _p(s["u"]), _p(s["v"]), _p(s["C"]),
           _p(s["F"]), _p(s["tau"]), _p(s["mass"]), _p(s["vol"]), _p(s["yield_stress"]), _p(self.x), _p(self.grid_m),
           _p(self.grid_mv), _p(self.grid_v), self._cc, C.c_int(len(self.colliders)),
           C.c_double(self.time), C.c_double(dt), C.c_int(nthreads))
        for c in self.colliders:
            move_collider(c, self.time, dt)
        self.time = self.time + dt

    # ---- reduced-order sampled substep (mpm_solver_warp.py:207-227) --------
    def _x(self):
        return np.ascontiguousarray(self.s["x0"] + self.s["u"])

    def mark_cells(self, samples, lo=-2, hi=2):
        """update_neighboring_cells (lo, hi = -2, 2) / ..._for_grid_v (-2, 4)."""
        ng = int(self.params["n_grid"])
        flags = np.zeros(ng ** 3, np.int32)
        smp = np.ascontiguousarray(samples, dtype=np.int32)
        getattr(lib(), "oracle_mark_cells" + _sfx(self.dtype))(C.byref(self._cp), _p(self._x()), _p(smp),
                                                                C.c_int64(smp.size), C.c_int(lo), C.c_int(hi),
                                                                _p(flags))
        return flags

    def step_reduced(self, samples, density, dt: float):
        """p2g2p_reduced: zero grids, flag cells, p2g_sample_all, grid update on
        flagged nodes, colliders (moving ones advanced), time += dt."""
        s = self.s
        self.flags = self.mark_cells(samples)
        self.grid_m[:] = 0
        self.grid_mv[:] = 0
        self.grid_v[:] = 0
        dens = np.ascontiguousarray(density, dtype=self.dtype)
        if any(c["kind"] in (2, 3, 5, 6) for c in self.colliders):
            self._cc = make_colliders(self.colliders)
        L = lib()
        getattr(L, "oracle_p2g_sample_all" + _sfx(self.dtype))(
            C.byref(self._cp), C.c_int64(self.n), _p(self._x()), _p(s["v"]), _p(s["F"]), _p(s["mass"]), _p(dens),
            _p(self.flags), _p(self.grid_m), _p(self.grid_mv), C.c_double(dt))
        getattr(L, "oracle_grid_update_sample" + _sfx(self.dtype))(
            C.byref(self._cp), _p(self.flags), _p(self.grid_m), _p(self.grid_mv), _p(self.grid_v), self._cc,
            C.c_int(len(self.colliders)), C.c_double(self.time), C.c_double(dt))
        for c in self.colliders:
            move_collider(c, self.time, dt)
        self.time = self.time + dt

    def g2p_sample(self, samples, dt: float):
        s = self.s
        smp = np.ascontiguousarray(samples, dtype=np.int32)
        getattr(lib(), "oracle_g2p_sample" + _sfx(self.dtype))(
            C.byref(self._cp), _p(self.grid_v), _p(s["x0"]), _p(s["u"]), _p(s["v"]), _p(s["F"]),
            _p(s["yield_stress"]), _p(smp), C.c_int64(smp.size), C.c_double(dt))

    def run(self, n_steps: int, dt: float, nthreads: int = 1):
        for _ in range(n_steps):
            self.step(dt, nthreads)
        return self
