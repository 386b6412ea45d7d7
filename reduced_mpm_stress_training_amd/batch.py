"""Many independent sand scenes advanced together (mpm_batch_* in
include/mpm_hip.h).

The reference steps one scene per solver instance and process
(generate_training_data/generate_sand.py:17-32); generating training data in
bulk means many such scenes.  SceneBatch puts the scenes' bins, grids and
particles one after the other so that each kernel of a substep -- grid
update, fused G2P2G, the periodic bank-order sort -- is one launch for every
scene: small scenes share the GPU instead of each leaving most CUs idle, and
a 300-substep run is 600 launches instead of 600 per scene.  Each scene keeps
its own grid size, material parameters, colliders and bin capacity; all
advance with the same dt.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L
from . import scenes as S

OVF_BLOCK = 4096      # slots per block of the batch's overflow run


def _bin_capacity(scene: S.SandScene) -> int:
    """Twice the fullest bin of the initial positions, at least 256 slots."""
    ng = scene.n_grid
    nb = (ng + 7) // 8
    inv_dx = scene.n_grid / scene.grid_lim
    b = np.clip((scene.x0.astype(np.float32) * np.float32(inv_dx) - np.float32(0.5)).astype(np.int64), 0, ng - 3) >> 3
    idx = (b[:, 0] * nb + b[:, 1]) * nb + b[:, 2]
    fullest = int(np.bincount(idx, minlength=nb ** 3).max()) if len(idx) else 0
    return max(256, ((2 * fullest + 63) // 64) * 64)


def _params(scene: S.SandScene) -> L.MpmParams:
    P = S.scene_params(scene)
    prm = L.MpmParams()
    prm.n_grid, prm.material = int(P["n_grid"]), int(P.get("material", 2))
    prm.dx, prm.inv_dx = float(P["dx"]), float(P["inv_dx"])
    prm.mu, prm.lam, prm.alpha = float(P["mu"]), float(P["lam"]), float(P["alpha"])
    for d in range(3):
        prm.gravity[d] = float(P["gravity"][d])
    return prm


class SceneBatch:
    def __init__(self, scenes: list, device="cuda:0"):
        if not scenes:
            raise ValueError("SceneBatch needs at least one scene")
        lib = L.load()
        self.device = torch.device(device)
        self.scenes = list(scenes)
        n = [sc.n_particles for sc in self.scenes]
        self.offsets = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
        self.n_particles = int(self.offsets[-1])
        cols = []
        descs = (L.MpmScene * len(self.scenes))()
        for i, sc in enumerate(self.scenes):
            descs[i].params = _params(sc)
            descs[i].n_particles = sc.n_particles
            cl = S.scene_colliders(sc)
            descs[i].n_colliders = len(cl)
            descs[i].bin_cap = _bin_capacity(sc)
            for c in cl:
                mc = L.MpmCollider()
                mc.kind, mc.surface_type = c["kind"], c.get("surface_type", 0)
                for d in range(3):
                    mc.point[d] = float(c.get("point", (0.0, 0.0, 0.0))[d])
                    mc.normal[d] = float(c.get("normal", (0.0, 0.0, 0.0))[d])
                mc.start_time, mc.end_time, mc.friction = float(c["start_time"]), float(c["end_time"]), 0.0
                cols.append(mc)
        self._descs = descs
        col_arr = (L.MpmCollider * max(1, len(cols)))(*cols)
        nbytes = int(lib.mpm_batch_workspace_bytes(descs, len(self.scenes)))
        if nbytes < 0:
            raise L.MpmError("mpm_batch_workspace_bytes: invalid scenes")
        self._ws = torch.zeros(nbytes + 256, dtype=torch.uint8, device=self.device)
        self._ws_ptr = (self._ws.data_ptr() + 255) & ~255
        self.batch = L.MpmBatch()
        L.check(lib.mpm_batch_setup(descs, len(self.scenes), col_arr, self._ws_ptr, self.batch), "mpm_batch_setup")
        # the batch's flat state: every scene's particles, PID plane = flat index
        N = self.n_particles
        planes = torch.zeros((L.NPLANES, max(N, 1)), dtype=torch.float32, device=self.device)
        x0 = np.concatenate([sc.x0 for sc in self.scenes]).astype(np.float32)
        v0 = np.concatenate([sc.v0 for sc in self.scenes]).astype(np.float32)
        vol = np.concatenate([sc.vol for sc in self.scenes]).astype(np.float32)
        mass = np.concatenate([sc.mass for sc in self.scenes]).astype(np.float32)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.device)
        planes[L.P_X0:L.P_X0 + 3, :N] = t(x0.T)
        planes[L.P_V:L.P_V + 3, :N] = t(v0.T)
        for d in range(3):
            planes[L.P_F + 4 * d, :N] = 1.0
        planes[L.P_MASS, :N] = t(mass)
        planes[L.P_VOL, :N] = t(vol)
        planes[L.P_PID, :N] = torch.arange(N, dtype=torch.int32, device=self.device).view(torch.float32)
        self.planes = planes
        # bins: per-scene capacities (in the scene table), one overflow run
        ovf = OVF_BLOCK * max(1, -(-max(4096, N // 16) // OVF_BLOCK))
        stride = (int(self.batch.n_slots) + ovf) * L.NPLANES
        self._planes2 = torch.empty((2, stride), dtype=torch.float32, device=self.device)
        self._count = torch.zeros((2, 2 * int(self.batch.n_bins) + 1), dtype=torch.int32, device=self.device)
        b = L.MpmBins()
        b.nb, b.bin_cap, b.overflow_cap, b.cur, b.epoch, b.since_sort, b.stride = 0, OVF_BLOCK, ovf, 0, 0, 0, stride
        b.wg_threads, b.sort_period = 0, 0
        b.planes[0] = self._planes2[0].data_ptr()
        b.planes[1] = self._planes2[1].data_ptr()
        b.count = self._count.data_ptr()
        self.bins = b
        self.time = 0.0
        self._pending_dt = None
        self._graph = None
        stream = self._stream()
        L.check(lib.mpm_batch_insert(self.batch, self._flat(), self.bins, stream), "mpm_batch_insert")
        L.check(lib.mpm_batch_clear_grid(self.batch, self.bins, stream), "mpm_batch_clear_grid")

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def set_launch(self, wg_threads: int = 0, sort_period: int = 0):
        """Threads per bin workgroup (256, 384, 512 or 768; 0 = 768) and substeps
        between bank-order sorts (0 = 16, < 0 = never); speed only."""
        if wg_threads not in (0, 256, 384, 512, 768):
            raise ValueError("wg_threads must be 0, 256, 384, 512 or 768")
        self.bins.wg_threads, self.bins.sort_period = int(wg_threads), int(sort_period)

    def _flat(self) -> L.MpmState:
        st = L.MpmState()
        st.n, st.stride, st.planes = self.n_particles, self.planes.shape[1], self.planes.data_ptr()
        return st

    def status(self):
        off = int(L.load().mpm_batch_status_offset(self.batch))
        return self._ws[self._ws_ptr - self._ws.data_ptr() + off:][:256].view(torch.int32).cpu()

    def run(self, n_substeps: int, dt: float):
        """n_substeps substeps of every scene (p2g2p, mpm_solver_warp.py:156-203)."""
        lib = L.load()
        stream = self._stream()
        if self._pending_dt != dt:
            if self._pending_dt is not None:
                L.check(lib.mpm_batch_clear_grid(self.batch, self.bins, stream), "mpm_batch_clear_grid")
            L.check(lib.mpm_batch_p2g(self.batch, self.bins, float(dt), stream), "mpm_batch_p2g")
            self._pending_dt = dt
        L.check(lib.mpm_batch_run(self.batch, self.bins, float(self.time), float(dt), int(n_substeps), stream),
                "mpm_batch_run")
        for _ in range(int(n_substeps)):
            self.time = self.time + dt

    def capture(self, dt: float, n_substeps: int | None = None):
        """Captures a run of n_substeps (default: twice the sort period) as
        a HIP graph (mpm_batch_graph_create); run_captured replays it."""
        lib = L.load()
        if n_substeps is None:
            period = self.bins.sort_period if self.bins.sort_period != 0 else 16
            n_substeps = 2 * period if period > 0 else 64
        self.release_graph()
        stream = self._stream()
        if self._pending_dt != dt:
            if self._pending_dt is not None:
                L.check(lib.mpm_batch_clear_grid(self.batch, self.bins, stream), "mpm_batch_clear_grid")
            L.check(lib.mpm_batch_p2g(self.batch, self.bins, float(dt), stream), "mpm_batch_p2g")
            self._pending_dt = dt
        g = ctypes.c_void_p()
        L.check(lib.mpm_batch_graph_create(self.batch, self.bins, float(dt), int(n_substeps), ctypes.byref(g)),
                "mpm_batch_graph_create")
        self._graph, self._graph_len, self._graph_dt = g, int(n_substeps), dt
        return self._graph_len

    def run_captured(self, n_replays: int):
        """n_replays replays of the captured run (n_replays * its length
        substeps); eager run() calls may come between."""
        if self._graph is None:
            raise RuntimeError("capture() first")
        lib = L.load()
        L.check(lib.mpm_batch_graph_launch(self._graph, self.batch, self.bins, float(self.time), int(n_replays),
                                           self._stream()), "mpm_batch_graph_launch")
        for _ in range(int(n_replays) * self._graph_len):
            self.time = self.time + self._graph_dt

    def release_graph(self):
        if getattr(self, "_graph", None) is not None:
            L.load().mpm_graph_destroy(self._graph)
        self._graph = None

    def __del__(self):
        try:
            self.release_graph()
        except Exception:
            pass

    def export(self, i: int):
        """(u, v, F_disp, stress, C) of scene i, in its particle order."""
        lib = L.load()
        L.check(lib.mpm_batch_extract(self.batch, self.bins, self._flat(), self._stream()), "mpm_batch_extract")
        lo, hi = int(self.offsets[i]), int(self.offsets[i + 1])
        p = self.planes[:, lo:hi]
        n = hi - lo
        u = p[L.P_U:L.P_U + 3].t().contiguous()
        v = p[L.P_V:L.P_V + 3].t().contiguous()
        F = p[L.P_F:L.P_F + 9].t().contiguous().view(n, 3, 3)
        tau = p[L.P_TAU:L.P_TAU + 9].t().contiguous().view(n, 3, 3)
        C = p[L.P_C:L.P_C + 9].t().contiguous().view(n, 3, 3)
        return u, v, F - torch.eye(3, device=self.device), tau, C


def uneven_scenes(n_scenes=64, seed=2, n_min=2_000, n_max=200_000, grid_min=32, grid_max=96, ppc=6.0):
    """Independent sand columns of uneven size: particle counts log-uniform
    in [n_min, n_max], column dimensions U[0.2, 0.5] per axis (base at the
    floor, centred), friction angle U[20, 40] deg, E U[1e5, 5e5] Pa, nu 0.3,
    particles i.i.d. uniform in the column; each scene's grid (clamped to
    [grid_min, grid_max]) gives about `ppc` particles per cell."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_scenes):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        dims = rng.uniform(0.2, 0.5, size=3)
        vol_box = float(np.prod(dims))
        ng = int(np.clip(round((n / ppc / vol_box) ** (1.0 / 3.0)), grid_min, grid_max))
        lo = np.array([0.5 - dims[0] / 2, 0.5 - dims[1] / 2, 0.05])
        hi = lo + dims
        x0 = lo + rng.uniform(size=(n, 3)) * dims
        sc = S.SandScene(name=f"batch{k}", n_grid=ng, x0=x0.astype(np.float32),
                         vol=np.full(n, vol_box / n, np.float32), v0=np.zeros((n, 3), np.float32),
                         E=float(rng.uniform(1e5, 5e5)), friction_angle=float(rng.uniform(20.0, 40.0)),
                         substeps=300)
        out.append(sc)
    return out
