"""Differentiable sand rollouts on MI355X: the forward substeps (mpm_step_ordered)
and their hand-written adjoints (mpm_adjoint_substep, include/mpm_hip.h)
behind a torch.autograd.Function, with the state checkpointed every K
substeps and the substeps between checkpoints recomputed on the way back --
or, where device memory holds them, every substep's state and grids kept by
the forward pass so that the way back only runs the adjoint kernels.

The reference never differentiates its solver (it records no wp.Tape,
engine_stress/mpm_solver_warp.py); the parameters follow its conventions:
E, nu -> mu, lam (mpm_solver_warp.py:41-42) and the friction angle in degrees
-> alpha with pi = 3.14159265 (mpm_solver_warp.py:113-116).

    x_final = sand_rollout(v0, E, nu, phi, scene=..., n_substeps=..., dt=...)

is differentiable in v0 (the initial particle velocities, (N, 3)) and in the
three material scalars.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib as L
from . import scenes as S

_PI = 3.14159265


def material(E: float, nu: float, phi: float):
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    s = math.sin(phi / 180.0 * _PI)
    alpha = math.sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s)
    return mu, lam, alpha


def material_jacobian(E: float, nu: float, phi: float):
    """d(mu, lam, alpha) / d(E, nu, phi) as a 3x3 list of rows."""
    dmu_dE = 1.0 / (2.0 * (1.0 + nu))
    dmu_dnu = -E / (2.0 * (1.0 + nu) ** 2)
    dlam_dE = nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    dlam_dnu = E * (1.0 + 2.0 * nu * nu) / ((1.0 + nu) ** 2 * (1.0 - 2.0 * nu) ** 2)
    s = math.sin(phi / 180.0 * _PI)
    dalpha_ds = math.sqrt(2.0 / 3.0) * 2.0 * 3.0 / (3.0 - s) ** 2
    dalpha_dphi = dalpha_ds * math.cos(phi / 180.0 * _PI) * _PI / 180.0
    return [[dmu_dE, dmu_dnu, 0.0], [dlam_dE, dlam_dnu, 0.0], [0.0, 0.0, dalpha_dphi]]


class Rollout:
    """n_substeps of a scene on the flat state with checkpoints every K.

    At every checkpoint the state is counting-sorted by bin
    (mpm_sort_by_bin), and the substeps up to the next checkpoint run on
    that slot order (mpm_step_ordered), so that their P2G and the adjoint's
    grid scatter accumulate per chunk of particles in LDS.  The adjoint of a
    sort is its inverse permutation (mpm_permute), so gradients come back in
    the caller's particle order.

    keep: "checkpoints" keeps the sorted state at each checkpoint and the
    backward pass replays each segment (mpm_adjoint_segment); "all" keeps
    every substep's state and grids from the forward pass itself
    (mpm_record_segment, then mpm_adjoint_recorded on the way back: no
    replay); "auto" takes "all" when that record needs at most half of the
    free device memory (1M particles on 128^3 for 200 substeps: ~45 GB)."""

    KEEP = ("auto", "all", "checkpoints")

    def __init__(self, scene: S.SandScene, n_substeps: int, dt: float, checkpoint_every: int = 20,
                 device="cuda:0", keep: str = "auto"):
        self.scene, self.n_substeps, self.dt, self.K = scene, int(n_substeps), float(dt), int(checkpoint_every)
        if self.K < 1:
            raise ValueError("checkpoint_every must be >= 1")
        if keep not in self.KEEP:
            raise ValueError(f"keep must be one of {self.KEEP}")
        self.keep = keep
        self.device = torch.device(device)
        self.n = scene.n_particles
        lib = L.load()
        cols = S.scene_colliders(scene)
        self._cols = (L.MpmCollider * max(1, len(cols)))()
        for i, c in enumerate(cols):
            self._cols[i].kind, self._cols[i].surface_type = c["kind"], c.get("surface_type", 0)
            for d in range(3):
                self._cols[i].point[d] = float(c.get("point", (0.0, 0.0, 0.0))[d])
                self._cols[i].normal[d] = float(c.get("normal", (0.0, 0.0, 0.0))[d])
            self._cols[i].start_time, self._cols[i].end_time = float(c["start_time"]), float(c["end_time"])
        self._ncol = len(cols)
        P = S.scene_params(scene)
        self._prm = L.MpmParams()
        self._prm.n_grid, self._prm.material = int(P["n_grid"]), 2
        self._prm.dx, self._prm.inv_dx = float(P["dx"]), float(P["inv_dx"])
        for d in range(3):
            self._prm.gravity[d] = float(P["gravity"][d])
        self._ws_step_bytes = int(lib.mpm_workspace_bytes(self._prm, self.n))
        self._ws_step = torch.zeros(self._ws_step_bytes + 256, dtype=torch.uint8, device=self.device)
        # grid sets (accumulator, velocities, block marks) of the recorded
        # and replayed segments: zeroed once, kept consistent by the library
        self._set_bytes = int(lib.mpm_step_grids_bytes(self._prm))
        self._ws_run_bytes = int(lib.mpm_run_ordered_workspace_bytes(self._prm))
        self._ws_run = torch.zeros(self._ws_run_bytes + 256, dtype=torch.uint8, device=self.device)
        self._ws_adj = torch.zeros(int(lib.mpm_adjoint_workspace_bytes(self._prm)) + 256, dtype=torch.uint8,
                                   device=self.device)
        self._order_bytes = int(lib.mpm_order_bytes(self._prm, self.n))
        self._perm_off = int(lib.mpm_order_perm_offset(self._prm, self.n))
        # initial state in the caller's order (v filled per call)
        st = S.scene_state(scene, np.float32)
        planes = torch.zeros((L.NPLANES, self.n), dtype=torch.float32, device=self.device)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.device)
        planes[L.P_X0:L.P_X0 + 3] = t(st["x0"].T)
        planes[L.P_F:L.P_F + 9] = t(st["F"].reshape(self.n, 9).T)
        planes[L.P_MASS] = t(st["mass"])
        planes[L.P_VOL] = t(st["vol"])
        planes[L.P_PID] = torch.arange(self.n, dtype=torch.int32, device=self.device).view(torch.float32)
        self._init = planes
        self._init_sorted = None   # (planes, order) of the initial state's sort, made once
        self._checkpoints = {}
        self._final = None
        self._ws_seg = []
        self._seg_states = []
        # the record of keep="all": per substep k a state and a grid
        # workspace, per segment an order buffer (reused across calls)
        self._tape_states, self._tape_grids, self._tape_orders = [], [], []
        self._tape_last = None
        self._recorded = False

    def record_bytes(self) -> int:
        """Device bytes of the keep="all" record (states and grids of every substep)."""
        return self.n_substeps * (L.NPLANES * self.n * 4 + self._set_bytes + 256)

    def _records(self) -> bool:
        if self.keep != "auto":
            return self.keep == "all"
        free, _ = torch.cuda.mem_get_info(self.device)
        held = sum(t.numel() * t.element_size() for t in self._tape_states + self._tape_grids)
        return self.record_bytes() <= 0.5 * (free + held)

    @staticmethod
    def _aligned(buf: torch.Tensor) -> int:
        return (buf.data_ptr() + 255) & ~255

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _state(self, planes: torch.Tensor) -> L.MpmState:
        st = L.MpmState()
        st.n, st.stride, st.planes = self.n, planes.shape[1], planes.data_ptr()
        return st

    def _set_material(self, E, nu, phi):
        self.mu, self.lam, self.alpha = material(E, nu, phi)
        self._prm.mu, self._prm.lam, self._prm.alpha = self.mu, self.lam, self.alpha

    def _sort(self, planes, dst=None, order=None):
        """(planes sorted by bin, their order buffer)."""
        dst = torch.empty_like(planes) if dst is None else dst
        order = torch.empty(self._order_bytes + 256, dtype=torch.uint8, device=self.device) if order is None else order
        L.check(L.load().mpm_sort_by_bin(self._prm, self._state(planes), self._state(dst), self._aligned(order),
                                         self._stream()), "mpm_sort_by_bin")
        return dst, order

    def _sorted_start(self, v0):
        """The initial state in the sort's slot order with v0 in place, and the
        sort's order buffer.  The sort keys are positions, which no call
        changes, so the caller-order initial state is sorted once; each call
        gathers v0 into the sorted slots by the pid plane."""
        if self._init_sorted is None:
            self._init_sorted = self._sort(self._init)
        sp, so = self._init_sorted
        planes = sp.clone()
        v = v0.to(self.device, torch.float32).t().contiguous()
        planes[L.P_V:L.P_V + 3] = self._permute(v, 3, sp[L.P_PID].data_ptr(), scatter=False)
        return planes, so

    def _copy_order(self, dst, src):
        """An order buffer's contents (from its 256-byte aligned start) into another."""
        a, b = self._aligned(src) - src.data_ptr(), self._aligned(dst) - dst.data_ptr()
        dst[b:b + self._order_bytes].copy_(src[a:a + self._order_bytes])

    def _step(self, planes, k, order, ws=None):
        ws = self._ws_step if ws is None else ws
        L.check(L.load().mpm_step_ordered(self._prm, self._state(planes), self._aligned(order), self._cols,
                                          self._ncol, self._aligned(ws), self._ws_step_bytes,
                                          float(k * self.dt), self.dt, self._stream()), "mpm_step_ordered")

    def _permute(self, src, n_planes, idx_ptr, scatter):
        # idx is a permutation of the n slots (a sort's or the pid plane), so
        # every slot of dst is written
        dst = torch.empty((n_planes, self.n), dtype=torch.float32, device=self.device)
        L.check(L.load().mpm_permute(src.data_ptr(), src.shape[1], dst.data_ptr(), self.n, n_planes, idx_ptr, self.n,
                                     int(scatter), self._stream()), "mpm_permute")
        return dst

    def forward(self, v0: torch.Tensor, E: float, nu: float, phi: float) -> torch.Tensor:
        """Positions x = x0 + u after n_substeps, in the caller's particle
        order; keeps the checkpoints (state and slot order)."""
        self._set_material(E, nu, phi)
        self._material_args = (float(E), float(nu), float(phi))
        start, order0 = self._sorted_start(v0)
        self._checkpoints = {}
        lib = L.load()
        self._recorded = self._records()
        if self._recorded:
            return self._forward_recorded(start, order0)
        self._release_record()
        planes = start
        for c0 in range(0, self.n_substeps, self.K):
            if c0 == 0:
                # a copy: the substeps count their failures into the order buffer
                order = torch.empty_like(order0)
                self._copy_order(order, order0)
            else:
                planes, order = self._sort(planes)
            self._checkpoints[c0] = (planes.clone(), order)
            L.check(lib.mpm_run_ordered(self._prm, self._state(planes), self._aligned(order), self._cols,
                                        self._ncol, self._aligned(self._ws_run), self._ws_run_bytes, c0, self.dt,
                                        min(self.K, self.n_substeps - c0), self._stream()), "mpm_run_ordered")
        return self._finish(planes)

    def _finish(self, planes):
        self._final = planes
        # x = x0 + u back to the caller's order (scatter by the pid plane)
        x = planes[L.P_X0:L.P_X0 + 3] + planes[L.P_U:L.P_U + 3]
        return self._permute(x, 3, planes[L.P_PID].data_ptr(), scatter=True).t().contiguous()

    def _release_record(self):
        self._tape_states, self._tape_grids, self._tape_orders = [], [], []
        self._tape_last = None

    def _forward_recorded(self, planes, order0):
        """forward() keeping every substep: per segment the sort into the
        record's state of its first substep (the first segment's state and
        order: the sorted start, copied), then mpm_record_segment."""
        lib = L.load()
        n_seg = (self.n_substeps + self.K - 1) // self.K
        while len(self._tape_states) < self.n_substeps:
            self._tape_states.append(torch.empty((L.NPLANES, self.n), dtype=torch.float32, device=self.device))
        while len(self._tape_grids) < self.n_substeps:
            self._tape_grids.append(torch.zeros(self._set_bytes + 256, dtype=torch.uint8, device=self.device))
        while len(self._tape_orders) < n_seg:
            self._tape_orders.append(torch.empty(self._order_bytes + 256, dtype=torch.uint8, device=self.device))
        if self._tape_last is None:
            self._tape_last = torch.empty((L.NPLANES, self.n), dtype=torch.float32, device=self.device)
        last = self._tape_last
        for s_i, c0 in enumerate(range(0, self.n_substeps, self.K)):
            nk = min(self.K, self.n_substeps - c0)
            if s_i == 0:
                self._tape_states[0].copy_(planes)
                self._copy_order(self._tape_orders[0], order0)
            else:
                self._sort(planes, dst=self._tape_states[c0], order=self._tape_orders[s_i])
            states, grids = self._segment_args(c0, nk)
            # the last substep's P2G of the next state goes straight into the
            # next segment's first grid (the sort does not change the grid)
            nxt = self._tape_grids[c0 + nk] if c0 + nk < self.n_substeps else self._ws_run
            L.check(lib.mpm_record_segment(self._prm, states, self._aligned(self._tape_orders[s_i]), grids, self._cols,
                                           self._ncol, c0, self.dt, nk, int(s_i > 0), self._state(last),
                                           self._aligned(nxt), nxt.numel() - 256, self._stream()),
                    "mpm_record_segment")
            planes = last
        return self._finish(planes)

    def _segment_args(self, c0, nk):
        states = (L.MpmState * nk)(*[self._state(self._tape_states[c0 + j]) for j in range(nk)])
        grids = (ctypes.c_void_p * nk)(*[self._aligned(self._tape_grids[c0 + j]) for j in range(nk)])
        return states, grids

    def backward(self, x_bar: torch.Tensor):
        """Adjoint of forward() for the position adjoint x_bar (N, 3) in the
        caller's order: returns (v0_bar (N, 3), [E_bar, nu_bar, phi_bar])."""
        lib = L.load()
        stream = self._stream()
        # x_bar to the final slot order (gather by the pid plane)
        xb = x_bar.to(self.device, torch.float32).t().contiguous()
        adj = torch.zeros((L.ADJ_NPLANES, self.n), dtype=torch.float32, device=self.device)
        if self.n_substeps == 0:
            return adj[L.ADJ_V:L.ADJ_V + 3].t().contiguous(), [0.0, 0.0, 0.0]
        adj[L.ADJ_X:L.ADJ_X + 3] = self._permute(xb, 3, self._final[L.P_PID].data_ptr(), scatter=False)
        nxt = torch.empty_like(adj)
        pb = torch.zeros(3, dtype=torch.float64, device=self.device)
        if self._recorded:
            for s_i in range((self.n_substeps + self.K - 1) // self.K - 1, -1, -1):
                c0 = s_i * self.K
                nk = min(self.K, self.n_substeps - c0)
                states, grids = self._segment_args(c0, nk)
                order = self._tape_orders[s_i]
                L.check(lib.mpm_adjoint_recorded(self._prm, states, self._aligned(order), grids, self._cols,
                                                 self._ncol, self._aligned(self._ws_adj), c0, self.dt, nk,
                                                 adj.data_ptr(), nxt.data_ptr(), self.n, pb.data_ptr(), stream),
                        "mpm_adjoint_recorded")
                adj = self._unsort(adj, order, c0)
            return self._param_grads(adj, pb)
        # a segment's replayed states and, per replayed substep, the grids it
        # leaves for its adjoint (one native call per segment)
        while len(self._seg_states) < self.K - 1:
            self._seg_states.append(torch.empty((L.NPLANES, self.n), dtype=torch.float32, device=self.device))
        while len(self._ws_seg) < self.K:
            self._ws_seg.append(torch.zeros(self._set_bytes + 256, dtype=torch.uint8, device=self.device))
        for c0 in sorted(self._checkpoints, reverse=True):
            nk = min(self.K, self.n_substeps - c0)
            cp, order = self._checkpoints[c0]
            states = (L.MpmState * nk)()
            states[0] = self._state(cp)
            for j in range(1, nk):
                states[j] = self._state(self._seg_states[j - 1])
            grids = (ctypes.c_void_p * nk)(*[self._aligned(w) for w in self._ws_seg[:nk]])
            L.check(lib.mpm_adjoint_segment(self._prm, states, self._aligned(order), grids, self._cols, self._ncol,
                                            self._aligned(self._ws_adj), c0, self.dt, nk, adj.data_ptr(),
                                            nxt.data_ptr(), self.n, pb.data_ptr(), stream), "mpm_adjoint_segment")
            # adjoint of the sort at c0: its inverse permutation
            adj = self._unsort(adj, order, c0)
        return self._param_grads(adj, pb)

    def _unsort(self, adj, order, c0):
        """Adjoint of the sort at substep c0: its inverse permutation.  At
        c0 = 0 only the initial-velocity planes are needed (in the caller's
        order); the others are dropped."""
        if c0 > 0:
            return self._permute(adj, L.ADJ_NPLANES, self._aligned(order) + self._perm_off, scatter=True)
        return self._permute(adj[L.ADJ_V:L.ADJ_V + 3], 3, self._aligned(order) + self._perm_off, scatter=True)

    def _param_grads(self, v_planes, pb):
        v0_bar = v_planes.t().contiguous()
        mu_b, lam_b, al_b = (float(v) for v in pb.cpu())
        J = material_jacobian(*self._material_args)
        E_b = mu_b * J[0][0] + lam_b * J[1][0]
        nu_b = mu_b * J[0][1] + lam_b * J[1][1]
        phi_b = al_b * J[2][2]
        return v0_bar, [E_b, nu_b, phi_b]


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, params, rollout):
        E, nu, phi = (float(p) for p in params.detach().cpu())
        ctx.rollout = rollout
        return rollout.forward(v0.detach(), E, nu, phi)

    @staticmethod
    def backward(ctx, x_bar):
        v0_bar, pb = ctx.rollout.backward(x_bar)
        return v0_bar, torch.tensor(pb, dtype=torch.float64), None


def sand_rollout(v0: torch.Tensor, params: torch.Tensor, rollout: Rollout) -> torch.Tensor:
    """Differentiable x_final(v0, params = [E, nu, friction angle in deg])."""
    return _RolloutFn.apply(v0, params, rollout)
