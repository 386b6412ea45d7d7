"""MPM_Simulator_HIP: drop-in for the reference's MPM_Simulator_WARP.

Mirrors the Python API of engine_stress/mpm_solver_warp.py (class at :24) --
method names, keyword arguments, argument meaning and exception types -- and
runs every substep through the HIP library libmpm_hip.so (C-ABI in
include/mpm_hip.h) on PyTorch-ROCm device memory.

Deviations from the reference (each deliberate):
  * state lives in solver-owned device memory ("planes", include/mpm_hip.h);
    loaders copy from the given tensors and exports return fresh tensors in
    the caller's particle order instead of aliasing solver memory
    (mpm_solver_warp.py:229-266, 305-310 alias);
  * F starts as the identity (the reference allocates it as zeros,
    mpm_solver_warp.py:56, and sets it to I only in load_from_sampling :149),
    and load_uvF_from_tensor sets F = F_disp + I (the reference sets F_disp
    only, :245);
  * extensions: set_parameters(E=, nu=, gravity=(gx,gy,gz)) and
    load_vol_from_tensor (the reference has no volume loader);
  * the per-step debug prints, host syncs and the blocking input() on a CFL
    violation (:173-192) run only with debug=True, and the CFL violation then
    raises RuntimeError; a missing sampling file raises FileNotFoundError
    instead of exit() (:130-132);
  * snapshots are written as .npz with the reference's h5 dataset names when
    h5py is not importable;
  * load_from_sampling keeps the file's f_tensor as the exported F_disp until
    the first substep (the reference stores it in particle_F_disp, :146,
    while F starts as I);
  * failures surface: particles that leave the grid are held at their last
    position and counted (the reference indexes past its grid), NaN/Inf in
    the grid sums raise FloatingPointError at the next export or health
    check, and health() returns the device counters;
  * set_rotation_on_disk raises ValueError when |normal_x| <= 0.1 (the
    reference's axis search never ends there, :611-614), and a node on the
    disk's own axis gets v = 0 (the reference divides 0 by 0).
"""
from __future__ import annotations

import math
import os
import types
from ctypes import c_float as C_float

import numpy as np
import torch

from . import _lib as L
from .scenes import alpha_from_friction_angle, lame_from_E_nu

_MATERIALS = {"jelly": 0, "metal": 1, "sand": 2}


def _stream_handle(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _check_float_tensor(t: torch.Tensor, ncols=None):
    """Checks of torch2warp_vec3/mat33 (engine_stress/warp_utils.py:5-43)."""
    assert t.is_contiguous()
    if t.dtype != torch.float32 and t.dtype != torch.int32:
        raise RuntimeError("Error aliasing Torch tensor to Warp array. Torch tensor must be float32 or int32 type")
    if ncols is not None:
        assert t.shape[1] == ncols


_ALL_PLANES = (1 << L.NPLANES) - 1


def _planes(first, n):
    """Bitmask of planes first .. first + n - 1."""
    return ((1 << n) - 1) << first


def _f32v(v):
    """A Warp vec3 (float32) from three numbers."""
    return tuple(float(np.float32(x)) for x in v)


def _move_with_velocity(time, dt, c):
    """modify of the plane, cube and sphere colliders (mpm_solver_warp.py:560-563,
    688-691, 715-717): point += dt * velocity while active, evaluated as
    Warp does it (Python floats, stored back into a float32 vec3)."""
    if time >= c["start_time"] and time < c["end_time"]:
        c["point"] = _f32v([c["point"][d] + dt * c["velocity"][d] for d in range(3)])


def _move_on_circle(time, dt, c):
    """modify of set_column_moving_as_circle (mpm_solver_warp.py:594-603)."""
    if time >= c["start_time"] and time < c["end_time"]:
        p = c["point"]
        z = np.array([p[0] - 0.5, p[1], p[2] - 0.5], np.float32)
        dp = np.float32(z[0] / np.sqrt(np.float32(z @ z)))
        th = float(np.arccos(dp))
        if z[2] < 0:
            th = -th
        vel = (-c["v_scale"] * math.sin(th), 0.0, c["v_scale"] * math.cos(th))
        c["point"] = _f32v([p[d] + dt * vel[d] for d in range(3)])


class MPM_Simulator_HIP:
    """Explicit MLS/APIC-MPM sand simulator on MI355X
    (MPM_Simulator_WARP, engine_stress/mpm_solver_warp.py:24-720; the
    substep path only)."""

    def __init__(self, n_particles, n_grid=100, grid_lim=1.0, device="cuda:0"):
        self.initialize(n_particles, n_grid, grid_lim, device=device)

    # ------------------------------------------------------------------
    # initialize (mpm_solver_warp.py:28-87)
    # ------------------------------------------------------------------
    def initialize(self, n_particles, n_grid=100, grid_lim=1.0, device="cuda:0"):
        self.n_particles = int(n_particles)
        self.device = torch.device(device)
        m = types.SimpleNamespace()
        m.grid_lim = grid_lim
        m.n_grid = int(n_grid)
        m.grid_dim_x = m.grid_dim_y = m.grid_dim_z = m.n_grid
        m.dx, m.inv_dx = m.grid_lim / m.n_grid, float(m.n_grid / m.grid_lim)
        m.E = 1000
        m.nu = 0.3
        m.mu = m.E / 2.0 / (1.0 + m.nu)
        m.lam = m.E * m.nu / (1.0 + m.nu) / (1.0 - 2.0 * m.nu)
        m.yield_stress = 0.0
        m.hardening = 0
        m.xi = 0.0
        m.material = 0
        m.friction_angle = 25.0
        m.alpha = alpha_from_friction_angle(m.friction_angle)
        m.gravitational_accelaration = 0.0
        m.gravity = (0.0, 0.0, 0.0)
        self.mpm_model = m

        n = self.n_particles
        self.planes = torch.zeros((L.NPLANES, max(n, 1)), dtype=torch.float32, device=self.device)
        # F = I (deviation: the reference allocates zeros, :56)
        for k in (0, 4, 8):
            self.planes[L.P_F + k].fill_(1.0)
        self.planes[L.P_PID].copy_(torch.arange(max(n, 1), dtype=torch.int32, device=self.device).view(torch.float32))
        self.particle_density = torch.zeros(max(n, 1), dtype=torch.float32, device=self.device)
        self.particle_color = np.ones(shape=n, dtype=float)
        self.time = 0.0
        self.debug = False
        self.stage_timer = None      # list -> p2g2p records per-stage events into it
        self.colliders = []          # list of dicts, registration order
        # binned path (default): state moves into per-block bins at the first
        # substep and back to the flat caller-order planes when a loader runs
        self.use_bins = True
        # binned path: keep the summed P2G of every substep (the reference's
        # grid_v_in), costing one extra grid write per substep
        self.keep_grid_v_in = False
        self._record = None
        self._bins = None            # dict of tensors + ctypes struct
        self._launch = (0, 0)        # (wg_threads, sort_period) of the binned kernels; 0 = default
        self._loc = "flat"           # where the authoritative state lives
        self._flat_valid = _ALL_PLANES   # bit k: flat plane k mirrors the binned state
        self._fdisp_loaded = None        # F_disp of a sampling file, exported until the first substep
        self._pending_dt = None      # dt of the P2G held in the accumulator
        self._since_check = 0
        self._left_warned = 0        # particles that had left the grid at the last health check
        # reduced-order sampled step (p2g2p_reduced): sample indices and its
        # own workspace (grids + neighboring_cells flag grids)
        self.sample_particle_index_short = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._reduced_ws = None
        self._alloc_grid()

    def _alloc_grid(self):
        prm = self._params()
        nbytes = L.load().mpm_workspace_bytes(prm, self.n_particles)
        if nbytes < 0:
            L.check(-1, "mpm_workspace_bytes")
        self.workspace = torch.zeros(int(nbytes) + 256, dtype=torch.uint8, device=self.device)

    def _ws_ptr(self):
        p = self.workspace.data_ptr()
        return (p + 255) & ~255

    def _params(self) -> L.MpmParams:
        m = self.mpm_model
        prm = L.MpmParams()
        prm.n_grid = int(m.n_grid)
        prm.material = int(m.material)
        prm.dx = float(m.dx)
        prm.inv_dx = float(m.inv_dx)
        prm.mu = float(m.mu)
        prm.lam = float(m.lam)
        prm.alpha = float(m.alpha)
        g = self._gravity_vec()
        for i in range(3):
            prm.gravity[i] = float(g[i])
        prm.hardening = int(m.hardening)
        prm.xi = float(m.xi)
        return prm

    def _gravity_vec(self):
        m = self.mpm_model
        if getattr(m, "gravity_override", None) is not None:
            return m.gravity_override
        return (0.0, -float(m.gravitational_accelaration), 0.0)

    def _state(self) -> L.MpmState:
        st = L.MpmState()
        st.n = self.n_particles
        st.stride = self.planes.shape[1]
        st.planes = self.planes.data_ptr()
        return st

    def _colliders(self):
        n = len(self.colliders)
        if n > L.MAX_COLLIDERS:
            raise ValueError(f"at most {L.MAX_COLLIDERS} colliders are supported")
        arr = (L.MpmCollider * max(n, 1))()
        zero = (0.0, 0.0, 0.0)
        for i, c in enumerate(self.colliders):
            arr[i].kind = c["kind"]
            arr[i].surface_type = c.get("surface_type", 0)
            for d in range(3):
                arr[i].point[d] = float(c.get("point", zero)[d])
                arr[i].normal[d] = float(c.get("normal", zero)[d])
                arr[i].velocity[d] = float(c.get("velocity", zero)[d])
                arr[i].x_unit[d] = float(c.get("x_unit", zero)[d])
                arr[i].y_unit[d] = float(c.get("y_unit", zero)[d])
            arr[i].start_time = float(c["start_time"])
            arr[i].end_time = float(c["end_time"])
            arr[i].friction = float(c.get("friction", 0.0))
            for k in ("threshold", "radius", "height", "width", "v_scale"):
                setattr(arr[i], k, float(c.get(k, 0.0)))
            arr[i].reset = int(c.get("reset", 0))
        return arr, n

    def _moving(self):
        return any(c.get("modify") is not None for c in self.colliders)

    def _apply_modify(self, time, dt):
        """The colliders' host callbacks after a substep at `time` (the time
        before the increment), in registration order (mpm_solver_warp.py:184-185)."""
        for c in self.colliders:
            if c.get("modify") is not None:
                c["modify"](time, dt, c)

    # ------------------------------------------------------------------
    # set_parameters (mpm_solver_warp.py:89-124)
    # ------------------------------------------------------------------
    def set_parameters(self, device="cuda:0", **kwargs):
        self._ensure_flat()
        m = self.mpm_model
        if "material" in kwargs:
            if kwargs["material"] in _MATERIALS:
                m.material = _MATERIALS[kwargs["material"]]
            else:
                raise TypeError("Undefined material type")
        if "grid_lim" in kwargs:
            m.grid_lim = kwargs["grid_lim"]
        if "n_grid" in kwargs:
            m.n_grid = int(kwargs["n_grid"])
        m.grid_dim_x = m.grid_dim_y = m.grid_dim_z = m.n_grid
        m.dx, m.inv_dx = m.grid_lim / m.n_grid, float(m.n_grid / m.grid_lim)
        self._alloc_grid()
        if "mu" in kwargs:
            m.mu = kwargs["mu"]
        if "lam" in kwargs:
            m.lam = kwargs["lam"]
        if "E" in kwargs or "nu" in kwargs:          # extension: Young's modulus / Poisson ratio
            m.E = kwargs.get("E", m.E)
            m.nu = kwargs.get("nu", m.nu)
            m.mu, m.lam = lame_from_E_nu(m.E, m.nu)
        if "yield_stress" in kwargs:
            # set_value_to_float_array over the per-particle array (mpm_solver_warp.py:108-110)
            m.yield_stress = kwargs["yield_stress"]
            self.planes[L.P_YIELD, :self.n_particles].fill_(float(kwargs["yield_stress"]))
        if "hardening" in kwargs:
            m.hardening = kwargs["hardening"]
        if "xi" in kwargs:
            m.xi = kwargs["xi"]
        if "friction_angle" in kwargs:
            m.friction_angle = kwargs["friction_angle"]
            m.alpha = alpha_from_friction_angle(m.friction_angle)
        if "g" in kwargs:
            m.gravitational_accelaration = kwargs["g"]
        if "gravity" in kwargs:                       # extension: full gravity vector
            g = kwargs["gravity"]
            m.gravity_override = None if g is None else tuple(float(a) for a in g)
        if "density" in kwargs:
            n = self.n_particles
            self.particle_density[:n].fill_(float(kwargs["density"]))
            self.planes[L.P_MASS, :n] = self.particle_density[:n] * self.planes[L.P_VOL, :n]

    # ------------------------------------------------------------------
    # colliders (mpm_solver_warp.py:407-451, 499-525)
    # ------------------------------------------------------------------
    def add_surface_collider(self, point, normal, surface="sticky", friction=0.0, start_time=0.0,
                             end_time=999.0):
        point = list(point)
        normal_scale = 1.0 / math.sqrt(sum(x ** 2 for x in normal))
        normal = list(normal_scale * x for x in normal)
        if surface == "sticky" and friction != 0:
            raise ValueError("friction must be 0 on sticky surfaces.")
        stype = 0 if surface == "sticky" else (1 if surface == "slip" else 2)
        self.colliders.append(dict(kind=0, surface_type=stype, point=tuple(point), normal=tuple(normal),
                                   start_time=start_time, end_time=end_time, friction=friction))

    def add_surface_collider2(self, point, normal, surface="sticky", friction=0.0, start_time=0.0,
                              end_time=999.0):
        """Same collider as add_surface_collider (mpm_solver_warp.py:453-497)."""
        self.add_surface_collider(point, normal, surface, friction, start_time, end_time)

    def add_bounding_box(self, start_time=0.0, end_time=999.0):
        self.colliders.append(dict(kind=1, start_time=start_time, end_time=end_time))

    def set_velocity_on_infinite_plane(self, point, normal, velocity, threshold, start_time=0.0, end_time=999.0,
                                       reset=0):
        """Nodes within `threshold` of the plane take `velocity`; the plane
        moves with it (mpm_solver_warp.py:527-566)."""
        s = 1.0 / math.sqrt(sum(x ** 2 for x in normal))
        self.colliders.append(dict(kind=2, point=_f32v(point), normal=_f32v([s * x for x in normal]),
                                   velocity=_f32v(velocity), threshold=float(threshold), start_time=start_time,
                                   end_time=end_time, reset=int(reset), modify=_move_with_velocity))

    def set_column_moving_as_circle(self, R, radius, height, v_scale, start_time=0.0, end_time=999.0):
        """A column of nodes around `point` (initially (0.5, 0.069, 0.5 + R))
        moving on a circle about x = z = 0.5 (mpm_solver_warp.py:568-605)."""
        self.colliders.append(dict(kind=3, point=_f32v((0.5, 0.069, 0.5 + R)), radius=float(radius),
                                   height=float(height), v_scale=float(v_scale), R=float(R), start_time=start_time,
                                   end_time=end_time, modify=_move_on_circle))

    def set_rotation_on_disk(self, point, normal, radius, width, v_scale, start_time=0.0, end_time=999.0, reset=0):
        """Nodes in a disk take a rotation about its normal (mpm_solver_warp.py:607-664).
        Local axes as the reference builds them: x = (-(n_y + n_z)/n_x, 1, 1),
        y = n x x, both normalised.  The reference's axis search never ends
        when |n_x| <= 0.1 (its loop index is not incremented, :611-614); that
        case raises ValueError here."""
        s = 1.0 / math.sqrt(sum(x ** 2 for x in normal))
        n = [s * x for x in normal]
        if not abs(n[0]) > 1e-1:
            raise ValueError("set_rotation_on_disk: the reference's local axes need |normal[0]| > 0.1 "
                             "(its axis search loops forever otherwise)")
        ax = [1.0, 1.0, 1.0]
        ax[0] = -(n[0] + n[1] + n[2] - n[0]) / n[0]
        ay = list(np.cross(np.array(n), np.array(ax)))
        ax = [x / math.sqrt(sum(y ** 2 for y in ax)) for x in ax]
        ay = [x / math.sqrt(sum(y ** 2 for y in ay)) for x in ay]
        self.colliders.append(dict(kind=4, point=_f32v(point), normal=_f32v(n), x_unit=_f32v(ax), y_unit=_f32v(ay),
                                   radius=float(radius), width=float(width), v_scale=float(v_scale),
                                   start_time=start_time, end_time=end_time, reset=int(reset)))

    def set_velocity_on_cube(self, point, width, velocity, start_time=0.0, end_time=999.0, reset=0):
        """Nodes within `width` of `point` in every axis take `velocity`; the
        cube moves with it (mpm_solver_warp.py:666-693)."""
        self.colliders.append(dict(kind=5, point=_f32v(point), width=float(width), velocity=_f32v(velocity),
                                   start_time=start_time, end_time=end_time, reset=int(reset),
                                   modify=_move_with_velocity))

    def add_sphere_leader(self, center, radius, velocity, start_time=0, end_time=999):
        """Nodes within `radius` of the center take `velocity`; the sphere
        moves with it (mpm_solver_warp.py:695-720)."""
        self.colliders.append(dict(kind=6, point=_f32v(center), radius=float(radius), velocity=_f32v(velocity),
                                   start_time=start_time, end_time=end_time, modify=_move_with_velocity))

    # ------------------------------------------------------------------
    # p2g2p (mpm_solver_warp.py:156-203)
    # ------------------------------------------------------------------
    def p2g2p(self, step, dt, print_norm=False, device="cuda:0"):
        self._fdisp_loaded = None
        lib = L.load()
        prm = self._params()
        cols, ncol = self._colliders()
        stream = _stream_handle(self.device)
        t0 = self.time
        if self.use_bins:
            self._bins_run(1, float(dt))          # advances self.time
        else:
            self._ensure_flat()
            st = self._state()
            if self.stage_timer is None:
                ws = self.workspace.numel() - 256
                rc = lib.mpm_step(prm, st, cols, ncol, self._ws_ptr(), ws, float(self.time), float(dt), stream)
                L.check(rc, "mpm_step")
            else:
                self._timed_stages(lib, prm, st, cols, ncol, float(dt), stream)
            self.time = self.time + dt
        self._apply_modify(t0, dt)
        if self.debug:
            v = self.export_uvFstress_from_warp()[1]
            vmax = float(torch.max(torch.abs(v)))
            print("max particle v: ", vmax)
            print("max allowed  v: ", self.mpm_model.dx / dt)
            if vmax > self.mpm_model.dx / dt:
                raise RuntimeError(f"CFL violated: max |v| {vmax} > dx/dt {self.mpm_model.dx / dt}")
        if print_norm:
            u, v, F_disp, stress = self.export_uvFstress_from_warp()
            print("at mpm step", step, end=" ")
            print("norm of u, v, F_disp, stress, respectively: ",
                  "{:.4f}".format(float(torch.linalg.norm(u))), "{:.4f}".format(float(torch.linalg.norm(v))),
                  "{:.4f}".format(float(torch.linalg.norm(F_disp))), "{:.4f}".format(float(torch.linalg.norm(stress))))

    # ------------------------------------------------------------------
    # reduced-order sampled step (mpm_solver_warp.py:207-227, mpm_utils.py:318-471)
    # ------------------------------------------------------------------
    def load_sample_particle_index_from_tensor(self, index, device="cuda:0"):
        """The sample particles of p2g2p_reduced / g2p_sample (the reference's
        mpm_state.sample_particle_index_short, int32 particle indices)."""
        idx = torch.as_tensor(index).reshape(-1).to(device=self.device, dtype=torch.int32).contiguous()
        if idx.numel() and (int(idx.min()) < 0 or int(idx.max()) >= self.n_particles):
            raise ValueError("sample indices must lie in [0, n_particles)")
        self.sample_particle_index_short = idx

    def _reduced_workspace(self):
        lib = L.load()
        nbytes = int(lib.mpm_reduced_workspace_bytes(self._params()))
        if nbytes < 0:
            L.check(-1, "mpm_reduced_workspace_bytes")
        if self._reduced_ws is None or self._reduced_ws.numel() < nbytes + 256:
            self._reduced_ws = torch.zeros(nbytes + 256, dtype=torch.uint8, device=self.device)
        return (self._reduced_ws.data_ptr() + 255) & ~255, nbytes

    def p2g2p_reduced(self, step, dt, device="cuda:0"):
        """p2g2p_reduced (mpm_solver_warp.py:214-227): zero the grids, flag
        the cells around the sample particles, P2G of the particles in flagged
        cells (PIC, F = F_disp + I, stress from F, volume m / density), grid
        normalization on flagged nodes, the colliders, time += dt.  The caller
        then advances the samples with g2p_sample."""
        self._ensure_flat()
        self._fdisp_loaded = None
        lib = L.load()
        ws, nbytes = self._reduced_workspace()
        cols, ncol = self._colliders()
        smp = self.sample_particle_index_short
        t0 = self.time
        L.check(lib.mpm_step_reduced(self._params(), self._state(), smp.data_ptr(), smp.numel(),
                                     self.particle_density.data_ptr(), cols, ncol, ws, nbytes, float(self.time),
                                     float(dt), _stream_handle(self.device)), "mpm_step_reduced")
        self._apply_modify(t0, dt)
        self.time = self.time + dt

    def g2p_sample(self, dt, device="cuda:0"):
        """g2p_sample (mpm_utils.py:423-471) on the sample particles, from the
        grid of the last p2g2p_reduced: v, x, F (return map) and u; C and the
        stress are left as they were."""
        self._ensure_flat()
        if self._reduced_ws is None:
            raise RuntimeError("g2p_sample needs a p2g2p_reduced first")
        self._fdisp_loaded = None
        lib = L.load()
        ws, _ = self._reduced_workspace()
        smp = self.sample_particle_index_short
        L.check(lib.mpm_g2p_sample(self._params(), self._state(), smp.data_ptr(), smp.numel(), ws, float(dt),
                                   _stream_handle(self.device)), "mpm_g2p_sample")

    def experiment_for_C(self, device="cuda:0"):
        """update_neighboring_cells_for_grid_v over the samples; returns the
        flag grid neighboring_cells_grid_v (n, n, n) int32
        (mpm_solver_warp.py:207-209)."""
        self._ensure_flat()
        lib = L.load()
        ws, _ = self._reduced_workspace()
        smp = self.sample_particle_index_short
        L.check(lib.mpm_mark_cells_grid_v(self._params(), self._state(), smp.data_ptr(), smp.numel(), ws,
                                          _stream_handle(self.device)), "mpm_mark_cells_grid_v")
        return self._reduced_views()["cells_v"]

    def _reduced_views(self):
        ng = self.mpm_model.n_grid
        nn = ng ** 3
        ws, _ = self._reduced_workspace()
        off = ws - self._reduced_ws.data_ptr()
        b = self._reduced_ws
        return dict(acc=b[off:off + 16 * nn].view(torch.float32).view(ng, ng, ng, 4),
                    vel=b[off + 16 * nn:off + 32 * nn].view(torch.float32).view(ng, ng, ng, 4),
                    cells=b[off + 32 * nn:off + 36 * nn].view(torch.int32).view(ng, ng, ng),
                    cells_v=b[off + 36 * nn:off + 40 * nn].view(torch.int32).view(ng, ng, ng))

    @property
    def neighboring_cells(self):
        """The flag grid of the last p2g2p_reduced (mpm_state.neighboring_cells)."""
        return self._reduced_views()["cells"]

    def _timed_stages(self, lib, prm, st, cols, ncol, dt, stream):
        """The substep as its stage launches with a torch.cuda.Event between
        stages (same kernels as mpm_step, on the same stream)."""
        ng = self.mpm_model.n_grid
        acc = self._ws_ptr()
        gv = acc + 16 * ng ** 3
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
        ev[0].record()
        L.check(lib.mpm_zero_grid(prm, acc, stream), "mpm_zero_grid")
        ev[1].record()
        L.check(lib.mpm_p2g(prm, st, acc, dt, stream), "mpm_p2g")
        ev[2].record()
        L.check(lib.mpm_grid_update(prm, acc, gv, cols, ncol, float(self.time), dt, stream), "mpm_grid_update")
        ev[3].record()
        L.check(lib.mpm_g2p(prm, st, gv, dt, stream), "mpm_g2p")
        ev[4].record()
        self.stage_timer.append(ev)

    def run(self, n_substeps, dt):
        """n_substeps consecutive substeps (p2g2p without prints) in one
        library call, without host synchronisation (one call per substep
        while a collider moves: its host callback runs between substeps)."""
        if self.use_bins and not self._moving() and self.stage_timer is None and not self.debug:
            self._bins_run(int(n_substeps), float(dt))
        else:
            for k in range(int(n_substeps)):
                self.p2g2p(k, dt)

    # ------------------------------------------------------------------
    # binned path
    # ------------------------------------------------------------------
    def _flat_state(self) -> L.MpmState:
        return self._state()

    def _alloc_bins(self):
        """Bins of 8^3 cells sized from the current positions: capacity twice
        the fullest bin (at least 256 slots), overflow run n/16."""
        ng = self.mpm_model.n_grid
        nb = (ng + 7) // 8
        n = self.n_particles
        x = self.planes[L.P_X0:L.P_X0 + 3, :n] + self.planes[L.P_U:L.P_U + 3, :n]
        b = torch.clamp((x * float(self.mpm_model.inv_dx) - 0.5).to(torch.int64), 0, ng - 3) >> 3
        idx = (b[0] * nb + b[1]) * nb + b[2]
        fullest = int(torch.bincount(idx, minlength=nb ** 3).max()) if n > 0 else 0
        cap = max(256, ((2 * fullest + 63) // 64) * 64)
        ovf = cap * max(1, -(-max(4096, n // 16) // cap))      # whole bins
        stride = (nb ** 3 * cap + ovf) * L.NPLANES              # bin-major plane sets
        old = self._bins
        if old is not None and old["nb"] == nb and old["cap"] == cap and old["ovf"] == ovf:
            return old
        planes = torch.zeros((2, stride), dtype=torch.float32, device=self.device)
        count = torch.zeros((2, 2 * nb ** 3 + 1), dtype=torch.int32, device=self.device)
        st = L.MpmBins()
        st.nb, st.bin_cap, st.overflow_cap, st.cur, st.epoch, st.stride = nb, cap, ovf, 0, 0, stride
        st.wg_threads, st.sort_period = self._launch
        st.planes[0] = planes[0].data_ptr()
        st.planes[1] = planes[1].data_ptr()
        st.count = count.data_ptr()
        prm = self._params()
        nbytes = L.load().mpm_bins_workspace_bytes(prm)
        ws = torch.zeros(int(nbytes) + 256, dtype=torch.uint8, device=self.device)
        self._bins = dict(nb=nb, cap=cap, ovf=ovf, stride=stride, planes=planes, count=count, struct=st, ws=ws,
                          ng=ng)
        return self._bins

    def set_launch(self, wg_threads: int = 0, sort_period: int = 0):
        """Launch parameters of the binned kernels: threads per bin workgroup
        (256, 384, 512 or 768; 0 = 768) and substeps between bank-order sorts (0 = 16,
        < 0 = never).  They change speed only (see tune.py)."""
        if wg_threads not in (0, 256, 384, 512, 768):
            raise ValueError("wg_threads must be 0, 256, 384, 512 or 768")
        self._launch = (int(wg_threads), int(sort_period))
        if self._bins is not None:
            self._bins["struct"].wg_threads, self._bins["struct"].sort_period = self._launch

    def _bins_ws_ptr(self):
        p = self._bins["ws"].data_ptr()
        return (p + 255) & ~255

    def _bins_status(self):
        """Device status counters (host sync)."""
        off = self._bins_ws_ptr() - self._bins["ws"].data_ptr()
        start = off + int(L.load().mpm_bins_status_offset(self._params()))
        return self._bins["ws"][start:start + 256].view(torch.int32).cpu()

    def health(self) -> dict:
        """The binned path's device counters (a host sync): particles dropped
        because bins and overflow were full, particles whose scatter left
        their bin's tile (the spill path), particles that left the grid (held
        at their last position: the reference indexes past the grid there),
        and grid nodes whose P2G sum was NaN or Inf -- these accumulate over
        the run -- and the particles in the overflow run now (bins past their
        capacity: which particles land there depends on the order of the
        atomics, so two runs agree bit for bit only while it stays empty).
        Zeros before the first binned substep."""
        keys = ("dropped", "outside_tile", "left_grid", "nonfinite_nodes")
        if self._bins is None:
            return dict.fromkeys(keys + ("overflow",), 0)
        st = self._bins_status()
        out = {k: int(st[i]) for i, k in enumerate(keys)}
        B = self._bins
        out["overflow"] = min(int(B["count"][B["struct"].cur, B["nb"] ** 3]), B["ovf"])
        return out

    def _check_bins_health(self):
        """Run every 256 substeps and at every export: raises on dropped
        particles and on NaN/Inf in the grid (FloatingPointError), warns once
        per increase of the particles that left the grid."""
        st = self._bins_status()
        if int(st[0]) > 0:
            raise RuntimeError(f"{int(st[0])} particles were dropped: bins and overflow run full")
        if int(st[3]) > 0:
            raise FloatingPointError(f"{int(st[3])} grid node sums were NaN or Inf: a particle's state is not "
                                     "finite (check dt against dx / max|v|)")
        left = int(st[2])
        if left > self._left_warned:
            import warnings
            warnings.warn(f"{left} particles left the grid and are held at their last position", RuntimeWarning,
                          stacklevel=3)
        self._left_warned = left
        cnt = self._bins["count"][self._bins["struct"].cur, self._bins["nb"] ** 3].item()
        if cnt > self._bins["ovf"] // 2:
            # re-bin with larger bins at the next substep
            self._ensure_flat()
        self._since_check = 0

    def _bins_run(self, n_substeps, dt):
        lib = L.load()
        prm = self._params()
        stream = _stream_handle(self.device)
        if self._loc == "flat":
            B = self._alloc_bins()
            ws = self._bins_ws_ptr()
            L.check(lib.mpm_bins_clear_grid(prm, B["struct"], ws, stream), "mpm_bins_clear_grid")
            L.check(lib.mpm_bins_insert(prm, self._flat_state(), B["struct"], ws, stream), "mpm_bins_insert")
            L.check(lib.mpm_bins_p2g(prm, B["struct"], ws, dt, stream), "mpm_bins_p2g")
            self._pending_dt = dt
            self._loc = "bins"
        B = self._bins
        ws = self._bins_ws_ptr()
        if self._pending_dt != dt:
            L.check(lib.mpm_bins_clear_grid(prm, B["struct"], ws, stream), "mpm_bins_clear_grid")
            L.check(lib.mpm_bins_p2g(prm, B["struct"], ws, dt, stream), "mpm_bins_p2g")
            self._pending_dt = dt
        cols, ncol = self._colliders()
        if self.keep_grid_v_in:
            ng = self.mpm_model.n_grid
            if self._record is None or self._record.shape[0] != ng:
                self._record = torch.zeros((ng, ng, ng, 4), dtype=torch.float32, device=self.device)
            B["struct"].record_acc = self._record.data_ptr()
        else:
            B["struct"].record_acc = None
            self._record = None
        if self.stage_timer is None:
            L.check(lib.mpm_bins_run(prm, B["struct"], cols, ncol, ws, float(self.time), dt, int(n_substeps),
                                     stream), "mpm_bins_run")
            for _ in range(int(n_substeps)):
                self.time = self.time + dt
        else:
            # the two kernels of each substep with a torch.cuda.Event between them
            for _ in range(int(n_substeps)):
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
                ev[0].record()
                L.check(lib.mpm_bins_grid_update(prm, B["struct"], cols, ncol, ws, float(self.time), dt, stream),
                        "mpm_bins_grid_update")
                ev[1].record()
                L.check(lib.mpm_bins_g2p2g(prm, B["struct"], ws, dt, stream), "mpm_bins_g2p2g")
                ev[2].record()
                self.stage_timer.append(ev)
                self.time = self.time + dt
        self._flat_valid = 0
        self._fdisp_loaded = None
        self._since_check += int(n_substeps)
        if self._since_check >= 256:
            self._check_bins_health()

    def profile_kernels(self, n_substeps, dt):
        """Runs n_substeps on the binned path with HIP events around each
        kernel (waits after every substep); returns the average per substep
        (grid_update_ms, g2p2g_ms, sort_ms)."""
        lib = L.load()
        self._bins_run(0, float(dt))       # bins the state and computes the pending P2G
        B = self._bins
        cols, ncol = self._colliders()
        gm, km, sm = C_float(), C_float(), C_float()
        L.check(lib.mpm_bins_profile(self._params(), B["struct"], cols, ncol, self._bins_ws_ptr(), float(self.time),
                                     float(dt), int(n_substeps), _stream_handle(self.device), gm, km, sm),
                "mpm_bins_profile")
        for _ in range(int(n_substeps)):
            self.time = self.time + float(dt)
        self._flat_valid = 0
        self._fdisp_loaded = None
        return gm.value, km.value, sm.value

    def _sync_flat(self, planes=_ALL_PLANES):
        """Make the flat planes named by the bitmask `planes` mirror the binned
        state (no change of owner); only the planes not yet copied move."""
        miss = planes & ~self._flat_valid
        if self._loc == "bins" and miss:
            lib = L.load()
            L.check(lib.mpm_bins_extract_planes(self._params(), self._bins["struct"], self._flat_state(), miss,
                                                _stream_handle(self.device)), "mpm_bins_extract_planes")
            self._flat_valid |= miss
            self._check_bins_health()

    def _ensure_flat(self):
        """Move the authoritative state to the flat planes (before a loader,
        a parameter change or a step on the stage path)."""
        self._sync_flat()
        self._loc = "flat"
        self._pending_dt = None

    # ------------------------------------------------------------------
    # torch loaders (mpm_solver_warp.py:229-266) -- copies, not aliases
    # ------------------------------------------------------------------
    def _import(self, plane, ncomp, t, mode=0):
        self._ensure_flat()
        if t.numel() != self.n_particles * ncomp:
            raise ValueError(f"expected {self.n_particles} x {ncomp} values, got shape {tuple(t.shape)}")
        t = t.to(device=self.device, dtype=torch.float32).contiguous()
        lib = L.load()
        L.check(lib.mpm_import(self._state(), plane, ncomp, t.data_ptr(), mode, _stream_handle(self.device)),
                "mpm_import")

    def load_x_initial_from_tensor(self, tensor_x_iniital, device="cuda:0"):
        _check_float_tensor(tensor_x_iniital, 3)
        self._import(L.P_X0, 3, tensor_x_iniital)

    def load_mass_from_tensor(self, tensor_mass, device="cuda:0"):
        _check_float_tensor(tensor_mass)
        self._import(L.P_MASS, 1, tensor_mass.reshape(-1).contiguous())

    def load_vol_from_tensor(self, tensor_vol, device="cuda:0"):
        """Extension: the reference has no volume loader."""
        _check_float_tensor(tensor_vol)
        self._import(L.P_VOL, 1, tensor_vol.reshape(-1).contiguous())

    def load_uvF_from_tensor(self, tensor_u, tensor_v, tensor_F, clone=True, device="cuda:0"):
        tensor_F = torch.reshape(tensor_F, (-1, 3, 3)).contiguous()
        _check_float_tensor(tensor_u, 3)
        _check_float_tensor(tensor_v, 3)
        _check_float_tensor(tensor_F, 3)
        self._import(L.P_U, 3, tensor_u)
        self._import(L.P_V, 3, tensor_v)
        self._import(L.P_F, 9, tensor_F, mode=2)       # F = F_disp + I
        self._fdisp_loaded = None

    def load_C_from_tensor(self, tensor_C, clone=True, device="cuda:0"):
        tensor_C = torch.reshape(tensor_C, (-1, 3, 3)).contiguous()
        _check_float_tensor(tensor_C, 3)
        self._import(L.P_C, 9, tensor_C)

    def load_uvstress_from_tensor(self, tensor_u, tensor_v, tensor_stress, clone=True, device="cuda:0"):
        tensor_stress = torch.reshape(tensor_stress, (-1, 3, 3)).contiguous()
        _check_float_tensor(tensor_u, 3)
        _check_float_tensor(tensor_v, 3)
        _check_float_tensor(tensor_stress, 3)
        self._import(L.P_U, 3, tensor_u)
        self._import(L.P_V, 3, tensor_v)
        self._import(L.P_TAU, 9, tensor_stress)

    def load_velocity_from_tensor(self, tensor_v, device="cuda:0"):
        _check_float_tensor(tensor_v, 3)
        self._import(L.P_V, 3, tensor_v)

    def load_F_from_tensor(self, tensor_F, device="cuda:0"):
        """Sets the elastic deformation gradient F itself ((N,9) or (N,3,3))."""
        tensor_F = torch.reshape(tensor_F, (-1, 3, 3)).contiguous()
        _check_float_tensor(tensor_F, 3)
        self._import(L.P_F, 9, tensor_F)
        self._fdisp_loaded = None

    # ------------------------------------------------------------------
    # export (mpm_solver_warp.py:305-310)
    # ------------------------------------------------------------------
    def _export(self, plane, ncomp, mode=0):
        need = _planes(L.P_X0, 3) | _planes(L.P_U, 3) if mode == 3 else _planes(plane, ncomp)
        self._sync_flat(need | _planes(L.P_PID, 1))
        out = torch.empty((self.n_particles, ncomp), dtype=torch.float32, device=self.device)
        lib = L.load()
        L.check(lib.mpm_export(self._state(), plane, ncomp, out.data_ptr(), mode, _stream_handle(self.device)),
                "mpm_export")
        return out

    def export_uvFstress_from_warp(self, device="cuda:0"):
        """(u (N,3), v (N,3), F_disp (N,3,3), stress (N,3,3)) in the caller's
        particle order, as fresh tensors."""
        self._sync_flat(_planes(L.P_U, 3) | _planes(L.P_V, 3) | _planes(L.P_F, 9) | _planes(L.P_TAU, 9) |
                        _planes(L.P_PID, 1))
        u = self._export(L.P_U, 3)
        v = self._export(L.P_V, 3)
        if self._fdisp_loaded is not None:
            F_disp = self._fdisp_loaded.clone()
        else:
            F_disp = self._export(L.P_F, 9, mode=1).view(-1, 3, 3)
        stress = self._export(L.P_TAU, 9).view(-1, 3, 3)
        return u, v, F_disp, stress

    export_uvFstress = export_uvFstress_from_warp

    def export_x(self):
        return self._export(L.P_X0, 3, mode=3)

    def export_x_initial(self):
        return self._export(L.P_X0, 3)

    def export_C(self):
        return self._export(L.P_C, 9).view(-1, 3, 3)

    def export_F(self):
        return self._export(L.P_F, 9).view(-1, 3, 3)

    def export_mass(self):
        return self._export(L.P_MASS, 1).view(-1)

    def export_yield_stress(self):
        """The per-particle yield stress (metal; mpm_model.yield_stress of
        the reference, warp_utils.py:75)."""
        return self._export(L.P_YIELD, 1).view(-1)

    def load_yield_stress_from_tensor(self, tensor_ys, device="cuda:0"):
        """Extension: per-particle yield stress (the reference fills one value
        with set_parameters(yield_stress=))."""
        _check_float_tensor(tensor_ys)
        self._import(L.P_YIELD, 1, tensor_ys.reshape(-1).contiguous())

    def export_vol(self):
        return self._export(L.P_VOL, 1).view(-1)

    # grid views (reference: mpm_state.grid_m / grid_v_in / grid_v_out)
    def _grid_views(self):
        """(accumulator, velocity) grids, (ng, ng, ng, 4) each.  On the
        binned path the P2G lives in per-bin tiles: the accumulator view is
        the record of the last substep's summed P2G (keep_grid_v_in = True),
        or None."""
        ng = self.mpm_model.n_grid
        nn = ng ** 3
        if self._loc == "bins" and self._bins is not None:
            ws = self._bins["ws"]
            off = self._bins_ws_ptr() - ws.data_ptr()
            lib = L.load()
            # blocks the particles have left still hold their last velocities
            L.check(lib.mpm_bins_settle_grid(self._params(), self._bins["struct"], self._bins_ws_ptr(),
                                             _stream_handle(self.device)), "mpm_bins_settle_grid")
            vel_off = int(lib.mpm_bins_velocity_offset(self._params()))
            vel = ws[off + vel_off:off + vel_off + nn * 16].view(torch.float32).view(ng, ng, ng, 4)
            return self._record, vel
        else:
            ws = self.workspace
            off = self._ws_ptr() - ws.data_ptr()
            vel_off = nn * 16
        acc = ws[off:off + nn * 16].view(torch.float32).view(ng, ng, ng, 4)
        vel = ws[off + vel_off:off + vel_off + nn * 16].view(torch.float32).view(ng, ng, ng, 4)
        return acc, vel

    @property
    def grid_m(self):
        """Node masses of the last substep's grid update."""
        return self._grid_views()[1][..., 3]

    @property
    def grid_v_in(self):
        """The last substep's P2G momentum (grid_v_in).  On the binned path it
        is recorded only with keep_grid_v_in = True set before stepping."""
        acc = self._grid_views()[0]
        if acc is None:
            raise RuntimeError("grid_v_in on the binned path needs keep_grid_v_in = True before the substep "
                               "(or use_bins = False)")
        return acc[..., :3]

    @property
    def grid_v_out(self):
        return self._grid_views()[1][..., :3]

    # ------------------------------------------------------------------
    # samplings and snapshots (mpm_solver_warp.py:129-153, 312-404)
    # ------------------------------------------------------------------
    def load_from_sampling(self, sampling_h5, n_grid=100, grid_lim=1.0, device="cuda:0"):
        """load_from_sampling (mpm_solver_warp.py:129-153): positions, masses,
        volumes and densities of a sampling file (path = os.getcwd() + name,
        as the reference builds it); the solver is re-initialised for its
        particle count, v = 0, F = I.  The file's f_tensor is the reference's
        F_disp until the first substep (:148): exports and a frame saved
        before stepping show it, and the first substep starts from F = I as
        the reference's does."""
        from . import snapshot
        d = snapshot.read_sampling(os.getcwd() + sampling_h5)   # FileNotFoundError if absent
        x, q = d["x"].T, d["q"].T
        self.dim, n = q.shape[1], q.shape[0]
        self.initialize(n, n_grid, grid_lim, device=device)
        print("Sampling particles are loaded from h5 file. Simulator is re-initialized for the correct n_particles")
        dev = self.device
        f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)
        self.load_x_initial_from_tensor(f32(x))
        self.load_mass_from_tensor(f32(np.squeeze(d["masses"], 0)))
        self.load_vol_from_tensor(f32(np.squeeze(d["particle_volume"], 0)))
        self.particle_density[:n] = f32(np.squeeze(d["particle_density"], 0))
        if "f_tensor" in d:
            self._fdisp_loaded = f32(snapshot.f_tensor_rows(d["f_tensor"], n)).view(n, 3, 3)
        print("Particle position, F, mass, and density are loaded from initial sampling.")

    def load_uvFstress_from_h5(self, h5filename, sample_point=0, device="cuda:0"):
        """Restart loader (mpm_solver_warp.py:268-303): u, v and the stress of a
        saved frame (path = os.getcwd() + name).  The stress comes back from
        its Voigt rows [xx, xy, xz, yy, yz, zz] as a symmetric 3x3; the
        frame's f_tensor is read but not loaded, as in the reference (:301)."""
        from . import snapshot
        path = os.getcwd() + h5filename
        d = snapshot.read_frame(snapshot.resolve_frame_path(path))
        q = np.ascontiguousarray(d["q"].T, dtype=np.float32)
        v = np.ascontiguousarray(d["v"].T, dtype=np.float32)
        stress_in = snapshot.stress_from_voigt(d["stress"].T)
        dev = self.device
        f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)
        self.load_uvstress_from_tensor(f32(q), f32(v), f32(stress_in.reshape(-1, 9)))
        print("stress in:", np.linalg.norm(stress_in))

    def save_data_at_frame(self, dir_name, frame, include_parameter="friction_angle", save_to_h5=True,
                           save_to_obj=True, save_visual_F=False):
        from . import snapshot
        snapshot.save_frame(self, dir_name, frame, include_parameter=include_parameter, save_to_h5=save_to_h5,
                            save_to_obj=save_to_obj, save_visual_F=save_visual_F)


# drop-in name of the reference class
MPM_Simulator_WARP = MPM_Simulator_HIP
