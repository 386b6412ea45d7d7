"""Seeded synthetic sand scenes and material-parameter helpers.

The reference loads its initial particles from h5 samplings that are not
shipped (generate_training_data/generate_sand.py:18 reads
``initial_data/.../sand_column_large170_initial_sampling.h5``), so scenes
here are generated with ``numpy.random.default_rng(seed)`` in the layout the
solver consumes: x0 (N,3), vol (N,), v (N,3), all float32.

Material parameters follow MPM_Simulator_WARP:
  E, nu -> mu, lam        engine_stress/mpm_solver_warp.py:39-42
  friction angle -> alpha engine_stress/mpm_solver_warp.py:113-116
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def lame_from_E_nu(E: float, nu: float):
    """mu = E/(2(1+nu)), lam = E nu/((1+nu)(1-2nu)) (mpm_solver_warp.py:41-42)."""
    mu = E / 2.0 / (1.0 + nu)
    lam = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu)
    return mu, lam


def alpha_from_friction_angle(phi_deg: float) -> float:
    """alpha = sqrt(2/3) * 2 sin(phi) / (3 - sin(phi)), phi in degrees with the
    reference's pi literal 3.14159265 (mpm_solver_warp.py:115-116)."""
    s = math.sin(phi_deg / 180.0 * 3.14159265)
    return math.sqrt(2.0 / 3.0) * 2.0 * s / (3.0 - s)


@dataclass
class SandScene:
    """A sand column: particles, material and boundary description."""
    name: str
    n_grid: int
    x0: np.ndarray
    vol: np.ndarray
    v0: np.ndarray
    density: float = 2200.0
    E: float = 2.0e5
    nu: float = 0.3
    friction_angle: float = 30.0
    gravity: tuple = (0.0, 0.0, -9.8)
    dt: float = 1.0e-4
    substeps: int = 100
    floor_z: float = 0.05
    floor_start: float = 0.0      # the sticky floor acts from this time on
    grid_lim: float = 1.0
    extra: dict = field(default_factory=dict)

    @property
    def n_particles(self) -> int:
        return int(self.x0.shape[0])

    @property
    def mass(self) -> np.ndarray:
        return (self.density * self.vol).astype(np.float32)


def column_lattice(n_grid=32, lo=(0.35, 0.35, 0.05), hi=(0.65, 0.65, 0.55), per_axis=2,
                   seed=0, name="column_lattice") -> SandScene:
    """Jittered sub-lattice seeding: per_axis^3 particles per grid cell inside
    the box [lo, hi]; each particle jittered by U(-0.5,0.5)*dx/per_axis.
    The 32^3 CPU-runnable column: 8 particles/cell in
    [0.35,0.65]^2 x [0.05,0.55]."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / n_grid
    h = dx / per_axis
    axes = [np.arange(lo[d] + 0.5 * h, hi[d] - 1e-12, h) for d in range(3)]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
    g = g + rng.uniform(-0.5, 0.5, size=g.shape) * h
    n = g.shape[0]
    vol = np.full(n, h ** 3)
    return SandScene(name=name, n_grid=n_grid, x0=g.astype(np.float32), vol=vol.astype(np.float32),
                     v0=np.zeros((n, 3), np.float32))


def column_uniform(n_particles=1_000_000, n_grid=128, lo=(0.3, 0.3, 0.05), hi=(0.7, 0.7, 0.6),
                   v_std=0.1, seed=0, name="column_uniform", substeps=500) -> SandScene:
    """i.i.d. uniform particles in the box [lo, hi] (seed), velocities
    N(0, v_std) per component (seed+1), vol = box volume / N.  The headline
    workload: 1M particles on a 128^3 grid, 500 substeps."""
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo); hi = np.asarray(hi)
    x0 = lo + rng.uniform(size=(n_particles, 3)) * (hi - lo)
    rng_v = np.random.default_rng(seed + 1)
    v0 = rng_v.normal(0.0, v_std, size=(n_particles, 3)) if v_std > 0 else np.zeros((n_particles, 3))
    vol = np.full(n_particles, float(np.prod(hi - lo)) / n_particles)
    return SandScene(name=name, n_grid=n_grid, x0=x0.astype(np.float32), vol=vol.astype(np.float32),
                     v0=v0.astype(np.float32), substeps=substeps)


def scene_params(scene: SandScene) -> dict:
    """Model scalars in the oracle/solver layout (float64 values)."""
    mu, lam = lame_from_E_nu(scene.E, scene.nu)
    return dict(n_grid=scene.n_grid, dx=scene.grid_lim / scene.n_grid,
                inv_dx=float(scene.n_grid / scene.grid_lim), mu=mu, lam=lam,
                alpha=alpha_from_friction_angle(scene.friction_angle),
                gravity=tuple(scene.gravity), material=2)


def scene_colliders(scene: SandScene) -> list:
    """Sticky floor at z = floor_z (normal +z) then the bounding box, the
    order generate_sand.py:24-26 registers them in."""
    return [dict(kind=0, surface_type=0, point=(0.0, 0.0, scene.floor_z), normal=(0.0, 0.0, 1.0),
                 start_time=float(scene.floor_start), end_time=999.0),
            dict(kind=1, start_time=0.0, end_time=999.0)]


def scene_state(scene: SandScene, dtype=np.float32) -> dict:
    """Initial particle state: u = 0, C = 0, F = I, tau = 0."""
    n = scene.n_particles
    F = np.tile(np.eye(3), (n, 1, 1))
    return dict(x0=scene.x0.astype(dtype), u=np.zeros((n, 3), dtype), v=scene.v0.astype(dtype),
                C=np.zeros((n, 3, 3), dtype), F=F.astype(dtype), tau=np.zeros((n, 3, 3), dtype),
                mass=scene.mass.astype(dtype), vol=scene.vol.astype(dtype))


def make_simulator(scene: SandScene, device="cuda:0"):
    """Build an MPM_Simulator_HIP for a scene through the reference API:
    initialize, torch loaders, set_parameters, colliders (generate_sand.py:17-26
    call order)."""
    import torch
    from .mpm_solver import MPM_Simulator_HIP
    sim = MPM_Simulator_HIP(scene.n_particles, n_grid=scene.n_grid, grid_lim=scene.grid_lim, device=device)
    dev = sim.device
    sim.load_x_initial_from_tensor(torch.from_numpy(np.ascontiguousarray(scene.x0, dtype=np.float32)).to(dev))
    sim.load_vol_from_tensor(torch.from_numpy(np.ascontiguousarray(scene.vol, dtype=np.float32)).to(dev))
    sim.load_velocity_from_tensor(torch.from_numpy(np.ascontiguousarray(scene.v0, dtype=np.float32)).to(dev))
    sim.set_parameters(material="sand", E=scene.E, nu=scene.nu, friction_angle=scene.friction_angle,
                       gravity=scene.gravity, density=scene.density)
    for c in scene_colliders(scene):
        if c["kind"] == 0:
            sim.add_surface_collider(c["point"], c["normal"], "sticky", 0.0, c["start_time"], c["end_time"])
        else:
            sim.add_bounding_box(c["start_time"], c["end_time"])
    return sim
