"""GPU parity of the grid colliders: every collider kind of the reference
(surface, bounding box, velocity plane, circling column, rotating disk,
velocity cube, sphere leader: engine_stress/mpm_solver_warp.py:407-720)
through mpm_grid_update against the oracle's grid update, and a binned run
with moving colliders (their host callbacks between substeps) against the
FP64 oracle.

Tolerances: the grid update is elementwise FP32 arithmetic, so nodes agree
with the FP32 oracle to 1e-6 relative (acos/sin/cos of the circle and disk
kinds included, whose FP32 results may differ by an ulp); the run is held to
the column tolerance of tests/test_gpu_parity.py.
"""
import ctypes as Ct

import numpy as np
import pytest
import torch

from conftest import rel_l2
from reduced_mpm_stress_training_amd import _lib as L
from reduced_mpm_stress_training_amd import scenes as S
from reduced_mpm_stress_training_amd.mpm_solver import MPM_Simulator_HIP

pytestmark = pytest.mark.gpu


def _all_kinds(sim):
    sim.add_surface_collider((0.0, 0.0, 0.1), (0.0, 0.0, 1.0))
    sim.set_velocity_on_infinite_plane((0.5, 0.3, 0.5), (1.0, 2.0, 2.0), (0.3, -0.1, 0.2), 0.06, 0.0, 1.0)
    sim.set_column_moving_as_circle(0.2, 0.12, 0.3, 0.4)
    sim.set_rotation_on_disk((0.5, 0.53, 0.47), (1.0, 0.2, 0.1), 0.3, 0.08, 2.0)
    sim.set_velocity_on_cube((0.7, 0.6, 0.4), 0.08, (1.0, 0.0, -1.0))
    sim.add_sphere_leader((0.3, 0.7, 0.6), 0.1, (0.0, 2.0, 0.0))
    sim.add_bounding_box()


def _oracle_cols(sim):
    return [{k: v for k, v in c.items() if k != "modify"} for c in sim.colliders]


@pytest.mark.parametrize("time", [0.0, 1.0 + 5e-4])
def test_grid_update_every_collider_kind(oracle, time):
    ng = 32
    sim = MPM_Simulator_HIP(1, n_grid=ng)
    _all_kinds(sim)
    for c in sim.colliders:
        c["reset"] = 1
    rng = np.random.default_rng(3)
    nn = ng ** 3
    m = rng.uniform(0.5, 2.0, nn).astype(np.float32)
    m[rng.uniform(size=nn) < 0.2] = 0.0                        # empty nodes: colliders still act on them
    v = rng.normal(0, 1.0, (nn, 3)).astype(np.float32)
    acc = np.concatenate([v * m[:, None], m[:, None]], 1).astype(np.float32)
    dev = torch.device("cuda:0")
    acc_t = torch.from_numpy(acc).to(dev)
    out_t = torch.zeros_like(acc_t)
    cols, ncol = sim._colliders()
    dt = 1e-4
    L.check(L.load().mpm_grid_update(sim._params(), acc_t.data_ptr(), out_t.data_ptr(), cols, ncol, time, dt,
                                     torch.cuda.current_stream().cuda_stream), "mpm_grid_update")
    torch.cuda.synchronize()
    got = out_t.cpu().numpy()[:, :3]
    P = dict(n_grid=ng, dx=1.0 / ng, inv_dx=float(ng), mu=1.0, lam=1.0, alpha=0.0, gravity=(0.0, 0.0, 0.0))
    want = oracle.grid_update(P, acc[:, 3].copy(), np.ascontiguousarray(acc[:, :3]), _oracle_cols(sim), time, dt)
    err = np.abs(got - want).max(axis=1) / np.maximum(np.abs(want).max(axis=1), 1.0)
    print(time, err.max(), (np.abs(want - v * (m[:, None] > 0)).max(axis=1) > 0).sum())
    assert err.max() < 1e-6
    # every kind touched some nodes
    assert (np.abs(got) > 0).any()


def test_moving_colliders_run(oracle):
    """The 32^3 column hit by a sphere leader and a velocity plane moving
    through it, 40 substeps on the binned path (one library call per
    substep, the colliders moved by their host callbacks in between)."""
    sc = S.column_lattice()
    sim = S.make_simulator(sc)
    sim.add_sphere_leader((0.2, 0.5, 0.3), 0.08, (6.0, 0.0, 0.0))
    sim.set_velocity_on_infinite_plane((0.5, 0.36, 0.5), (0.0, 1.0, 0.0), (0.0, 1.5, 0.0), 0.03, 0.0, 0.003)
    n = 40
    sim.run(n, sc.dt)
    u, v, Fd, tau = sim.export_uvFstress_from_warp()
    g = dict(u=u.cpu().numpy(), v=v.cpu().numpy(), F=(Fd + torch.eye(3, device=Fd.device)).cpu().numpy(),
             tau=tau.cpu().numpy())
    cols = _oracle_cols(sim)
    # the oracle starts from the registered (unmoved) collider positions
    cols[2]["point"] = (0.2, 0.5, 0.3)
    cols[3]["point"] = (0.5, 0.36, 0.5)
    assert abs(sim.colliders[2]["point"][0] - (0.2 + 6.0 * sc.dt * n)) < 1e-4

    def run(dtype, order=None):
        st = S.scene_state(sc, np.float64)
        if order is not None:
            st = {k: a[order] for k, a in st.items()}
        o = oracle.OracleScene(S.scene_params(sc), st, cols, dtype).run(n, sc.dt)
        out = {k: o.s[k].reshape(sc.n_particles, -1) for k in ("u", "v", "F", "tau")}
        if order is not None:
            inv = np.argsort(order)
            out = {k: a[inv] for k, a in out.items()}
        return out, o

    o64, oo = run(np.float64)
    o32, _ = run(np.float32)
    o32r, _ = run(np.float32, np.arange(sc.n_particles)[::-1].copy())
    assert np.allclose(oo.colliders[2]["point"], sim.colliders[2]["point"], atol=1e-7)
    x0 = sc.x0.astype(np.float64)
    g["x"], o64["x"], o32["x"], o32r["x"] = (x0 + a["u"] for a in (g, o64, o32, o32r))
    # the obstacles moved particles
    assert np.abs(o64["v"]).max() > 1.0
    report = {}
    for k in ("x", "u", "v", "F", "tau"):
        gk = g[k].reshape(sc.n_particles, -1)
        report[k] = (rel_l2(gk, o64[k]), rel_l2(o32[k], o64[k]), rel_l2(o32r[k], o32[k]))
    print(report)
    for k, (e, e32, s32) in report.items():
        assert np.isfinite(g[k]).all() and e <= max(1e-5, 3 * e32 + s32), (k, e, e32, s32)
