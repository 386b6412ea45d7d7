"""Pins of the oracle's grid colliders (oracle/mpm_oracle.c oracle_grid_update)
by closed forms: the reference's velocity colliders and moving obstacles
(engine_stress/mpm_solver_warp.py:527-720), their time windows and reset
branches, registration order, and the host callbacks that move them.
"""
import numpy as np

NG = 16
DX = 1.0 / NG
P = dict(n_grid=NG, dx=DX, inv_dx=float(NG), mu=1.0, lam=1.0, alpha=0.0, gravity=(0.0, 0.0, 0.0))
IDX = np.indices((NG, NG, NG)).transpose(1, 2, 3, 0).reshape(-1, 3)
POS = IDX * DX


def _grid(seed=0):
    rng = np.random.default_rng(seed)
    gm = rng.uniform(0.5, 2.0, NG ** 3)
    v = rng.normal(0, 1.0, (NG ** 3, 3))
    return gm, v * gm[:, None], v


def _run(oracle, cols, time=0.0, dt=1e-3):
    gm, gmv, v = _grid()
    return oracle.grid_update(P, gm, gmv, cols, time, dt), v


def test_velocity_plane_and_reset(oracle):
    n = np.array([1.0, 2.0, 2.0]) / 3.0
    c = dict(kind=2, point=(0.5, 0.5, 0.5), normal=tuple(n), velocity=(0.3, -0.1, 0.2), threshold=0.1,
             start_time=0.0, end_time=0.5, reset=1)
    out, v = _run(oracle, [c])
    slab = np.abs((POS - 0.5) @ n) <= 0.1
    assert slab.any() and (~slab).any()
    assert np.allclose(out[slab], [0.3, -0.1, 0.2]) and np.allclose(out[~slab], v[~slab])
    # inactive with reset: the slab is zeroed until end + 15 dt, then left alone
    out, v = _run(oracle, [c], time=0.5 + 10 * 1e-3)
    assert np.all(out[slab] == 0.0) and np.allclose(out[~slab], v[~slab])
    out, v = _run(oracle, [c], time=0.5 + 16 * 1e-3)
    assert np.allclose(out, v)
    # inactive without reset: untouched
    out, v = _run(oracle, [dict(c, reset=0)], time=0.6)
    assert np.allclose(out, v)


def test_column_on_circle(oracle):
    c = dict(kind=3, point=(0.5, 0.069, 0.7), radius=0.15, height=0.3, v_scale=0.4, start_time=0.0, end_time=1.0)
    out, v = _run(oracle, [c])
    off = POS - np.array(c["point"])
    inside = (off[:, 0] ** 2 + off[:, 2] ** 2 < 0.15 ** 2) & (np.abs(off[:, 1]) < 0.3)
    Z = np.array([0.0, 0.069, 0.2])
    th = np.arccos(Z[0] / np.linalg.norm(Z))         # Z.z > 0: no sign flip
    assert inside.any()
    assert np.allclose(out[inside], [-0.4 * np.sin(th), 0.0, 0.4 * np.cos(th)])
    assert np.allclose(out[~inside], v[~inside])


def test_rotating_disk_is_a_rotation_field(oracle):
    """normal (1,0,0): local x = (0,1,1)/sqrt2, y = n x x; the disk's nodes
    get v_scale (-(r.y) x + (r.x) y) with r = (0, dy, dz): tangential, of
    length v_scale |r|."""
    n = np.array([1.0, 0.0, 0.0])
    ax = np.array([0.0, 1.0, 1.0]) / np.sqrt(2.0)
    ay = np.cross(n, ax)
    c = dict(kind=4, point=(0.5, 0.53, 0.47), normal=tuple(n), x_unit=tuple(ax), y_unit=tuple(ay), radius=0.3,
             width=0.1, v_scale=2.0, start_time=0.0, end_time=1.0, reset=1)
    out, v = _run(oracle, [c])
    off = POS - np.array(c["point"])
    r = off.copy()
    r[:, 0] = 0.0
    rl = np.linalg.norm(r, axis=1)
    disk = (np.abs(off @ n) < 0.1) & (rl < 0.3)
    want = 2.0 * (-(r @ ay)[:, None] * ax + (r @ ax)[:, None] * ay)
    assert disk.sum() > 20
    assert np.abs(out[disk] - want[disk]).max() < 1e-12
    assert np.abs(np.sum(out[disk] * r[disk], 1)).max() < 1e-12
    assert np.allclose(np.linalg.norm(out[disk], axis=1), 2.0 * rl[disk])
    assert np.allclose(out[~disk], v[~disk])
    # inactive with reset: every node 0 until end + 1.5 dt
    out, _ = _run(oracle, [c], time=1.0 + 1e-3)
    assert np.all(out == 0.0)
    out, v = _run(oracle, [c], time=1.0 + 2e-3)
    assert np.allclose(out, v)


def test_cube_sphere_and_registration_order(oracle):
    cube = dict(kind=5, point=(0.4, 0.5, 0.5), width=0.1, velocity=(1.0, 0.0, -1.0), start_time=0.0, end_time=1.0)
    sph = dict(kind=6, point=(0.5, 0.5, 0.5), radius=0.15, velocity=(0.0, 2.0, 0.0), start_time=0.0, end_time=1.0)
    in_cube = np.all(np.abs(POS - np.array(cube["point"])) < 0.1, axis=1)
    in_sph = np.linalg.norm(POS - np.array(sph["point"]), axis=1) < 0.15
    both = in_cube & in_sph
    assert both.any() and (in_cube & ~in_sph).any()
    out, v = _run(oracle, [cube, sph])
    assert np.allclose(out[in_sph], [0, 2, 0]) and np.allclose(out[in_cube & ~in_sph], [1, 0, -1])
    out, v = _run(oracle, [sph, cube])
    assert np.allclose(out[in_cube], [1, 0, -1]) and np.allclose(out[in_sph & ~in_cube], [0, 2, 0])
    rest = ~(in_cube | in_sph)
    assert np.allclose(out[rest], v[rest])
    # the cube's reset zeroes every node within 1.5 dt after its window
    out, _ = _run(oracle, [dict(cube, reset=1)], time=1.0 + 1e-3)
    assert np.all(out == 0.0)


def test_host_callbacks_move_the_colliders(oracle):
    """move_collider: kinds 2, 5, 6 advance by dt * velocity while active (in
    float32, like a Warp vec3); the circle keeps its distance to the axis."""
    c = dict(kind=6, point=(0.3, 0.4, 0.5), radius=0.1, velocity=(0.5, -0.25, 0.125), start_time=0.0, end_time=0.05)
    t, dt, n_act = 0.0, 1e-3, 0
    for _ in range(80):
        oracle.move_collider(c, t, dt)
        n_act += t < 0.05          # active at the substeps whose (accumulated) time is in the window
        t += dt
    assert abs(c["point"][0] - (0.3 + 0.5 * dt * n_act)) < 1e-5
    assert abs(c["point"][1] - (0.4 - 0.25 * dt * n_act)) < 1e-5
    # the circle: cos t = Z.x / |Z| with |Z| including Z.y (a reference quirk:
    # the step is not quite tangential), integrated here in float64
    circ = dict(kind=3, point=(0.5, 0.069, 0.7), v_scale=0.4, start_time=0.0, end_time=1.0)
    p = np.array(circ["point"])
    for k in range(200):
        oracle.move_collider(circ, k * dt, dt)
        z = p - [0.5, 0.0, 0.5]
        th = np.arccos(z[0] / np.linalg.norm(z)) * (1 if z[2] >= 0 else -1)
        p = p + dt * np.array([-0.4 * np.sin(th), 0.0, 0.4 * np.cos(th)])
    assert np.abs(np.array(circ["point"]) - p).max() < 2e-6
    assert abs(np.hypot(p[0] - 0.5, p[2] - 0.5) - 0.2) < 2e-3
