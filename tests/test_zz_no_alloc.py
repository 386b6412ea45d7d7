"""mpm_step_noalloc (the substep with its grids in the library's static device
array) against mpm_step (the same substep on a caller workspace), through the
C-ABI, on two calls in a row -- the second with a smaller grid and fewer
particles, so a stale grid from the first call would show.

Tolerance: both run the same kernels; the P2G's float atomics add in no fixed
order, so the two differ by FP32 rounding only.  Positions x = x0 + u agree
within 1e-6 relative L2 and u within 2 ulp of x; v, C, F, tau within
max(1e-5, 3 e32 + s32) relative L2, the column bound of
tests/test_gpu_parity.py (e32: the FP32 oracle's error against the FP64
oracle on the same scene and substeps, s32: the FP32 oracle's spread when its
P2G sums run in reverse order -- what a different atomic order alone moves).
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from reduced_mpm_stress_training_amd import _lib as L
from reduced_mpm_stress_training_amd import scenes as S

pytestmark = pytest.mark.gpu


def _random_sim(n_grid, per_axis, seed):
    sc = S.column_lattice(n_grid=n_grid, per_axis=per_axis, seed=seed)
    sim = S.make_simulator(sc)
    sim.use_bins = False
    rng = np.random.default_rng(seed)
    n = sc.n_particles
    dev = sim.device
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)
    sim.load_velocity_from_tensor(t(rng.normal(0, 0.2, (n, 3))))
    sim.load_F_from_tensor(t((np.eye(3)[None] + 0.01 * rng.normal(size=(n, 3, 3))).reshape(n, 9)))
    sim.load_C_from_tensor(t(rng.normal(0, 1.0, (n, 9))))
    return sc, sim


def _state(sim, planes):
    st = L.MpmState()
    st.n = sim.n_particles
    st.stride = planes.shape[1]
    st.planes = planes.data_ptr()
    return st


def _bounds(oracle, sc, sim, n_steps):
    """max(1e-5, 3 e32 + s32) per field from the CPU oracles on the state the
    GPU starts from."""
    n = sc.n_particles
    pl = sim.planes.cpu().numpy()
    st = dict(x0=pl[L.P_X0:L.P_X0 + 3, :n].T, u=pl[L.P_U:L.P_U + 3, :n].T, v=pl[L.P_V:L.P_V + 3, :n].T,
              C=pl[L.P_C:L.P_C + 9, :n].T.reshape(n, 3, 3), F=pl[L.P_F:L.P_F + 9, :n].T.reshape(n, 3, 3),
              tau=pl[L.P_TAU:L.P_TAU + 9, :n].T.reshape(n, 3, 3), mass=pl[L.P_MASS, :n], vol=pl[L.P_VOL, :n])

    def run(dtype, order=None):
        s = st if order is None else {k: a[order] for k, a in st.items()}
        o = oracle.OracleScene(S.scene_params(sc), s, S.scene_colliders(sc), dtype).run(n_steps, sc.dt)
        out = {k: o.s[k].reshape(n, -1) for k in ("v", "C", "F", "tau")}
        if order is not None:
            inv = np.argsort(order)
            out = {k: a[inv] for k, a in out.items()}
        return out

    o64, o32, o32r = run(np.float64), run(np.float32), run(np.float32, np.arange(n)[::-1].copy())
    return {k: max(1e-5, 3 * rel_l2(o32[k], o64[k]) + rel_l2(o32r[k], o32[k])) for k in o64}


def _compare(sim, a, b, bound):
    n = sim.n_particles
    x_a = a[L.P_X0:L.P_X0 + 3, :n] + a[L.P_U:L.P_U + 3, :n]
    x_b = b[L.P_X0:L.P_X0 + 3, :n] + b[L.P_U:L.P_U + 3, :n]
    assert rel_l2(x_b.cpu().numpy(), x_a.cpu().numpy()) <= 1e-6
    # u = x - x0 is a few 1e-5 after three substeps, where one ulp of x is a
    # few percent of it: u is held to 2 ulp of x
    ulp = np.spacing(np.abs(x_a.cpu().numpy()).astype(np.float32))
    du = np.abs(b[L.P_U:L.P_U + 3, :n].cpu().numpy() - a[L.P_U:L.P_U + 3, :n].cpu().numpy())
    assert np.all(du <= 2 * ulp)
    for name, first, ncomp in (("v", L.P_V, 3), ("C", L.P_C, 9), ("F", L.P_F, 9), ("tau", L.P_TAU, 9)):
        ga = a[first:first + ncomp, :n].cpu().numpy()
        gb = b[first:first + ncomp, :n].cpu().numpy()
        assert np.isfinite(gb).all()
        assert rel_l2(gb, ga) <= bound[name], (name, rel_l2(gb, ga), bound[name])


def test_noalloc_step_matches_mpm_step(oracle):
    lib = L.load()
    assert lib.mpm_noalloc_max_grid() == 64
    stream = torch.cuda.current_stream().cuda_stream
    for n_grid, per_axis, seed in ((48, 2, 0), (16, 2, 1)):     # the second call: smaller grid, fewer particles
        sc, sim = _random_sim(n_grid, per_axis, seed)
        bound = _bounds(oracle, sc, sim, 3)
        prm = sim._params()
        cols, ncol = sim._colliders()
        ref = sim.planes.clone()
        out = sim.planes.clone()
        ws = torch.zeros(int(lib.mpm_workspace_bytes(prm, sim.n_particles)) + 256, dtype=torch.uint8,
                         device=sim.device)
        wp = (ws.data_ptr() + 255) & ~255
        for k in range(3):
            t = float(k * sc.dt)
            L.check(lib.mpm_step(prm, _state(sim, ref), cols, ncol, wp, ws.numel() - 256, t, sc.dt, stream),
                    "mpm_step")
            L.check(lib.mpm_step_noalloc(prm, _state(sim, out), cols, ncol, t, sc.dt, stream), "mpm_step_noalloc")
        torch.cuda.synchronize()
        _compare(sim, ref, out, bound)
        assert float(torch.abs(out[L.P_U:L.P_U + 3, :sim.n_particles]).max()) > 0


def test_noalloc_rejects_grids_beyond_its_scratch():
    lib = L.load()
    sc, sim = _random_sim(16, 2, 2)
    prm = sim._params()
    prm.n_grid = 65
    rc = lib.mpm_step_noalloc(prm, _state(sim, sim.planes), None, 0, 0.0, 1e-4,
                              torch.cuda.current_stream().cuda_stream)
    assert rc < 0 and b"static scratch" in lib.mpm_last_error()
