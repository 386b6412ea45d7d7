"""Batches of independent scenes (mpm_batch_*): every scene of a batch
evolves as it does alone on the single-scene binned path.

Two GPU runs of the same scene differ only in the order of their float
atomics, so the bound is the spread already used between two single-scene
runs (u 1e-4, v 2e-5, C 2e-4, F 1e-6, tau 2e-2 relative L2)."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from reduced_mpm_stress_training_amd import batch as B
from reduced_mpm_stress_training_amd import scenes as S

pytestmark = pytest.mark.gpu

SPREAD = dict(u=1e-4, v=2e-5, C=2e-4, F=1e-6, tau=2e-2)


def _scenes():
    a = S.column_lattice(n_grid=32, per_axis=2, seed=0)
    b = S.column_lattice(n_grid=24, lo=(0.3, 0.4, 0.05), hi=(0.6, 0.6, 0.4), per_axis=2, seed=1)
    b.E, b.friction_angle = 3.5e5, 38.0
    c = S.column_uniform(n_particles=30_000, n_grid=40, lo=(0.35, 0.3, 0.05), hi=(0.6, 0.7, 0.5), v_std=0.3, seed=3)
    c.E, c.friction_angle = 1.2e5, 22.0
    return [a, b, c]


@pytest.mark.parametrize("launch", [(0, 0), (256, 16)])
def test_batch_matches_single_scenes(launch):
    scenes = _scenes()
    batch = B.SceneBatch(scenes)
    batch.set_launch(*launch)
    dt, n_steps = 1e-4, 70          # crosses two bank-order sorts
    batch.run(n_steps, dt)
    torch.cuda.synchronize()
    assert int(batch.status()[0]) == 0
    for i, sc in enumerate(scenes):
        sim = S.make_simulator(sc)
        sim.run(n_steps, dt)
        u1, v1, Fd1, t1 = sim.export_uvFstress_from_warp()
        C1 = sim.export_C()
        u2, v2, Fd2, t2, C2 = batch.export(i)
        torch.cuda.synchronize()
        ref = dict(u=u1, v=v1, C=C1.reshape(-1, 9), F=(Fd1 + torch.eye(3, device=Fd1.device)).reshape(-1, 9),
                   tau=t1.reshape(-1, 9))
        got = dict(u=u2, v=v2, C=C2.reshape(-1, 9), F=(Fd2 + torch.eye(3, device=Fd2.device)).reshape(-1, 9),
                   tau=t2.reshape(-1, 9))
        for k in SPREAD:
            e = rel_l2(got[k].cpu().numpy(), ref[k].cpu().numpy())
            assert np.isfinite(got[k].cpu().numpy()).all()
            assert e < SPREAD[k], (i, k, e)


def test_uneven_batch_runs_and_conserves_particles():
    scenes = B.uneven_scenes(n_scenes=12, seed=7, n_min=2_000, n_max=40_000, grid_min=32, grid_max=64)
    batch = B.SceneBatch(scenes)
    batch.run(40, 1e-4)
    torch.cuda.synchronize()
    st = batch.status()
    assert int(st[0]) == 0
    cnt = batch._count[batch.bins.cur].cpu().numpy().astype(np.int64)
    nb = int(batch.batch.n_bins)
    assert int(cnt[nb + 1:].sum()) + int(cnt[nb]) == batch.n_particles
    batch.planes[3].fill_(float("nan"))      # plane U: every particle must come back
    for i in range(len(scenes)):
        u, v, Fd, tau, C = batch.export(i)
        assert bool(torch.isfinite(u).all() and torch.isfinite(tau).all()), i


def test_autotune_returns_a_timed_candidate():
    from reduced_mpm_stress_training_amd.tune import autotune
    scenes = B.uneven_scenes(n_scenes=6, seed=3, n_min=2_000, n_max=20_000, grid_min=32, grid_max=48)
    cands = ((512, 32), (256, 16))
    best = autotune(lambda: B.SceneBatch(scenes), 1e-4, trial_substeps=4, candidates=cands)
    assert (best["launch"]["wg_threads"], best["launch"]["sort_period"]) in cands
    assert set(best["trials"]) == {"512x32", "256x16"}
    assert best["ms_per_substep"] == min(best["trials"].values()) > 0.0


def _compare_batches(a, b, n_scenes):
    for i in range(n_scenes):
        ga, gb = a.export(i), b.export(i)
        torch.cuda.synchronize()
        for k, x, y in zip(("u", "v", "F", "tau", "C"), ga, gb):
            if k not in SPREAD:
                continue
            e = rel_l2(y.reshape(len(y), -1).cpu().numpy(), x.reshape(len(x), -1).cpu().numpy())
            assert np.isfinite(y.cpu().numpy()).all(), (i, k)
            assert e < SPREAD[k], (i, k, e)


def test_captured_run_matches_eager_run():
    """The batched run captured as a HIP graph (mpm_batch_graph_*) replayed
    several times, with eager work and synchronizations between replays and
    eager substeps after them, against the same batch run eagerly; every
    replay is checked.  The colliders start at t = 0.012 so that the device
    time base matters."""
    scenes = _scenes()
    for sc in scenes:
        sc.floor_start = 0.012
    dt = 1e-4
    eager, graph = B.SceneBatch(scenes), B.SceneBatch(scenes)
    # 64 substeps per replay (four sort periods), so that the third replay
    # crosses the colliders' start
    k = graph.capture(dt, 64)
    assert k == 64
    for _ in range(3):
        eager.run(k, dt)
        graph.run_captured(1)
        torch.cuda.synchronize()
        assert abs(graph.time - eager.time) < 1e-12
        float(torch.randn(4096, device="cuda").sum())     # unrelated eager work between replays
        _compare_batches(eager, graph, len(scenes))
    eager.run(k, dt)              # eager substeps continue from a replay's state
    graph.run(k, dt)
    _compare_batches(eager, graph, len(scenes))
    eager.run(2 * k, dt)
    graph.run_captured(2)
    torch.cuda.synchronize()
    _compare_batches(eager, graph, len(scenes))
    assert int(graph.status()[0]) == 0
    graph.release_graph()


def test_captured_run_of_odd_length_matches_eager_run():
    """A graph of 17 substeps (one sort; 17 + 1 launches of the plane sets'
    swap keep the state on its plane set): each replay renumbers the pending
    epoch from odd to even, so the grid-block list of the pending P2G moves to
    the other parity's list at every replay.  Every replay against the eager
    run."""
    scenes = _scenes()
    dt = 1e-4
    eager, graph = B.SceneBatch(scenes), B.SceneBatch(scenes)
    k = graph.capture(dt, 17)
    assert k == 17
    for _ in range(3):
        eager.run(k, dt)
        graph.run_captured(1)
        torch.cuda.synchronize()
        _compare_batches(eager, graph, len(scenes))
    eager.run(5, dt)              # eager substeps continue from a replay's state
    graph.run(5, dt)
    _compare_batches(eager, graph, len(scenes))
    assert int(graph.status()[0]) == 0
    graph.release_graph()
