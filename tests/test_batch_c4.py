"""The batched-scenes configuration at full size: 64 independent sand
columns (particle counts log-uniform in [2k, 200k], grids 32^3-96^3, column
dimensions U[0.2, 0.5], friction angle U[20, 40] deg, E U[1e5, 5e5] Pa,
seed 2) in one batch with autotuned launch parameters, 300 substeps replayed
as a HIP graph.

* every scene keeps all its particles (bin counts) and finite state;
* the smallest, the median and the largest scene of a batch against the
  FP64 CPU oracle after 10 substeps, and the smallest after all 300 graph
  substeps, within the column tolerance max(1e-5, 3 e32 + s32) relative L2
  (tests/test_gpu_parity.py; e32 the FP32 oracle's error, s32 its spread
  under a reordered P2G sum), positions x = x0 + u within 1e-5.
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from reduced_mpm_stress_training_amd import batch as B
from reduced_mpm_stress_training_amd import scenes as S
from reduced_mpm_stress_training_amd.tune import autotune

pytestmark = pytest.mark.gpu

DT = 1e-4
FIELDS = ("u", "v", "F", "tau", "C")


def _check_scene(oracle, sc, got, n_steps):
    n = sc.n_particles

    def run(dtype, order=None):
        st = S.scene_state(sc, np.float64)
        if order is not None:
            st = {k: a[order] for k, a in st.items()}
        o = oracle.OracleScene(S.scene_params(sc), st, S.scene_colliders(sc), dtype).run(n_steps, DT, nthreads=16)
        out = {k: o.s[k].reshape(n, -1) for k in FIELDS}
        if order is not None:
            inv = np.argsort(order)
            out = {k: a[inv] for k, a in out.items()}
        return out

    o64, o32 = run(np.float64), run(np.float32)
    o32r = run(np.float32, np.arange(n)[::-1].copy())
    g = {k: got[k].reshape(n, -1) for k in FIELDS}
    x0 = sc.x0.astype(np.float64)
    assert rel_l2(x0 + g["u"], x0 + o64["u"]) <= 1e-5, sc.name
    for k in FIELDS:
        e, e32, s32 = rel_l2(g[k], o64[k]), rel_l2(o32[k], o64[k]), rel_l2(o32r[k], o32[k])
        assert np.isfinite(g[k]).all(), (sc.name, k)
        assert e <= max(1e-5, 3 * e32 + s32), (sc.name, k, e, e32, s32)


def _export(batch, i):
    u, v, Fd, tau, C = batch.export(i)
    torch.cuda.synchronize()
    return dict(u=u.cpu().numpy(), v=v.cpu().numpy(), F=(Fd + torch.eye(3, device=Fd.device)).cpu().numpy(),
                tau=tau.cpu().numpy(), C=C.cpu().numpy())


def test_64_uneven_scenes_300_captured_substeps(oracle):
    scenes = B.uneven_scenes(n_scenes=64, seed=2, ppc=4.0)
    sizes = np.array([sc.n_particles for sc in scenes])
    assert sizes.min() >= 2000 and sizes.max() <= 200_000 and len(scenes) == 64
    order = np.argsort(sizes)
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    best = autotune(lambda: B.SceneBatch(scenes), DT, trial_substeps=8,
                    candidates=((768, 32), (512, 32), (256, 32), (256, 16)), rounds=1)
    launch = best["launch"]

    # the first 10 substeps of three scenes against the oracle
    b10 = B.SceneBatch(scenes)
    b10.set_launch(**launch)
    b10.run(10, DT)
    got10 = {i: _export(b10, i) for i in picks}
    del b10
    torch.cuda.empty_cache()
    for i in picks:
        _check_scene(oracle, scenes[i], got10[i], 10)

    # 300 substeps as graph replays (plus an eager remainder)
    bt = B.SceneBatch(scenes)
    bt.set_launch(**launch)
    k = bt.capture(DT)
    reps, rest = divmod(300, k)
    bt.run_captured(reps)
    if rest:
        bt.run(rest, DT)
    torch.cuda.synchronize()
    assert abs(bt.time - 300 * DT) < 1e-12
    assert int(bt.status()[0]) == 0
    cnt = bt._count[bt.bins.cur].cpu().numpy().astype(np.int64)
    nb = int(bt.batch.n_bins)
    assert int(cnt[nb + 1:].sum()) + int(cnt[nb]) == bt.n_particles
    for i in range(len(scenes)):
        g = _export(bt, i)
        assert all(np.isfinite(a).all() for a in g.values()), i
    i0 = picks[0]
    _check_scene(oracle, scenes[i0], _export(bt, i0), 300)
    bt.release_graph()
