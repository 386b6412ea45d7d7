"""The 32^3 sand column against the frozen fixtures of tests/golden/
(written by tests/golden/make_golden.py from the FP64 oracle).

CPU: the scene generator and the oracle still produce the fixture's numbers.
GPU: the HIP path after 10 and 100 substeps lies within max(1e-5, 3 e32 + s32)
(relative L2 on the fixture's 1024 sample particles) of the FP64 oracle's
values, e32 / s32 being the FP32 oracle's distance / order spread stored in
the fixture; positions x = x0 + u within 1e-5.
"""
import os

import numpy as np
import pytest

from conftest import rel_l2
from reduced_mpm_stress_training_amd import scenes as S

FIELDS = ("u", "v", "C", "F", "tau")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture(n_steps):
    return dict(np.load(os.path.join(GOLDEN, f"column32_{n_steps}.npz")))


@pytest.mark.parametrize("n_steps", [10, 100])
def test_oracle_reproduces_golden(oracle, n_steps):
    g = _fixture(n_steps)
    sc = S.column_lattice()
    assert sc.n_particles == int(g["n_particles"])
    assert abs(sc.x0.astype(np.float64).sum() - float(g["x0_sum"])) <= 1e-12 * abs(float(g["x0_sum"]))
    o = oracle.OracleScene(S.scene_params(sc), S.scene_state(sc, np.float64), S.scene_colliders(sc),
                           np.float64).run(n_steps, sc.dt, nthreads=1)
    idx = g["idx"]
    for k in FIELDS:
        a = np.asarray(o.s[k], np.float64).reshape(sc.n_particles, -1)
        assert rel_l2(a[idx], g[k]) <= 1e-12, k
        assert abs(a.sum() - float(g[k + "_sum"])) <= 1e-9 * max(1.0, np.abs(a).sum()), k


@pytest.mark.gpu
@pytest.mark.parametrize("n_steps", [10, 100])
def test_gpu_matches_golden(n_steps):
    import torch
    g = _fixture(n_steps)
    sc = S.column_lattice()
    sim = S.make_simulator(sc)
    sim.run(n_steps, sc.dt)
    u, v, Fd, tau = sim.export_uvFstress_from_warp()
    C = sim.export_C()
    torch.cuda.synchronize()
    n = sc.n_particles
    out = dict(u=u.cpu().numpy(), v=v.cpu().numpy(), C=C.cpu().numpy().reshape(n, 9),
               F=(Fd + torch.eye(3, device=Fd.device)).cpu().numpy().reshape(n, 9),
               tau=tau.cpu().numpy().reshape(n, 9))
    idx = g["idx"]
    x0 = sc.x0.astype(np.float64)[idx]
    assert rel_l2(x0 + out["u"][idx], x0 + g["u"]) <= 1e-5
    for k in FIELDS:
        e = rel_l2(out[k][idx].reshape(len(idx), -1), g[k])
        bound = max(1e-5, 3 * float(g[k + "_e32"]) + float(g[k + "_s32"]))
        assert e <= bound, (k, e, bound)
