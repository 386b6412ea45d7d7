"""Run-to-run reproducibility of the binned substep and its failure counters.

The reference accumulates P2G with float atomics (mpm_utils.py:224-227), so
two of its runs need not agree bit for bit.  Here every P2G sum is taken in
FP64 over FP32 contributions -- the bins' LDS tiles, the FP64 tiles of
split bins and the spill grid of particles that leave their bin's tile --
and the grid update adds the tiles in a fixed order: two runs of the same
input give the same bits (unless a node's sum needs more than 53 bits).  The
tests run each scene twice in one process and compare every exported field
bitwise, with the split and spill paths shown to be taken.

One path is not bit-reproducible: when a bin fills past its capacity, which
of its particles go to the overflow run (and the spill grid) depends on the
order of the atomics, and a bin's tile is rounded to FP32 before the spill
sum joins it.  The solver re-bins with larger bins when the overflow run is
half full, and health() reports its occupancy; the converging scene below
drives particles into it and the two runs are held to the FP32 parity
tolerance instead.

The failure counters (the workspace's status words, surfaced by
MPMSimulator.health() and checked at every export): a non-finite stress
loaded into one particle is reported as NaN/Inf grid sums and raises
FloatingPointError at export; a particle that falls out of the grid is
counted once and warned about (the reference has no such guard: its CFL
check, mpm_solver_warp.py:188-192, stops on fast particles only).
"""
import numpy as np
import pytest
import torch

from reduced_mpm_stress_training_amd import _lib as L
from reduced_mpm_stress_training_amd import scenes as S

pytestmark = pytest.mark.gpu


def _state(sim):
    out = sim.export_uvFstress_from_warp()
    torch.cuda.synchronize()
    return [a.cpu().numpy().copy() for a in out]


def _bits_equal(a, b):
    return all(np.array_equal(x.view(np.int32), y.view(np.int32)) for x, y in zip(a, b))


def _converging(n=60_000, seed=4):
    """A sand cube [0.2, 0.55]^3 on 32^3 with the velocity field
    v = -100 (x - c), c = 0.375 the centre of bin (1,1,1): the bins around c
    fill past their capacity (twice the fullest bin at binning time) and the
    excess runs through the overflow run and the spill grid."""
    rng = np.random.default_rng(seed)
    c = 0.375
    x0 = (c + rng.uniform(-0.175, 0.175, (n, 3))).astype(np.float32)
    v0 = (-100.0 * (x0 - c)).astype(np.float32)
    vol = np.full(n, 0.35 ** 3 / n, np.float32)
    return S.SandScene(name="converging", n_grid=32, x0=x0, vol=vol, v0=v0)


def _split_stamps(sim):
    """Bins whose tile stamp is negative: the fused kernel split them into
    parts.  The stamps follow the 64 status words in the workspace."""
    B = sim._bins
    lib = L.load()
    off = sim._bins_ws_ptr() - B["ws"].data_ptr() + int(lib.mpm_bins_status_offset(sim._params())) + 256
    nb3 = B["nb"] ** 3
    return int((B["ws"][off:off + 4 * nb3].view(torch.int32) < 0).sum())


def test_split_bins_are_bitwise_reproducible():
    """250k particles on 64^3 (bins split into parts over the CUs), 70
    substeps: past two bank-order sorts."""
    sc = S.column_uniform(n_particles=250_000, n_grid=64)
    runs, splits = [], []
    for _ in range(2):
        sim = S.make_simulator(sc)
        sim.run(70, sc.dt)
        torch.cuda.synchronize()
        splits.append(_split_stamps(sim))
        runs.append(_state(sim))
    print("split bins", splits)
    assert min(splits) > 0
    assert _bits_equal(*runs)


def test_spill_path_is_bitwise_reproducible():
    """A sand block translating at 350 m/s on 64^3 (dt 1e-4: 2.2 cells per
    substep): the particles in the last one or two cell layers of each bin
    land outside their bin's tile (about a sixth of them every substep) and
    scatter through the FP64 spill grid."""
    rng = np.random.default_rng(5)
    n = 40_000
    x0 = rng.uniform((0.1, 0.3, 0.3), (0.2, 0.5, 0.5), (n, 3)).astype(np.float32)
    v0 = np.zeros((n, 3), np.float32)
    v0[:, 0] = 350.0
    sc = S.SandScene(name="translating", n_grid=64, x0=x0, vol=np.full(n, 0.004 / n, np.float32), v0=v0)
    runs, health = [], []
    for _ in range(2):
        sim = S.make_simulator(sc)
        sim.run(20, sc.dt)
        health.append(sim.health())
        runs.append(_state(sim))
    print("health", health)
    assert health[0]["outside_tile"] >= n and health[0] == health[1]
    assert health[0]["overflow"] == 0 and health[0]["left_grid"] == 0
    assert all(np.isfinite(a).all() for a in runs[0])
    assert _bits_equal(*runs)


def test_overflow_path_is_reproducible_to_fp32_tolerance():
    sc = _converging()
    runs, ovf, health = [], [], []
    for _ in range(2):
        sim = S.make_simulator(sc)
        used = 0
        for _ in range(8):
            sim.run(5, sc.dt)
            used = max(used, sim.health()["overflow"])
        ovf.append(used)
        health.append(sim.health())
        runs.append(_state(sim))
    diff = [float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)) for a, b in zip(*runs)]
    print("overflow slots", ovf, "health", health, "run-to-run relative L2 (u, v, F, tau)", diff)
    assert min(ovf) > 0
    assert health[0]["dropped"] == 0 and health[0]["nonfinite_nodes"] == 0
    assert all(np.isfinite(a).all() for a in runs[0])
    assert max(diff) <= 1e-4


def test_nonfinite_state_is_reported_and_raises_at_export():
    sc = S.column_lattice()
    sim = S.make_simulator(sc)
    sim.run(2, sc.dt)
    assert sim.health()["nonfinite_nodes"] == 0
    u, v, _, tau = sim.export_uvFstress_from_warp()
    tau = tau.clone()
    tau[7, 1, 1] = float("nan")
    sim.load_uvstress_from_tensor(u.clone(), v.clone(), tau)
    sim.run(1, sc.dt)
    h = sim.health()
    print(h)
    assert h["nonfinite_nodes"] > 0 and h["dropped"] == 0
    with pytest.raises(FloatingPointError):
        sim.export_uvFstress_from_warp()


def test_particle_leaving_the_grid_is_counted_and_warned():
    """One particle in free fall without colliders leaves the grid through
    its floor: counted once, held at its last position, warned at export."""
    sc = S.SandScene(name="one", n_grid=32, x0=np.array([[0.5, 0.5, 0.6]], np.float32),
                     vol=np.array([1e-6], np.float32), v0=np.zeros((1, 3), np.float32))
    sim = S.make_simulator(sc)
    sim.colliders = []
    sim.run(60, 1e-2)
    h = sim.health()
    assert h["left_grid"] == 1 and h["nonfinite_nodes"] == 0
    with pytest.warns(RuntimeWarning, match="left the grid"):
        u = sim.export_uvFstress_from_warp()[0]
    z = 0.6 + float(u[0, 2])
    assert np.isfinite(z) and z < 1.5 / 32
