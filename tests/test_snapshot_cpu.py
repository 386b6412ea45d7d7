"""Snapshot and sampling file format (no GPU): the frame datasets of
save_data_at_frame (engine_stress/mpm_solver_warp.py:350-403), the Voigt
round trip of the restart loader load_uvFstress_from_h5 (:285-294), and the
sampling file that load_from_sampling reads (:134-152).
"""
import os

import numpy as np
import pytest

from reduced_mpm_stress_training_amd import snapshot as SN


def test_frame_layout_and_voigt_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    n = 7
    x0, u, v = (rng.normal(size=(n, 3)).astype(np.float32) for _ in range(3))
    Fd, C = (rng.normal(size=(n, 3, 3)).astype(np.float32) for _ in range(2))
    A = rng.normal(size=(n, 3, 3)).astype(np.float32)
    tau = A + np.swapaxes(A, 1, 2)                  # the stress is symmetric
    m = rng.uniform(size=n).astype(np.float32)
    d = SN.frame_dict(x0, u, v, Fd, tau, C, m, 0.25, 30.0)
    assert d["x"].shape == (3, n) and d["q"].shape == (3, n) and d["v"].shape == (3, n)
    assert d["time"].shape == (1, 1) and d["time"][0, 0] == 0.25
    assert d["masses"].shape == (1, n) and d["mu"].shape == (1, n) and np.all(d["mu"] == 30.0)
    # f_tensor and C: (9, N) row-major columns of the 3x3
    assert np.array_equal(d["f_tensor"][:, 2], Fd[2].reshape(9)) and np.array_equal(d["C"][:, 4], C[4].reshape(9))
    # stress: (6, N) rows xx, xy, xz, yy, yz, zz
    assert np.array_equal(d["stress"][:, 3], tau[3][[0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]])
    assert np.array_equal(SN.stress_from_voigt(d["stress"].T), tau)
    assert "mu" not in SN.frame_dict(x0, u, v, Fd, tau, C, m, 0.25, None)
    # written and read back unchanged (.npz twin of the h5 name without h5py)
    path = str(tmp_path / "h5_f_0000000003.npz")
    np.savez(path, **d)
    back = SN.read_frame(SN.resolve_frame_path(path[:-4] + ".h5") if SN.h5py is None else path)
    for k in d:
        assert np.array_equal(back[k], d[k]), k


def test_sampling_file_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    n = 11
    x = rng.uniform(size=(n, 3)).astype(np.float32)
    vol = np.full(n, 1e-6, np.float32)
    rho = np.full(n, 2200.0, np.float32)
    f = rng.normal(size=(n, 3, 3)).astype(np.float32)
    path = str(tmp_path / "sampling.h5")
    SN.write_sampling(path, x, rho * vol, vol, rho, f)
    d = SN.read_sampling(path)
    assert d["x"].shape == (3, n) and d["q"].shape == (3, n) and not d["q"].any()
    assert np.array_equal(d["x"].T, x)
    assert np.array_equal(np.squeeze(d["particle_volume"], 0), vol)
    assert np.array_equal(np.squeeze(d["masses"], 0), rho * vol)
    assert np.array_equal(SN.f_tensor_rows(d["f_tensor"], n), f.reshape(n, 9))
    # a frame-style (9, N) f_tensor reads the same
    assert np.array_equal(SN.f_tensor_rows(f.reshape(n, 9).T, n), f.reshape(n, 9))
    with pytest.raises(FileNotFoundError):
        SN.read_sampling(str(tmp_path / "missing.h5"))
