"""The file steps of the CROM label pipeline (no GPU): a reconstruction file
as the reference's test_step writes it (offline_training_u/CROMnet.py:93-128,
SimulationDataset.py:93-108), the prev_label dataset add_label_to_training
adds to a ground-truth frame (generate_training_data/add_label_to_training.py:
44-86), and the secondary network's training item (offline_training_xhat_and_
stress_and_mu/SimulationDataset.py:25-55).
"""
import numpy as np

from reduced_mpm_stress_training_amd import crom_pipeline as CP
from reduced_mpm_stress_training_amd import snapshot as SN


def test_label_files_and_secondary_item(tmp_path):
    rng = np.random.default_rng(0)
    n, lbl = 50, 16
    x0, u, v = (rng.normal(size=(n, 3)).astype(np.float32) for _ in range(3))
    A = rng.normal(size=(n, 3, 3)).astype(np.float32)
    tau = A + np.swapaxes(A, 1, 2)
    frame = SN.frame_dict(x0, u, v, rng.normal(size=(n, 3, 3)), tau, rng.normal(size=(n, 3, 3)),
                          rng.uniform(size=n), 0.01, 33.0)
    gt = str(tmp_path / "gt" / "sim_seq_33" / "h5_f_0000000004.npz")
    (tmp_path / "gt" / "sim_seq_33").mkdir(parents=True)
    np.savez(gt, **frame)
    label = rng.normal(size=lbl)
    rec = str(tmp_path / "rec" / "sim_seq_33" / "h5_f_0000000004.h5")
    CP.write_reconstruction(rec, x0[:5], u[:5], 0.01, label)
    r = SN.read_frame(SN.resolve_frame_path(rec))
    assert r["label"].shape == (lbl, 1) and r["label"].dtype == np.float64 and r["time"].shape == (1, 1)
    assert r["x"].shape == (3, 5) and r["q"].shape == (3, 5)
    out_path = str(tmp_path / "gt_and_xlabel" / "sim_seq_33" / "h5_f_0000000004.h5")
    out = CP.add_label_to_frame(gt, rec, out_path)
    back = SN.read_frame(SN.resolve_frame_path(out_path))
    assert back["prev_label"].shape == (lbl, n)
    assert np.array_equal(back["prev_label"], np.repeat(label.reshape(lbl, 1), n, 1))
    for k in frame:
        assert np.array_equal(back[k], frame[k]), k
    item = CP.secondary_item(out, every_n_points=10)
    assert item["encoder_input"].shape == (5, lbl + 4) and item["q"].shape == (5, 6)
    assert np.allclose(item["encoder_input"][:, :lbl], label[None])
    assert np.allclose(item["encoder_input"][:, lbl:lbl + 3], x0[:5])
    assert np.all(item["encoder_input"][:, -1] == 33.0)
    assert np.allclose(item["q"], tau[:5].reshape(5, 9)[:, SN.VOIGT])
