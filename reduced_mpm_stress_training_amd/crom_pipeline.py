"""The second stage of the CROM workflow: latents of the trained
displacement network become labels of the solver's frames, and a stress
network is trained on (label, X, friction angle) -> Voigt stress.

The reference runs it as three programs:

1. offline_training_u/run.py -mode reconstruct: the trained CROMnet encodes
   every frame; test_step (offline_training_u/CROMnet.py:93-128) writes one
   reconstruction file per frame with datasets x, q, time and label -- the
   frame's latent, (lbl, 1) -- (SimulationDataset.py:93-108);
2. generate_training_data/add_label_to_training.py:44-86 copies every
   ground-truth frame and adds prev_label = the frame's label repeated over
   its particles, (lbl, N) (:79-82);
3. offline_training_xhat_and_stress_and_mu trains the secondary network
   (CROMnet_secondary.py:126-136): input [prev_label, X, mu] per point
   (SimulationDataset.py:51-76, mu = the friction angle, unscaled), target
   the frame's stress (6, N) Voigt [xx, xy, xz, yy, yz, zz]
   (mpm_solver_warp.py:383), 1000 * MSE, Adam with the secondary schedule
   (learning-rate factors [10, 5, 2, 1] over [1000, 1000, 500, 200] epochs),
   on the first N/10 points of each frame (sparse_train_every_n_points = 10,
   run.py:32; SimulationDataset.py:34-43).

Here the files keep the reference's dataset names (.npz when h5py is not
installed, snapshot.py), and the networks are the library's BF16-MFMA
kernels (crom.py).
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import snapshot as SN
from .crom import Crom, CromDecoder, lr_factor

SECONDARY_LR = ((10.0, 5.0, 2.0, 1.0), (1000, 1000, 500, 200))   # CROMnet_secondary.py via run.py


# ---------------------------------------------------------------------------
# 1. labels: the encoder's latent of every frame (CROMnet.py:93-128)
# ---------------------------------------------------------------------------
@torch.no_grad()
def label_frames(crom: Crom, q_sample: torch.Tensor, batch: int = 16) -> torch.Tensor:
    """Latents (F, lbl) of frames q_sample (F, P, 3) -- the encoder's points
    of each frame (the first P particles)."""
    out = []
    for f0 in range(0, q_sample.shape[0], batch):
        out.append(crom.encoder(q_sample[f0:f0 + batch].to(crom.device, torch.float32).contiguous()))
    return torch.cat(out).reshape(q_sample.shape[0], crom.lbl)


def write_reconstruction(path: str, x: np.ndarray, q: np.ndarray, time: float, label: np.ndarray):
    """A reconstruction file as SimulationState.write_to_file writes it
    (SimulationDataset.py:93-108): x, q (3, P); time (1, 1) float64; label
    (lbl, 1) float64."""
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    arrs = {"x": np.asarray(x).T, "q": np.asarray(q).T, "time": np.array([[time]], np.float64),
            "label": np.asarray(label, np.float64).reshape(-1, 1)}
    if path.endswith(".h5") and SN.h5py is not None:
        with SN.h5py.File(path, "w") as f:
            for k, a in arrs.items():
                f.create_dataset(k, data=a)
    else:
        np.savez(path[:-3] + ".npz" if path.endswith(".h5") else path, **arrs)


# ---------------------------------------------------------------------------
# 2. prev_label (add_label_to_training.py:44-86)
# ---------------------------------------------------------------------------
def add_label_to_frame(ground_truth: str, reconstruction: str, out_path: str):
    """Copies every dataset of a ground-truth frame and adds prev_label: the
    reconstruction's label repeated over the frame's particles, (lbl, N)
    (add_label_to_training.py:69-82)."""
    gt = SN.read_frame(SN.resolve_frame_path(ground_truth))
    rec = SN.read_frame(SN.resolve_frame_path(reconstruction))
    out = dict(gt)
    out["prev_label"] = np.repeat(rec["label"], gt["x"].shape[1], 1)
    os.makedirs(os.path.dirname(out_path) or ".", exist_ok=True)
    if out_path.endswith(".h5") and SN.h5py is not None:
        with SN.h5py.File(out_path, "w") as f:
            for k, a in out.items():
                f.create_dataset(k, data=a)
    else:
        np.savez(out_path[:-3] + ".npz" if out_path.endswith(".h5") else out_path, **out)
    return out


def secondary_item(frame: dict, every_n_points: int = 10):
    """One training item of the secondary dataset (SimulationDataset.py:
    25-55): the first N / every_n_points points' (prev_label, x, mu) inputs
    and stress (6) targets, each (P, .)."""
    x, q = frame["x"].T, frame["stress"].T
    mu, lab = frame["mu"].T, frame["prev_label"].T
    if every_n_points != 1:
        k = int(x.shape[0] / every_n_points)
        x, q, mu, lab = x[:k], q[:k], mu[:k], lab[:k]
    return {"encoder_input": np.concatenate([lab, x, mu], 1).astype(np.float32), "q": q.astype(np.float32)}


# ---------------------------------------------------------------------------
# 3. the stress network (CROMnet_secondary.py:203-307)
# ---------------------------------------------------------------------------
class StressHead:
    """The secondary network: [label (lbl), X (3), mu (1)] -> Voigt stress
    (6) through the decoder trunk (5 x 128 ELU), X standardised and mu passed
    through as the reference's Prepare does (CROMnet_secondary.py:507-513),
    the output de-standardised with the stress statistics."""

    def __init__(self, lbl: int, device="cuda:0", seed: int = 0):
        self.net = CromDecoder(lbl, out_dim=6, mu_dim=1, device=device, seed=seed)
        self.lbl, self.device = lbl, torch.device(device)

    def set_preprop(self, X: torch.Tensor, stress: torch.Tensor):
        s = stress.reshape(-1, 6)
        self.net.set_preprop(X.mean(0), X.std(0), s.mean(0), s.std(0).clamp_min(1e-12))

    @staticmethod
    def inputs(labels: torch.Tensor, X: torch.Tensor, mu) -> torch.Tensor:
        """(B * P, lbl + 4) inputs [label_b, X_p, mu_b] of B frames' labels
        (B, lbl) over P points X (P, 3); mu a scalar or (B,)."""
        B, P = labels.shape[0], X.shape[0]
        mu = torch.as_tensor(mu, dtype=torch.float32, device=labels.device).reshape(-1)
        mu = mu.expand(B) if mu.numel() == 1 else mu
        return torch.cat([labels[:, None, :].expand(B, P, labels.shape[1]), X[None].expand(B, P, 3),
                          mu[:, None, None].expand(B, P, 1)], 2).reshape(B * P, labels.shape[1] + 4).contiguous()

    def predict(self, labels, X, mu) -> torch.Tensor:
        """Voigt stress (B, P, 6)."""
        return self.net(self.inputs(labels, X, mu)).view(labels.shape[0], X.shape[0], 6)

    def train_step(self, labels, X, mu, stress, lr: float) -> torch.Tensor:
        """One Adam step on 1000 * MSE against stress (B, P, 6)."""
        return self.net.train_step(self.inputs(labels, X, mu), stress.reshape(-1, 6), lr)


def train_stress_head(head: StressHead, labels: torch.Tensor, X: torch.Tensor, mu: torch.Tensor,
                      stress: torch.Tensor, steps: int, lr0: float = 1e-4, batch: int = 10, seed: int = 0):
    """Minibatches of `batch` frames (run.py's batch size 10), learning rate
    lr0 times the secondary schedule's factor, the schedule compressed onto
    `steps`.  labels (F, lbl), X (P, 3), mu (F,), stress (F, P, 6).
    Returns the last loss."""
    g = torch.Generator(device=labels.device).manual_seed(seed)
    total = sum(SECONDARY_LR[1])
    loss = None
    for it in range(steps):
        idx = torch.randint(0, labels.shape[0], (min(batch, labels.shape[0]),), device=labels.device, generator=g)
        lr = lr0 * lr_factor(int(it * total / steps), *SECONDARY_LR)
        loss = head.train_step(labels[idx], X, mu[idx], stress[idx], lr)
    return loss


def relative_l2(a: torch.Tensor, b: torch.Tensor) -> float:
    return float((a - b).norm() / b.norm().clamp_min(1e-30))
