"""CROM decoder on MI355X matrix cores (libmpm_hip.so, include/crom_hip.h).

CromDecoder is the reference's decoder network NetDec
(offline_training_u/CROMnet.py:218-301): Prepare standardizes the three
reference-position columns of the input [label, X] (:497-534), five Linear
layers of width 128 each followed by ELU(alpha=1) (:226-237, :457-473), a
Linear layer to out_dim, and invStandardizeQ (mean_q + y * std_q, :475-487).
With mu_dim = 1 and out_dim = 6 it is the stress head of
offline_training_xhat_and_stress_and_mu/CROMnet_secondary.py:203-307 (input
[label, X, mu], Voigt stress out).

Weights are FP32 torch tensors in nn.Linear layout; every product runs in the
library's BF16-MFMA kernels with FP32 accumulation.  There is no CPU path.
"""
from __future__ import annotations

import math

import torch

from . import _lib as L

HIDDEN = 128
N_LINEAR = 6


def _aligned_bytes(nbytes: int, device) -> tuple[torch.Tensor, int]:
    buf = torch.empty(nbytes + 256, dtype=torch.uint8, device=device)
    ptr = (buf.data_ptr() + 255) & ~255
    return buf, ptr


class CromDecoder:
    """NetDec with hidden width 128 (CROMnet.py:218-301)."""

    def __init__(self, lbl: int, out_dim: int = 3, mu_dim: int = 0, device="cuda:0", seed: int = 0):
        self.lbl, self.mu_dim, self.out_dim = int(lbl), int(mu_dim), int(out_dim)
        self.in_dim = self.lbl + 3 + self.mu_dim
        if not (4 <= self.in_dim <= 32) or not (1 <= self.out_dim <= 32):
            raise ValueError("CromDecoder: in_dim must be in [4, 32] and out_dim in [1, 32]")
        self.device = torch.device(device)
        lib = L.load()
        self.n_params = int(lib.crom_param_count(self.in_dim, self.out_dim))
        # flat FP32 parameters (per layer: weight [out][in], then bias), the
        # layout the gradient and Adam kernels use; weights / biases are views
        self.params = torch.empty(self.n_params, dtype=torch.float32, device=self.device)
        self.grads = torch.zeros_like(self.params)
        self.adam_m = torch.zeros_like(self.params)
        self.adam_v = torch.zeros_like(self.params)
        self.adam_step = 0
        g = torch.Generator().manual_seed(seed)
        dims = [self.in_dim] + [HIDDEN] * 5 + [self.out_dim]
        self.weights, self.biases = [], []
        off = 0
        for k in range(N_LINEAR):
            fan_in, fan_out = dims[k], dims[k + 1]
            w = self.params[off:off + fan_out * fan_in].view(fan_out, fan_in)
            off += fan_out * fan_in
            b = self.params[off:off + fan_out]
            off += fan_out
            # nn.init.xavier_uniform_ (CROMnet.py:275); nn.Linear's default bias init
            a = math.sqrt(6.0 / (fan_in + fan_out))
            w.copy_(((torch.rand((fan_out, fan_in), generator=g) * 2 - 1) * a).to(self.device))
            bb = 1.0 / math.sqrt(fan_in)
            b.copy_(((torch.rand((fan_out,), generator=g) * 2 - 1) * bb).to(self.device))
            self.weights.append(w)
            self.biases.append(b)
        assert off == self.n_params
        self.mean_x = torch.zeros(3, device=self.device)
        self.std_x = torch.ones(3, device=self.device)
        self.mean_q = torch.zeros(self.out_dim, device=self.device)
        self.std_q = torch.ones(self.out_dim, device=self.device)
        self._packed_buf, self._packed = _aligned_bytes(int(lib.crom_packed_bytes()), self.device)
        self._dirty = True
        self._bufs = {}

    # -- parameters ---------------------------------------------------------
    def set_preprop(self, mean_x, std_x, mean_q, std_q):
        """Prepare / invStandardizeQ constants (CROMnet.py:507-511, :483-484)."""
        t = lambda v, n: torch.as_tensor(v, dtype=torch.float32).reshape(n).to(self.device)
        self.mean_x, self.std_x = t(mean_x, 3), t(std_x, 3)
        self.mean_q, self.std_q = t(mean_q, self.out_dim), t(std_q, self.out_dim)
        self._dirty = True

    def load_linear_layers(self, layers):
        """Copy weights from torch nn.Linear-like objects dec0..dec4, enc."""
        with torch.no_grad():
            for k, layer in enumerate(layers):
                self.weights[k].copy_(layer.weight.detach().to(self.device, torch.float32))
                self.biases[k].copy_(layer.bias.detach().to(self.device, torch.float32))
        self._dirty = True

    def state_dict(self) -> dict:
        return {"params": self.params.detach().cpu().clone(), "mean_x": self.mean_x.cpu(), "std_x": self.std_x.cpu(),
                "mean_q": self.mean_q.cpu(), "std_q": self.std_q.cpu(), "lbl": self.lbl, "mu_dim": self.mu_dim,
                "out_dim": self.out_dim}

    def load_state_dict(self, sd: dict):
        self.params.copy_(sd["params"].to(self.device))
        self.set_preprop(sd["mean_x"], sd["std_x"], sd["mean_q"], sd["std_q"])

    def mark_updated(self):
        self._dirty = True

    def _net(self) -> L.CromNet:
        net = L.CromNet()
        net.in_dim, net.out_dim, net.x_col = self.in_dim, self.out_dim, self.lbl
        for k in range(N_LINEAR):
            assert self.weights[k].is_contiguous() and self.weights[k].dtype == torch.float32
            net.weight[k] = self.weights[k].data_ptr()
            net.bias[k] = self.biases[k].data_ptr()
        net.mean_x, net.std_x = self.mean_x.data_ptr(), self.std_x.data_ptr()
        net.mean_q, net.std_q = self.mean_q.data_ptr(), self.std_q.data_ptr()
        return net

    def pack(self):
        """Re-pack the BF16 operand fragments after a weight change."""
        lib = L.load()
        L.check(lib.crom_pack(self._net(), self._packed, torch.cuda.current_stream(self.device).cuda_stream),
                "crom_pack")
        self._dirty = False

    # -- evaluation ---------------------------------------------------------
    def forward(self, x: torch.Tensor, acts: torch.Tensor | None = None) -> torch.Tensor:
        """NetDec.forward (CROMnet.py:293-296): x (n, in_dim) -> (n, out_dim)."""
        if x.dim() != 2 or x.shape[1] != self.in_dim:
            raise ValueError(f"x must be (n, {self.in_dim})")
        x = x.to(self.device, torch.float32).contiguous()
        if self._dirty:
            self.pack()
        n = x.shape[0]
        y = torch.empty((n, self.out_dim), dtype=torch.float32, device=self.device)
        lib = L.load()
        L.check(lib.crom_forward(self._packed, x.data_ptr(), n, y.data_ptr(),
                                 acts.data_ptr() if acts is not None else None,
                                 torch.cuda.current_stream(self.device).cuda_stream), "crom_forward")
        return y

    __call__ = forward

    # -- training -----------------------------------------------------------
    def _buffer(self, name, nbytes):
        buf = self._bufs.get(name)
        if buf is None or buf[0].numel() < nbytes + 256:
            self._bufs[name] = _aligned_bytes(int(nbytes), self.device)
        return self._bufs[name][1]

    def backward(self, x: torch.Tensor, q: torch.Tensor, loss_weight: float = 1000.0, input_grad: bool = False):
        """Forward, then the gradient of  loss_weight * MSE(y, q)  (the
        reference's training loss, CROMnet.py:76) into self.grads.  Returns
        (y, dx or None)."""
        lib = L.load()
        x = x.to(self.device, torch.float32).contiguous()
        q = q.to(self.device, torch.float32).contiguous()
        n = x.shape[0]
        if q.shape != (n, self.out_dim):
            raise ValueError(f"q must be ({n}, {self.out_dim})")
        if self._dirty:
            self.pack()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        nb = int(lib.crom_acts_bytes(n))
        acts = self._buffer("acts", nb)
        dlt = self._buffer("grads_buf", nb)
        ws = self._buffer("wgrad_ws", int(lib.crom_wgrad_workspace_bytes(n)))
        y = torch.empty((n, self.out_dim), dtype=torch.float32, device=self.device)
        L.check(lib.crom_forward(self._packed, x.data_ptr(), n, y.data_ptr(), acts, stream), "crom_forward")
        dx = torch.empty((n, self.in_dim), dtype=torch.float32, device=self.device) if input_grad else None
        scale = 2.0 * loss_weight / float(n * self.out_dim)
        L.check(lib.crom_backward(self._packed, acts, y.data_ptr(), q.data_ptr(), n, scale, dlt,
                                  dx.data_ptr() if dx is not None else None, stream), "crom_backward")
        L.check(lib.crom_wgrad(acts, dlt, n, self.in_dim, self.out_dim, ws, self.grads.data_ptr(), stream),
                "crom_wgrad")
        return y, dx

    def adam_step_(self, lr: float, betas=(0.9, 0.999), eps: float = 1e-8):
        """torch.optim.Adam's update with the current self.grads
        (CROMnet.py:193-196), then re-pack."""
        lib = L.load()
        self.adam_step += 1
        L.check(lib.crom_adam(self.params.data_ptr(), self.grads.data_ptr(), self.adam_m.data_ptr(),
                              self.adam_v.data_ptr(), self.n_params, float(lr), float(betas[0]), float(betas[1]),
                              float(eps), self.adam_step, torch.cuda.current_stream(self.device).cuda_stream),
                "crom_adam")
        self._dirty = True

    def train_step(self, x: torch.Tensor, q: torch.Tensor, lr: float, loss_weight: float = 1000.0):
        """One optimizer step on (x, q); returns the loss (device scalar)."""
        y, _ = self.backward(x, q, loss_weight)
        loss = loss_weight * torch.mean((y - q.to(y.device)) ** 2)
        self.adam_step_(lr)
        return loss
