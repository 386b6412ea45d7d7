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
import time

import numpy as np

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

    def jacobian(self, x: torch.Tensor) -> torch.Tensor:
        """d q / d x per point, (n, out_dim, in_dim): NetDecFuncGrad
        (CROMnet.py:319-355), which the reference checks against autograd
        (Exporter.py:104-106).  One pass of the backward chain per output
        component c, with the output gradient e_c (crom_backward of the loss
        1/2 sum (y - q)^2 at q = y - e_c)."""
        lib = L.load()
        x = x.to(self.device, torch.float32).contiguous()
        n = x.shape[0]
        if x.dim() != 2 or x.shape[1] != self.in_dim:
            raise ValueError(f"x must be (n, {self.in_dim})")
        if self._dirty:
            self.pack()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        nb = int(lib.crom_acts_bytes(n))
        acts = self._buffer("acts", nb)
        dlt = self._buffer("grads_buf", nb)
        y = torch.empty((n, self.out_dim), dtype=torch.float32, device=self.device)
        L.check(lib.crom_forward(self._packed, x.data_ptr(), n, y.data_ptr(), acts, stream), "crom_forward")
        J = torch.empty((n, self.out_dim, self.in_dim), dtype=torch.float32, device=self.device)
        dx = torch.empty((n, self.in_dim), dtype=torch.float32, device=self.device)
        for c in range(self.out_dim):
            q = y.clone()
            q[:, c] -= 1.0
            L.check(lib.crom_backward(self._packed, acts, y.data_ptr(), q.data_ptr(), n, 1.0, dlt, dx.data_ptr(),
                                      stream), "crom_backward")
            J[:, c, :] = dx
        return J

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


# ---------------------------------------------------------------------------
# Encoder and the full network (CROMnet)
# ---------------------------------------------------------------------------
def conv_lengths(npoints: int, ks: int = 6, strides: int = 4, o_dim: int = 3):
    """Input lengths of NetEnc's Conv1d layers and the flattened length
    (CROMnet.py:373-384, util.py:201-203: stack while o_dim * L_out >= 32)."""
    lens = [int(npoints)]
    while True:
        l_out = (lens[-1] - (ks - 1) - 1) // strides + 1
        if o_dim * l_out >= 32:
            lens.append(l_out)
        else:
            break
    return lens


class CromEncoder:
    """NetEnc (CROMnet.py:361-438): standardizeQ, Conv1d(3->3, ks, stride) +
    ELU while 3 * L_out >= 32, flatten, Linear -> 32 + ELU, Linear -> lbl +
    ELU.  FP32 kernels (crom_conv_*, crom_dense_*)."""

    def __init__(self, npoints: int, lbl: int, ks: int = 6, strides: int = 4, device="cuda:0", seed: int = 0):
        self.npoints, self.lbl, self.ks, self.strides = int(npoints), int(lbl), int(ks), int(strides)
        self.device = torch.device(device)
        self.lens = conv_lengths(self.npoints, ks, strides)
        self.n_conv = len(self.lens) - 1
        if self.n_conv < 1:
            raise ValueError("CromEncoder: too few points for one Conv1d layer")
        self.flat = 3 * self.lens[-1]
        sizes = [3 * 3 * ks + 3] * self.n_conv + [32 * self.flat + 32, lbl * 32 + lbl]
        self.n_params = sum(sizes)
        self.params = torch.empty(self.n_params, dtype=torch.float32, device=self.device)
        self.grads = torch.zeros_like(self.params)
        self.adam_m = torch.zeros_like(self.params)
        self.adam_v = torch.zeros_like(self.params)
        self.adam_step = 0
        g = torch.Generator().manual_seed(seed)
        off = 0
        self.W, self.b = [], []
        shapes = [((3, 3, ks), 3 * ks, 3 * ks)] * self.n_conv + [((32, self.flat), self.flat, 32),
                                                                  ((lbl, 32), 32, lbl)]
        for shp, fan_in, fan_out in shapes:
            nw = int(np.prod(shp))
            w = self.params[off:off + nw].view(*shp)
            off += nw
            b = self.params[off:off + shp[0]]
            off += shp[0]
            # xavier_uniform_ (CROMnet.py:412-421); torch's default bias init
            a = math.sqrt(6.0 / (fan_in + fan_out)) if len(shp) == 2 else math.sqrt(6.0 / (3 * ks + 3 * ks))
            w.copy_(((torch.rand(shp, generator=g) * 2 - 1) * a).to(self.device))
            b.copy_(((torch.rand((shp[0],), generator=g) * 2 - 1) / math.sqrt(fan_in)).to(self.device))
            self.W.append(w)
            self.b.append(b)
        self.mean_q = torch.zeros(3, device=self.device)
        self.std_q = torch.ones(3, device=self.device)
        self._acts = None

    # enc10 acts on the flattened [position][channel] order of the reference
    # (CROMnet.py:427-429); the conv kernels produce [channel][position], so
    # the kernel multiplies with enc10's columns taken in that order.
    def _enc10_kernel_weight(self):
        L = self.lens[-1]
        return self.W[-2].view(32, L, 3).permute(0, 2, 1).contiguous().view(32, 3 * L)

    def set_preprop(self, mean_q, std_q):
        self.mean_q = torch.as_tensor(mean_q, dtype=torch.float32).reshape(3).to(self.device)
        self.std_q = torch.as_tensor(std_q, dtype=torch.float32).reshape(3).to(self.device)

    def forward(self, q: torch.Tensor) -> torch.Tensor:
        """q (B, npoints, 3) -> latents (B, lbl)."""
        lib = L.load()
        q = q.to(self.device, torch.float32).contiguous()
        B = q.shape[0]
        if q.shape[1:] != (self.npoints, 3):
            raise ValueError(f"q must be (B, {self.npoints}, 3)")
        stream = torch.cuda.current_stream(self.device).cuda_stream
        qis = (1.0 / self.std_q).contiguous()
        acts = [q]
        for k in range(self.n_conv):
            out = torch.empty((B, 3, self.lens[k + 1]), dtype=torch.float32, device=self.device)
            first = k == 0
            L.check(lib.crom_conv_forward(acts[-1].data_ptr(), B, self.lens[k], self.ks, self.strides,
                                          self.W[k].data_ptr(), self.b[k].data_ptr(), out.data_ptr(),
                                          self.mean_q.data_ptr() if first else None,
                                          qis.data_ptr() if first else None, stream), "crom_conv_forward")
            acts.append(out)
        w10 = self._enc10_kernel_weight()
        h = torch.empty((B, 32), dtype=torch.float32, device=self.device)
        L.check(lib.crom_dense_forward(acts[-1].data_ptr(), B, self.flat, 32, w10.data_ptr(), self.b[-2].data_ptr(),
                                       h.data_ptr(), stream), "crom_dense_forward")
        z = torch.empty((B, self.lbl), dtype=torch.float32, device=self.device)
        L.check(lib.crom_dense_forward(h.data_ptr(), B, 32, self.lbl, self.W[-1].data_ptr(), self.b[-1].data_ptr(),
                                       z.data_ptr(), stream), "crom_dense_forward")
        self._acts = (acts, h, z, w10, qis)
        return z

    __call__ = forward

    def backward(self, g_z: torch.Tensor):
        """Gradients of the last forward's parameters for d(loss)/d(latents)."""
        lib = L.load()
        acts, h, z, w10, qis = self._acts
        B = z.shape[0]
        stream = torch.cuda.current_stream(self.device).cuda_stream
        g_z = g_z.to(self.device, torch.float32).contiguous()
        sizes = [3 * 3 * self.ks + 3] * self.n_conv + [32 * self.flat + 32, self.lbl * 32 + self.lbl]
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        gW = lambda k: self.grads[offs[k]:offs[k] + sizes[k] - self.W[k].shape[0]]
        gB = lambda k: self.grads[offs[k] + sizes[k] - self.W[k].shape[0]:offs[k + 1]]
        g_h = torch.empty_like(h)
        L.check(lib.crom_dense_backward(h.data_ptr(), z.data_ptr(), g_z.data_ptr(), B, 32, self.lbl,
                                        self.W[-1].data_ptr(), g_h.data_ptr(), gW(self.n_conv + 1).data_ptr(),
                                        gB(self.n_conv + 1).data_ptr(), stream), "crom_dense_backward")
        g_flat = torch.empty_like(acts[-1])
        g10 = torch.empty((32, self.flat), dtype=torch.float32, device=self.device)
        L.check(lib.crom_dense_backward(acts[-1].data_ptr(), h.data_ptr(), g_h.data_ptr(), B, self.flat, 32,
                                        w10.data_ptr(), g_flat.data_ptr(), g10.data_ptr(),
                                        gB(self.n_conv).data_ptr(), stream), "crom_dense_backward")
        Lf = self.lens[-1]
        gW(self.n_conv).view(32, Lf, 3).copy_(g10.view(32, 3, Lf).permute(0, 2, 1))
        g_out = g_flat
        for k in range(self.n_conv - 1, -1, -1):
            first = k == 0
            g_in = None if first else torch.empty_like(acts[k])
            L.check(lib.crom_conv_backward(acts[k].data_ptr(), acts[k + 1].data_ptr(), g_out.data_ptr(), B,
                                           self.lens[k], self.ks, self.strides, self.W[k].data_ptr(),
                                           g_in.data_ptr() if g_in is not None else None, gW(k).data_ptr(),
                                           gB(k).data_ptr(), self.mean_q.data_ptr() if first else None,
                                           qis.data_ptr() if first else None, stream), "crom_conv_backward")
            g_out = g_in

    def adam_step_(self, lr: float, betas=(0.9, 0.999), eps: float = 1e-8):
        lib = L.load()
        self.adam_step += 1
        L.check(lib.crom_adam(self.params.data_ptr(), self.grads.data_ptr(), self.adam_m.data_ptr(),
                              self.adam_v.data_ptr(), self.n_params, float(lr), float(betas[0]), float(betas[1]),
                              float(eps), self.adam_step, torch.cuda.current_stream(self.device).cuda_stream),
                "crom_adam")


def lr_factor(epoch: int, learning_rates=(10.0, 5.0, 2.0, 1.0), epochs=(400, 200, 100, 50)) -> float:
    """LambdaLR factor of CROMnet.adaptiveLRfromRange (CROMnet.py:183-191,
    util.py:187-194; run.py:33-34 defaults)."""
    acc = np.concatenate([[0], np.cumsum(epochs)])
    for i in range(len(acc) - 1):
        if acc[i] <= epoch < acc[i + 1]:
            return float(learning_rates[i])
    if epoch == acc[-1]:
        return float(learning_rates[-1])
    raise ValueError("epoch outside the schedule")


class Crom:
    """CROMnet (CROMnet.py:20-153): encoder -> latent per frame -> decoder at
    every point of the frame, trained with 1000 * MSE (CROMnet.py:70-83) by
    Adam at initial_lr * LambdaLR factor."""

    def __init__(self, npoints: int, lbl: int = 16, ks: int = 6, strides: int = 4, device="cuda:0", seed: int = 0):
        self.encoder = CromEncoder(npoints, lbl, ks, strides, device=device, seed=seed)
        self.decoder = CromDecoder(lbl, out_dim=3, device=device, seed=seed + 1)
        self.lbl, self.npoints, self.device = lbl, npoints, torch.device(device)

    def set_preprop(self, mean_x, std_x, mean_q, std_q):
        self.encoder.set_preprop(mean_q, std_q)
        self.decoder.set_preprop(mean_x, std_x, mean_q, std_q)

    def _decoder_input(self, z: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        B, P = x.shape[0], x.shape[1]
        return torch.cat([z[:, None, :].expand(B, P, self.lbl), x], 2).reshape(B * P, self.lbl + 3)

    def forward(self, q: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        """CROMnet.forward (CROMnet.py:130-150): q, x (B, P, 3) -> q_hat."""
        z = self.encoder(q)
        return self.decoder(self._decoder_input(z, x.to(self.device))).view(x.shape[0], x.shape[1], 3)

    def train_step(self, q: torch.Tensor, x: torch.Tensor, lr: float) -> torch.Tensor:
        """One optimizer step on a batch of frames q, x (B, P, 3); returns the
        loss 1000 * MSE (device scalar)."""
        lib = L.load()
        q = q.to(self.device, torch.float32).contiguous()
        B, P = q.shape[0], q.shape[1]
        z = self.encoder(q)
        xin = self._decoder_input(z, x.to(self.device, torch.float32))
        y, dx = self.decoder.backward(xin, q.reshape(B * P, 3), input_grad=True)
        g_z = torch.empty((B, self.lbl), dtype=torch.float32, device=self.device)
        L.check(lib.crom_latent_grad(dx.data_ptr(), B, P, self.lbl + 3, self.lbl, g_z.data_ptr(),
                                     torch.cuda.current_stream(self.device).cuda_stream), "crom_latent_grad")
        self.encoder.backward(g_z)
        self.decoder.adam_step_(lr)
        self.encoder.adam_step_(lr)
        return 1000.0 * torch.mean((y - q.reshape(B * P, 3)) ** 2)


VOIGT = [0, 1, 2, 4, 5, 8]     # stress (N, 9) row-major -> [xx, xy, xz, yy, yz, zz] (mpm_solver_warp.py:383)


def solver_snapshots(scene, friction_angle: float, n_frames: int, substeps_per_frame: int, dt: float,
                     device="cuda:0", n_points: int | None = None, with_stress: bool = False):
    """Snapshots of one sand run by the MPM solver (the training data the
    reference writes with save_data_at_frame, mpm_solver_warp.py:312-404,
    generate_sand.py:27-32), frame 0 = initial: displacements q = u
    (frames, P, 3) of the first P = n_points particles (default all), and with
    with_stress also the Voigt stress (frames, P, 6)."""
    from . import scenes as S
    import copy
    sc = copy.copy(scene)
    sc.friction_angle = float(friction_angle)
    sim = S.make_simulator(sc, device=device)
    P = int(n_points or sc.n_particles)
    dev = torch.device(device)
    out = torch.empty((n_frames, P, 3), dtype=torch.float32, device=dev)
    st = torch.empty((n_frames, P, 6), dtype=torch.float32, device=dev) if with_stress else None
    for f in range(n_frames):
        u, _, _, tau = sim.export_uvFstress_from_warp()
        out[f] = u[:P]
        if with_stress:
            st[f] = tau[:P].reshape(P, 9)[:, VOIGT]
        if f + 1 < n_frames:
            sim.run(substeps_per_frame, dt)
    return (out, st) if with_stress else out


# ---------------------------------------------------------------------------
# CROM-reduced runs: a sampled MPM run + the decoder, checked against the full run
# ---------------------------------------------------------------------------
def sampled_scene(scene, n_sample: int, n_grid: int):
    """The first n_sample particles of a (pre-shuffled) scene, each standing
    for N / n_sample of them (volume scaled, so mass and volume are
    conserved), on an n_grid^3 grid -- the points the encoder reads
    (SimulationDataset.py:36-43 trains on the first N/10 points of a frame)."""
    import copy
    sc = copy.copy(scene)
    scale = scene.n_particles / float(n_sample)
    sc.x0 = np.ascontiguousarray(scene.x0[:n_sample])
    sc.v0 = np.ascontiguousarray(scene.v0[:n_sample])
    sc.vol = np.ascontiguousarray(scene.vol[:n_sample] * scale).astype(np.float32)
    sc.n_grid = int(n_grid)
    return sc


class ReducedRun:
    """A sand run on the CROM-reduced path: the MPM solver advances only the
    encoder's sample of particles (on a coarser grid); every frame the
    encoder turns the sample's displacements into a latent and the decoder
    reconstructs the displacement of all N particles.  The plain MPM step is
    the fallback and the judge of the error, in one of two ways:

    * verify=True: the full-order run advances in lockstep, each frame's
      relative L2 error against it is recorded, and from the first frame
      whose error exceeds `tol` the frames are the full run's own (the
      measurement mode: it costs the full run);
    * check_every=K (no lockstep run): every K frames the decoded
      displacements of the sample points are compared with the sampled MPM
      run's own; from the first check above `self_tol` the full-order run is
      replayed from frame 0 to that frame once and continues in place of the
      reduced one.  This catches the decoder leaving its training manifold;
      the sampled run's own drift from the full run is what verify measures.

        rr = ReducedRun(crom, scene, friction_angle, n_sample=N // 10, coarse_grid=64)
        out = rr.run(n_frames, substeps_per_frame, dt, check_every=5, self_tol=0.1)
        out["frames"]  (n_frames, N, 3) displacements; out["source"] "crom" | "mpm" per frame

    With a stress_head (crom_pipeline.StressHead) each reduced frame's latent
    also gives the Voigt stress of all N particles (out["stress"]), and
    verify records its relative L2 error against the full run's stress
    (out["stress_errors"]).
    """

    def __init__(self, crom: Crom, scene, friction_angle: float, n_sample: int | None = None,
                 coarse_grid: int | None = None, device="cuda:0", stress_head=None):
        self.crom, self.scene, self.phi = crom, scene, float(friction_angle)
        self.stress_head = stress_head
        self.n_sample = int(n_sample or crom.npoints)
        if self.n_sample != crom.npoints:
            raise ValueError("n_sample must equal the encoder's number of points")
        self.coarse_grid = int(coarse_grid or max(16, scene.n_grid // 2))
        self.device = torch.device(device)
        self.X = torch.from_numpy(np.ascontiguousarray(scene.x0, dtype=np.float32)).to(self.device)

    def _sim(self, scene):
        import copy
        from . import scenes as S
        sc = copy.copy(scene)
        sc.friction_angle = self.phi
        return S.make_simulator(sc, device=self.device)

    def decode(self, u_sample: torch.Tensor, with_latent: bool = False):
        """Displacements of all N particles from the sample's (n_sample, 3)
        (and the latent (1, lbl) with with_latent)."""
        z = self.crom.encoder(u_sample[None].contiguous())
        n = self.X.shape[0]
        xin = torch.cat([z.expand(n, self.crom.lbl), self.X], 1).contiguous()
        q = self.crom.decoder(xin)
        return (q, z) if with_latent else q

    def run(self, n_frames: int, substeps_per_frame: int, dt: float, verify: bool = False,
            tol: float | None = None, check_every: int | None = None, self_tol: float | None = None) -> dict:
        coarse = self._sim(sampled_scene(self.scene, self.n_sample, self.coarse_grid))
        full = self._sim(self.scene) if verify else None
        frames = torch.empty((n_frames, self.X.shape[0], 3), dtype=torch.float32, device=self.device)
        head = self.stress_head
        stress = (torch.empty((n_frames, self.X.shape[0], 6), dtype=torch.float32, device=self.device)
                  if head is not None else None)
        source, errors, checks, stress_errors = [], [], [], []
        fallen = False
        t_reduced = t_full = 0.0
        P = self.n_sample
        for f in range(n_frames):
            if f > 0:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                if not fallen:
                    coarse.run(substeps_per_frame, dt)
                    torch.cuda.synchronize()
                t1 = time.perf_counter()
                if full is not None:
                    full.run(substeps_per_frame, dt)
                    torch.cuda.synchronize()
                t2 = time.perf_counter()
                t_reduced += t1 - t0
                if verify:
                    t_full += t2 - t1
                else:
                    t_reduced += t2 - t1       # the fallen-back run's own steps
            q = s_hat = None
            if not fallen:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                u_s = coarse.export_uvFstress_from_warp()[0]
                q, z = self.decode(u_s, with_latent=True)
                if head is not None:
                    s_hat = head.predict(z, self.X, self.phi)[0]
                if check_every and f > 0 and f % check_every == 0:
                    den = float(u_s.norm())
                    e = float((q[:P] - u_s).norm()) / den if den > 0 else 0.0
                    checks.append((f, e))
                    if self_tol is not None and e > self_tol and not verify:
                        # the full-order run replayed to this frame, once
                        fallen = True
                        full = self._sim(self.scene)
                        full.run(f * substeps_per_frame, dt)
                torch.cuda.synchronize()
                t_reduced += time.perf_counter() - t0
            u_true = tau_true = None
            if full is not None:
                u_true, _, _, tau_true = full.export_uvFstress_from_warp()
                tau_true = tau_true.reshape(-1, 9)[:, VOIGT] if head is not None else None
            if verify and u_true is not None:
                den = float(u_true.norm())
                err = float((q - u_true).norm()) / den if (den > 0 and q is not None) else 0.0
                errors.append(err)
                if s_hat is not None:
                    den = float(tau_true.norm())
                    stress_errors.append(float((s_hat - tau_true).norm()) / den if den > 0 else 0.0)
                if tol is not None and not fallen and err > tol:
                    fallen = True
            if fallen:
                frames[f] = u_true
                if stress is not None:
                    stress[f] = tau_true
                source.append("mpm")
            else:
                frames[f] = q
                if stress is not None:
                    stress[f] = s_hat
                source.append("crom")
        return {"frames": frames, "stress": stress, "source": source, "errors": errors if verify else None,
                "stress_errors": stress_errors if (verify and head is not None) else None, "checks": checks,
                "reduced_seconds": t_reduced, "full_seconds": t_full if verify else None}
