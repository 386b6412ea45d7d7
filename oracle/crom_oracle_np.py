"""FP64 NumPy restatement of the CROM network (decoder, encoder, Adam) with
hand-written backward passes.

TEST INFRASTRUCTURE ONLY: this module is a checker.  Only tests/ may import
it; the product package never does.

The GPU kernels (libmpm_hip.so, include/crom_hip.h) are checked against it;
the module itself is pinned on the CPU by central finite differences of its
own forward passes and by torch autograd (tests/test_crom_oracle.py).

`bf16=True` reproduces the decoder kernels' roundings: the prepared input,
every weight matrix, every hidden activation and every back-propagated
pre-activation gradient rounded to BF16 (round to nearest even), all sums in
FP64 (the kernels accumulate in FP32).

Reference citations (zeshunzong/reduced_mpm_stress_training):
  NetDec.forward             offline_training_u/CROMnet.py:293-296
  NetDecFuncGrad (Jacobian)  offline_training_u/CROMnet.py:319-355
  five Linear+ELU layers     offline_training_u/CROMnet.py:226-237
  Prepare (standardize X)    offline_training_u/CROMnet.py:497-534
  Activation (ELU, alpha=1)  offline_training_u/CROMnet.py:457-473
  invStandardizeQ            offline_training_u/CROMnet.py:475-487
  loss = 1000 * MSE          offline_training_u/CROMnet.py:70-83
  NetEnc                     offline_training_u/CROMnet.py:361-438
  Conv1d stack length        offline_training_u/util.py:201-203
  standardizeQ               offline_training_u/CROMnet.py:440-455
  Adam (torch.optim.Adam)    offline_training_u/CROMnet.py:193-196
"""
from __future__ import annotations

import numpy as np


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def elu(z):
    """ELU with alpha = 1 (CROMnet.py:471-472)."""
    return np.where(z > 0, z, np.expm1(np.minimum(z, 0.0)))


def elu_grad_from_output(a):
    """d ELU / dz expressed through a = ELU(z): 1 for z > 0, else a + 1."""
    return np.where(a > 0, 1.0, a + 1.0)


# ---------------------------------------------------------------------------
# decoder (NetDec)
# ---------------------------------------------------------------------------
def decoder_forward(W, b, x, lbl, mean_x, std_x, mean_q, std_q, bf16_ops=False):
    """NetDec: x (n, lbl + 3 [+ mu]) -> q (n, out).  W[k] (out_k, in_k), b[k].
    Returns (q, cache)."""
    r = bf16 if bf16_ops else (lambda t: np.asarray(t, np.float64))
    x = np.array(x, dtype=np.float64)
    x[:, lbl:lbl + 3] = (x[:, lbl:lbl + 3] - np.asarray(mean_x, np.float64)) / np.asarray(std_x, np.float64)
    Wr = [r(w) for w in W]
    acts = [r(x)]
    for k in range(len(W) - 1):
        acts.append(r(elu(acts[-1] @ Wr[k].T + np.asarray(b[k], np.float64))))
    y = acts[-1] @ Wr[-1].T + np.asarray(b[-1], np.float64)
    q = np.asarray(mean_q, np.float64) + y * np.asarray(std_q, np.float64)
    return q, dict(acts=acts, W=Wr, std_q=np.asarray(std_q, np.float64), std_x=np.asarray(std_x, np.float64),
                   lbl=lbl, r=r)


def decoder_backward(cache, g_q):
    """Gradients of a scalar with d/dq = g_q: ([gW], [gb], g_x)."""
    acts, W, r = cache["acts"], cache["W"], cache["r"]
    n_lin = len(W)
    dz = r(np.asarray(g_q, np.float64) * cache["std_q"])
    gW, gb = [None] * n_lin, [None] * n_lin
    g = None
    for k in range(n_lin - 1, -1, -1):
        gW[k] = dz.T @ acts[k]
        gb[k] = dz.sum(0)
        g = dz @ W[k]
        if k > 0:
            dz = r(g * elu_grad_from_output(acts[k]))
    g_x = g.copy()
    lbl = cache["lbl"]
    g_x[:, lbl:lbl + 3] /= cache["std_x"]
    return gW, gb, g_x


def decoder_jacobian(W, b, x, lbl, mean_x, std_x, mean_q, std_q, bf16_ops=False):
    """d q / d x per point, (n, out, in): one backward pass per output
    component (NetDecFuncGrad, CROMnet.py:319-355)."""
    q, cache = decoder_forward(W, b, x, lbl, mean_x, std_x, mean_q, std_q, bf16_ops)
    n, out = q.shape
    J = np.empty((n, out, np.asarray(x).shape[1]))
    for c in range(out):
        g = np.zeros_like(q)
        g[:, c] = 1.0
        J[:, c, :] = decoder_backward(cache, g)[2]
    return J


def decoder_loss_grads(W, b, x, q_target, lbl, mean_x, std_x, mean_q, std_q, loss_weight=1000.0,
                       bf16_ops=False):
    """loss_weight * MSE(q, q_target) (CROMnet.py:76) and its gradients."""
    q, cache = decoder_forward(W, b, x, lbl, mean_x, std_x, mean_q, std_q, bf16_ops)
    t = np.asarray(q_target, np.float64)
    loss = loss_weight * np.mean((q - t) ** 2)
    gW, gb, gx = decoder_backward(cache, 2.0 * loss_weight * (q - t) / t.size)
    return loss, gW, gb, gx


# ---------------------------------------------------------------------------
# encoder (NetEnc)
# ---------------------------------------------------------------------------
def conv1d(s, Wc, bc, stride):
    """Conv1d(3 -> 3): s (B, 3, L_in), Wc (3, 3, ks) -> (B, 3, L_out)."""
    B, C, L_in = s.shape
    ks = Wc.shape[2]
    L_out = (L_in - ks) // stride + 1
    idx = np.arange(L_out)[:, None] * stride + np.arange(ks)[None, :]       # (L_out, ks)
    patches = s[:, :, idx]                                                    # (B, C, L_out, ks)
    return np.einsum("bclk,ock->bol", patches, Wc) + bc[None, :, None]


def encoder_forward(convs, dense, q, mean_q, std_q, stride):
    """NetEnc: q (B, P, 3) -> z (B, lbl).  convs: [(W (3, 3, ks), b (3))];
    dense: [(W10 (32, flat), b10), (W11 (lbl, 32), b11)], flat in the
    reference's [position][channel] order (CROMnet.py:427-429)."""
    s = (np.asarray(q, np.float64) - np.asarray(mean_q, np.float64)) / np.asarray(std_q, np.float64)
    s = s.transpose(0, 2, 1)                       # (B, 3, P)
    acts = [s]
    for Wc, bc in convs:
        acts.append(elu(conv1d(acts[-1], np.asarray(Wc, np.float64), np.asarray(bc, np.float64), stride)))
    B = s.shape[0]
    flat = acts[-1].transpose(0, 2, 1).reshape(B, -1)
    (W10, b10), (W11, b11) = [(np.asarray(w, np.float64), np.asarray(c, np.float64)) for w, c in dense]
    h = elu(flat @ W10.T + b10)
    z = elu(h @ W11.T + b11)
    return z, dict(acts=acts, flat=flat, h=h, z=z, convs=convs, W10=W10, W11=W11, stride=stride,
                   std_q=np.asarray(std_q, np.float64))


def encoder_backward(cache, g_z):
    """Gradients of a scalar with d/dz = g_z: ([(gW, gb) per conv], [(gW10, gb10), (gW11, gb11)], g_q)."""
    acts, h, z, stride = cache["acts"], cache["h"], cache["z"], cache["stride"]
    dz = np.asarray(g_z, np.float64) * elu_grad_from_output(z)
    gW11, gb11 = dz.T @ h, dz.sum(0)
    dh = (dz @ cache["W11"]) * elu_grad_from_output(h)
    gW10, gb10 = dh.T @ cache["flat"], dh.sum(0)
    g_flat = dh @ cache["W10"]
    B = g_flat.shape[0]
    g = g_flat.reshape(B, acts[-1].shape[2], 3).transpose(0, 2, 1)     # (B, 3, L)
    conv_grads = [None] * len(cache["convs"])
    for k in range(len(cache["convs"]) - 1, -1, -1):
        Wc = np.asarray(cache["convs"][k][0], np.float64)
        ks = Wc.shape[2]
        dzc = g * elu_grad_from_output(acts[k + 1])                     # (B, 3, L_out)
        s = acts[k]
        L_out = dzc.shape[2]
        idx = np.arange(L_out)[:, None] * stride + np.arange(ks)[None, :]
        patches = s[:, :, idx]                                          # (B, C, L_out, ks)
        conv_grads[k] = (np.einsum("bol,bclk->ock", dzc, patches), dzc.sum((0, 2)))
        gin = np.zeros_like(s)
        contrib = np.einsum("bol,ock->bclk", dzc, Wc)                   # (B, C, L_out, ks)
        for j in range(ks):
            np.add.at(gin, (slice(None), slice(None), idx[:, j]), contrib[:, :, :, j])
        g = gin
    g_q = g.transpose(0, 2, 1) / cache["std_q"]
    return conv_grads, [(gW10, gb10), (gW11, gb11)], g_q


# ---------------------------------------------------------------------------
# Adam (torch.optim.Adam, no weight decay, no amsgrad)
# ---------------------------------------------------------------------------
def adam_step(p, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update at step t >= 1 (in place on p, m, v)."""
    m *= beta1
    m += (1.0 - beta1) * g
    v *= beta2
    v += (1.0 - beta2) * g * g
    mhat = m / (1.0 - beta1 ** t)
    vhat = v / (1.0 - beta2 ** t)
    p -= lr * mhat / (np.sqrt(vhat) + eps)
    return p
