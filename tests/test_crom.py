"""CROM decoder MLP (NetDec, offline_training_u/CROMnet.py:218-301) on the
BF16 MFMA kernels against a plain PyTorch FP32 reference of the same network.

Tolerances: against the reference evaluated with the kernel's BF16 roundings
(weights, prepared inputs and every hidden activation rounded to BF16, FP32
accumulation) relative L2 <= 5e-3 (accumulation order and rare one-ulp BF16
flips); against the pure FP32 reference <= 3e-2 (BF16 has 8 mantissa bits).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def torch_reference(dec, x, bf16=False):
    """The reference's NetDec forward in FP32 on the CPU (CROMnet.py:293-296,
    Prepare :515-534, ELU :471-472, invStandardizeQ :486-487)."""
    rnd = (lambda t: t.to(torch.bfloat16).to(torch.float32)) if bf16 else (lambda t: t)
    x = x.detach().cpu().to(torch.float32).clone()
    mx, sx = dec.mean_x.cpu(), dec.std_x.cpu()
    x[:, dec.lbl:dec.lbl + 3] = (x[:, dec.lbl:dec.lbl + 3] - mx) / sx
    a = rnd(x)
    for k in range(6):
        w, b = rnd(dec.weights[k].cpu()), dec.biases[k].cpu()
        z = a @ w.t() + b
        a = rnd(torch.nn.functional.elu(z)) if k < 5 else z
    return dec.mean_q.cpu() + a * dec.std_q.cpu()


def _rel(a, b):
    a, b = a.double(), b.double()
    return float((a - b).norm() / b.norm())


@pytest.mark.parametrize("lbl,mu_dim,out_dim,n", [(16, 0, 3, 1000), (16, 1, 6, 4133), (8, 0, 3, 32), (16, 0, 3, 200000)])
def test_forward_matches_torch(lbl, mu_dim, out_dim, n):
    from reduced_mpm_stress_training_amd.crom import CromDecoder
    torch.manual_seed(0)
    dec = CromDecoder(lbl, out_dim=out_dim, mu_dim=mu_dim, seed=1)
    dec.set_preprop(mean_x=[0.5, 0.4, 0.3], std_x=[0.1, 0.2, 0.15],
                    mean_q=torch.linspace(-0.01, 0.01, out_dim), std_q=torch.linspace(0.02, 0.05, out_dim))
    x = torch.randn(n, dec.in_dim)
    x[:, lbl:lbl + 3] = torch.rand(n, 3)
    y = dec(x.cuda()).cpu()
    torch.cuda.synchronize()
    y_bf = torch_reference(dec, x, bf16=True)
    y_32 = torch_reference(dec, x, bf16=False)
    e_bf, e_32 = _rel(y - dec.mean_q.cpu(), y_bf - dec.mean_q.cpu()), _rel(y - dec.mean_q.cpu(), y_32 - dec.mean_q.cpu())
    print(n, e_bf, e_32)
    assert torch.isfinite(y).all()
    assert e_bf <= 5e-3 and e_32 <= 3e-2, (e_bf, e_32)
