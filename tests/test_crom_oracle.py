"""Pins of the CROM oracle (oracle/crom_oracle_np.py) on the CPU.

The oracle's hand-written backward passes are checked against central finite
differences of its own forward passes (FP64, relative step 1e-6: agreement
to 1e-6) and against torch autograd in FP64 (an independent implementation:
agreement to 1e-9).  Its BF16 rounding is checked bit for bit against
torch's float32 -> bfloat16 conversion, its Adam against torch.optim.Adam.
A dropped ELU derivative, a transposed weight or a wrong stride in the
Conv1d backward fails these.
"""
import numpy as np
import torch

from oracle import crom_oracle_np as O


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _decoder(rng, dims):
    W = [rng.normal(0, 1.0 / np.sqrt(dims[k]), (dims[k + 1], dims[k])) for k in range(len(dims) - 1)]
    b = [rng.normal(0, 0.1, dims[k + 1]) for k in range(len(dims) - 1)]
    return W, b


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.normal(0, 10, 100000), [1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0]])
    a = a.astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(O.bf16(a), ref)
    assert O.bf16(np.float32(1 + 2 ** -8)) == 1.0                  # tie to even
    assert O.bf16(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6      # tie to even (up)


def test_decoder_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    lbl, dims, n = 2, [5, 7, 6, 3], 9
    W, b = _decoder(rng, dims)
    x = rng.normal(size=(n, dims[0]))
    t = rng.normal(size=(n, 3))
    mx, sx, mq, sq = [0.1, -0.2, 0.3], [0.5, 1.5, 0.8], [0.01, 0.0, -0.02], [0.2, 0.3, 0.1]
    f = lambda W_, b_, x_: O.decoder_loss_grads(W_, b_, x_, t, lbl, mx, sx, mq, sq)[0]
    _, gW, gb, gx = O.decoder_loss_grads(W, b, x, t, lbl, mx, sx, mq, sq)
    h = 1e-6
    for k in range(len(W)):
        for (i, j) in [(0, 0), (1, 2), (W[k].shape[0] - 1, W[k].shape[1] - 1)]:
            Wp = [w.copy() for w in W]; Wp[k][i, j] += h
            Wm = [w.copy() for w in W]; Wm[k][i, j] -= h
            fd = (f(Wp, b, x) - f(Wm, b, x)) / (2 * h)
            assert abs(fd - gW[k][i, j]) <= 1e-6 * max(1.0, abs(fd)), (k, i, j, fd, gW[k][i, j])
        bp = [c.copy() for c in b]; bp[k][1] += h
        bm = [c.copy() for c in b]; bm[k][1] -= h
        fd = (f(W, bp, x) - f(W, bm, x)) / (2 * h)
        assert abs(fd - gb[k][1]) <= 1e-6 * max(1.0, abs(fd))
    for (i, j) in [(0, 0), (3, lbl), (n - 1, lbl + 2)]:      # a latent and two position columns
        xp = x.copy(); xp[i, j] += h
        xm = x.copy(); xm[i, j] -= h
        fd = (f(W, b, xp) - f(W, b, xm)) / (2 * h)
        assert abs(fd - gx[i, j]) <= 1e-6 * max(1.0, abs(fd)), (i, j, fd, gx[i, j])


def test_decoder_matches_torch_autograd():
    rng = np.random.default_rng(2)
    lbl, n = 16, 60
    dims = [19, 128, 128, 128, 128, 128, 3]
    W, b = _decoder(rng, dims)
    x = rng.normal(size=(n, 19))
    t = 0.05 * rng.normal(size=(n, 3))
    mx, sx, mq, sq = [0.5, 0.4, 0.3], [0.1, 0.2, 0.15], [0.0, 0.01, 0.0], [0.05, 0.04, 0.06]
    loss, gW, gb, gx = O.decoder_loss_grads(W, b, x, t, lbl, mx, sx, mq, sq)
    Wt = [torch.tensor(w, requires_grad=True) for w in W]
    bt = [torch.tensor(c, requires_grad=True) for c in b]
    xt = torch.tensor(x, requires_grad=True)
    xs = torch.cat([xt[:, :lbl], (xt[:, lbl:lbl + 3] - torch.tensor(mx, dtype=torch.float64)) /
                    torch.tensor(sx, dtype=torch.float64), xt[:, lbl + 3:]], 1)
    a = xs
    for k in range(5):
        a = torch.nn.functional.elu(a @ Wt[k].T + bt[k])
    q = torch.tensor(mq, dtype=torch.float64) + (a @ Wt[5].T + bt[5]) * torch.tensor(sq, dtype=torch.float64)
    lt = 1000.0 * torch.mean((q - torch.tensor(t)) ** 2)
    lt.backward()
    assert abs(float(lt.detach()) - loss) <= 1e-12 * abs(loss)
    for k in range(6):
        assert _rel(gW[k], Wt[k].grad.numpy()) <= 1e-9
        assert _rel(gb[k], bt[k].grad.numpy()) <= 1e-9
    assert _rel(gx, xt.grad.numpy()) <= 1e-9


def _encoder(rng, P, lbl, ks=6, stride=4):
    lens = [P]
    while 3 * ((lens[-1] - ks) // stride + 1) >= 32:
        lens.append((lens[-1] - ks) // stride + 1)
    convs = [(rng.normal(0, 0.4, (3, 3, ks)), rng.normal(0, 0.1, 3)) for _ in range(len(lens) - 1)]
    flat = 3 * lens[-1]
    dense = [(rng.normal(0, 1 / np.sqrt(flat), (32, flat)), rng.normal(0, 0.1, 32)),
             (rng.normal(0, 1 / np.sqrt(32), (lbl, 32)), rng.normal(0, 0.1, lbl))]
    return convs, dense, lens


def test_encoder_backward_matches_finite_differences_and_autograd():
    rng = np.random.default_rng(3)
    P, B, lbl, stride = 300, 3, 4, 4
    convs, dense, lens = _encoder(rng, P, lbl)
    assert len(lens) == 3           # 300 -> 74 -> 18
    q = 0.02 * rng.normal(size=(B, P, 3))
    mq, sq = [0.001, -0.002, 0.0], [0.02, 0.03, 0.01]
    gz = rng.normal(size=(B, lbl))
    f = lambda c_, d_, q_: float(np.sum(O.encoder_forward(c_, d_, q_, mq, sq, stride)[0] * gz))
    z, cache = O.encoder_forward(convs, dense, q, mq, sq, stride)
    cg, dg, gq = O.encoder_backward(cache, gz)
    h = 1e-6
    for k in range(len(convs)):
        for idx in [(0, 0, 0), (2, 1, 5), (1, 2, 3)]:
            cp = [(w.copy(), c.copy()) for w, c in convs]; cp[k][0][idx] += h
            cm = [(w.copy(), c.copy()) for w, c in convs]; cm[k][0][idx] -= h
            fd = (f(cp, dense, q) - f(cm, dense, q)) / (2 * h)
            assert abs(fd - cg[k][0][idx]) <= 1e-6 * max(1.0, abs(fd)), (k, idx, fd, cg[k][0][idx])
    for (p, c) in [(0, 0), (151, 1), (P - 1, 2)]:
        qp = q.copy(); qp[1, p, c] += h
        qm = q.copy(); qm[1, p, c] -= h
        fd = (f(convs, dense, qp) - f(convs, dense, qm)) / (2 * h)
        assert abs(fd - gq[1, p, c]) <= 1e-6 * max(1.0, abs(fd)), (p, c, fd, gq[1, p, c])
    # torch autograd, FP64 (NetEnc as torch.nn.functional calls)
    ct = [(torch.tensor(w, requires_grad=True), torch.tensor(c, requires_grad=True)) for w, c in convs]
    dt = [(torch.tensor(w, requires_grad=True), torch.tensor(c, requires_grad=True)) for w, c in dense]
    qt = torch.tensor(q, requires_grad=True)
    s = ((qt - torch.tensor(mq, dtype=torch.float64)) / torch.tensor(sq, dtype=torch.float64)).transpose(1, 2)
    for w, c in ct:
        s = torch.nn.functional.elu(torch.nn.functional.conv1d(s, w, c, stride=stride))
    s = s.transpose(1, 2).reshape(B, -1)
    s = torch.nn.functional.elu(s @ dt[0][0].T + dt[0][1])
    zt = torch.nn.functional.elu(s @ dt[1][0].T + dt[1][1])
    (zt * torch.tensor(gz)).sum().backward()
    assert _rel(z, zt.detach().numpy()) <= 1e-12
    for k in range(len(convs)):
        assert _rel(cg[k][0], ct[k][0].grad.numpy()) <= 1e-9
        assert _rel(cg[k][1], ct[k][1].grad.numpy()) <= 1e-9
    for k in range(2):
        assert _rel(dg[k][0], dt[k][0].grad.numpy()) <= 1e-9
        assert _rel(dg[k][1], dt[k][1].grad.numpy()) <= 1e-9
    assert _rel(gq, qt.grad.numpy()) <= 1e-9


def test_adam_matches_torch():
    rng = np.random.default_rng(4)
    p0 = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(4)]
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    pt = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-2)
    for t, g in enumerate(grads, 1):
        O.adam_step(p, g, m, v, t, 1e-2)
        pt.grad = torch.tensor(g)
        opt.step()
    assert _rel(p, pt.detach().numpy()) <= 1e-13


def test_decoder_jacobian_matches_autograd():
    """The oracle's Jacobian d q / d x against torch.autograd.functional.jacobian
    (FP64), the reference's own check of NetDecFuncGrad (Exporter.py:104-106)."""
    rng = np.random.default_rng(5)
    lbl, n = 16, 7
    dims = [20, 128, 128, 128, 128, 128, 6]     # stress head: [label, X, mu] -> 6
    W, b = _decoder(rng, dims)
    x = rng.normal(size=(n, 20))
    mx, sx, mq, sq = [0.5, 0.4, 0.3], [0.1, 0.2, 0.15], np.zeros(6), np.full(6, 0.05)
    J = O.decoder_jacobian(W, b, x, lbl, mx, sx, mq, sq)
    Wt = [torch.tensor(w) for w in W]
    bt = [torch.tensor(c) for c in b]

    def f(xi):
        xs = torch.cat([xi[:lbl], (xi[lbl:lbl + 3] - torch.tensor(mx, dtype=torch.float64)) /
                        torch.tensor(sx, dtype=torch.float64), xi[lbl + 3:]])
        a = xs
        for k in range(5):
            a = torch.nn.functional.elu(a @ Wt[k].T + bt[k])
        return torch.tensor(mq) + (a @ Wt[5].T + bt[5]) * torch.tensor(sq)
    for i in range(n):
        Jt = torch.autograd.functional.jacobian(f, torch.tensor(x[i])).numpy()
        assert _rel(J[i], Jt) <= 1e-10
