"""Independent FP64 NumPy restatement of the MPM sand substep.

TEST INFRASTRUCTURE ONLY: this module is a checker.  Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product package never does.

It restates the same reference functions as ``oracle/mpm_oracle.c`` but by a
different route, so that the two pin each other:

* the 3x3 SVD is LAPACK's (``numpy.linalg.svd``), sign-fixed to rotations;
* P2G is a vectorised scatter with ``numpy.add.at`` instead of a loop;
* G2P is a vectorised gather (einsum) instead of a per-particle loop.

Reference citations (zeshunzong/reduced_mpm_stress_training):
  update_x_from_u            engine_stress/mpm_utils.py:4-7
  B-spline weights           engine_stress/mpm_utils.py:196-209, 244-257
  compute_dweight            engine_stress/mpm_utils.py:54-57
  p2g_apic_stress_given      engine_stress/mpm_utils.py:190-227
  grid_normalization_and_gravity  engine_stress/mpm_utils.py:229-237
  surface collider           engine_stress/mpm_solver_warp.py:429-449
  bounding box               engine_stress/mpm_solver_warp.py:505-523
  g2p_apic_and_stress        engine_stress/mpm_utils.py:241-306
  sand_return_mapping        engine_stress/mpm_utils.py:87-112
  kirchoff_stress_drucker_prager engine_stress/mpm_utils.py:45-52
  von_mises_return_mapping   engine_stress/mpm_utils.py:59-84
  kirchoff_stress_StVK       engine_stress/mpm_utils.py:36-43
  kirchoff_stress_FCR        engine_stress/mpm_utils.py:29-34
  G2P material switch        engine_stress/mpm_utils.py:284-304
  update_u_from_x            engine_stress/mpm_utils.py:12-15
  p2g2p (order)              engine_stress/mpm_solver_warp.py:156-203
"""
from __future__ import annotations

import numpy as np


def svd3_rot(A: np.ndarray):
    """SVD of a stack of 3x3 matrices with U, V rotations (det +1) and the
    sign of det A carried by the smallest singular value (the convention of
    Warp's wp.svd3 for mpm_utils.py:92,297)."""
    U, s, Vt = np.linalg.svd(A)
    s = s.copy()
    dU = np.linalg.det(U)
    dV = np.linalg.det(Vt)
    fixU = dU < 0
    U[fixU, :, 2] *= -1.0
    s[fixU, 2] *= -1.0
    fixV = dV < 0
    Vt[fixV, 2, :] *= -1.0
    s[fixV, 2] *= -1.0
    return U, s, np.swapaxes(Vt, -1, -2)


def weights(x: np.ndarray, inv_dx: float):
    """Quadratic B-spline weights (mpm_utils.py:196-209).  Returns base (N,3)
    int, fx (N,3), w (N,3 axes,3 nodes), dw (N,3,3)."""
    gp = x * inv_dx
    base = np.trunc(gp - 0.5).astype(np.int64)
    fx = gp - base
    w = np.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2], axis=-1)
    dw = np.stack([fx - 1.5, -2.0 * (fx - 1.0), fx - 0.5], axis=-1)
    return base, fx, w, dw


def _stencil(w, dw, fx, inv_dx):
    """Per-node weight (N,27), gradient (N,27,3), offset (N,27,3) in grid units."""
    off = np.array([(i, j, k) for i in range(3) for j in range(3) for k in range(3)])
    wi = w[:, 0, off[:, 0]]
    wj = w[:, 1, off[:, 1]]
    wk = w[:, 2, off[:, 2]]
    weight = wi * wj * wk
    grad = np.stack([dw[:, 0, off[:, 0]] * wj * wk,
                     wi * dw[:, 1, off[:, 1]] * wk,
                     wi * wj * dw[:, 2, off[:, 2]]], axis=-1) * inv_dx
    dpos = off[None, :, :] - fx[:, None, :]
    return off, weight, grad, dpos


def sand_return_map(Ftr, mu, lam, alpha):
    """sand_return_mapping, mpm_utils.py:87-112."""
    U, s, V = svd3_rot(Ftr)
    eps = np.log(np.maximum(np.abs(s), 1e-14))
    tr = eps.sum(-1)
    eh = eps - tr[:, None] / 3.0
    ehn = np.linalg.norm(eh, axis=-1)
    dg = ehn + (3.0 * lam + 2.0 * mu) / (2.0 * mu) * tr * alpha
    Fe = Ftr.copy()
    b2 = (dg > 0) & (tr > 0)
    b3 = (dg > 0) & (tr <= 0)
    Fe[b2] = U[b2] @ np.swapaxes(V[b2], -1, -2)
    if b3.any():
        H = eps[b3] - eh[b3] * (dg[b3] / ehn[b3])[:, None]
        Fe[b3] = (U[b3] * np.exp(H)[:, None, :]) @ np.swapaxes(V[b3], -1, -2)
    return Fe


def dp_stress(F, mu, lam):
    """kirchoff_stress_drucker_prager (mpm_utils.py:45-52) on a fresh SVD of F
    (mpm_utils.py:297)."""
    U, s, V = svd3_rot(F)
    ls = np.log(s).sum(-1)
    c = 2.0 * mu * np.log(s) / s + lam * ls[:, None] / s
    return (U * c[:, None, :]) @ np.swapaxes(V, -1, -2) @ np.swapaxes(F, -1, -2)


def von_mises(Ftr, mu, lam, ys, hardening=0, xi=0.0):
    """von_mises_return_mapping (mpm_utils.py:59-84) on a stack of F_tr with
    per-matrix yield stresses ys; returns (F_e, ys after hardening)."""
    U, s, V = svd3_rot(Ftr)
    eps = np.log(np.maximum(s, 0.01))
    temp = eps.sum(-1) / 3.0
    t = 2.0 * mu * eps + lam * eps.sum(-1)[:, None]
    cond = t - t.sum(-1)[:, None] / 3.0
    yielded = np.linalg.norm(cond, axis=-1) > ys
    Fe = Ftr.copy()
    ys = np.array(ys, dtype=np.float64, copy=True)
    if yielded.any():
        eh = eps[yielded] - temp[yielded, None]
        ehn = np.linalg.norm(eh, axis=-1) + 1e-6
        dg = ehn - ys[yielded] / (2.0 * mu)
        e_new = eps[yielded] - (dg / ehn)[:, None] * eh
        Fe[yielded] = (U[yielded] * np.exp(e_new)[:, None, :]) @ np.swapaxes(V[yielded], -1, -2)
        if hardening == 1:
            ys[yielded] = ys[yielded] + 2.0 * mu * xi * dg
    return Fe, ys


def stvk_stress(F, mu, lam):
    """kirchoff_stress_StVK (mpm_utils.py:36-43) on a fresh SVD of F:
    U diag(2 mu ln s + lam sum ln s) V^T, s clamped at 0.01."""
    U, s, V = svd3_rot(F)
    ls = np.log(np.maximum(s, 0.01))
    c = 2.0 * mu * ls + lam * ls.sum(-1)[:, None]
    return (U * c[:, None, :]) @ np.swapaxes(V, -1, -2)


def fcr_stress(F, mu, lam):
    """kirchoff_stress_FCR (mpm_utils.py:29-34): 2 mu (F - U V^T) F^T +
    lam J (J - 1) I with J = det F (mpm_utils.py:292)."""
    U, s, V = svd3_rot(F)
    R = U @ np.swapaxes(V, -1, -2)
    J = np.linalg.det(F)
    return 2.0 * mu * (F - R) @ np.swapaxes(F, -1, -2) + (lam * J * (J - 1.0))[:, None, None] * np.eye(3)


def p2g(params, x, v, C, tau, mass, vol, dt):
    """p2g_apic_stress_given (mpm_utils.py:190-227) -> grid (m, mv)."""
    ng = params["n_grid"]
    dx, inv_dx = params["dx"], params["inv_dx"]
    base, fx, w, dw = weights(x, inv_dx)
    off, weight, grad, dpos = _stencil(w, dw, fx, inv_dx)
    dpos = dpos * dx
    nodes = base[:, None, :] + off[None, :, :]
    flat = (nodes[..., 0] * ng + nodes[..., 1]) * ng + nodes[..., 2]
    wm = weight * mass[:, None]
    mom = wm[..., None] * (v[:, None, :] + np.einsum("pab,pnb->pna", C, dpos))
    force = np.einsum("pab,pnb->pna", -vol[:, None, None] * tau, grad)
    grid_m = np.zeros(ng ** 3)
    grid_mv = np.zeros((ng ** 3, 3))
    np.add.at(grid_m, flat.ravel(), wm.ravel())
    np.add.at(grid_mv, flat.ravel(), (mom + dt * force).reshape(-1, 3))
    return grid_m, grid_mv


def grid_update(params, grid_m, grid_mv, colliders, time, dt):
    """grid_normalization_and_gravity (mpm_utils.py:229-237) followed by the
    colliders in registration order (mpm_solver_warp.py:182-185)."""
    ng = params["n_grid"]
    g = np.asarray(params["gravity"], dtype=np.float64)
    gv = np.zeros((ng ** 3, 3))
    act = grid_m > 1e-15
    gv[act] = grid_mv[act] * (1.0 / grid_m[act])[:, None] + dt * g
    gv = gv.reshape(ng, ng, ng, 3)
    idx = np.indices((ng, ng, ng)).transpose(1, 2, 3, 0)
    for c in colliders:
        if not (time >= c["start_time"] and time < c["end_time"]):
            continue
        if c["kind"] == 0:
            off = idx * params["dx"] - np.asarray(c["point"])
            below = (off @ np.asarray(c["normal"])) < 0.0
            gv[below] = 0.0
        else:
            pad = 3
            for d in range(3):
                lo = (idx[..., d] < pad) & (gv[..., d] < 0)
                gv[..., d][lo] = 0.0
                hi = (idx[..., d] >= ng - pad) & (gv[..., d] > 0)
                gv[..., d][hi] = 0.0
    return gv.reshape(-1, 3)


def g2p(params, grid_v, x, F, dt, ys=None):
    """g2p_apic_and_stress (mpm_utils.py:241-306) with the material switch
    of :284-304 (params["material"]: 0 jelly, 1 metal, 2 sand).  Returns
    x, v, C, F, tau and the yield stresses (metal)."""
    ng = params["n_grid"]
    inv_dx = params["inv_dx"]
    base, fx, w, dw = weights(x, inv_dx)
    off, weight, grad, dpos = _stencil(w, dw, fx, inv_dx)
    nodes = base[:, None, :] + off[None, :, :]
    flat = (nodes[..., 0] * ng + nodes[..., 1]) * ng + nodes[..., 2]
    g = grid_v[flat]                                          # (N,27,3)
    v_new = np.einsum("pn,pna->pa", weight, g)
    C_new = np.einsum("pna,pnb,pn->pab", g, dpos, weight * inv_dx * 4.0)
    gradv = np.einsum("pna,pnb->pab", g, grad)
    x_new = x + dt * v_new
    Ftr = (np.eye(3)[None] + gradv * dt) @ F
    mat = params.get("material", 2)
    mu, lam = params["mu"], params["lam"]
    if ys is None:
        ys = np.zeros(x.shape[0])
    if mat == 2:
        Fe = sand_return_map(Ftr, mu, lam, params["alpha"])
        tau = dp_stress(Fe, mu, lam)
    elif mat == 1:
        Fe, ys = von_mises(Ftr, mu, lam, ys, params.get("hardening", 0), params.get("xi", 0.0))
        tau = stvk_stress(Fe, mu, lam)
    else:
        Fe = Ftr
        tau = fcr_stress(Fe, mu, lam)
    return x_new, v_new, C_new, Fe, tau, ys


def step(params, state, colliders, time, dt):
    """One p2g2p substep (mpm_solver_warp.py:156-203).  ``state`` is a dict of
    float64 arrays x0,u,v (N,3), C,F,tau (N,3,3), mass,vol (N,); returns a new
    dict.  ``colliders`` is a list of dicts (kind, point, normal, start_time,
    end_time)."""
    x = state["x0"] + state["u"]
    gm, gmv = p2g(params, x, state["v"], state["C"], state["tau"], state["mass"], state["vol"], dt)
    gv = grid_update(params, gm, gmv, colliders, time, dt)
    x_new, v, C, F, tau, ys = g2p(params, gv, x, state["F"], dt, state.get("yield_stress"))
    out = dict(state)
    out.update(u=x_new - state["x0"], v=v, C=C, F=F, tau=tau, yield_stress=ys)
    return out
