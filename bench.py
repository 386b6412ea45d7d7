"""Headline benchmark: forward MPM sand substeps at 1M particles on one GPU.

Workload: sand column collapse, 1,000,000 particles i.i.d. uniform in
[0.3,0.7]^2 x [0.05,0.6] (seed 0), v ~ N(0, 0.1) per component (seed 1),
128^3 grid on the unit cube, Drucker-Prager sand E=2e5 Pa, nu=0.3, 30 deg,
density 2200, gravity -9.8 z, dt=1e-4, sticky floor at z=0.05 + bounding
box, FP32.  A step is one substep (p2g2p) over all particles; the state is
resident in HBM before the timed region starts.

Prints one JSON line with the metric, the roofline of the dominant kernel
(algorithmic bytes per launch / its average duration, HIP events on the
launch stream), the metric's second half -- forward + adjoint of a 200-substep
differentiable rollout of the same 1M-particle column (`fwd_adjoint`, with its
roofline at 700 B per particle-substep) -- and the CPU baseline (the FP32 CPU
restatement of the same substep, OpenMP over the host cores allotted to the
process, timed on a bounded sample; plus the 32^3 column on the same threads
and the FP64 numpy oracle on one core).

Exactly --steps substeps are timed.  The bank-order sort of the bins runs
every sort-period substeps (32), so a short window holds none or one of
them; the reported ms_per_step carries steps / period sorts (the window's
own sorts replaced by that share, from the sort's measured duration), and
`timing` gives the raw window.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "particle-substeps/s (fwd; fwd+adjoint) at 1M particles; % HBM roofline"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle and launch (FP32 state that must cross HBM)
ALG_BYTES = {
    # whole substep: read x0,u,v,C,F,tau,m,vol (152) + write u,v,C,F,tau (132)
    "step": 284,
    # the fused kernel does all particle work of a substep (G2P of this
    # substep + P2G of the next), so it carries the whole substep's bytes
    "g2p2g": 284,
    # stage kernels of the unbinned path: P2G reads x0,u,v,C,tau,m,vol;
    # G2P reads x0,u,F and writes u,v,C,F,tau
    "p2g": 116,
    "g2p": 192,
    # forward 284 + adjoint: read the forward state 152, read the adjoint of
    # the next state 132, write the adjoint of this state 132
    "fwd_adjoint": 700,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed substeps (default: column 500, adjoint 200, batch 300; crom: training steps, 1500)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--particles", type=int, default=None, help="default 1M (adjoint: 250k)")
    ap.add_argument("--n-grid", type=int, default=None, help="default 128 (adjoint: 64)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="budget of the CPU baseline sample (0 disables it)")
    ap.add_argument("--stage-steps", type=int, default=20,
                    help="substeps timed per kernel for the roofline")
    ap.add_argument("--unbinned", action="store_true",
                    help="time the unbinned stage kernels (atomic P2G, grid update, G2P) instead")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (reference for the HBM byte counters)")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="column: substeps of the forward + adjoint rollout measured beside the forward (0 = skip)")
    ap.add_argument("--keep", default="auto", choices=["auto", "all", "checkpoints"],
                    help="adjoint: keep every substep from the forward pass, or replay segments from checkpoints")
    ap.add_argument("--workload", default="column", choices=["column", "batch", "adjoint", "crom"],
                    help="column: the 1M-particle headline; batch: 64 uneven scenes advanced together; "
                         "adjoint: forward + adjoint of a differentiable rollout")
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--wg-threads", type=int, default=0, choices=[0, 256, 384, 512, 768],
                    help="threads per bin workgroup of the fused kernel (0 = library default)")
    ap.add_argument("--sort-period", type=int, default=0, help="substeps between bank-order sorts (0 = default)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batch: use --wg-threads/--sort-period instead of timing the candidates first")
    args = ap.parse_args()
    adj = args.workload == "adjoint"
    if args.particles is None:
        args.particles = 250_000 if adj else 1_000_000
    if args.n_grid is None:
        args.n_grid = 64 if adj else 128
    if args.steps is None:
        args.steps = {"column": 500, "adjoint": 200, "batch": 300, "crom": 1500}[args.workload]
    return args


def bench_batch(args):
    """64 independent sand columns of uneven size (particle counts
    log-uniform in [2k, 200k], grids 32^3-96^3, column dimensions U[0.2,0.5],
    friction angle U[20,40] deg, E U[1e5,5e5] Pa, seed 2), 300 substeps at
    dt = 1e-4: the batch (one launch per kernel for all scenes) against the
    same scenes run one after the other on the single-scene path."""
    import torch
    from reduced_mpm_stress_training_amd import batch as B
    from reduced_mpm_stress_training_amd import scenes as S
    steps = args.steps
    dt = 1e-4
    scenes = B.uneven_scenes(n_scenes=args.scenes, seed=2, ppc=4.0)
    total = sum(sc.n_particles for sc in scenes)
    print(f"[bench] batch: {len(scenes)} scenes, {total} particles, grids "
          f"{min(sc.n_grid for sc in scenes)}..{max(sc.n_grid for sc in scenes)}", file=sys.stderr, flush=True)
    tuned = None
    if not args.no_autotune:
        from reduced_mpm_stress_training_amd.tune import autotune
        tuned = autotune(lambda: B.SceneBatch(scenes), dt)
        print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
    bt = B.SceneBatch(scenes)
    launch = tuned["launch"] if tuned else {"wg_threads": args.wg_threads, "sort_period": args.sort_period}
    bt.set_launch(**launch)
    bt.run(args.warmup, dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bt.run(steps, dt)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    ok = int(bt.status()[0]) == 0
    del bt
    torch.cuda.empty_cache()
    # the same run replayed as a HIP graph (mpm_batch_graph_*), runs of 2 x sort period substeps
    bg = B.SceneBatch(scenes)
    bg.set_launch(**launch)
    bg.run(args.warmup, dt)
    k = bg.capture(dt)
    n_rep = max(1, steps // k)
    bg.run_captured(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bg.run_captured(n_rep)
    torch.cuda.synchronize()
    ms_graph = (time.perf_counter() - t0) * 1e3 / (n_rep * k)
    ok = ok and int(bg.status()[0]) == 0
    bg.release_graph()
    del bg
    torch.cuda.empty_cache()
    print(f"[bench] batch {t_batch * 1e3 / steps:.3f} ms/substep; sequential ...", file=sys.stderr, flush=True)
    t_seq = 0.0
    for sc in scenes:
        sim = S.make_simulator(sc)
        sim.run(args.warmup, dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.run(steps, dt)
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t0
        del sim
    value = total * steps / t_batch
    out = {
        "metric": "particle-substeps/s, batched uneven scenes",
        "value": value, "unit": "particle-substeps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": t_batch * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "uneven sand scenes in one batch", "n_scenes": len(scenes), "n_particles": total,
                   "grids": "32^3-96^3", "dt": dt, "parallelism": "single GPU"},
        "sequential": {"value": total * steps / t_seq, "ms_per_step": t_seq * 1e3 / steps,
                       "speedup_of_batch": t_seq / t_batch},
        "captured_graph_ms_per_step": ms_graph, "no_particles_dropped": ok, "launch": launch, "autotune_ms_per_substep": tuned["trials"] if tuned else None,
    }
    print(json.dumps(out), flush=True)


def calibrate_copy():
    """A 1 GiB fill and three 1 GiB copies: under rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE their per-dispatch readings give the counters' scale against
    known byte counts."""
    import torch
    n = 256 * 1024 * 1024
    src = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read_per_copy": 4 * n, "bytes_written_per_copy": 4 * n,
                      "bytes_written_per_fill": 4 * n}), flush=True)


TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "g2p2g_1M_128_hbm_counters.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the headline workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy
    calibration recorded there); None if not recorded."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == "k_" + kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def _cpu_threads():
    """Threads of the CPU baselines: the host cores allotted to this process
    (OMP_NUM_THREADS, the box's share for one GPU), else every core in the
    affinity mask."""
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    use = int(share) if share and share.isdigit() and int(share) > 0 else visible
    return max(1, min(use, visible)), visible


def _time_oracle(o, dt, budget_s, nthreads):
    o.step(dt, nthreads=nthreads)          # first touch / warm-up
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(dt, nthreads=nthreads)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or el / steps * (steps + 1) > 2 * budget_s:
            return steps, el


def cpu_baseline(scene, budget_s):
    """FP32 CPU restatement (oracle/, OpenMP) on the same workload for as many
    substeps as fit in the budget (at least 1); beside it the same C
    restatement on the 32^3 column, and the FP64 numpy oracle (one process)
    on that column."""
    from oracle import oracle_lib
    from oracle import mpm_oracle_np as NP
    from reduced_mpm_stress_training_amd import scenes as S
    cores, visible = _cpu_threads()
    oracle_lib.build()
    o = oracle_lib.OracleScene(S.scene_params(scene), S.scene_state(scene, np.float32),
                               S.scene_colliders(scene), np.float32)
    steps, el = _time_oracle(o, scene.dt, budget_s, cores)
    out = {"value": scene.n_particles * steps / el, "unit": "particle-substeps/s", "cores": cores,
           "host_cores_visible": visible, "kind": "port",
           "sample": f"{steps} substeps of the same {scene.n_particles}-particle, {scene.n_grid}^3 workload "
                     f"({el:.1f} s), FP32 C restatement with OpenMP atomics, {cores} threads"}
    c1 = S.column_lattice()
    o1 = oracle_lib.OracleScene(S.scene_params(c1), S.scene_state(c1, np.float32), S.scene_colliders(c1),
                                np.float32)
    s1, e1 = _time_oracle(o1, c1.dt, min(3.0, budget_s), cores)
    out["column_32"] = {"value": c1.n_particles * s1 / e1, "unit": "particle-substeps/s", "cores": cores,
                        "sample": f"{s1} substeps of the 32^3 column ({c1.n_particles} particles, {e1:.1f} s)"}
    P = S.scene_params(c1)
    st = S.scene_state(c1, np.float64)
    cols = S.scene_colliders(c1)
    NP.step(P, st, cols, 0.0, c1.dt)
    k, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < min(5.0, budget_s) or k == 0:
        st = NP.step(P, st, cols, k * c1.dt, c1.dt)
        k += 1
    e2 = time.perf_counter() - t0
    out["fp64_numpy_column_32"] = {"value": c1.n_particles * k / e2, "unit": "particle-substeps/s", "cores": 1,
                                   "sample": f"{k} substeps of the 32^3 column, FP64 numpy oracle ({e2:.1f} s)"}
    return out


def measure_adjoint(n, ng, steps, keep="auto", fwd_ms=None):
    """Differentiable rollout: n particles i.i.d. in [0.3,0.7]^2 x [0.05,0.6]
    (v ~ N(0, 0.1)) on an ng^3 grid, `steps` substeps of dt = 1e-4,
    checkpoints every 20; loss = mean |x_final - x_target|^2 against a target
    rollout at E = 3e5 Pa, friction angle 35 deg; gradient with respect to
    (E, nu, friction angle, initial velocities) by the adjoint kernels.
    A step is one forward substep plus its adjoint: with keep "all" (the
    default "auto" takes it where half the free HBM holds the record) the
    forward keeps every substep's state and grids and the adjoint runs on
    them; with keep "checkpoints" each segment is replayed first.  One
    gradient evaluation runs untimed first (the rollout's buffers are then
    allocated, as in any fitting loop after its first iteration); the next one
    is timed.  fwd_ms: the binned forward's ms per substep on the same scene
    (measured here when not given)."""
    import torch
    from reduced_mpm_stress_training_amd import adjoint as A
    from reduced_mpm_stress_training_amd import scenes as S
    sc = S.column_uniform(n_particles=n, n_grid=ng)
    ro = A.Rollout(sc, steps, 1e-4, checkpoint_every=20, keep=keep)
    v0 = torch.from_numpy(sc.v0).cuda()
    with torch.no_grad():
        x_t = ro.forward(v0, 3.0e5, 0.3, 35.0).clone()
    torch.cuda.synchronize()

    def grad_eval():
        v = v0.clone().requires_grad_(True)
        params = torch.tensor([2.0e5, 0.3, 30.0], dtype=torch.float64, requires_grad=True)
        x = A.sand_rollout(v, params, ro)
        loss = torch.mean(torch.sum((x - x_t) ** 2, 1))
        loss.backward()
        return v, params

    grad_eval()
    torch.cuda.synchronize()
    v = v0.clone().requires_grad_(True)
    params = torch.tensor([2.0e5, 0.3, 30.0], dtype=torch.float64, requires_grad=True)
    t0 = time.perf_counter()
    x = A.sand_rollout(v, params, ro)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss = torch.mean(torch.sum((x - x_t) ** 2, 1))
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g = params.grad.tolist()
    finite = bool(torch.isfinite(v.grad).all()) and all(map(lambda z: z == z, g))
    if fwd_ms is None:
        # the forward alone on the production (binned) path, same scene and substeps
        sim = S.make_simulator(sc)
        sim.run(10, 1e-4)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        sim.run(steps, 1e-4)
        torch.cuda.synchronize()
        fwd_ms = (time.perf_counter() - t3) * 1e3 / steps
        del sim
    ms = (t2 - t0) * 1e3 / steps
    out = {"value": n * steps / (t2 - t0), "unit": "particle-substeps/s", "ms_per_step": ms,
           "forward_ms_per_step": (t1 - t0) * 1e3 / steps, "backward_ms_per_step": (t2 - t1) * 1e3 / steps,
           "binned_forward_ms_per_step": fwd_ms, "fwd_adjoint_over_binned_forward": ms / fwd_ms,
           "config": {"workload": "differentiable sand rollout, forward + adjoint", "n_particles": n, "n_grid": ng,
                      "substeps": steps, "checkpoint_every": 20,
                      "keep": "all" if ro._recorded else "checkpoints",
                      "record_gb": ro.record_bytes() / 1e9 if ro._recorded else 0.0},
           "roofline": {"bound": "hbm", "achieved": ALG_BYTES["fwd_adjoint"] * n / (ms / 1e3) / 1e9,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ALG_BYTES["fwd_adjoint"] * n / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                        "bytes_per_particle": ALG_BYTES["fwd_adjoint"]},
           "grad": {"E": g[0], "nu": g[1], "friction_angle": g[2]}, "finite": finite}
    del ro
    torch.cuda.empty_cache()
    return out


def bench_adjoint(args):
    """The forward + adjoint rollout alone (measure_adjoint), as its own JSON line."""
    r = measure_adjoint(args.particles, args.n_grid, args.steps, args.keep)
    out = {"metric": "particle-substeps/s (fwd+adjoint)", "value": r["value"], "unit": r["unit"], "n_gpus": 1,
           "steps": args.steps, "warmup": args.steps, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": dict(r["config"], parallelism="single GPU")}
    out.update({k: v for k, v in r.items() if k not in ("value", "unit", "ms_per_step", "config")})
    print(json.dumps(out), flush=True)


def bench_crom(args):
    """CROM on solver snapshots: 2000 snapshots = 50 sand columns of 1M
    particles (128^3, the headline column) run by the MPM kernels at friction
    angles drawn from U[20, 40] deg (seed 3), 40 frames of 10 substeps each,
    the first N/10 particles of every frame kept (SimulationDataset.py:36-43).
    CROMnet (latent 16, decoder 19 -> 128 x 5 -> 3, BF16 MFMA) trains on the
    displacements, batches of 10 frames; its encoder labels every frame; the
    stress network [label, X, friction angle] -> 6 trains on the solver's
    Voigt stress.  A held-out column (another angle) is encoded frame by
    frame and both networks are evaluated at all 1M particles: the line's
    value is the decoder's points/s, with the stress network's, and the
    relative L2 errors of displacement and stress against the full MPM run.
    Then the CROM-reduced run of the held-out angle: MPM on the N/10 sample
    (half-resolution grid) + encoder + both networks per frame, (a) checked
    frame by frame against the full run in lockstep, (b) on its own with the
    self-check every 5 frames and the full-MPM fallback."""
    import torch
    from reduced_mpm_stress_training_amd import crom as CR
    from reduced_mpm_stress_training_amd import crom_pipeline as CP
    from reduced_mpm_stress_training_amd import scenes as S
    rng = np.random.default_rng(3)
    n, dt, spf, frames, n_runs = args.particles, 1e-4, 10, 40, 50
    P = n // 10                                    # first N/10 points (SimulationDataset.py:36-43)
    sc = S.column_uniform(n_particles=n, n_grid=args.n_grid)
    angles = [float(a) for a in rng.uniform(20.0, 40.0, size=n_runs + 1)]
    t0 = time.perf_counter()
    q_tr = torch.empty((n_runs * frames, P, 3), device="cuda")
    s_tr = torch.empty((n_runs * frames, P, 6), device="cuda")
    for r in range(n_runs):
        q, st = CR.solver_snapshots(sc, angles[r], frames, spf, dt, n_points=P, with_stress=True)
        q_tr[r * frames:(r + 1) * frames], s_tr[r * frames:(r + 1) * frames] = q, st
        del q, st
        print(f"[bench] crom snapshots: run {r + 1}/{n_runs} (angle {angles[r]:.2f})", file=sys.stderr, flush=True)
    held_q, held_s = CR.solver_snapshots(sc, angles[-1], frames, spf, dt, with_stress=True)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    mu_tr = torch.tensor([angles[r] for r in range(n_runs) for _ in range(frames)], device="cuda")
    X = torch.from_numpy(sc.x0).cuda()
    qf = q_tr.reshape(-1, 3)
    net = CR.Crom(P, lbl=16, seed=0)
    net.set_preprop(X[:P].mean(0), X[:P].std(0), qf.mean(0), qf.std(0).clamp_min(1e-12))
    xb = X[None, :P].expand(10, P, 3).contiguous()
    steps = args.steps
    g = torch.Generator(device="cuda").manual_seed(0)
    t0 = time.perf_counter()
    for it in range(steps):
        idx = torch.randint(0, q_tr.shape[0], (10,), device="cuda", generator=g)
        loss = net.train_step(q_tr[idx], xb, 1e-4 * CR.lr_factor(int(it * 750 / steps)))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    print(f"[bench] crom: displacement network trained ({steps} steps)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    labels = CP.label_frames(net, q_tr)
    head = CP.StressHead(16, seed=1)
    head.set_preprop(X[:P], s_tr)
    loss_s = CP.train_stress_head(head, labels, X[:P], mu_tr, s_tr, steps=steps)
    torch.cuda.synchronize()
    t_train_s = time.perf_counter() - t0
    print(f"[bench] crom: stress network trained ({steps} steps)", file=sys.stderr, flush=True)
    # held-out column: encode each frame from its first N/10 points, evaluate
    # both networks at all N particles
    num = {"u": 0.0, "s": 0.0}
    den = {"u": 0.0, "s": 0.0}
    t_dec = t_str = 0.0
    for f in range(1, frames):
        z = CP.label_frames(net, held_q[f:f + 1, :P])
        xin = torch.cat([z.expand(n, 16), X], 1).contiguous()
        sin = CP.StressHead.inputs(z, X, angles[-1])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        qh = net.decoder(xin)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        sh = head.net(sin)
        torch.cuda.synchronize()
        t_dec += t2 - t1
        t_str += time.perf_counter() - t2
        num["u"] += float(((qh - held_q[f]) ** 2).sum())
        den["u"] += float((held_q[f] ** 2).sum())
        num["s"] += float(((sh - held_s[f]) ** 2).sum())
        den["s"] += float((held_s[f] ** 2).sum())
    rel_u, rel_s = (num["u"] / den["u"]) ** 0.5, (num["s"] / den["s"]) ** 0.5
    ms_dec, ms_str = t_dec * 1e3 / (frames - 1), t_str * 1e3 / (frames - 1)
    del held_q, held_s, q_tr, s_tr
    # MPM time of one frame (10 substeps) on the same column
    sim = S.make_simulator(sc)
    sim.run(10, dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sim.run(100, dt)
    torch.cuda.synchronize()
    ms_frame = (time.perf_counter() - t1) * 1e3 / 10
    del sim
    # the CROM-reduced run of the held-out angle
    coarse = max(16, args.n_grid // 2)
    rr = CR.ReducedRun(net, sc, angles[-1], n_sample=P, coarse_grid=coarse, stress_head=head)
    rr.run(2, spf, dt)                       # warm-up of the sampled run's buffers
    red = rr.run(frames, spf, dt, verify=True)
    errs, serrs = red["errors"][1:], red["stress_errors"][1:]
    solo = rr.run(frames, spf, dt, check_every=5, self_tol=0.1)
    n_crom = solo["source"].count("crom")
    flop = 2 * (19 * 128 + 4 * 128 * 128 + 128 * 3)
    flop_s = 2 * (20 * 128 + 4 * 128 * 128 + 128 * 6)
    out = {
        "metric": "CROM decoder points/s (inference at 1M particles)", "value": n / (ms_dec / 1e3),
        "unit": "points/s", "n_gpus": 1, "steps": steps, "warmup": 0, "ms_per_step": ms_dec,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "CROM decoder + stress network trained on 2000 solver snapshots", "n_particles": n,
                   "n_grid": args.n_grid, "latent": 16, "hidden": 128, "snapshot_runs": n_runs,
                   "frames_per_run": frames, "substeps_per_frame": spf, "points_per_frame": P,
                   "friction_angles": "U[20, 40] deg, seed 3", "train_steps": steps},
        "decoder_tflops": flop * n / (ms_dec / 1e3) / 1e12,
        "stress_network": {"points_per_s": n / (ms_str / 1e3), "ms_at_1M": ms_str,
                           "tflops": flop_s * n / (ms_str / 1e3) / 1e12},
        "mpm_ms_per_frame_10_substeps": ms_frame, "decode_speedup_over_mpm_frame": ms_frame / ms_dec,
        "held_out": {"angle": angles[-1], "displacement_relative_l2": rel_u, "stress_relative_l2": rel_s},
        "final_train_loss": float(loss), "final_stress_loss": float(loss_s),
        "reduced_run": {"sample": P, "coarse_grid": coarse,
                        "ms_per_frame": red["reduced_seconds"] * 1e3 / (frames - 1),
                        "full_mpm_ms_per_frame": red["full_seconds"] * 1e3 / (frames - 1),
                        "speedup": red["full_seconds"] / red["reduced_seconds"],
                        "displacement_relative_l2_mean": float(np.mean(errs)),
                        "displacement_relative_l2_max": float(np.max(errs)),
                        "stress_relative_l2_mean": float(np.mean(serrs)),
                        "stress_relative_l2_max": float(np.max(serrs))},
        "reduced_run_self_checked": {"check_every": 5, "self_tol": 0.1, "frames_reduced": n_crom,
                                     "frames_fallback": frames - n_crom, "checks": solo["checks"],
                                     "ms_per_frame": solo["reduced_seconds"] * 1e3 / (frames - 1)},
        "snapshot_seconds": t_data, "train_seconds": t_train, "stress_train_seconds": t_train_s,
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the substep is not sharded "
              "across GPUs in this build", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    if args.workload == "crom":
        bench_crom(args)
        return
    import torch
    from reduced_mpm_stress_training_amd import scenes as S

    scene = S.column_uniform(n_particles=args.particles, n_grid=args.n_grid)
    sim = S.make_simulator(scene)
    sim.set_launch(args.wg_threads, args.sort_period)
    sim.use_bins = not args.unbinned
    n = scene.n_particles
    dt = scene.dt
    print(f"[bench] workload ready: {n} particles, {scene.n_grid}^3 grid", file=sys.stderr, flush=True)

    sim.run(args.warmup, dt)
    torch.cuda.synchronize()
    period = args.sort_period or 16
    since0 = int(sim._bins["struct"].since_sort) if sim.use_bins and sim._bins is not None else 0
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    t_start.record()
    sim.run(args.steps, dt)
    t_end.record()
    torch.cuda.synchronize()
    ms_total = t_start.elapsed_time(t_end)
    ms_measured = ms_total / args.steps
    print(f"[bench] timed {args.steps} substeps: {ms_measured:.4f} ms/substep", file=sys.stderr, flush=True)

    # per-kernel durations: HIP events on the launch stream around each kernel
    if args.unbinned:
        sim.stage_timer = []
        sim.run(args.stage_steps, dt)
        torch.cuda.synchronize()
        names = ["zero_grid", "p2g", "grid_update", "g2p"]
        dur = {nm: 0.0 for nm in names}
        for ev in sim.stage_timer:
            for i, nm in enumerate(names):
                dur[nm] += ev[i].elapsed_time(ev[i + 1])
        dur = {k: v / len(sim.stage_timer) for k, v in dur.items()}
        sim.stage_timer = None
        sorts_in_window, sort_call_ms, ms_step = 0, 0.0, ms_measured
    else:
        # events recorded by the library itself between its own launches, over
        # a whole number of sort periods (so the sort's average is per call / period)
        names = ["grid_update", "g2p2g", "sort"]
        n_prof = max(args.stage_steps, period)
        n_prof = -(-n_prof // period) * period
        gm, km, sm = sim.profile_kernels(n_prof, dt)
        dur = {"grid_update": gm, "g2p2g": km, "sort": sm}
        # The bank-order sort runs every `period` substeps; a timed window of
        # K substeps holds (since + K) // period of them.  The step time
        # reported carries exactly K / period sorts: the measured window with
        # its own sorts replaced by the sort's share.
        sorts_in_window = (since0 + args.steps) // period
        sort_call_ms = sm * n_prof / max(1, n_prof // period)
        ms_step = (ms_total - sorts_in_window * sort_call_ms + args.steps / period * sort_call_ms) / args.steps
    value = n / (ms_step / 1e3)
    dom = max((k for k in names if k in ALG_BYTES), key=lambda k: dur[k])
    achieved = ALG_BYTES[dom] * n / (dur[dom] / 1e3) / 1e9
    step_gbs = ALG_BYTES["step"] * n / (ms_step / 1e3) / 1e9

    u, v, Fd, tau = sim.export_uvFstress_from_warp()
    finite = bool(torch.isfinite(u).all() and torch.isfinite(tau).all())
    del sim, u, v, Fd, tau
    torch.cuda.empty_cache()

    fwd_adj = None
    if args.adjoint_steps > 0 and not args.unbinned:
        print("[bench] forward + adjoint ...", file=sys.stderr, flush=True)
        fwd_adj = measure_adjoint(n, scene.n_grid, args.adjoint_steps, "auto", fwd_ms=ms_step)

    cpu = None
    if args.cpu_seconds > 0:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(scene, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "sand column collapse, forward substep", "n_particles": n,
                   "path": "unbinned stage kernels" if args.unbinned else "binned G2P2G",
                   "launch": {"wg_threads": args.wg_threads or 768, "sort_period": args.sort_period or 16},
                   "n_grid": scene.n_grid, "box": "[0.3,0.7]^2x[0.05,0.6]", "dt": dt, "E": scene.E,
                   "nu": scene.nu, "friction_angle": scene.friction_angle, "density": scene.density,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom) if (not args.unbinned and args.particles == 1_000_000
                                                           and args.n_grid == 128) else None,
                     "bytes_per_particle": ALG_BYTES[dom],
                     "stage_ms": dur, "step_achieved_gbs": step_gbs, "step_frac": step_gbs / HBM_PEAK_GBS},
        "timing": {"ms_per_step_measured": ms_measured, "sorts_in_window": sorts_in_window,
                   "sort_ms_per_call": sort_call_ms, "sort_period": period,
                   "note": "ms_per_step carries steps/sort_period bank-order sorts: the timed window with its own "
                           "sorts replaced by that share"},
        "fwd_adjoint": fwd_adj,
        "cpu_baseline": cpu,
        "finite": finite,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
