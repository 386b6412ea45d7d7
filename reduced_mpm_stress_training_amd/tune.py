"""Launch-parameter autotuner of the binned substep.

The fused G2P2G kernel has two free launch parameters that change speed but
not results (beyond the order of the FP64 tile sums): the threads per bin
workgroup (256 puts two bins on a CU at once, 768 runs one bin with three
waves per SIMD, 512 with two) and the period of the bank-order sort (sorting costs a pass over
the state, not sorting lets LDS bank conflicts grow as particles drift).
Which is best depends on the scenes -- bin occupancy, particle counts, how
fast particles move -- so `autotune` times each candidate on a throwaway
copy of the workload and returns the fastest.

    best = autotune(lambda: SceneBatch(scenes), dt)
    batch = SceneBatch(scenes); batch.set_launch(**best["launch"])

The factory must build an object with `set_launch(wg_threads, sort_period)`
and `run(n_substeps, dt)` (SceneBatch, MPM_Simulator_HIP).  Candidates run
round-robin, twice, so that each sees early and later states of the
throwaway run; a candidate's figure is its faster round.
"""
from __future__ import annotations

import time

import torch

CANDIDATES = ((768, 16), (768, 32), (512, 16), (256, 16), (256, 32), (256, 64))


def autotune(factory, dt: float, trial_substeps: int = 32, warmup_substeps: int = 4,
             candidates=CANDIDATES, rounds: int = 2) -> dict:
    sim = factory()
    sim.run(warmup_substeps, dt)
    torch.cuda.synchronize()
    best = {}
    for _ in range(rounds):
        for wg, sp in candidates:
            sim.set_launch(wg, sp)
            sim.run(2, dt)           # the new launch shape's first launches
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sim.run(trial_substeps, dt)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / trial_substeps
            best[(wg, sp)] = min(ms, best.get((wg, sp), float("inf")))
    del sim
    torch.cuda.empty_cache()
    (wg, sp), ms = min(best.items(), key=lambda kv: kv[1])
    return {"launch": {"wg_threads": wg, "sort_period": sp}, "ms_per_substep": ms,
            "trials": {f"{k[0]}x{k[1]}": v for k, v in best.items()}}
