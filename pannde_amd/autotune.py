"""Autotuner for the RSG kernels.

Sweeps, for each of the two kernels independently, the workgroup size {64, 128, 256, 512},
the entities per thread {1, 2, 4} and LDS staging on/off, timing each candidate with the
library's HIP-event kernel timers on a representative job, and returns the fastest
choice as an ``rsg_tuning`` dict.  Every candidate computes bit-identical results, so the
choice only affects speed.  Members of a batch may differ in their tuning (each launch
covers the members of one kernel configuration), but members of one configuration share
launches, so tune once on a representative instance and create the batch's members with
the result.
"""
from __future__ import annotations

import itertools
import time
from typing import Dict, List, Optional, Tuple

from . import rsg

BLOCKS = (64, 128, 256, 512)
PER_THREAD = (1, 2, 4)
LDS = (1, -1)


def _time_candidate(case, precision, block, cpt, lds, steps, warmup):
    tuning = dict(block_velocity=block, block_stress=block, per_thread_velocity=cpt, per_thread_stress=cpt,
                  lds_velocity=lds, lds_stress=lds)
    s = rsg.DeviceSolver.from_case(case, precision, tuning=tuning)
    try:
        s.solve(warmup)
        s.synchronize()
        s.kernel_times()
        s.set_timing(True)
        s.solve(steps)
        s.synchronize()
        vms, sms, n = s.kernel_times()
        return vms / n, sms / n
    finally:
        s.close()


def tune(case, precision=rsg.RSG_FP32, steps=20, warmup=3, blocks=BLOCKS, per_thread=PER_THREAD, lds=LDS,
         log=None) -> Tuple[Dict, List[Dict]]:
    """Returns (best tuning dict, list of measured candidates)."""
    results = []
    for block, cpt, l in itertools.product(blocks, per_thread, lds):
        t0 = time.perf_counter()
        vms, sms = _time_candidate(case, precision, block, cpt, l, steps, warmup)
        results.append(dict(block=block, per_thread=cpt, lds=l, velocity_ms=vms, stress_ms=sms))
        if log:
            log("autotune: block %3d x %d per thread, lds %2d: velocity %.4f ms, stress %.4f ms (%.1f s)"
                % (block, cpt, l, vms, sms, time.perf_counter() - t0))
    bv = min(results, key=lambda r: r["velocity_ms"])
    bs = min(results, key=lambda r: r["stress_ms"])
    best = dict(block_velocity=bv["block"], per_thread_velocity=bv["per_thread"], lds_velocity=bv["lds"],
                block_stress=bs["block"], per_thread_stress=bs["per_thread"], lds_stress=bs["lds"])
    return best, results


# Choices that change a batch's launches: workgroup size and entities per thread of each
# kernel (LDS staging stays on: faster on every mesh measured here).
BATCH_CANDIDATES = [
    dict(block_velocity=bv, per_thread_velocity=cv, block_stress=bs, per_thread_stress=cs)
    for (bv, cv) in ((256, 1), (256, 2), (512, 1)) for (bs, cs) in ((256, 1), (256, 2))
]


def time_batched(case, precision, tuning, steps_per_block=25, blocks=4, repeats=3) -> float:
    """Seconds per step of ``case`` (or a list of cases, stepped together) in a batch with
    ``tuning``: the batched kernels and their graph-replayed K-step blocks, as a server runs
    them (median of ``repeats`` runs of ``blocks`` blocks after one warm-up block)."""
    group = list(case) if isinstance(case, (list, tuple)) else [case]
    solvers = [rsg.DeviceSolver.from_case(c, precision, tuning=tuning) for c in group]
    b = rsg.Batch(precision, steps_per_block=steps_per_block)
    try:
        slots = [b.add(s, (1 + blocks * repeats) * steps_per_block) for s in solvers]
        b.run(1)
        b.synchronize()
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            b.run(blocks)
            b.synchronize()
            times.append(time.perf_counter() - t0)
        for slot in slots:
            b.remove(slot)
        times.sort()
        return times[len(times) // 2] / (blocks * steps_per_block)
    finally:
        b.close()
        for s in solvers:
            s.close()


def tune_batched(case, precision=rsg.RSG_FP32, candidates=None, log=None, **kw) -> Tuple[Dict, List[Dict]]:
    """The fastest of ``candidates`` (default BATCH_CANDIDATES) for batched solves of meshes
    like ``case`` (a case, or a list of cases batched together); returns (best tuning,
    measured candidates)."""
    results = []
    for cand in candidates or BATCH_CANDIDATES:
        sec = time_batched(case, precision, cand, **kw)
        results.append(dict(cand, seconds_per_step=sec))
        if log:
            log("autotune: batched %s: %.4f ms/step" % (cand, sec * 1e3))
    best = min(results, key=lambda r: r["seconds_per_step"])
    return {k: v for k, v in best.items() if k != "seconds_per_step"}, results
