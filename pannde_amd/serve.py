"""Served inspection stream: parametric RSG solves on one GPU with continuous batching.

Requests (a cracked Ti-6Al-4V bracket or a cracked, holed CFRP plate, each with its own
crack length / depth / orientation, transducer position, mesh size and step count)
arrive as a seeded Poisson process.  The server

* admits arrived requests in order while their device footprint fits the HBM budget
  (fields + coefficients + connectivity + LDS plans, estimated from the mesh size before
  the solve is created) and the resident cell count stays under ``max_resident_cells``;
* builds each admitted job and creates its solver on host threads while the GPU works;
* folds ready solves into the one running batch at K-step boundaries (rsg_batch_add),
  where they share every launch with the solves already resident, and retires each solve
  when its own step count is reached (its fields stay resident until then), reading its
  results back off the loop that feeds the GPU;
* re-runs a seeded fraction of the requests in FP64 (a second batch on its own stream)
  and states each checked result's relative L2 sensor-trace error against it;
* isolates failures: a request whose preparation raises (a mesh the spacing rule rejects,
  device memory exhausted) or whose result is not finite (traces or field energy NaN or
  infinite) is reported with ``status`` "failed" and its error, and a request cancelled with
  ``cancel(rid)`` leaves the batch at the next K-step boundary with ``status`` "cancelled";
  either way its budget is released and the other requests run on unchanged.

Latency is the time from a request's arrival to its result being on the host.  Nothing
here is in the reference, which runs one solve per process
(applications/source/CaseRunners/TransducerModel.cpp); every solve's arithmetic is
rsg_step's, so a served result is bit-identical to the same request solved alone.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import dataclasses
import json
import math
import sys
import threading
import time
from typing import Dict, List, Optional

import numpy as np

from . import cases, rsg

# Device bytes per cell and per node of a solver (fields, coefficients, links, LDS plans,
# energy weights), from rsg_create's allocations; a 5 % margin and 4 MiB per solve for
# the amplitude and trace buffers are added on top.
BYTES_PER_CELL = {rsg.RSG_FP64: 303, rsg.RSG_FP32: 183}
BYTES_PER_NODE = {rsg.RSG_FP64: 119, rsg.RSG_FP32: 91}


def estimate_device_bytes(n_cells: int, n_nodes: int, precision: int) -> int:
    return int(1.05 * (BYTES_PER_CELL[precision] * n_cells + BYTES_PER_NODE[precision] * n_nodes)) + (4 << 20)


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float  # seconds after the stream starts
    kind: str  # "bracket" or "plate"
    cells: int  # target cell count
    nsteps: int
    crack_length: float
    crack_depth: float
    crack_angle: float
    xd: tuple  # transducer position on the inspection face, fractions of its extent
    seed: int
    spot_check: bool


def request_stream(seed=9, n=1000, rate=2.0, cells=(0.5e6, 8.0e6), steps=(1000, 4000), spot_fraction=0.05,
                   kinds=("bracket", "plate")) -> List[Request]:
    """Seeded Poisson arrivals at ``rate`` per second; mesh sizes log-uniform in ``cells``,
    steps uniform in ``steps``, crack length U(0.5, 5) mm, depth U(0.2, 2) mm, orientation
    U(0, 180) deg, transducer uniform on the inspection face."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        kind = kinds[int(rng.integers(0, len(kinds)))]
        target = int(math.exp(rng.uniform(math.log(cells[0]), math.log(cells[1]))))
        out.append(Request(rid=i, arrival=t, kind=kind, cells=target,
                           nsteps=int(rng.integers(steps[0], steps[1] + 1)),
                           crack_length=float(rng.uniform(0.5e-3, 5.0e-3)),
                           crack_depth=float(rng.uniform(0.2e-3, 2.0e-3)),
                           crack_angle=float(rng.uniform(0.0, 180.0)),
                           xd=(float(rng.uniform()), float(rng.uniform())),
                           seed=int(rng.integers(0, 2 ** 31 - 1)), spot_check=bool(rng.uniform() < spot_fraction)))
    return out


def mesh_dims(req: Request):
    """Grid dimensions for the request's target cell count."""
    if req.kind == "bracket":  # two legs of leg x w x w cells, leg = 16/3 w as in 512 x 96 x 96
        w = max(12, int(round((req.cells * 3 / 32) ** (1 / 3))))
        return dict(leg=max(2 * w, int(round(w * 16 / 3))), width=w, height=w)
    n = max(24, int(round(math.sqrt(req.cells / 64))))  # plate: n x n x 64 cells
    return dict(nx=n, ny=n)


def build_case(req: Request) -> cases.Case:
    dx = 5.0e-5
    d = mesh_dims(req)
    u, v = req.xd
    if req.kind == "bracket":
        L, W = d["leg"] * dx, d["width"] * dx
        return cases.ti_bracket(seed=req.seed, dx=dx, nsteps=req.nsteps, crack_length=req.crack_length,
                                crack_depth=req.crack_depth, crack_angle=req.crack_angle,
                                xd_center=((0.1 + 0.35 * u) * L, (0.2 + 0.6 * v) * W), **d)
    lx, ly, lz = d["nx"] * dx, d["ny"] * dx, 64 * dx
    rng = np.random.default_rng(req.seed)
    crack = cases.Crack((rng.uniform(0.2, 0.8) * lx, rng.uniform(0.2, 0.8) * ly, lz), req.crack_angle,
                        req.crack_length, min(req.crack_depth, 0.8 * lz))
    return cases.cfrp_plate(seed=req.seed, nsteps=req.nsteps, renumber=False, crack=crack,
                            xd_center=((0.25 + 0.5 * u) * lx, (0.25 + 0.5 * v) * ly), **d)


def rel_l2_traces(tr, ref):
    """Per-sensor relative L2 error over all steps and components (sensors whose reference
    norm is below 1e-6 of the largest are left out)."""
    num = np.sqrt(((tr - ref) ** 2).sum(axis=(0, 2)))
    den = np.sqrt((ref ** 2).sum(axis=(0, 2)))
    keep = den > 1e-6 * max(den.max(), 1e-300)
    return num[keep] / den[keep]


# Per-instance tuning for batched solves when the server does not tune: two nodes per
# velocity thread halve the workgroups of every batched velocity launch, and with them the
# per-workgroup item and descriptor loads (measured on 64 instances of 32^3..128^3 cells:
# +3 % in FP32, +2 % in FP64).  Results do not depend on the tuning.
BATCH_TUNING = {"per_thread_velocity": 2}

# Size classes of the autotuned server: a request's class is its kind and whether its target
# cell count is below SIZE_SPLIT; each class is tuned once, on a representative request of
# REPRESENTATIVE_CELLS[kind][bucket] cells.
SIZE_SPLIT = 2_000_000
REPRESENTATIVE_CELLS = {"bracket": (1_000_000, 3_000_000), "plate": (1_000_000, 3_000_000)}
_TUNED: Dict = {}  # (kind, bucket, precision) -> tuning, shared by the servers of a process


def size_class(req: "Request"):
    return req.kind, int(req.cells >= SIZE_SPLIT)


def tune_classes(classes, precision, log=None):
    """Autotuned batch tuning for each (kind, bucket) in ``classes`` (autotune.tune_batched
    on a representative request), cached per process; returns {class: tuning}."""
    from . import autotune

    out = {}
    for kind, bucket in sorted(classes):
        key = (kind, bucket, precision)
        if key not in _TUNED:
            rep = Request(rid=-1, arrival=0.0, kind=kind, cells=REPRESENTATIVE_CELLS[kind][bucket], nsteps=100,
                          crack_length=2e-3, crack_depth=1e-3, crack_angle=30.0, xd=(0.5, 0.5), seed=12345,
                          spot_check=False)
            best, _ = autotune.tune_batched(build_case(rep), precision)
            _TUNED[key] = best
            if log:
                log("serve: tuned %s %s cells: %s" % (kind, "<2M" if bucket == 0 else ">=2M", best))
        out[(kind, bucket)] = _TUNED[key]
    return out


@dataclasses.dataclass
class _Job:
    req: Request
    case: cases.Case
    solver: rsg.DeviceSolver
    bytes: int
    slot: int = -1
    admitted: float = 0.0
    ready: float = 0.0
    started: float = 0.0


class Server:
    """Continuous-batching server over one rsg batch (plus one FP64 batch for spot checks)."""

    def __init__(self, precision=rsg.RSG_FP32, hbm_budget=96 << 30, max_resident_cells=48_000_000,
                 steps_per_block=50, prep_threads=6, keep_traces=False, log=None, tuning=BATCH_TUNING,
                 on_block=None):
        """tuning: the solvers' rsg_tuning dict, or "auto": each size class of the stream is
        autotuned for batched solves when serve() starts (tune_classes; the time is reported
        as tuning_seconds and is not part of the served clock)."""
        self.precision = precision
        self.tuning = tuning
        self.class_tuning = {}
        self.hbm_budget = int(hbm_budget)
        self.max_resident_cells = int(max_resident_cells)
        self.K = int(steps_per_block)
        self.prep_threads = int(prep_threads)
        self.keep_traces = keep_traces
        self.log = log or (lambda *a: None)
        self.on_block = on_block  # called with the active request ids after each enqueued block
        self._cancel = set()
        self._cancel_lock = threading.Lock()

    def cancel(self, rid: int):
        """Asks the running serve() to drop request ``rid`` (thread-safe): a waiting request is
        never admitted, one being prepared is discarded when ready, an active one leaves the
        batch at the next K-step boundary.  Its result gets status "cancelled"."""
        with self._cancel_lock:
            self._cancel.add(int(rid))

    def _cancelled(self, rid: int) -> bool:
        with self._cancel_lock:
            return rid in self._cancel

    def _tuning_for(self, req: Request):
        if self.tuning == "auto":
            return self.class_tuning.get(size_class(req), BATCH_TUNING)
        return self.tuning

    def _prepare(self, req: Request, precision: int, case: Optional[cases.Case] = None):
        t0 = time.perf_counter()
        if case is None:
            case = build_case(req)
        s = rsg.DeviceSolver.from_case(case, precision, tuning=self._tuning_for(req))
        return case, s, time.perf_counter() - t0

    def serve(self, requests: List[Request], time_scale: float = 1.0) -> Dict:
        """Serves the requests (arrival times scaled by ``time_scale``) and returns the report.

        The loop keeps the GPU fed: it enqueues the next K-step block of every batch, then
        removes the solves that finished two blocks back (their last block has run by then, so
        removal does not wait), admits and folds in newcomers.  Reading a finished solve's
        traces and energies and tearing it down happen on threads of their own."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t_tune = time.perf_counter()
        if self.tuning == "auto":
            self.class_tuning = tune_classes({size_class(r) for r in reqs}, self.precision, log=self.log)
        t_tune = time.perf_counter() - t_tune
        pool = cf.ThreadPoolExecutor(max_workers=self.prep_threads)
        # finished solves are read back on threads of their own: those reads wait on the device,
        # which would otherwise hold preparation threads
        fetcher = cf.ThreadPoolExecutor(max_workers=2)
        main = rsg.Batch(self.precision, steps_per_block=self.K)
        check = rsg.Batch(rsg.RSG_FP64, steps_per_block=self.K)
        start = time.perf_counter()
        now = lambda: time.perf_counter() - start  # noqa: E731
        st = dict(nxt=0, used_bytes=0, resident_cells=0, cell_steps=0, blocks=0)
        preps = {}  # future -> (request, bytes, admitted time, is spot check)
        fetches = {}  # future -> job: a finished solve's results read back on a pool thread
        active: Dict[int, _Job] = {}  # main batch: slot -> job
        checking: Dict[int, _Job] = {}  # FP64 batch: slot -> job
        results: Dict[int, Dict] = {}
        spot_errors = []
        prep_seconds = []
        last_log = -1e9

        def admit():
            while st["nxt"] < len(reqs) and reqs[st["nxt"]].arrival * time_scale <= now():
                r = reqs[st["nxt"]]
                if self._cancelled(r.rid):
                    finish(r, "cancelled")
                    st["nxt"] += 1
                    continue
                nc = 2 * r.cells  # upper bound before the mesh exists (holes, rounding)
                est = estimate_device_bytes(nc, int(1.1 * nc), self.precision)
                busy = st["used_bytes"] > 0 or st["resident_cells"] > 0
                if busy and (st["used_bytes"] + est > self.hbm_budget
                             or st["resident_cells"] + r.cells > self.max_resident_cells):
                    return
                st["used_bytes"] += est
                st["resident_cells"] += r.cells
                preps[pool.submit(self._prepare, r, self.precision)] = (r, est, now(), False)
                st["nxt"] += 1

        def finish(r, status, error=None, **extra):
            """Records a request that ends without a valid result."""
            t_done = now()
            results[r.rid] = dict(rid=r.rid, kind=r.kind, cells=0, nsteps=r.nsteps, arrival=r.arrival * time_scale,
                                  admitted=t_done, ready=t_done, finished=t_done,
                                  latency=t_done - r.arrival * time_scale, status=status, error=error, **extra)

        def collect(timeout):
            if not preps and not fetches:
                return
            done, _ = cf.wait(list(preps) + list(fetches), timeout=timeout, return_when=cf.FIRST_COMPLETED)
            for f in done:
                if f in fetches:
                    fetched(fetches.pop(f), f.result())
                    continue
                r, est, t_adm, is_check = preps.pop(f)
                try:
                    case, s, secs = f.result()
                except Exception as e:  # this request only: release its share and go on
                    st["used_bytes"] -= est
                    if is_check:
                        results[r.rid]["spot_check_error"] = "%s: %s" % (type(e).__name__, e)
                        results[r.rid].pop("fp32_traces", None)
                    else:
                        st["resident_cells"] -= r.cells
                        finish(r, "failed", "%s: %s" % (type(e).__name__, e))
                    self.log("serve: request %d failed in preparation: %s" % (r.rid, e))
                    continue
                prep_seconds.append(secs)
                if not is_check and self._cancelled(r.rid):
                    st["used_bytes"] -= est
                    st["resident_cells"] -= r.cells
                    pool.submit(s.close)
                    finish(r, "cancelled")
                    continue
                job = _Job(req=r, case=case, solver=s, bytes=est, admitted=t_adm, ready=now())
                if is_check:
                    job.slot = check.add(s, r.nsteps)
                    checking[job.slot] = job
                else:
                    job.slot = main.add(s, r.nsteps)
                    job.started = now()
                    active[job.slot] = job

        def slot_left(batch, slot):
            done, total = batch.status(slot)
            return done < total

        def drop_cancelled():
            """Active requests cancelled since the last block leave the batch now."""
            for slot in [k for k, j in active.items() if self._cancelled(j.req.rid) and slot_left(main, k)]:
                job = active.pop(slot)
                main.remove(slot)  # waits for its lane's queued blocks
                st["used_bytes"] -= job.bytes
                st["resident_cells"] -= job.req.cells
                pool.submit(job.solver.close)
                finish(job.req, "cancelled")

        def enqueue(batch, jobs):
            """One K-step block if any member has steps left; returns the slots that finish in it."""
            if not any(batch.status(k)[0] < batch.status(k)[1] for k in jobs):
                return []
            before = {k for k in jobs if batch.status(k)[0] >= batch.status(k)[1]}
            batch.run(1)
            st["blocks"] += 1
            return [k for k in jobs if k not in before and batch.status(k)[0] >= batch.status(k)[1]]

        def fetch(job, is_check):
            """On a pool thread: the finished solve's traces and energies, then its tear-down.
            The host waits on the solver's own stream here, not in the loop that feeds the GPU.
            Returns (traces, error or None, time)."""
            try:
                tr = job.solver.traces()
                err = None
                # a diverged solve (non-finite traces or fields) is reported failed
                if not is_check and not (np.isfinite(tr).all() and all(math.isfinite(e) for e in job.solver.energy())):
                    err = "non-finite result (NaN or Inf in traces or fields)"
            except Exception as e:  # this request only
                tr, err = None, "%s: %s" % (type(e).__name__, e)
            finally:
                job.solver.close()
            return tr, err, now()

        def fetched(job, out):
            """A fetched result: recorded, its budget released (its memory is free by now)."""
            tr, error, t_done = out
            finite = error is None
            r = job.req
            st["cell_steps"] += job.case.n_cells * r.nsteps
            st["used_bytes"] -= job.bytes
            if job.slot < 0:  # an FP64 spot check: compared with the request's FP32 traces
                res = results[r.rid]
                if error is not None:
                    res["spot_check_error"] = error
                    res.pop("fp32_traces", None)
                    return
                err = rel_l2_traces(res.pop("fp32_traces"), tr)
                # None when no sensor saw the wave in the FP64 run (nothing to compare)
                res["spot_check_trace_rel_l2"] = float(err.max()) if err.size else None
                res["spot_check_sensors"] = int(err.size)
                if err.size:
                    spot_errors.append(res["spot_check_trace_rel_l2"])
                return
            st["resident_cells"] -= r.cells
            results[r.rid] = dict(rid=r.rid, kind=r.kind, cells=job.case.n_cells, nsteps=r.nsteps,
                                  arrival=r.arrival * time_scale, admitted=job.admitted, ready=job.ready,
                                  finished=t_done, latency=t_done - r.arrival * time_scale,
                                  status="ok" if finite else "failed", error=error)
            if not finite:
                self.log("serve: request %d failed: %s" % (r.rid, error))
            if self.keep_traces:
                results[r.rid]["traces"] = tr
            if r.spot_check and finite:  # the FP64 re-run reuses the case
                est = estimate_device_bytes(job.case.n_cells, job.case.n_nodes, rsg.RSG_FP64)
                st["used_bytes"] += est
                preps[pool.submit(self._prepare, r, rsg.RSG_FP64, job.case)] = (r, est, t_done, True)
                results[r.rid]["fp32_traces"] = tr

        def retire_main(slot):
            job = active.pop(slot)
            main.remove(slot)  # its last block ran one block ago: usually no wait
            fetches[fetcher.submit(fetch, job, False)] = job

        def retire_check(slot):
            job = checking.pop(slot)
            check.remove(slot)
            job.slot = -1
            fetches[fetcher.submit(fetch, job, True)] = job

        # the slots that finished in each of the last two enqueued blocks, older first: a solve is
        # retired once two more blocks have been enqueued after the one it finished in, so that
        # its removal finds that block done while the later ones still run
        pending_main, pending_check = [[], []], [[], []]
        while st["nxt"] < len(reqs) or preps or active or checking or fetches:
            if now() - last_log >= 15.0:
                last_log = now()
                self.log("serve: t=%.0f s, %d done, %d active, %d preparing, %d checking, %d waiting"
                         % (last_log, len(results) - len(checking), len(active), len(preps), len(checking),
                            len(reqs) - st["nxt"]))
            admit()
            collect(0)
            drop_cancelled()
            blocks0 = st["blocks"]
            new_main = enqueue(main, active)
            if self.on_block and st["blocks"] > blocks0:
                self.on_block([j.req.rid for j in active.values()])
            new_check = enqueue(check, checking)
            for slot in pending_main.pop(0):
                retire_main(slot)
            for slot in pending_check.pop(0):
                retire_check(slot)
            pending_main.append(new_main)
            pending_check.append(new_check)
            if not new_main and not new_check and not any(pending_main) and not any(pending_check):
                if not any(main.status(k)[0] < main.status(k)[1] for k in active) and \
                        not any(check.status(k)[0] < check.status(k)[1] for k in checking):
                    # nothing on the GPU: wait for a solver to be ready or the next arrival
                    wait = None
                    if st["nxt"] < len(reqs):
                        wait = max(0.0, reqs[st["nxt"]].arrival * time_scale - now())
                    if preps or fetches:
                        collect(wait if wait is not None else 1.0)
                    elif wait is not None:
                        time.sleep(min(wait, 0.05))
        main.synchronize()
        check.synchronize()
        elapsed = now()
        main.close()
        check.close()
        pool.shutdown()
        fetcher.shutdown()
        ok = [r for r in reqs if results[r.rid].get("status", "ok") == "ok"]
        lat = np.array([results[r.rid]["latency"] for r in ok]) if ok else np.zeros(1)
        served_cell_steps = sum(results[r.rid]["cells"] * results[r.rid]["nsteps"] for r in ok)
        report = dict(
            requests=len(reqs), completed=len(ok), elapsed_s=elapsed, solves_per_hour=len(ok) / elapsed * 3600.0,
            latency_p50_s=float(np.percentile(lat, 50)), latency_p99_s=float(np.percentile(lat, 99)),
            latency_mean_s=float(lat.mean()),
            gcell_steps_per_s=served_cell_steps / elapsed / 1e9,
            gcell_steps_per_s_with_checks=st["cell_steps"] / elapsed / 1e9,
            blocks=st["blocks"],
            prep_seconds_mean=float(np.mean(prep_seconds)) if prep_seconds else 0.0,
            spot_checks=sum(1 for r in ok if r.spot_check),
            spot_check_max_trace_rel_l2=max(spot_errors) if spot_errors else None,
            tuning=({"%s_%s" % (k, "lt2M" if b == 0 else "ge2M"): t for (k, b), t in self.class_tuning.items()}
                    if self.tuning == "auto" else self.tuning),
            tuning_seconds=t_tune,
            failed=[r.rid for r in reqs if results[r.rid].get("status") == "failed"],
            cancelled=[r.rid for r in reqs if results[r.rid].get("status") == "cancelled"],
            results=[results[r.rid] for r in reqs])
        return report


def main(argv=None):
    ap = argparse.ArgumentParser(description="Serve a seeded stream of parametric RSG inspection solves")
    ap.add_argument("--requests", type=int, default=1000)
    ap.add_argument("--rate", type=float, default=2.0, help="arrivals per second")
    ap.add_argument("--seed", type=int, default=9)
    ap.add_argument("--min-cells", type=float, default=0.5e6)
    ap.add_argument("--max-cells", type=float, default=8.0e6)
    ap.add_argument("--min-steps", type=int, default=1000)
    ap.add_argument("--max-steps", type=int, default=4000)
    ap.add_argument("--spot-fraction", type=float, default=0.05)
    ap.add_argument("--hbm-budget-gib", type=float, default=96.0)
    ap.add_argument("--max-resident-cells", type=float, default=48e6)
    ap.add_argument("--steps-per-block", type=int, default=50)
    ap.add_argument("--prep-threads", type=int, default=6)
    ap.add_argument("--no-autotune", action="store_true", help="fixed batch tuning instead of tuning per size class")
    ap.add_argument("--out", default=None, help="JSON report path")
    a = ap.parse_args(argv)
    reqs = request_stream(seed=a.seed, n=a.requests, rate=a.rate, cells=(a.min_cells, a.max_cells),
                          steps=(a.min_steps, a.max_steps), spot_fraction=a.spot_fraction)
    srv = Server(hbm_budget=int(a.hbm_budget_gib * (1 << 30)), max_resident_cells=int(a.max_resident_cells),
                 steps_per_block=a.steps_per_block, prep_threads=a.prep_threads,
                 tuning=BATCH_TUNING if a.no_autotune else "auto",
                 log=lambda *m: print(*m, file=sys.stderr, flush=True))
    rep = srv.serve(reqs)
    summary = {k: v for k, v in rep.items() if k != "results"}
    print(json.dumps(summary), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(rep, f)


if __name__ == "__main__":
    main()
