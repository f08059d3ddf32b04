#!/usr/bin/env python3
"""Throughput of the RSG elastodynamic time step on one MI355X.

Workload: the orthotropic CFRP laminate plate, 256x256x64 cells at 0.05 mm, 16 plies
[0/45/90/-45]2s, five seeded through-holes (cells removed), node ids seeded-permuted then
RCM-renumbered, 2 MHz 5-cycle burst on a 2 mm top-face disc, 16 sensors; FP64
DeviceSolver (the reference's arithmetic, bit-identical results).  One "step" is one
Model::solve(1): velocity update + stress update + source + sensor recording, with every
input resident in HBM.

Printed: one JSON line (stdout).  value = Gcell-steps/s of the FP64 DeviceSolver over the
timed steps; roofline = the stress kernel's algorithmic bytes per launch over its mean
launch duration (HIP events on the solver's stream, same timed region); cpu_baseline =
the reference solver itself (oracle/_ref/ref_rsg, built from the reference headers) under
mpirun on this host's cores, on a bounded sample of the same job; fp32 = the mixed-
precision form on the same job with its sensor-trace error against FP64; bracket = the
cracked Ti-6Al-4V L-bracket (9.4 M cells, 4000 steps) in both forms with the FP32 trace error;
batched = 64 independent instances (32^3..128^3 cells, 1500 steps, 50-step graph blocks)
advanced together in FP32, every instance's traces checked against the same batch run in
FP64; served = a bounded sample (the first 120 requests) of the seeded inspection stream
(cracked brackets and plates, 0.5M-8M cells, 1000-4000 steps, Poisson arrivals at 4 per
second) through the continuous-batching server, with solves/hr, p50/p99 latency and the
FP64 spot checks; served_saturated = the same requests arriving at 40 per second, faster
than the server retires them, so solves/hr is its capacity (latency then includes the
queue).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# The reference's data model (per-entity coefficients, 32-bit connectivity; each element
# touched once per pass), FP64 bytes per cell (stress kernel) and per node (velocity kernel):
#   stress: C 21x8 + S read+write 6x8x2 + spacings 3x8 + vertices 8x4 + V gathered 3x8 = 344
#   velocity: V read+write 3x8x2 + dt/rho 8 + spacings 3x8 + cell links 8x4 + S gathered 6x8 = 160
#   FP32: stress 188, velocity 96.  This layout moves fewer bytes than that model, so the model is
#   reported separately and the roofline uses the layout's own algorithmic bytes.
REFERENCE_MODEL_BYTES = {"stress": {"f64": 344, "f32": 188}, "velocity": {"f64": 160, "f32": 96}}
# Every 32nd step's two launches are bracketed by HIP events inside the timed region (about 60
# samples per kernel in a 2000-step run): events between every launch slowed the stream by 4 %
# (FP64) to 8 % (FP32).
TIMING_STRIDE = 32
# the sources the counter traffic in profiles/ was measured with
KERNEL_SOURCES = ("pannde_amd/csrc/rsg_kernels.hip", "pannde_amd/csrc/rsg_internal.h")
# Untimed first steps of the drop-in driver leg (code-object loading on first launch).
DROPIN_UNTIMED = 10
# The CPU baseline uses at most this many MPI ranks: the host's CPU share for one GPU.
CPU_SHARE = 16


def layout_bytes(tuning, nc, nn, elt):
    """Algorithmic bytes per launch of this layout (each array element the kernel needs read
    or written once per pass): (stress, velocity).  Per cell: the link record (16 B packed
    10-bit links with the key; 16 B 16-bit links + a 4-byte key; 32 B int32 links + key), the
    per-cell coefficients when there are no class keys (21 C + 3 spacings), S read and written;
    plus every node's V gathered once.  Per node likewise, with dt/rho + 3 spacings, V read and
    written, and every cell's S gathered once.  Gather lists and class tables are not counted."""
    def record(links, keyed):
        return 16 if links == 2 else (16 if links == 1 else 32) + (4 if keyed else 0)

    keyed_c, keyed_n = tuning["cell_classes"] >= 0, tuning["node_classes"] >= 0
    stress = nc * (record(tuning["links_stress"], keyed_c) + (0 if keyed_c else 24 * elt) + 12 * elt) + nn * 3 * elt
    velocity = nn * (record(tuning["links_velocity"], keyed_n) + (0 if keyed_n else 4 * elt) + 6 * elt) + nc * 6 * elt
    return stress, velocity


def kernel_sources_digest():
    import hashlib

    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def physical_cores():
    """Physical cores of this host (lscpu), or None."""
    import subprocess

    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], stdout=subprocess.PIPE, text=True, timeout=30).stdout
    except Exception:
        return None
    cores = {tuple(l.split(",")) for l in out.splitlines() if l and not l.startswith("#")}
    return len(cores) or None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_l2(a, b):
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return float(d / n) if n > 0 else float(d)


def run_timed(rsg, case, precision, warmup, steps):
    s = rsg.DeviceSolver.from_case(case, precision)
    s.solve(warmup)
    s.synchronize()
    s.kernel_times()  # reset
    s.set_timing(True, stride=TIMING_STRIDE)
    t0 = time.perf_counter()
    s.solve(steps)
    s.synchronize()
    t1 = time.perf_counter()
    s.set_timing(False)
    vms, sms, launches = s.kernel_times()
    traces = s.traces()
    V, S = s.get_state()
    dev_bytes = s.device_bytes()
    tuning = s.tuning()
    s.close()
    return dict(seconds=t1 - t0, vel_ms=vms, str_ms=sms, launches=launches, traces=traces, V=V, S=S,
                dev_bytes=dev_bytes, tuning=tuning)


def create_all(rsg, case_list, precision, threads=8, tuning=None):
    import concurrent.futures as cf

    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda c: rsg.DeviceSolver.from_case(c, precision, tuning=tuning), case_list))


def batched(rsg, cases, seed, steps_per_block=50, lanes=None):
    """All instances in one batch for their full step count, FP32 timed; the same batch in
    FP64 for the per-trace error.  lanes: the batch's scheduler lanes (None: the library's
    default)."""
    from pannde_amd import serve

    from pannde_amd import autotune

    inst = cases.batch_instances(seed=seed)
    nsteps = inst[0].nsteps
    cells = sum(c.n_cells for c in inst)
    out = {"instances": len(inst), "cells": cells, "steps": nsteps, "steps_per_block": steps_per_block}
    # the autotuner's batched choice on 8 instances spread over the size range, for every instance
    t0 = time.perf_counter()
    by_size = sorted(inst, key=lambda c: c.n_cells)
    group = [by_size[(2 * k + 1) * len(by_size) // 16] for k in range(8)]
    tuning, _ = autotune.tune_batched(group, rsg.RSG_FP32)
    out["tuning"] = tuning
    out["tuning_seconds"] = round(time.perf_counter() - t0, 2)
    log("bench: batched tuning on 8 instances (%d cells): %s" % (sum(c.n_cells for c in group), tuning))
    traces = {}
    for prec, tag in ((rsg.RSG_FP32, "fp32"), (rsg.RSG_FP64, "fp64")):
        t0 = time.perf_counter()
        solvers = create_all(rsg, inst, prec, tuning=tuning)
        t_create = time.perf_counter() - t0
        b = rsg.Batch(prec, steps_per_block=steps_per_block, lanes=lanes)
        for s in solvers:
            b.add(s, nsteps)
        b.synchronize()
        t0 = time.perf_counter()
        left = b.run(10 ** 6)
        b.synchronize()
        dt = time.perf_counter() - t0
        assert left == 0
        traces[tag] = [s.traces() for s in solvers]
        elt = 4 if prec == rsg.RSG_FP32 else 8
        step_bytes = sum(sum(layout_bytes(s.tuning(), c.n_cells, c.n_nodes, elt)) for s, c in zip(solvers, inst))
        for slot in list(b.members):
            b.remove(slot)
        b.close()
        for s in solvers:
            s.close()
        gbs = step_bytes * nsteps / dt / 1e9
        out[tag] = {"value": round(cells * nsteps / dt / 1e9, 4), "seconds": round(dt, 3),
                    "create_seconds": round(t_create, 2),
                    # both kernels' algorithmic bytes over the wall clock of the whole batch (launch gaps
                    # and graph set-up included; measured without a profiler)
                    "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        log("bench: batched %s: %d instances, %d cells, %d steps in %.3f s -> %.3f Gcell-steps/s (create %.1f s)"
            % (tag, len(inst), cells, nsteps, dt, out[tag]["value"], t_create))
    errs = [float(serve.rel_l2_traces(a, b).max()) for a, b in zip(traces["fp32"], traces["fp64"])]
    out["max_trace_rel_l2_fp32_vs_fp64"] = float("%.3e" % max(errs))
    return out


def bracket(rsg, cases):
    """The cracked Ti-6Al-4V L-bracket (2 x 512x96x96 cells, 4000 steps) in FP64 and FP32:
    rates and the per-trace error of the mixed-precision form."""
    from pannde_amd import serve

    case = cases.ti_bracket(seed=5)
    # the first steps of the run are untimed (first launches, table uploads); every step is
    # still part of the run, so the traces compared below cover all of them
    untimed = 100
    out = {"cells": case.n_cells, "nodes": case.n_nodes, "steps": case.nsteps,
           "crack_duplicated_nodes": case.meta["n_dup"], "untimed_first_steps": untimed}
    tr = {}
    for prec, tag in ((rsg.RSG_FP64, "fp64"), (rsg.RSG_FP32, "fp32")):
        r = run_timed(rsg, case, prec, untimed, case.nsteps - untimed)
        tr[tag] = r
        out[tag] = round(case.n_cells * (case.nsteps - untimed) / r["seconds"] / 1e9, 4)
    err = serve.rel_l2_traces(tr["fp32"]["traces"], tr["fp64"]["traces"])
    out["max_trace_rel_l2_fp32_vs_fp64"] = float("%.3e" % err.max())
    out["wavefield_rel_l2_fp32_vs_fp64"] = float("%.3e" % max(rel_l2(tr["fp32"]["V"], tr["fp64"]["V"]),
                                                              rel_l2(tr["fp32"]["S"], tr["fp64"]["S"])))
    log("bench: bracket %d cells, %d steps: FP64 %.3f, FP32 %.3f Gcell-steps/s, trace err %.2e"
        % (case.n_cells, case.nsteps, out["fp64"], out["fp32"], err.max()))
    return out


def served(rsg, n, rate, seed):
    from pannde_amd import serve

    reqs = serve.request_stream(seed=seed, n=n, rate=rate)
    srv = serve.Server(prep_threads=10, log=log, tuning="auto")
    rep = srv.serve(reqs)
    keys = ("requests", "completed", "elapsed_s", "solves_per_hour", "latency_p50_s", "latency_p99_s",
            "latency_mean_s", "gcell_steps_per_s", "gcell_steps_per_s_with_checks", "blocks", "prep_seconds_mean",
            "spot_checks", "spot_check_max_trace_rel_l2", "tuning", "tuning_seconds", "failed")
    out = {k: (float("%.4g" % rep[k]) if isinstance(rep[k], float) else rep[k]) for k in keys}
    out.update(arrival_rate_per_s=rate, seed=seed, hbm_budget_gib=srv.hbm_budget / (1 << 30),
               max_resident_cells=srv.max_resident_cells, steps_per_block=srv.K)
    log("bench: served %d requests in %.1f s: %.0f solves/hr, p50 %.2f s, p99 %.2f s, %.2f Gcell-steps/s"
        % (n, rep["elapsed_s"], rep["solves_per_hour"], rep["latency_p50_s"], rep["latency_p99_s"],
           rep["gcell_steps_per_s"]))
    return out


def dropin(rsg, case, steps, fast_rate):
    """The C++ drop-in (pannde_amd/rsg_driver: PanNDE::Model over the C ABI) driven the way
    TransducerModel drives HostRSGTDUT: time = model->solve() once per step
    (TransducerModel.cpp:68-73).  Two source forms: the library's signal table, and
    HostNormalXDZ hidden behind an opaque MultiVariate (evaluated on the host every step,
    one channel per support cell).  fast_rate: the same job's resident solve(N) rate."""
    import subprocess
    import tempfile

    from pannde_amd.case_file import write_case_file

    exe = os.path.join(ROOT, "pannde_amd", "rsg_driver")
    out = {"cells": case.n_cells, "steps": steps, "untimed_first_steps": DROPIN_UNTIMED}
    d = tempfile.mkdtemp(prefix="rsgdrop_", dir="/tmp")
    try:
        cp, op = os.path.join(d, "case.bin"), os.path.join(d, "out.bin")
        write_case_file(cp, case, nsteps=steps)
        for mode, flags in (("table", []), ("opaque", ["--opaque"])):
            # the first steps load the kernels' code objects: the timer starts after them
            r = subprocess.run([exe, cp, op, "--time-from", str(DROPIN_UNTIMED)] + flags, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True, timeout=900)
            if r.returncode != 0:
                raise RuntimeError("rsg_driver failed: %s" % r.stdout[-2000:])
            kv = dict(x.split("=", 1) for x in r.stdout.split() if "=" in x)
            secs, timed = float(kv["seconds"]), int(kv["timed_steps"])
            rate = case.n_cells * timed / secs / 1e9
            out[mode] = {"value": round(rate, 4), "ms_per_step": round(secs * 1e3 / timed, 4),
                         "overhead_us_per_step": round((secs / timed - case.n_cells / fast_rate / 1e9) * 1e6, 1)}
    finally:
        import shutil

        shutil.rmtree(d, ignore_errors=True)
    out["resident_solve_n_value"] = round(fast_rate, 4)
    return out


def cpu_baseline(case, ranks, steps, warmup=10, repeat=3):
    """The reference solver (oracle/_ref/ref_rsg, compiled from the reference headers) under
    mpirun -n ranks: warmup untimed steps, then the median of `repeat` windows of `steps` steps.
    Falls back to the oracle restatement on one core when the binary is missing."""
    from oracle import oracle_rsg as orc

    sample = ("%s mesh (%d cells): %d warm-up steps, then the median of %d windows of %d steps"
              % (case.name, case.n_cells, warmup, repeat, steps))
    if orc.have_reference():
        try:
            _, _, _, _, tim = orc.run_reference(case, nsteps=steps, ranks=ranks, warmup=warmup, repeat=repeat,
                                                workdir="/tmp")
            v = case.n_cells * steps / tim["seconds"] / 1e9
            return dict(value=v, unit="Gcell-steps/s", cores=ranks, kind="reference",
                        sample=sample + ", reference HostRSGTDUT under mpirun -n %d" % ranks,
                        windows_s=[round(w, 4) for w in tim.get("windows", [])])
        except Exception as e:  # fall through to the port
            log("cpu baseline: reference binary failed (%s); timing the oracle port" % e)
    t0 = time.perf_counter()
    orc.run_oracle(case, nsteps=steps)
    dt = time.perf_counter() - t0
    return dict(value=case.n_cells * steps / dt / 1e9, unit="Gcell-steps/s", cores=1, kind="port",
                sample="%s mesh (%d cells), %d steps, oracle restatement, one core (includes its set-up)"
                % (case.name, case.n_cells, steps))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-ranks", type=int, default=0, help="0: the host's physical cores, at most 16")
    ap.add_argument("--cpu-steps", type=int, default=3, help="steps per timed window of the plate's CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-bracket", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--dropin-steps", type=int, default=200, help="plate steps of the drop-in driver leg")
    ap.add_argument("--batch-seed", type=int, default=4)
    ap.add_argument("--batch-lanes", type=int, default=0, help="scheduler lanes of the batch (0: library default)")
    ap.add_argument("--serve-requests", type=int, default=120)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-saturated-rate", type=float, default=40.0,
                    help="arrival rate of the capacity leg (0: skip it)")
    ap.add_argument("--serve-seed", type=int, default=9)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the DeviceSolver holds the whole mesh on one GPU; multi-GPU runs are not part of "
              "this benchmark (requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from pannde_amd import cases, rsg

    log("bench: building the CFRP plate job (seed %d)" % args.seed)
    case = cases.cfrp_plate(seed=args.seed, nsteps=args.warmup + args.steps)
    nc, nn = case.n_cells, case.n_nodes
    log("bench: %d cells, %d nodes, dt %.4e s" % (nc, nn, case.dt))

    log("bench: FP64 DeviceSolver, %d warm-up + %d timed steps" % (args.warmup, args.steps))
    r64 = run_timed(rsg, case, rsg.RSG_FP64, args.warmup, args.steps)
    value = nc * args.steps / r64["seconds"] / 1e9
    ms_step = r64["seconds"] * 1e3 / args.steps
    str_avg = r64["str_ms"] / max(1, r64["launches"]) * 1e-3
    vel_avg = r64["vel_ms"] / max(1, r64["launches"]) * 1e-3
    tun = r64["tuning"]
    str_bytes, vel_bytes = layout_bytes(tun, nc, nn, 8)
    str_gbs = str_bytes / str_avg / 1e9
    vel_gbs = vel_bytes / vel_avg / 1e9
    step_gbs = (str_bytes + vel_bytes) / (str_avg + vel_avg) / 1e9
    log("bench: FP64 %.3f Gcell-steps/s, %.4f ms/step; stress %.1f us (%.0f GB/s), velocity %.1f us (%.0f GB/s)"
        % (value, ms_step, str_avg * 1e6, str_gbs, vel_avg * 1e6, vel_gbs))

    # measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE, calibrated), valid only for
    # the kernel sources it was measured with
    traffic, traffic_note = None, "no counter profile for this workload"
    tpath = os.path.join(ROOT, "profiles", "rsg_kernels_traffic_cfrp_plate.json")
    if os.path.exists(tpath) and args.seed == 2:
        with open(tpath) as f:
            prof = json.load(f)
        if prof.get("kernel_sources_sha256") == kernel_sources_digest():
            traffic = prof["hbm_bytes_per_launch"].get("stress_f64")
            traffic_note = "rocprofv3 counters at commit %s, same kernel sources" % prof.get("commit")
        else:
            traffic_note = "counter profile is of other kernel sources (commit %s): not used" % prof.get("commit")
    layout = {"coefficients": ("class keys: %d cell materials, %d node materials" % (tun["cell_classes"],
                                                                                   tun["node_classes"])
                               if tun["cell_classes"] >= 0 else "per-entity arrays"),
              "links": {2: "10-bit block-local packed with the key (LDS staging)",
                        1: "16-bit block-local (LDS staging)", 0: "32-bit global"}[tun["links_stress"]],
              "cells_per_thread_stress": tun["per_thread_stress"]}
    model_str = REFERENCE_MODEL_BYTES["stress"]["f64"] * nc
    model_vel = REFERENCE_MODEL_BYTES["velocity"]["f64"] * nn

    out = {
        "metric": "Gcell-steps/s FP64 vs FP32 @<=1e-4 trace err; served solves/hr + p50/p99 lat",
        "value": round(value, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded CFRP plate job generated in-process)",
        "config": {"workload": "cfrp_plate_256x256x64_dx0.05mm_16ply_5holes_rcm_2MHz", "cells": nc,
                   "nodes": nn, "seed": args.seed, "solver": "DeviceSolver FP64 (HostRSGTDUT arithmetic)",
                   # the served legs run a bounded sample of the 1000-request stream; the whole
                   # stream is profiled separately
                   "served_sample": "first %d requests of the seeded 1000-request stream (seed %d)"
                                    % (args.serve_requests, args.serve_seed)},
        "roofline": {"bound": "hbm", "achieved": round(str_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(str_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "stress_kernel<double>",
                     "bytes_per_launch": str_bytes,
                     "bytes_model": "this layout's algorithmic bytes: per cell the 16-byte link+key record, "
                                    "S read and written (96 B); every node's V gathered once (24 B)",
                     "mean_launch_us": round(str_avg * 1e6, 2),
                     "traffic_note": traffic_note,
                     "traffic_over_algorithmic": round(traffic / str_bytes, 3) if traffic else None,
                     "traffic_frac": round(traffic / str_avg / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                     "velocity_kernel": {"achieved": round(vel_gbs, 1), "frac": round(vel_gbs / HBM_PEAK_GBS, 4),
                                         "bytes_per_launch": vel_bytes, "mean_launch_us": round(vel_avg * 1e6, 2)},
                     "full_step_frac": round(step_gbs / HBM_PEAK_GBS, 4),
                     "layout": layout,
                     # the reference's data model (per-entity coefficients and 32-bit links), which this
                     # layout does not move: a rate on it exceeds what HBM can carry
                     "reference_data_model": {"stress_bytes_per_launch": model_str,
                                              "velocity_bytes_per_launch": model_vel,
                                              "stress_gbs_equivalent": round(model_str / str_avg / 1e9, 1),
                                              "note": "not a roofline: this layout moves fewer bytes"}},
        "device_bytes": r64["dev_bytes"],
    }

    if not args.no_fp32:
        log("bench: FP32 DeviceSolver on the same job")
        r32 = run_timed(rsg, case, rsg.RSG_FP32, args.warmup, args.steps)
        v32 = nc * args.steps / r32["seconds"] / 1e9
        tr_err = [rel_l2(r32["traces"][:, s, 2], r64["traces"][:, s, 2]) for s in range(r64["traces"].shape[1])]
        wf_err = max(rel_l2(r32["V"], r64["V"]), rel_l2(r32["S"], r64["S"]))
        s32 = r32["str_ms"] / max(1, r32["launches"]) * 1e-3
        out["fp32"] = {"value": round(v32, 4), "ratio_vs_fp64": round(v32 / value, 3),
                       "max_trace_rel_l2_vs_fp64": float("%.3e" % max(tr_err)),
                       "wavefield_rel_l2_vs_fp64": float("%.3e" % wf_err),
                       "stress_kernel_gbs": round(layout_bytes(r32["tuning"], nc, nn, 4)[0] / s32 / 1e9, 1)}
        log("bench: FP32 %.3f Gcell-steps/s (x%.2f), trace err %.2e, wavefield err %.2e"
            % (v32, v32 / value, max(tr_err), wf_err))

    if not args.no_bracket:
        log("bench: cracked L-bracket, FP64 and FP32")
        out["bracket"] = bracket(rsg, cases)

    if not args.no_batched:
        log("bench: batched FP32 instances (seed %d)" % args.batch_seed)
        out["batched"] = batched(rsg, cases, args.batch_seed, lanes=args.batch_lanes or None)
        out["batched"]["ratio_fp32_batched_vs_fp64_plain"] = round(out["batched"]["fp32"]["value"] / value, 3)

    if not args.no_served:
        log("bench: served stream, %d requests at %.2f per second" % (args.serve_requests, args.serve_rate))
        out["served"] = served(rsg, args.serve_requests, args.serve_rate, args.serve_seed)
        out["served"]["ratio_served_vs_fp64_plain"] = round(out["served"]["gcell_steps_per_s"] / value, 3)
        if args.serve_saturated_rate > 0:
            log("bench: served capacity, %d requests at %.2f per second (arrivals outpace the server)"
                % (args.serve_requests, args.serve_saturated_rate))
            out["served_saturated"] = served(rsg, args.serve_requests, args.serve_saturated_rate, args.serve_seed)

    if not args.no_dropin:
        log("bench: C++ drop-in driver, one solve() per step (plate %d steps, Al block %d steps)"
            % (args.dropin_steps, 500))
        out["dropin"] = {"plate": dropin(rsg, case, args.dropin_steps, value)}
        c1 = cases.al_block(seed=1)
        r1 = run_timed(rsg, c1, rsg.RSG_FP64, 10, c1.nsteps)
        out["dropin"]["al_block"] = dropin(rsg, c1, c1.nsteps, c1.n_cells * c1.nsteps / r1["seconds"] / 1e9)
        log("bench: drop-in %s" % json.dumps(out["dropin"]))

    if not args.no_cpu_baseline:
        phys = physical_cores()
        ranks = args.cpu_ranks or max(1, min(CPU_SHARE, phys or os.cpu_count() or 1))
        log("bench: CPU baseline, reference solver under mpirun -n %d (host: %s physical cores), %d warm-up "
            "steps + median of 3 windows of %d steps" % (ranks, phys, 10, args.cpu_steps))
        cb = cpu_baseline(case, ranks, args.cpu_steps)
        cb["host_physical_cores"] = phys
        cb["cores_note"] = ("ranks = the host's physical cores, capped at the %d-core CPU share a one-GPU job has "
                            "on this host" % CPU_SHARE)
        cb["gpu_over_cpu"] = round(value / cb["value"], 1)
        # configuration 1 as the reference configures it: the Al block 40x40x20 under mpirun -n 4
        c1 = cases.al_block(seed=1)
        c1_steps = 150
        log("bench: CPU baseline, Al block %d cells under mpirun -n 4, median of 3 windows of %d steps"
            % (c1.n_cells, c1_steps))
        cb1 = cpu_baseline(c1, 4, c1_steps)
        r1 = run_timed(rsg, c1, rsg.RSG_FP64, 10, c1.nsteps)
        cb["al_block_40x40x20_n4"] = {"value": round(cb1["value"], 6), "cores": cb1["cores"], "kind": cb1["kind"],
                                      "sample": cb1["sample"], "windows_s": cb1.get("windows_s"),
                                      "device_fp64_value": round(c1.n_cells * c1.nsteps / r1["seconds"] / 1e9, 4),
                                      "device_steps": c1.nsteps}
        out["cpu_baseline"] = {k: (round(v, 6) if isinstance(v, float) else v) for k, v in cb.items()}
        log("bench: CPU baseline %.5f Gcell-steps/s on %d cores (%s); Al block -n 4: %.5f"
            % (cb["value"], cb["cores"], cb["kind"], cb1["value"]))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
