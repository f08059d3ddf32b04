"""The workloads at their full configured sizes on the GPU.

* CFRP plate 256x256x64 (3.4 M cells) and the cracked Ti-6Al-4V L-bracket 2 x 512x96x96
  (9.4 M cells): the FP64 DeviceSolver equals the oracle bit for bit after a few steps
  (state, time and sensor traces), source active.
* Batched: the 64 instances of 32^3..128^3 cells for their 1500 steps in 50-step graph blocks,
  in FP64 and FP32: every instance equals the same instance stepped alone, bit for bit, and
  every FP32 sensor trace is within 1e-4 relative L2 of the FP64 one.
* Served: the first 8 requests of the seeded inspection stream at its configured sizes
  (0.5M-8M cells, 1000-4000 steps): each served result equals the request solved alone, and
  the FP64 spot checks hold the FP32 traces within 1e-4.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle_rsg as orc
from pannde_amd import cases, rsg, serve

pytestmark = pytest.mark.gpu


def device_run(case, nsteps, precision=rsg.RSG_FP64, tuning=None):
    s = rsg.DeviceSolver.from_case(case, precision, tuning=tuning)
    try:
        t = s.solve(nsteps)
        V, S = s.get_state()
        return V, S, t, s.traces()
    finally:
        s.close()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("workload,nsteps", [("plate", 6), ("bracket", 4)])
def test_full_size_fp64_matches_oracle(workload, nsteps):
    case = cases.cfrp_plate(seed=2, nsteps=nsteps) if workload == "plate" else cases.ti_bracket(seed=5, nsteps=nsteps)
    print("%s: %d cells, %d nodes, %d steps" % (case.name, case.n_cells, case.n_nodes, nsteps), flush=True)
    V, S, t, tr = device_run(case, nsteps)
    V0, S0, t0, tr0 = orc.run_oracle(case, nsteps=nsteps)
    assert np.abs(V0).max() > 0 and np.abs(S0).max() > 0  # the source has acted
    assert t == t0
    assert np.array_equal(V, V0) and np.array_equal(S, S0)
    assert np.array_equal(tr, tr0)


@pytest.mark.timeout(900)
def test_served_requests_fp64_match_oracle():
    """Requests of the inspection stream as the server builds them (a cracked, holed plate in the
    generator's own numbering and a cracked bracket, transducer and crack placed per request):
    the FP64 DeviceSolver equals the oracle bit for bit after a few steps, source active."""
    reqs = serve.request_stream(seed=9, n=1000, rate=4.0)[:24]
    picks = [min((r for r in reqs if r.kind == kind), key=lambda r: r.cells) for kind in ("plate", "bracket")]
    for r in picks:
        case = dataclasses.replace(serve.build_case(r), nsteps=5)
        print("%s request %d: %d cells, %d nodes, %d duplicated nodes" % (r.kind, r.rid, case.n_cells, case.n_nodes,
                                                                          case.meta.get("n_dup", 0)), flush=True)
        assert case.meta.get("n_dup", 0) > 0  # the crack is open
        V, S, t, tr = device_run(case, case.nsteps)
        V0, S0, t0, tr0 = orc.run_oracle(case, nsteps=case.nsteps)
        assert np.abs(V0).max() > 0 and np.abs(S0).max() > 0
        assert t == t0
        assert np.array_equal(V, V0) and np.array_equal(S, S0)
        assert np.array_equal(tr, tr0)


@pytest.mark.timeout(1200)
def test_batch_config_equals_alone_and_fp32_within_tolerance():
    inst = cases.batch_instances(seed=4)
    assert len(inst) == 64 and inst[0].nsteps == 1500
    traces = {}
    for prec in (rsg.RSG_FP64, rsg.RSG_FP32):
        solvers = [rsg.DeviceSolver.from_case(c, prec, tuning=serve.BATCH_TUNING) for c in inst]
        b = rsg.Batch(prec, steps_per_block=50)
        slots = [b.add(s, c.nsteps) for s, c in zip(solvers, inst)]
        assert b.run(10 ** 6) == 0
        b.synchronize()
        traces[prec] = []
        for s, slot, c in zip(solvers, slots, inst):
            b.remove(slot)
            V, S = s.get_state()
            t, tr = s.time, s.traces()
            s.close()
            V0, S0, t0, tr0 = device_run(c, c.nsteps, prec)
            assert t == t0, c.name
            assert np.array_equal(V, V0) and np.array_equal(S, S0), c.name
            assert np.array_equal(tr, tr0), c.name
            traces[prec].append(tr)
        b.close()
        print("precision %d: 64 instances batched == alone" % prec, flush=True)
    worst = max(float(serve.rel_l2_traces(a, r).max()) for a, r in zip(traces[rsg.RSG_FP32], traces[rsg.RSG_FP64]))
    print("worst FP32 trace rel L2 vs FP64: %.3e" % worst)
    assert worst <= 1e-4


@pytest.mark.timeout(1200)
def test_stream_prefix_at_configured_sizes():
    reqs = serve.request_stream(seed=9, n=1000, rate=4.0)[:8]
    # two of them re-run in FP64 (the stream's own 5 % may pick none of the first 8)
    reqs = [dataclasses.replace(r, spot_check=r.spot_check or r.rid in (1, 4)) for r in reqs]
    assert min(r.cells for r in reqs) >= 0.5e6 and max(r.nsteps for r in reqs) <= 4000
    srv = serve.Server(keep_traces=True, prep_threads=8)
    rep = srv.serve(reqs)
    assert rep["requests"] == len(reqs) and not rep.get("failed")
    res = {r["rid"]: r for r in rep["results"]}
    for r in reqs:
        got = res[r.rid]
        case = serve.build_case(r)
        _, _, _, tr = device_run(case, r.nsteps, rsg.RSG_FP32)
        assert np.array_equal(got["traces"], tr), r.rid
        if r.spot_check:
            assert got["spot_check_sensors"] > 0
            assert 0 < got["spot_check_trace_rel_l2"] <= 1e-4, (r.rid, got["spot_check_trace_rel_l2"])
        print("request %d: %s %d cells %d steps: served == alone" % (r.rid, r.kind, got["cells"], r.nsteps),
              flush=True)
