"""Served stream on the GPU: requests that arrive while others run are admitted as the
budget frees up, folded into the running batch at K-step boundaries, and each result is
bit-identical to the same request solved alone; the FP64 spot checks hold the FP32
results within 1e-4 relative L2 per sensor trace."""
import numpy as np
import pytest

from pannde_amd import rsg, serve

pytestmark = pytest.mark.gpu


def small_stream(n=10, seed=21, spot=0.3):
    return serve.request_stream(seed=seed, n=n, rate=40.0, cells=(15e3, 60e3), steps=(60, 200), spot_fraction=spot)


def alone_traces(req):
    case = serve.build_case(req)
    s = rsg.DeviceSolver.from_case(case, rsg.RSG_FP32)
    try:
        s.solve(req.nsteps)
        return s.traces()
    finally:
        s.close()


def test_served_results_equal_alone_and_spot_checks_pass():
    reqs = small_stream()
    assert any(r.spot_check for r in reqs)
    # a resident-cell cap of about three requests forces arrivals to wait for retirements
    srv = serve.Server(steps_per_block=16, max_resident_cells=150_000, prep_threads=3, keep_traces=True)
    rep = srv.serve(reqs)
    assert rep["requests"] == len(reqs)
    res = {r["rid"]: r for r in rep["results"]}
    admitted = sorted(r["admitted"] for r in rep["results"])
    finished = sorted(r["finished"] for r in rep["results"])
    # some request was admitted only after an earlier one had finished (overlap + admission)
    assert any(a > finished[0] for a in admitted)
    for r in reqs:
        got = res[r.rid]
        assert got["finished"] >= got["ready"] >= got["admitted"] >= got["arrival"] - 1e-9
        assert got["traces"].shape[0] == r.nsteps
        assert np.abs(got["traces"]).max() > 0, r.rid  # the wave reached the sensors
        ref = alone_traces(r)
        assert np.array_equal(got["traces"], ref), r.rid
        if r.spot_check:
            assert got["spot_check_sensors"] > 0
            assert 0 < got["spot_check_trace_rel_l2"] <= 1e-4, (r.rid, got["spot_check_trace_rel_l2"])
    assert rep["spot_checks"] == sum(r.spot_check for r in reqs)
    assert rep["latency_p99_s"] >= rep["latency_p50_s"] > 0


def test_byte_budget_admits_one_at_a_time():
    reqs = small_stream(n=4, seed=22, spot=0.0)
    # a budget below any single request still admits one request at a time
    srv = serve.Server(hbm_budget=1 << 20, steps_per_block=16, prep_threads=2, keep_traces=True)
    rep = srv.serve(reqs)
    rows = sorted(rep["results"], key=lambda r: r["admitted"])
    for a, b in zip(rows, rows[1:]):
        assert b["admitted"] >= a["finished"] - 1e-9
    for r in reqs:
        got = [x for x in rep["results"] if x["rid"] == r.rid][0]
        assert np.array_equal(got["traces"], alone_traces(r))


class _Faulty(serve.Server):
    """Host-side injections: request `bad_prep` raises in preparation, request `bad_state`
    starts from a state holding a NaN (written with set_state), request `cancel_ready` is
    cancelled while it is prepared."""

    def __init__(self, bad_prep, bad_state, cancel_ready, **kw):
        super().__init__(**kw)
        self.bad_prep, self.bad_state, self.cancel_ready = bad_prep, bad_state, cancel_ready

    def _prepare(self, req, precision, case=None):
        if req.rid == self.bad_prep and precision != rsg.RSG_FP64:
            raise RuntimeError("injected preparation failure")
        case, s, secs = super()._prepare(req, precision, case)
        if req.rid == self.bad_state and precision != rsg.RSG_FP64:
            V = np.zeros((3, s.n_nodes))
            V[0, s.n_nodes // 2] = np.nan
            s.set_state(V, None)
        if req.rid == self.cancel_ready:
            self.cancel(req.rid)
        return case, s, secs


def test_failures_and_cancellations_leave_the_other_requests_unchanged():
    reqs = small_stream(n=9, seed=23, spot=0.0)
    state = {"cancelled_active": None}

    def on_block(active):
        # cancel the first request seen active with another block still to come
        if state["cancelled_active"] is None:
            for rid in active:
                if rid not in (1, 3, 5, 7):
                    srv.cancel(rid)
                    state["cancelled_active"] = rid
                    return

    srv = _Faulty(bad_prep=1, bad_state=3, cancel_ready=5, steps_per_block=16, prep_threads=3, keep_traces=True,
                  on_block=on_block)
    srv.cancel(7)  # cancelled before it arrives: never admitted
    rep = srv.serve(reqs)
    res = {r["rid"]: r for r in rep["results"]}
    assert res[1]["status"] == "failed" and "injected" in res[1]["error"]
    assert res[3]["status"] == "failed" and "non-finite" in res[3]["error"]
    assert res[5]["status"] == "cancelled" and res[7]["status"] == "cancelled"
    ca = state["cancelled_active"]
    assert ca is not None and res[ca]["status"] == "cancelled"
    assert sorted(rep["failed"]) == [1, 3] and sorted(rep["cancelled"]) == sorted([5, 7, ca])
    ok = [r for r in reqs if r.rid not in (1, 3, 5, 7, ca)]
    assert rep["completed"] == len(ok) >= 3
    for r in ok:
        assert res[r.rid]["status"] == "ok"
        assert np.array_equal(res[r.rid]["traces"], alone_traces(r)), r.rid
