"""The autotuner returns a valid per-kernel choice; the tuned solver's results are the
default solver's, bit for bit (every candidate computes the same arithmetic)."""
import numpy as np
import pytest

from conftest import load_golden
from pannde_amd import autotune, rsg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [rsg.RSG_FP64, rsg.RSG_FP32])
def test_autotune_picks_a_valid_tuning(precision):
    case, _ = load_golden("cfrp_holes_rcm")
    best, table = autotune.tune(case, precision, steps=4, warmup=1, blocks=(128, 256), per_thread=(1, 2),
                                lds=(1, -1))
    assert len(table) == 8 and all(r["velocity_ms"] > 0 and r["stress_ms"] > 0 for r in table)
    assert best["block_velocity"] in (128, 256) and best["per_thread_stress"] in (1, 2)
    assert best["lds_velocity"] in (1, -1) and best["lds_stress"] in (1, -1)
    out = []
    for tuning in (best, None):
        s = rsg.DeviceSolver.from_case(case, precision, tuning=tuning)
        s.solve(case.nsteps)
        out.append((s.get_state(), s.traces()))
        s.close()
    (V0, S0), T0 = out[0]
    (V1, S1), T1 = out[1]
    assert np.array_equal(V0, V1) and np.array_equal(S0, S1) and np.array_equal(T0, T1)


def test_batched_autotuner_and_autotuned_server(monkeypatch):
    """tune_batched times candidates on the batched kernels and returns one of them; a server
    with tuning="auto" tunes each size class of its stream and serves every request with the
    result of solving it alone."""
    from pannde_amd import serve

    case, _ = load_golden("cfrp_holes_rcm")
    cands = autotune.BATCH_CANDIDATES[:3]
    best, table = autotune.tune_batched(case, rsg.RSG_FP32, candidates=cands, steps_per_block=8, blocks=2,
                                        repeats=2)
    assert len(table) == 3 and all(r["seconds_per_step"] > 0 for r in table)
    assert best in cands
    # small representatives keep the per-class tuning short here
    monkeypatch.setattr(serve, "REPRESENTATIVE_CELLS", {"bracket": (60_000, 200_000), "plate": (60_000, 200_000)})
    monkeypatch.setattr(serve, "_TUNED", {})
    reqs = serve.request_stream(seed=24, n=6, rate=40.0, cells=(15e3, 60e3), steps=(60, 120), spot_fraction=0.0)
    srv = serve.Server(tuning="auto", steps_per_block=16, prep_threads=3, keep_traces=True)
    rep = srv.serve(reqs)
    kinds = {r.kind for r in reqs}
    assert set(rep["tuning"]) == {k + "_lt2M" for k in kinds}
    assert all(t in autotune.BATCH_CANDIDATES for t in rep["tuning"].values())
    assert rep["tuning_seconds"] > 0 and rep["completed"] == len(reqs)
    res = {r["rid"]: r for r in rep["results"]}
    for r in reqs:
        s = rsg.DeviceSolver.from_case(serve.build_case(r), rsg.RSG_FP32)
        s.solve(r.nsteps)
        assert np.array_equal(res[r.rid]["traces"], s.traces()), r.rid
        s.close()
