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
