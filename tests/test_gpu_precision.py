"""Mixed precision against the FP64 DeviceSolver on the same jobs, at full size.

The FP32 form keeps fields and updates in FP32 and does the stiffness and inverse-density
precompute (dt*C, 0.25/spacing) in FP64 before rounding once; sensor traces and the
energy are accumulated in FP64.  Bar: the relative L2 error of every sensor trace (all
steps, three components) is at most 1e-4 against the FP64 run; traces whose FP64 norm is
below 1e-6 of the largest (sensors the wave has not reached) are left out of the ratio.
The final wavefield and the energy are held to 1e-4 as well (measured values print:
about 1e-6 for the block and the plate, 6e-6 for the cracked bracket after 4000 steps).
"""
import numpy as np
import pytest

from pannde_amd import cases, rsg

pytestmark = pytest.mark.gpu

TRACE_TOL = 1e-4
FIELD_TOL = 1e-4


def run(case, precision):
    s = rsg.DeviceSolver.from_case(case, precision)
    try:
        s.solve(case.nsteps)
        V, S = s.get_state()
        return V, S, s.traces(), s.energy()
    finally:
        s.close()


def trace_errors(tr, ref):
    num = np.sqrt(((tr - ref) ** 2).sum(axis=(0, 2)))
    den = np.sqrt((ref ** 2).sum(axis=(0, 2)))
    keep = den > 1e-6 * den.max()
    return num[keep] / den[keep], int(keep.sum())


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


WORKLOADS = {
    "al_block_40x40x20_500": lambda: cases.al_block(),
    "cfrp_plate_256x256x64_2000": lambda: cases.cfrp_plate(),
    "ti_bracket_2x512x96x96_crack_4000": lambda: cases.ti_bracket(),
}


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_fp32_traces_and_fields_against_fp64(name):
    case = WORKLOADS[name]()
    V64, S64, tr64, e64 = run(case, rsg.RSG_FP64)
    V32, S32, tr32, e32 = run(case, rsg.RSG_FP32)
    assert tr32.shape == tr64.shape and tr64.shape[0] == case.nsteps
    err, n = trace_errors(tr32, tr64)
    ev, es = rel_l2(V32, V64), rel_l2(S32, S64)
    ek = abs(e32[0] - e64[0]) / e64[0]
    eu = abs(e32[1] - e64[1]) / e64[1]
    print("%s: %d traces, trace rel L2 max %.3e median %.3e; V %.3e S %.3e; energy %.3e %.3e"
          % (name, n, err.max(), np.median(err), ev, es, ek, eu), flush=True)
    assert n >= 4
    assert err.max() <= TRACE_TOL
    assert ev <= FIELD_TOL and es <= FIELD_TOL
    assert ek <= FIELD_TOL and eu <= FIELD_TOL
