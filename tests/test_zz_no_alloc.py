"""rsg_energy_noalloc (the allocation-free energy over a static device work area) through the
C ABI: the same values as rsg_energy, bit for bit, on two calls in a row, the second with a
smaller mesh, in both precisions."""
import ctypes

import pytest

from pannde_amd import cases, rsg

pytestmark = pytest.mark.gpu


def energy_noalloc(s):
    k, u = ctypes.c_double(0), ctypes.c_double(0)
    rc = rsg.lib().rsg_energy_noalloc(s._h, ctypes.byref(k), ctypes.byref(u))
    assert rc == rsg.RSG_OK, rsg.last_error()
    return k.value, u.value


@pytest.mark.parametrize("precision", [rsg.RSG_FP64, rsg.RSG_FP32])
def test_noalloc_energy_equals_energy(precision):
    big = cases.al_block(seed=41, nx=20, ny=16, nz=10, nsteps=90, n_sensors=4)
    small = cases.al_block(seed=42, nx=9, ny=7, nz=5, nsteps=40, n_sensors=2)
    for case in (big, small):  # the second call has the smaller input
        s = rsg.DeviceSolver.from_case(case, precision)
        s.solve(case.nsteps)
        ref = s.energy()
        got = energy_noalloc(s)
        again = energy_noalloc(s)
        s.close()
        print(case.name, ref, got)
        assert ref[0] > 0 and ref[1] > 0
        assert got == ref and again == ref
