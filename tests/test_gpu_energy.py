"""rsg_energy on the device against the numpy checker (tests/energy_ref.py), closed forms,
and conservation after the source stops.

Bar: relative 1e-12 against the checker in FP64 (only the summation order differs);
closed forms to 1e-12 (FP64) and 1e-6 (FP32, whose state and dt*C are rounded to FP32)."""
import math

import numpy as np
import pytest

import energy_ref
from conftest import load_golden
from pannde_amd import cases, rsg

pytestmark = pytest.mark.gpu

FORMS = [None, dict(coeff_classes=-1), dict(lds_stress=-1, lds_velocity=-1)]


def close(a, b, tol):
    return abs(a - b) <= tol * max(abs(a), abs(b))


@pytest.mark.parametrize("tuning", FORMS, ids=["keys_packed", "per_entity_arrays", "keys_direct"])
def test_energy_matches_checker_after_steps(tuning):
    case = cases.al_block(seed=7, nx=12, ny=10, nz=6, nsteps=60, n_sensors=4)
    s = rsg.DeviceSolver.from_case(case, tuning=tuning)
    s.solve(case.nsteps)
    V, S = s.get_state()
    k, u = s.energy()
    s.close()
    k0, u0 = energy_ref.energy(case, V, S)
    print("kinetic %.6e / %.6e, strain %.6e / %.6e" % (k, k0, u, u0))
    assert k0 > 0 and u0 > 0
    assert close(k, k0, 1e-12) and close(u, u0, 1e-12)


@pytest.mark.parametrize("tuning", FORMS, ids=["keys_packed", "per_entity_arrays", "keys_direct"])
def test_energy_with_nodal_anisotropic_stiffness(tuning):
    """The reference's own state (fixture) on NODE-placed C with shear couplings: a
    mis-ordered upper triangle or a missing /8 average changes the strain energy."""
    case, ref = load_golden("water_column_nodec")
    s = rsg.DeviceSolver.from_case(case, tuning=tuning)
    s.set_state(ref["V"], ref["S"])
    k, u = s.energy()
    s.close()
    k0, u0 = energy_ref.energy(case, ref["V"], ref["S"])
    assert close(k, k0, 1e-12) and close(u, u0, 1e-12), (k, k0, u, u0)


@pytest.mark.parametrize("precision,tol", [(rsg.RSG_FP64, 1e-12), (rsg.RSG_FP32, 1e-6)])
def test_energy_closed_forms(precision, tol):
    c = cases.al_block(seed=1, nx=9, ny=7, nz=5, nsteps=1, n_sensors=2)
    lam, mu, rho = cases.AL["lam"], cases.AL["mu"], cases.AL["rho"]
    L = np.ptp(c.x) * np.ptp(c.y) * np.ptp(c.z)
    s = rsg.DeviceSolver.from_case(c, precision)
    V = np.tile(np.array([[1.0], [-2.0], [0.5]]), (1, c.n_nodes))
    S = np.zeros((6, c.n_cells))
    s.set_state(V, S)
    k, u = s.energy()
    assert close(k, 0.5 * rho * L * 5.25, tol) and u == 0.0, k
    S[5] = 2.0e6
    s.set_state(np.zeros((3, c.n_nodes)), S)
    k, u = s.energy()
    assert k == 0.0 and close(u, L * 2.0e6 ** 2 / (2 * mu), tol), u
    E = mu * (3 * lam + 2 * mu) / (lam + mu)
    S[:] = 0.0
    S[0] = 3.0e6
    s.set_state(None, S)
    assert close(s.energy()[1], L * 3.0e6 ** 2 / (2 * E), tol)
    s.close()


def test_energy_is_conserved_after_the_source_stops():
    case = cases.al_block(seed=3, nx=14, ny=12, nz=8, nsteps=1, n_sensors=2)
    pulse = len(case.force[0].signal_t)
    s = rsg.DeviceSolver.from_case(case)
    s.solve(pulse + 5)
    e = []
    for _ in range(6):
        s.solve(37)
        k, u = s.energy()
        e.append(k + u)
    s.close()
    e = np.array(e)
    assert e.max() / e.min() - 1 < 0.05, e


def test_energy_fp32_against_fp64():
    case = cases.al_block(seed=8, nx=16, ny=12, nz=8, nsteps=150, n_sensors=4)
    out = {}
    for prec in (rsg.RSG_FP64, rsg.RSG_FP32):
        s = rsg.DeviceSolver.from_case(case, prec)
        s.solve(case.nsteps)
        out[prec] = s.energy()
        s.close()
    for a, b in zip(out[rsg.RSG_FP32], out[rsg.RSG_FP64]):
        assert math.isfinite(a) and close(a, b, 1e-4), out
