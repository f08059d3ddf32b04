"""Pins the numpy energy checker (tests/energy_ref.py) on the CPU: closed forms for the mass
and the strain energy, and conservation of their sum on oracle states once the source has
stopped (the property that fixes the kinetic weight against the strain weight)."""
import math

import numpy as np

import energy_ref
from oracle import oracle_rsg as orc
from pannde_amd import cases


def test_node_masses_sum_to_the_body_mass():
    c = cases.al_block(seed=1, nx=7, ny=5, nz=4, nsteps=1, n_sensors=2)
    mass, vol, _ = energy_ref.weights(c)
    L = [np.ptp(c.x), np.ptp(c.y), np.ptp(c.z)]
    assert math.isclose(mass.sum(), cases.AL["rho"] * L[0] * L[1] * L[2], rel_tol=1e-12)
    assert math.isclose(vol.sum(), L[0] * L[1] * L[2], rel_tol=1e-12)
    # a rigid translation carries 1/2 M |v|^2
    V = np.tile(np.array([[1.0], [-2.0], [0.5]]), (1, c.n_nodes))
    k, u = energy_ref.energy(c, V, np.zeros((6, c.n_cells)))
    assert math.isclose(k, 0.5 * mass.sum() * 5.25, rel_tol=1e-12) and u == 0.0


def test_strain_energy_closed_forms():
    c = cases.al_block(seed=1, nx=6, ny=5, nz=4, nsteps=1, n_sensors=2)
    lam, mu = cases.AL["lam"], cases.AL["mu"]
    E = mu * (3 * lam + 2 * mu) / (lam + mu)
    vol = np.ptp(c.x) * np.ptp(c.y) * np.ptp(c.z)
    V = np.zeros((3, c.n_nodes))
    S = np.zeros((6, c.n_cells))
    S[0] = 3.0e6  # uniaxial Sxx: energy density s^2 / 2E
    assert math.isclose(energy_ref.energy(c, V, S)[1], vol * 3.0e6 ** 2 / (2 * E), rel_tol=1e-12)
    S[:] = 0.0
    S[5] = 2.0e6  # pure shear Sxy: tau^2 / 2 mu (C66 = mu acts on the engineering shear rate)
    assert math.isclose(energy_ref.energy(c, V, S)[1], vol * 2.0e6 ** 2 / (2 * mu), rel_tol=1e-12)
    S[:] = 0.0
    S[:3] = 1.5e6  # hydrostatic: p^2 / 2K per unit volume times 3 ... = 3 p^2 / (2 (3 lam + 2 mu))
    assert math.isclose(energy_ref.energy(c, V, S)[1], vol * 3 * 1.5e6 ** 2 / (2 * (3 * lam + 2 * mu)),
                        rel_tol=1e-12)


def energy_history(case, start, blocks, stride, kinetic_scale=1.0):
    V, S, t, _ = orc.run_oracle(case, nsteps=start)
    w = energy_ref.weights(case)
    out = []
    for _ in range(blocks):
        V, S, t, _ = orc.run_oracle(case, nsteps=stride, V0=V, S0=S, t0=t)
        k, u = energy_ref.energy(case, V, S, w)
        out.append(kinetic_scale * k + u)
    return np.array(out)


def test_energy_is_conserved_after_the_source_stops():
    case = cases.al_block(seed=3, nx=14, ny=12, nz=8, nsteps=1, n_sensors=2)
    pulse = len(case.force[0].signal_t)
    e = energy_history(case, pulse + 5, 6, 37)
    spread = e.max() / e.min() - 1
    assert spread < 0.05, e
    # the check discriminates: doubling the kinetic weight breaks it
    e2 = energy_history(case, pulse + 5, 6, 37, kinetic_scale=2.0)
    assert e2.max() / e2.min() - 1 > 0.15, e2
