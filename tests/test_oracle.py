"""Pins the CPU oracle (oracle/cpu_rsg.c) before anything is compared against it.

* the reference's own known-answer tests for the stencil, the field averaging, the
  signal and the source window (modules/HostSolver/testing/include/HostRSGTDUTTests.hpp,
  modules/PanNDE/testing/include/FieldTests.hpp, modules/HostData/testing/include/
  HostUnivariateTests.hpp, modules/Stubs/testing/include/HostNormalXDZTests.hpp);
* the golden fixtures: outputs of the reference solver itself (tests/golden);
* closed forms and invariants of the scheme (rigid translation, exact linearity,
  uniform strain rate, mirror symmetry);
* planted mistakes: each of five plausible slips in the oracle fails these checks.
"""
import math

import numpy as np
import pytest

from conftest import golden_names, load_golden
from oracle import oracle_rsg as orc
from pannde_amd import cases


def plate_builder(length, width, thickness, nlength, nthickness):
    """The mesh of Stubs::PlateBuilder (PlateBuilder.hpp:72-170): square xy cells of
    length/nlength, nthickness cells through z, numbering kx outer, kz inner."""
    ds = length / nlength
    ny = int(math.ceil(width / ds))
    x, y, z, cells, _ = cases.structured_mesh(nlength, ny, nthickness, ds, ds, thickness / nthickness)
    return x, y, z, cells


def dd(x, y, z, cells, node_field=None, cell_field=None):
    nf = np.zeros(x.shape[0]) if node_field is None else node_field
    cf = np.zeros(cells.shape[0]) if cell_field is None else cell_field
    return orc.dd_probe(x, y, z, cells, nf, cf)


def close4ulp(a, b):
    return a == b or abs(a - b) <= 4 * np.spacing(max(abs(a), abs(b)))


# ---------------------------------------------------------------- RightHex KATs

def test_unit_cell_mesh():  # HostRSGTDUTTests.hpp:117-156
    x, y, z, cells = plate_builder(1.0, 1.0, 0.25, 1, 1)
    assert cells.shape == (1, 8) and x.shape[0] == 8
    expect = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, .25), (1, 0, .25), (1, 1, .25), (0, 1, .25)]
    for v, e in enumerate(expect):
        n = cells[0, v]
        assert (x[n], y[n], z[n]) == e


def test_derivatives_of_coordinates():  # :161-184
    x, y, z, cells = plate_builder(1.0, 1.0, 0.25, 1, 1)
    for k, f in enumerate((x, y, z)):
        dcc = dd(x, y, z, cells, node_field=f)[0][0]
        for a in range(3):
            assert dcc[a] == (1.0 if a == k else 0.0)


def test_derivatives_of_distance():  # :185-200
    x, y, z, cells = plate_builder(1.0, 1.0, 0.25, 1, 1)
    f = np.sqrt(x * x + y * y + z * z)
    dcc = dd(x, y, z, cells, node_field=f)[0][0]
    ddx = 0.25 * (math.sqrt(2.) + (math.sqrt(33.) / 4.) - (1. / 4.)) / 1.
    ddz = 0.25 * ((math.sqrt(17.) / 2. + math.sqrt(33.) / 4. + 1. / 4.) - (2. + math.sqrt(2.))) / 0.25
    assert close4ulp(dcc[0], ddx) and close4ulp(dcc[1], ddx) and close4ulp(dcc[2], ddz)


def test_derivatives_of_node_index():  # :201-214
    x, y, z, cells = plate_builder(1.0, 1.0, 0.25, 1, 1)
    dcc = dd(x, y, z, cells, node_field=np.arange(8, dtype=float))[0][0]
    assert list(dcc) == [4.0, 2.0, 4.0]


def test_eight_pack_node_spacing_and_cell_centres():  # :215-256
    x, y, z, cells = plate_builder(1.0, 1.0, 0.25, 2, 2)
    assert cells.shape[0] == 8 and x.shape[0] == 27
    assert (x[13], y[13], z[13]) == (0.5, 0.5, 0.125)
    for idx in range(3):
        centre = np.array([sum((x, y, z)[idx][cells[c, v]] / 8. for v in range(8)) for c in range(8)])
        _, dnc, dn, _ = dd(x, y, z, cells, cell_field=centre)
        assert list(dn[13]) == [0.5, 0.5, 0.125]
        for a in range(3):
            assert close4ulp(dnc[13][a], 1.0 if a == idx else 0.0)


def test_node_spacing_is_mirrored_at_faces():
    """nodePrep mirrors the missing bound: a face node of a uniform grid gets the
    interior spacing, not half of it (RightHex.hpp:290-301)."""
    x, y, z, cells, _ = cases.structured_mesh(3, 2, 2, 0.5, 0.25, 0.125)
    _, _, dn, dc = dd(x, y, z, cells)
    assert np.all(dn[:, 0] == 0.5) and np.all(dn[:, 1] == 0.25) and np.all(dn[:, 2] == 0.125)
    assert np.all(dc == np.array([0.5, 0.25, 0.125]))


def test_degenerate_mesh_raises_dx():
    x, y, z, cells, _ = cases.structured_mesh(1, 1, 1, 1.0, 1.0, 1.0)
    x = np.zeros_like(x)  # every node at x = 0
    with pytest.raises(ValueError, match="DX"):
        dd(x, y, z, cells)


def test_slot_convention():  # MeshTests.hpp:125-151 (slot s = octant of the cell)
    x, y, z, cells = plate_builder(1.0, 1.0, 0.25, 2, 2)
    links = orc.build_links(x.shape[0], cells)
    for s in range(8):
        c = links[13, s]
        cx = sum(x[cells[c, v]] for v in range(8)) / 8
        cy = sum(y[cells[c, v]] for v in range(8)) / 8
        cz = sum(z[cells[c, v]] for v in range(8)) / 8
        octant = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)][s]
        assert (int(cx > x[13]), int(cy > y[13]), int(cz > z[13])) == octant


# ---------------------------------------------------------------- field / signal / window KATs

def test_cell_to_node_average_counts_missing_cells_as_zero():  # FieldTests.hpp:216-224
    x, y, z, cells, _ = cases.structured_mesh(3, 2, 2, 1.0, 1.0, 1.0)
    links = orc.build_links(x.shape[0], cells)
    ones = cases.Field(cases.FIELD_CELL, np.ones(cells.shape[0]))
    const = cases.Field(cases.FIELD_CONSTANT, np.ones(1))
    for n in range(x.shape[0]):
        cnt = int(np.sum(links[n] != -1))
        assert orc.field_at_node(ones, links, n) == cnt / 8.0
        assert orc.field_at_node(const, links, n) == cnt / 8.0
    nodef = cases.Field(cases.FIELD_NODE, x + 2 * y + 3 * z)
    for c in range(cells.shape[0]):
        assert orc.field_at_cell(nodef, cells, c) == sum(nodef.data[cells[c]]) / 8.0


def test_univariate_linear_and_held():  # HostUnivariateTests.hpp:52-61
    k = np.arange(101)
    t, v = 0.01 * k, -10. * k
    got = orc.signal_eval(t, v, [0.01, 0.001, 0.0001, 0.015, 1., 1.1, 11.])
    want = [-10., -1., -0.1, -15., -1000., -1000., -1000.]
    np.testing.assert_allclose(got, want, rtol=1e-6)  # EXPECT_FLOAT_EQ


def test_univariate_repeated_key_keeps_first():
    got = orc.signal_eval([0., 1., 1., 2.], [0., 5., 9., 7.], [1.0, 1.5])
    assert list(got) == [5.0, 6.0]


def xdz_signal():  # HostNormalXDZTests.hpp:22-36
    nt, dt = 1001, 0.001
    k = np.arange(nt)
    resp = 0.5 * (1. - np.cos(2. * np.pi * k / (nt - 1))) * np.sin(2. * np.pi * 5. * k / (nt - 1))
    return k * dt, resp


def test_normal_xdz_window():  # HostNormalXDZTests.hpp:58-68
    center, ds, radius = (1.0, -1.0, 0.5), 0.01, 0.5
    t, v = xdz_signal()
    s = orc.signal_eval(t, v, [1., 10., 0.25])
    assert orc.disc_check(center, radius, ds, (1., -1., .5)) and s[0] == 0.0 and s[1] == 0.0
    assert not orc.disc_check(center, radius, ds, (1.5, -1., .5))  # dist == radius is outside
    assert orc.disc_check(center, radius, ds, (1., -1., .5)) and s[2] > 0
    assert orc.disc_check(center, radius, ds, (1., -1., .5 + 0.0075))  # |dz| <= 0.75 ds
    assert not orc.disc_check(center, radius, ds, (1., -1., .5 + 0.0076))


def test_write_time_schedule():  # TransducerModel.cpp:68-73
    dt = 0.1
    wt = np.arange(4) * 10 * dt
    steps = orc.schedule(dt, wt)
    assert steps[0] == 0 and steps.sum() == sum(steps)
    t = 0.0
    for k, w in enumerate(wt):
        n = 0
        while t < w:
            t += dt
            n += 1
        assert steps[k] == n


def test_write_time_schedule_exact_cases():
    """What exact arithmetic fixes: with dt a power of two and write times on multiples of dt
    (the case makers' arange(NT+1) * interval * dt), time accumulates exactly, so the loop
    `while (time < write_time) solve()` takes exactly `interval` steps per write and none
    before the first (write time 0).  A non-strict comparison would take one more step at
    every write."""
    for dt, interval, nt in ((0.125, 8, 4), (2.0 ** -20, 37, 6), (1.0, 1, 3)):
        wt = np.arange(nt + 1) * interval * dt
        steps = orc.schedule(dt, wt)
        assert list(steps) == [0] + [interval] * nt


# ---------------------------------------------------------------- golden fixtures

@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_fixture_bitwise(name):
    case, ref = load_golden(name)
    V, S, t, tr = orc.run_oracle(case)
    assert t == ref["time"]
    assert np.array_equal(V, ref["V"]) and np.array_equal(S, ref["S"])
    assert np.array_equal(tr, ref["traces"])
    assert np.abs(V).max() > 0 and np.abs(S).max() > 0


@pytest.mark.skipif(not orc.have_reference(), reason="oracle/_ref/ref_rsg not built (reference tree absent)")
@pytest.mark.parametrize("seed", [101, 102, 103])
def test_oracle_matches_reference_on_random_jobs(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = rng.integers(4, 9, size=3)
    case = cases.al_block(seed=seed, nx=int(nx), ny=int(ny), nz=int(nz), nsteps=15, n_sensors=3,
                          rho_type=int(rng.integers(0, 3)))
    # a randomly perturbed, non-uniform C and a moment form half of the time
    for f in case.cij:
        f.data = f.data * (1 + 0.05 * rng.standard_normal(f.data.shape))
    if seed % 2:
        wc = cases.water_column_block(seed=seed, nx=int(nx), ny=int(ny), nz=int(nz), nsteps=15)
        case.moment, case.moment_form = wc.moment, 1
    V, S, t, tr = orc.run_oracle(case)
    V2, S2, t2, tr2, _ = orc.run_reference(case, ranks=1 + seed % 3)
    assert t == t2 and np.array_equal(V, V2) and np.array_equal(S, S2) and np.array_equal(tr, tr2)


# ---------------------------------------------------------------- invariants and closed forms

def quiet(case):
    case.force, case.moment = [], []
    return case


def test_rigid_translation_is_exact():
    case = quiet(cases.al_block(nx=6, ny=5, nz=4, nsteps=10, n_sensors=2))
    V0 = np.zeros((3, case.n_nodes))
    V0[0], V0[1], V0[2] = 0.3, -1.25, 2.0
    V, S, _, _ = orc.run_oracle(case, V0=V0)
    assert np.array_equal(V, V0) and not np.any(S)


def test_linearity_is_bitwise():
    case = cases.al_block(nx=7, ny=6, nz=4, nsteps=20, n_sensors=3)
    V1, S1, _, tr1 = orc.run_oracle(case)
    case.force[0].signal_v = 2.0 * case.force[0].signal_v
    V2, S2, _, tr2 = orc.run_oracle(case)
    assert np.array_equal(V2, 2 * V1) and np.array_equal(S2, 2 * S1) and np.array_equal(tr2, 2 * tr1)


def test_uniform_strain_rate_gives_c_times_strain():
    case = quiet(cases.al_block(nx=5, ny=4, nz=3, nsteps=1, n_sensors=1))
    rng = np.random.default_rng(7)
    c6 = rng.uniform(1e9, 5e9, size=(6, 6))
    c6 = c6 + c6.T
    case.cij = cases.cij_cell_fields(np.broadcast_to(c6, (case.n_cells, 6, 6)))
    g = rng.uniform(-1, 1, size=(3, 3))  # V_k = sum_i g[i,k] x_i
    X = np.stack([case.x, case.y, case.z])
    V0 = g.T @ X
    _, S, _, _ = orc.run_oracle(case, V0=V0)
    # the stress update runs after the velocity update, which changes nothing here only
    # in the interior; compare interior cells of the single step against C:grad V
    eps = np.array([g[0, 0], g[1, 1], g[2, 2], g[1, 2] + g[2, 1], g[0, 2] + g[2, 0], g[0, 1] + g[1, 0]])
    want = case.dt * (c6 @ eps)
    _, _, _, _ = dd(case.x, case.y, case.z, case.cells)
    np.testing.assert_allclose(S[:, 0], want, rtol=1e-12)
    np.testing.assert_allclose(S, np.repeat(want[:, None], case.n_cells, axis=1), rtol=1e-12)


def test_mirror_symmetry_of_a_centred_source():
    case = cases.al_block(nx=8, ny=8, nz=4, nsteps=30, n_sensors=1)
    V, S, _, _ = orc.run_oracle(case)
    xs = case.x.max() - case.x
    key = {(round(a / 1e-5), round(b / 1e-5), round(c / 1e-5)): n for n, (a, b, c) in
           enumerate(zip(case.x, case.y, case.z))}
    mirror = np.array([key[(round(a / 1e-5), round(b / 1e-5), round(c / 1e-5))] for a, b, c in
                       zip(xs, case.y, case.z)])
    scale = np.abs(V[2]).max()
    assert scale > 0
    np.testing.assert_allclose(V[2][mirror], V[2], atol=1e-12 * scale)
    np.testing.assert_allclose(V[0][mirror], -V[0], atol=1e-12 * scale)


# ---------------------------------------------------------------- planted mistakes

@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5])
def test_planted_mistakes_are_caught(mutation):
    """Each planted slip (1 slot rule, 2 the /8 rule, 3 Syz read for Sxy in the x-velocity update,
    4 dropped 0.125 source weight, 5 flipped y difference) must break a check."""
    L = orc.lib()
    L.orc_set_mutation(mutation)
    try:
        caught = False
        for name in golden_names():
            case, ref = load_golden(name)
            V, S, _, _ = orc.run_oracle(case)
            if not (np.array_equal(V, ref["V"]) and np.array_equal(S, ref["S"])):
                caught = True
                break
        assert caught, "mutation %d slipped through every fixture" % mutation
    finally:
        L.orc_set_mutation(0)
