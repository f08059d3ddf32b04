"""Parity of the FP64 DeviceSolver (HIP, through the C ABI) with the reference.

* every golden fixture: fields, time and sensor traces bit-identical to what the
  reference solver computed (tests/golden, made by oracle/_ref/ref_rsg);
* seeded jobs of every flavour: bit-identical to the oracle;
* the Al-block workload (40x40x20 cells, 500 steps) in full: bit-identical to the oracle;
* the CFRP-plate workload at full size through size-independent properties: the input
  node numbering does not change one bit of the result, doubling the source doubles every
  field exactly, and a zero source leaves the zero state untouched.
The tolerance of the north-star (1e-10 relative max-norm) is asserted too, but the
device reproduces the reference's operation order, so equality is exact.
"""
import os

import numpy as np
import pytest

from conftest import golden_names, load_golden
from oracle import oracle_rsg as orc
from pannde_amd import cases, rsg

pytestmark = pytest.mark.gpu


def relmax(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


def run_device(case, nsteps=None, precision=rsg.RSG_FP64, V0=None, S0=None):
    s = rsg.DeviceSolver.from_case(case, precision)
    try:
        if V0 is not None or S0 is not None:
            s.set_state(V0, S0)
        t = s.solve(case.nsteps if nsteps is None else nsteps)
        V, S = s.get_state()
        return V, S, t, s.traces()
    finally:
        s.close()


def assert_same(dev, ref):
    V, S, t, tr = dev
    V0, S0, t0, tr0 = ref
    assert t == t0
    for k in range(3):
        assert relmax(V[k], V0[k]) <= 1e-10
    for k in range(6):
        assert relmax(S[k], S0[k]) <= 1e-10
    assert np.array_equal(V, V0), "V differs: max %.3e" % np.abs(V - V0).max()
    assert np.array_equal(S, S0), "S differs: max %.3e" % np.abs(S - S0).max()
    assert np.array_equal(tr, tr0)


@pytest.mark.parametrize("name", golden_names())
def test_device_matches_reference_fixture(name):
    case, ref = load_golden(name)
    assert_same(run_device(case), (ref["V"], ref["S"], ref["time"], ref["traces"]))


TUNINGS = [
    dict(block_velocity=64, block_stress=64),
    dict(block_velocity=128, block_stress=512),
    dict(block_velocity=512, block_stress=128),
    dict(per_thread_velocity=2, per_thread_stress=2),
    dict(per_thread_velocity=4, per_thread_stress=4, block_velocity=128, block_stress=128),
    dict(lds_velocity=-1, lds_stress=-1),
    dict(lds_velocity=-1, lds_stress=1, per_thread_stress=2),
    dict(lds_velocity=1, lds_stress=-1, per_thread_velocity=2, per_thread_stress=4),
    dict(tile=(1 << 20, 1, 1)),
    dict(tile=(16, 4, 4), c_layout=-1, xcd_remap=-1),
    dict(coeff_classes=-1),
    dict(coeff_classes=-1, lds_velocity=-1, per_thread_stress=2, c_layout=-1),
]


@pytest.mark.parametrize("tuning", TUNINGS, ids=[str(sorted(t.items())) for t in TUNINGS])
@pytest.mark.parametrize("name", ["cfrp_holes_rcm", "water_column_moment"])
def test_tuning_does_not_change_results(name, tuning):
    case, ref = load_golden(name)
    s = rsg.DeviceSolver.from_case(case, tuning=tuning)
    try:
        t = s.solve(case.nsteps)
        V, S = s.get_state()
        assert_same((V, S, t, s.traces()), (ref["V"], ref["S"], ref["time"], ref["traces"]))
    finally:
        s.close()


def test_coefficient_class_tables_in_use():
    """Plies and homogeneous blocks share material rows; per-cell random C falls back to
    per-cell arrays; the fallback can also be asked for."""
    plate, _ = load_golden("cfrp_holes_rcm")  # 16 plies in 4 orientations, [0/45/90/-45]2s
    s = rsg.DeviceSolver.from_case(plate)
    t = s.tuning()
    s.close()
    # node materials: dt/rho at interior, face, edge, corner and hole-rim nodes (the /8 rule)
    assert t["coeff_classes"] == 1 and t["cell_classes"] == 4 and 0 < t["node_classes"] <= 8
    s = rsg.DeviceSolver.from_case(plate, tuning=dict(coeff_classes=-1))
    t = s.tuning()
    s.close()
    assert t["coeff_classes"] == -1 and t["cell_classes"] == -1 and t["node_classes"] == -1
    rng = np.random.default_rng(5)
    case = cases.al_block(seed=5, nx=30, ny=30, nz=20, nsteps=20, n_sensors=4)
    for f in case.cij:
        f.data = f.data * (1 + 0.05 * rng.standard_normal(f.data.shape))
    s = rsg.DeviceSolver.from_case(case)
    t = s.tuning()
    s.solve(case.nsteps)
    dev = s.get_state()
    s.close()
    assert t["cell_classes"] == -1 and t["node_classes"] > 0  # 18000 distinct rows > 18000/64
    ref = orc.run_oracle(case)
    assert np.array_equal(dev[0], ref[0]) and np.array_equal(dev[1], ref[1])


FORMS = [  # (tuning, expected velocity links, expected stress links)
    (dict(), 2, 2),                                          # keys packed with 10-bit links
    (dict(coeff_classes=-1), 1, 1),                          # per-entity arrays, 16-bit links
    (dict(lds_velocity=-1, lds_stress=-1), 0, 0),            # keys, direct gathers
    (dict(block_stress=512, per_thread_stress=2), 2, 1),     # staged sets above 1022: keys + 16-bit
]


@pytest.mark.parametrize("form", FORMS, ids=[str(f[0]) for f in FORMS])
def test_link_and_coefficient_forms_agree(form):
    tuning, lv, ls = form
    case = cases.al_block(seed=9, nx=40, ny=40, nz=20, nsteps=60, n_sensors=6)
    s = rsg.DeviceSolver.from_case(case, tuning=tuning)
    t = s.tuning()
    s.solve(case.nsteps)
    dev = (*s.get_state(), s.time, s.traces())
    s.close()
    assert (t["links_velocity"], t["links_stress"]) == (lv, ls)
    assert_same(dev, orc.run_oracle(case))


@pytest.mark.parametrize("tuning", [dict(block_stress=1024), dict(block_velocity=96), dict(per_thread_stress=3)])
def test_unsupported_tuning_is_rejected(tuning):
    case, _ = load_golden("al_block_cellrho")
    with pytest.raises(rsg.RSGError, match="workgroup size|per thread"):
        rsg.DeviceSolver.from_case(case, tuning=tuning)


@pytest.mark.parametrize("seed", [201, 202, 203, 204])
def test_device_matches_oracle_on_seeded_jobs(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(v) for v in rng.integers(5, 14, size=3))
    if seed % 2:
        case = cases.water_column_block(seed=seed, nx=nx, ny=ny, nz=nz, nsteps=30)
    else:
        case = cases.al_block(seed=seed, nx=nx, ny=ny, nz=nz, nsteps=30, n_sensors=5,
                              rho_type=int(rng.integers(0, 3)))
    # C perturbed per cell: on the meshes of at most 1024 cells (seeds 201, 202) every cell is
    # its own material class, so each wavefront of the stress kernel spans up to 64 materials
    for f in case.cij:
        f.data = f.data * (1 + 0.05 * rng.standard_normal(f.data.shape))
    V0 = 1e-9 * rng.standard_normal((3, case.n_nodes))
    S0 = 1e3 * rng.standard_normal((6, case.n_cells))
    assert_same(run_device(case, V0=V0, S0=S0), orc.run_oracle(case, V0=V0, S0=S0))


def test_al_block_workload_full():
    """40x40x20 aluminium block, 5 MHz burst, 500 steps, 16 top-face sensors."""
    case = cases.al_block(seed=1)
    assert case.n_cells == 32000 and case.n_nodes == 35301
    dev = run_device(case)
    assert np.abs(dev[3]).max() > 0
    assert_same(dev, orc.run_oracle(case))


def test_split_solve_calls_equal_one_call():
    case = cases.al_block(seed=3, nx=12, ny=10, nz=6, nsteps=300, n_sensors=4)
    s = rsg.DeviceSolver.from_case(case)
    for n in (1, 7, 255, 37):
        s.solve(n)
    V, S = s.get_state()
    tr = s.traces()
    s.close()
    V1, S1, _, tr1 = run_device(case)
    assert np.array_equal(V, V1) and np.array_equal(S, S1) and np.array_equal(tr, tr1)


@pytest.mark.parametrize("stride", [1, 4])
def test_kernel_timing_counts_sampled_steps_and_leaves_results(stride):
    """Timed runs (every step or every 4th) count the timed steps per call of up to 256
    steps and compute the same results as untimed ones."""
    case = cases.al_block(seed=5, nx=10, ny=9, nz=6, nsteps=300, n_sensors=2)
    s = rsg.DeviceSolver.from_case(case)
    s.set_timing(True, stride=stride)
    s.solve(10)
    s.solve(290)  # one block of 256 steps and one of 34
    vms, sms, n = s.kernel_times()
    assert n == sum(-(-k // stride) for k in (10, 256, 34)) and vms > 0 and sms > 0
    with pytest.raises(rsg.RSGError, match="stride"):
        s.set_timing(True, stride=-1)
    s.set_timing(False)
    V, S = s.get_state()
    s.close()
    V1, S1, _, _ = run_device(case)
    assert np.array_equal(V, V1) and np.array_equal(S, S1)


def test_caller_supplied_amplitudes_equal_library_signal():
    case = cases.al_block(seed=4, nx=10, ny=9, nz=5, nsteps=60, n_sensors=3)
    d = case.force[0]
    s = rsg.DeviceSolver.from_case(case)
    sup = rsg.disc_support((case.x, case.y, case.z, case.cells), 0, d.center, d.radius, d.ds)
    s.set_channels([(rsg.RSG_FORCE, 2, sup, None, None)])
    t = 0.0
    amps = []
    for _ in range(case.nsteps):
        amps.append(0.125 * (orc.signal_eval(d.signal_t, d.signal_v, [t])[0] * 1.0))
        t += case.dt
    s.solve(case.nsteps, np.array(amps).reshape(-1, 1))
    V, S = s.get_state()
    s.close()
    V0, S0, _, _ = orc.run_oracle(case)
    assert np.array_equal(V, V0) and np.array_equal(S, S0)


def test_state_round_trip_and_errors():
    case = cases.water_column_block(nx=6, ny=5, nz=4, nsteps=5)
    s = rsg.DeviceSolver.from_case(case)
    rng = np.random.default_rng(9)
    V0, S0 = rng.standard_normal((3, case.n_nodes)), rng.standard_normal((6, case.n_cells))
    s.set_state(V0, S0)
    V, S = s.get_state()
    assert np.array_equal(V, V0) and np.array_equal(S, S0)
    with pytest.raises(rsg.RSGError):
        s.set_sensors([case.n_nodes + 1])
    s.close()
    plain = cases.al_block(nx=6, ny=5, nz=4, nsteps=5)
    p = rsg.DeviceSolver.from_case(plain)
    with pytest.raises(rsg.RSGError, match="moment_form"):
        p.set_sources([], case.moment)
    p.close()


@pytest.fixture(scope="module")
def plate_pair():
    """The CFRP plate workload with the seeded RCM numbering and with the natural
    case-maker numbering (same mesh, same holes)."""
    a = cases.cfrp_plate(seed=2, nsteps=40)
    b = cases.cfrp_plate(seed=2, nsteps=40, renumber=False)
    return a, b


def test_plate_numbering_invariance_full_size(plate_pair):
    a, b = plate_pair
    assert a.n_cells == b.n_cells and a.n_nodes == b.n_nodes and a.n_cells > 3_000_000
    Va, Sa, ta, tra = run_device(a)
    Vb, Sb, tb, trb = run_device(b)
    key = lambda x, y, z: np.lexsort((np.round(z / 1e-6), np.round(y / 1e-6), np.round(x / 1e-6)))
    ia, ib = key(a.x, a.y, a.z), key(b.x, b.y, b.z)
    assert ta == tb and np.abs(Va).max() > 0
    assert np.array_equal(Va[:, ia], Vb[:, ib])
    assert np.array_equal(Sa, Sb)  # cells keep the case-maker order in both


def test_plate_linearity_and_zero_source_full_size(plate_pair):
    a, _ = plate_pair
    V1, S1, _, tr1 = run_device(a, nsteps=25)
    a.force[0].signal_v = 2.0 * a.force[0].signal_v
    V2, S2, _, tr2 = run_device(a, nsteps=25)
    a.force[0].signal_v = 0.5 * a.force[0].signal_v
    assert np.array_equal(V2, 2 * V1) and np.array_equal(S2, 2 * S1) and np.array_equal(tr2, 2 * tr1)
    a.force[0].signal_v = 0.0 * a.force[0].signal_v
    V0, S0, _, _ = run_device(a, nsteps=5)
    assert not np.any(V0) and not np.any(S0)


def test_velocity_workgroups_fall_back_when_staged_sets_are_large(plate_pair):
    """On meshes of 3 Mi nodes or more the velocity kernel takes 512-node workgroups, unless
    their staged cell sets overflow LDS; then 256-node workgroups.  Bricks of one whole
    layer (x fastest) make 512 consecutive nodes touch about 1500 cells, over the 64 KiB of
    FP64 LDS.  Results stay bit-identical."""
    a, _ = plate_pair
    s = rsg.DeviceSolver.from_case(a)
    try:
        t = s.tuning()
        assert t["block_velocity"] == 512 and t["links_velocity"] == 2  # packed LDS links
    finally:
        s.close()
    ref = run_device(a, nsteps=10)
    s = rsg.DeviceSolver.from_case(a, tuning=dict(tile=(256, 256, 1)))
    try:
        t = s.tuning()
        assert t["block_velocity"] == 256 and t["links_velocity"] in (1, 2)  # still LDS-staged
        tm = s.solve(10)
        V, S = s.get_state()
        dev = (V, S, tm, s.traces())
    finally:
        s.close()
    assert_same(dev, ref)


def test_jittered_planes_and_per_cell_materials():
    """A 10x10x10 block whose node planes are jittered (10 spacing classes per axis, about 11
    node spacings per axis) and whose C is perturbed per cell (1000 materials): keys that
    vary in every field within a wavefront.  Results stay bit-identical to the oracle."""
    rng = np.random.default_rng(7)
    case = cases.al_block(seed=7, nx=10, ny=10, nz=10, nsteps=20, n_sensors=4)
    dx = 1.0e-4
    for coord in (case.x, case.y, case.z):
        plane = np.rint(coord / dx).astype(np.int64)
        coord += dx * 0.02 * rng.random(plane.max() + 1)[plane]
    for f in case.cij:
        f.data = f.data * (1 + 0.05 * rng.standard_normal(f.data.shape))
    s = rsg.DeviceSolver.from_case(case)
    try:
        assert s.tuning()["cell_classes"] == 1000
    finally:
        s.close()
    assert_same(run_device(case), orc.run_oracle(case))


def _with_duplicated_cells(case, ids):
    """The case with copies of cells ``ids`` appended: each copy's vertices claim node slots
    its original holds, and the later cell keeps the slot (HexMesh::buildLinks)."""
    import dataclasses

    ids = np.asarray(ids)

    def grow(f):
        if f.type != cases.FIELD_CELL:
            return f
        return cases.Field(f.type, np.concatenate([f.data, f.data[ids]]))

    return dataclasses.replace(case, cells=np.concatenate([case.cells, case.cells[ids]]).astype(np.int32),
                               rho=grow(case.rho), cij=[grow(f) for f in case.cij])


@pytest.mark.parametrize("maker", ["water_column", "al_block"])
def test_node_slots_claimed_twice_keep_the_later_cell(maker):
    """Cells that duplicate others (every vertex slot claimed twice): the later cell wins
    each slot, as in the reference's loop, and stress sources reach both copies; bitwise
    against the oracle.  Also the path of the source tables for such meshes."""
    if maker == "water_column":
        case = cases.water_column_block(seed=8, nx=14, ny=12, nz=7, nsteps=150)
    else:
        case = cases.al_block(seed=8, nx=14, ny=12, nz=7, nsteps=150, n_sensors=6)
    top = np.nonzero(np.isclose(case.z[case.cells].max(axis=1), case.z.max()))[0]  # cells on the top face
    ids = list(top[len(top) // 2 - 3:len(top) // 2 + 3]) + [5, 17]  # under the source, and elsewhere
    dup = _with_duplicated_cells(case, ids)
    assert_same(run_device(dup), orc.run_oracle(dup))


@pytest.mark.parametrize("precision", [rsg.RSG_FP64, rsg.RSG_FP32])
def test_sources_and_sensors_replaced_on_a_live_handle(precision):
    """Table builds patch only the keys whose special flag changes: a handle whose sources and
    sensors are replaced several times (moment discs, then force discs and other sensor nodes,
    each change building the tables again) and then stepped equals, bit for bit, a fresh
    handle created with the final ones."""
    case = cases.water_column_block(seed=9, nx=16, ny=14, nz=7, nsteps=120)
    other = cases.al_block(seed=9, nx=16, ny=14, nz=7, nsteps=120, n_sensors=9)
    assert not np.array_equal(np.sort(case.sensors), np.sort(other.sensors))
    live = rsg.DeviceSolver(case.x, case.y, case.z, case.cells, case.rho, case.cij, case.dt, case.force,
                            case.moment, 1, case.sensors, precision)
    try:
        live.set_sources(other.force, case.moment[:1])
        live.set_sensors(other.sensors)
        live.set_sources(other.force, [])
        t = live.solve(case.nsteps)
        got = live.get_state() + (live.traces(),)
    finally:
        live.close()
    fresh = rsg.DeviceSolver(case.x, case.y, case.z, case.cells, case.rho, case.cij, case.dt, other.force, [], 1,
                             other.sensors, precision)
    try:
        t0 = fresh.solve(case.nsteps)
        ref = fresh.get_state() + (fresh.traces(),)
    finally:
        fresh.close()
    assert t == t0
    assert np.abs(ref[0]).max() > 0 and np.abs(ref[2]).max() > 0
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_table_amplitudes_uploaded_ahead_mix_with_caller_amplitudes():
    """Signal-table amplitudes are uploaded a block of steps ahead and read by later calls;
    calls with caller amplitudes in between (equal to the table's) and calls of one step at a
    time (a driver's solve() per step) give the oracle's result bit for bit."""
    case = cases.al_block(seed=4, nx=10, ny=9, nz=5, nsteps=90, n_sensors=3)
    d = case.force[0]
    t = 0.0
    amps = []
    for _ in range(case.nsteps):
        amps.append(0.125 * (orc.signal_eval(d.signal_t, d.signal_v, [t])[0] * 1.0))
        t += case.dt
    amps = np.array(amps).reshape(-1, 1)
    s = rsg.DeviceSolver.from_case(case)
    try:
        s.solve(23)
        s.solve(4, amps[23:27])
        for _ in range(10):
            s.solve(1)
        s.solve(3, amps[37:40])
        s.solve(case.nsteps - 40)
        V, S = s.get_state()
        tr = s.traces()
    finally:
        s.close()
    V0, S0, _, tr0 = orc.run_oracle(case)
    assert np.array_equal(V, V0) and np.array_equal(S, S0) and np.array_equal(tr, tr0)
