"""Batched instances (rsg_batch_*): every instance's result is the same, bit for bit, alone
and in a mix of instances of uneven sizes, flavours and step counts; instances join and
leave between K-step blocks; graph replay and direct launches agree."""
import numpy as np
import pytest

from conftest import load_golden
from oracle import oracle_rsg as orc
from pannde_amd import cases, rsg

pytestmark = pytest.mark.gpu


def mix():
    """Six instances: Al blocks of three shapes (CELL / NODE / CONSTANT density), a
    moment-source block, a holed CFRP plate and a three-transducer block; uneven steps."""
    out = []
    out.append((cases.al_block(seed=31, nx=12, ny=10, nz=6, nsteps=37, n_sensors=4), 37))
    out.append((cases.al_block(seed=32, nx=20, ny=9, nz=7, nsteps=100, n_sensors=3, rho_type=cases.FIELD_NODE), 100))
    out.append((cases.water_column_block(seed=33, nx=9, ny=8, nz=5, nsteps=64), 64))
    c, _ = load_golden("cfrp_holes_rcm")
    out.append((c, 120))
    c, _ = load_golden("three_transducers")
    out.append((c, 45))
    out.append((cases.al_block(seed=34, nx=7, ny=16, nz=5, nsteps=80, n_sensors=2, rho_type=cases.FIELD_CONSTANT), 80))
    return out


def alone(case, n, precision):
    s = rsg.DeviceSolver.from_case(case, precision)
    try:
        s.solve(n)
        V, S = s.get_state()
        return V, S, s.time, s.traces()
    finally:
        s.close()


@pytest.mark.parametrize("precision", [rsg.RSG_FP64, rsg.RSG_FP32])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_alone(precision, use_graph):
    jobs = mix()
    solo = [alone(c, n, precision) for c, n in jobs]
    b = rsg.Batch(precision, steps_per_block=16, use_graph=use_graph)
    # every other member keeps per-entity coefficient arrays: the two forms share launches
    solvers = [rsg.DeviceSolver.from_case(c, precision, tuning=dict(coeff_classes=-1) if i % 2 else None)
               for i, (c, _) in enumerate(jobs)]
    slots = [b.add(s, n) for s, (_, n) in zip(solvers, jobs)]
    left = b.run(1000)
    b.synchronize()
    assert left == 0
    for s, slot, (c, n), ref in zip(solvers, slots, jobs, solo):
        assert b.status(slot) == (n, n)
        V, S = s.get_state()
        assert s.time == ref[2]
        assert np.array_equal(V, ref[0]) and np.array_equal(S, ref[1]), c.name
        assert np.array_equal(s.traces(), ref[3]), c.name
    for slot in slots:
        b.remove(slot)
    b.close()


# (grid, group_bytes, xcd_run, lanes): one group / one instance per group / a budget that
# splits the mix unevenly; table order, one eighth per XCD, runs of 3 and 64 items;
# persistent grids; one to four lanes
SCHEDULES = [(0, 0, -1, 1), (0, 1, 0, 4), (0, 200_000, 3, 2), (-1, 1 << 27, 64, 3), (64, 0, 64, 2),
             (37, 1, 0, 3), (0, 1, 64, 1)]


@pytest.mark.parametrize("precision", [rsg.RSG_FP64, rsg.RSG_FP32])
def test_schedules_agree(precision):
    """Group budget, XCD order, grid form and lanes only reorder independent work."""
    jobs = mix()
    solo = [alone(c, n, precision) for c, n in jobs]
    for grid, group_bytes, xcd_run, lanes in SCHEDULES:
        b = rsg.Batch(precision, steps_per_block=9, grid=grid, group_bytes=group_bytes, xcd_run=xcd_run,
                      lanes=lanes)
        solvers = [rsg.DeviceSolver.from_case(c, precision) for c, _ in jobs]
        slots = [b.add(s, n) for s, (_, n) in zip(solvers, jobs)]
        assert b.run(1000) == 0
        b.synchronize()
        for s, (c, n), ref in zip(solvers, jobs, solo):
            V, S = s.get_state()
            tag = "%s grid=%d group=%d xcd=%d lanes=%d" % (c.name, grid, group_bytes, xcd_run, lanes)
            assert np.array_equal(V, ref[0]) and np.array_equal(S, ref[1]), tag
            assert np.array_equal(s.traces(), ref[3]), tag
        for slot in slots:
            b.remove(slot)
        b.close()
        for s in solvers:
            s.close()


def test_batch_option_validation():
    b = rsg.Batch(rsg.RSG_FP64)
    with pytest.raises(rsg.RSGError, match="group bytes"):
        b.set_option(rsg.RSG_BATCH_GROUP_BYTES, -1)
    with pytest.raises(rsg.RSGError, match="xcd run"):
        b.set_option(rsg.RSG_BATCH_XCD_RUN, -2)
    for bad in (0, 5):
        with pytest.raises(rsg.RSGError, match="lanes"):
            b.set_option(rsg.RSG_BATCH_LANES, bad)
    with pytest.raises(rsg.RSGError, match="unknown batch option"):
        b.set_option(99, 0)
    b.close()


def test_batch_fp64_matches_reference_fixture():
    case, ref = load_golden("water_column_moment")
    other, _ = load_golden("cfrp_holes_rcm")
    b = rsg.Batch(rsg.RSG_FP64, steps_per_block=7)
    s1, s2 = rsg.DeviceSolver.from_case(case), rsg.DeviceSolver.from_case(other)
    b.add(s2, 40)
    b.add(s1, case.nsteps)
    b.run(100)
    V, S = s1.get_state()
    assert np.array_equal(V, ref["V"]) and np.array_equal(S, ref["S"]) and s1.time == ref["time"]
    assert np.array_equal(s1.traces(), ref["traces"])


def test_join_and_leave_between_blocks():
    jobs = mix()
    b = rsg.Batch(rsg.RSG_FP64, steps_per_block=10)
    s0 = rsg.DeviceSolver.from_case(jobs[0][0])
    s1 = rsg.DeviceSolver.from_case(jobs[1][0])
    a0 = b.add(s0, 1000)
    a1 = b.add(s1, 55)
    b.run(2)  # 20 steps
    s3 = rsg.DeviceSolver.from_case(jobs[3][0])
    a3 = b.add(s3, 73)  # joins a running batch
    b.run(3)  # s0 at 50 steps
    b.remove(a0)  # leaves before it is finished
    done0, _ = (50, 1000)
    with pytest.raises(rsg.RSGError):
        b.status(a0)
    b.run(100)
    b.synchronize()
    assert b.status(a1) == (55, 55) and b.status(a3) == (73, 73)
    for s, (c, _), n in ((s0, jobs[0], done0), (s1, jobs[1], 55), (s3, jobs[3], 73)):
        ref = alone(c, n, rsg.RSG_FP64)
        V, S = s.get_state()
        assert np.array_equal(V, ref[0]) and np.array_equal(S, ref[1]), c.name
        assert np.array_equal(s.traces(), ref[3]), c.name
    # the detached handle steps alone again, continuing where the batch left it
    s0.solve(5)
    ref = alone(jobs[0][0], 55, rsg.RSG_FP64)
    V, S = s0.get_state()
    assert np.array_equal(V, ref[0]) and np.array_equal(S, ref[1])


def test_batched_handle_refuses_direct_use():
    case, _ = load_golden("al_block_cellrho")
    b = rsg.Batch(rsg.RSG_FP64)
    s = rsg.DeviceSolver.from_case(case)
    slot = b.add(s, 10)
    with pytest.raises(rsg.RSGError, match="batch"):
        s.solve(1)
    with pytest.raises(rsg.RSGError, match="batch"):
        b.add(s, 5)
    b.remove(slot)
    s.solve(1)


def test_batch_rejects_other_precision():
    case, _ = load_golden("al_block_cellrho")
    b = rsg.Batch(rsg.RSG_FP64)
    b.add(rsg.DeviceSolver.from_case(case), 5)
    with pytest.raises(rsg.RSGError, match="precision"):
        b.add(rsg.DeviceSolver.from_case(case, rsg.RSG_FP32), 5)


TUNED = [None, dict(block_stress=128), dict(per_thread_stress=2, per_thread_velocity=4, block_velocity=128),
         dict(lds_velocity=-1), dict(coeff_classes=-1, block_velocity=512), dict(per_thread_stress=4)]


@pytest.mark.parametrize("precision", [rsg.RSG_FP64, rsg.RSG_FP32])
def test_members_with_different_tunings(precision):
    """Each launch covers the members of one kernel configuration; results equal alone."""
    jobs = mix()
    solo = [alone(c, n, precision) for c, n in jobs]
    b = rsg.Batch(precision, steps_per_block=11)
    solvers = [rsg.DeviceSolver.from_case(c, precision, tuning=t) for (c, _), t in zip(jobs, TUNED)]
    for s, (_, n) in zip(solvers, jobs):
        b.add(s, n)
    assert b.run(1000) == 0
    b.synchronize()
    for s, (c, n), ref in zip(solvers, jobs, solo):
        V, S = s.get_state()
        assert np.array_equal(V, ref[0]) and np.array_equal(S, ref[1]), c.name
        assert np.array_equal(s.traces(), ref[3]), c.name
