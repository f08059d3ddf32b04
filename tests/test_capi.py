"""CPU checks of the C ABI library: it loads, exports every entry point the header
declares, and its host-only helpers (source window, signal) agree with the oracle.
No GPU work is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from oracle import oracle_rsg as orc
from pannde_amd import cases, rsg

HEADER = os.path.join(ROOT, "include", "pannde_rsg.h")


def declared_entry_points():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^int (rsg_\w+)\(", text, flags=re.M)))


def test_header_matches_binding_list():
    assert declared_entry_points() == sorted(rsg.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = rsg.lib()
    for name in declared_entry_points():
        assert hasattr(L, name), name


def test_disc_support_matches_oracle_window():
    case = cases.al_block(nx=12, ny=10, nz=5, n_sensors=1)
    d = case.force[0]
    got_cells = rsg.disc_support((case.x, case.y, case.z, case.cells), 0, d.center, d.radius, d.ds)
    want = []
    for c in range(case.n_cells):
        v = case.cells[c]
        cc = [0.0, 0.0, 0.0]
        for k in range(8):  # HostRSGTDUT::cellcenter order
            cc[0] += 0.125 * case.x[v[k]]
            cc[1] += 0.125 * case.y[v[k]]
            cc[2] += 0.125 * case.z[v[k]]
        if orc.disc_check(d.center, d.radius, d.ds, cc):
            want.append(c)
    assert list(got_cells) == want and len(want) > 0
    got_nodes = rsg.disc_support((case.x, case.y, case.z, case.cells), 1, d.center, d.radius, d.ds)
    want_nodes = [n for n in range(case.n_nodes)
                  if orc.disc_check(d.center, d.radius, d.ds, (case.x[n], case.y[n], case.z[n]))]
    assert list(got_nodes) == want_nodes and len(want_nodes) > 0


def test_signal_eval_matches_oracle_bitwise():
    t, v = cases.tone_burst(5e6, 3, 4.857e-9)
    rng = np.random.default_rng(0)
    q = np.concatenate([t[:50], rng.uniform(0, t[-1] * 1.2, 200), [t[-1], t[-1] * 2]])
    assert np.array_equal(rsg.signal_eval(t, v, q), orc.signal_eval(t, v, q))
    tt = np.array([0., 1., 1., 2.])
    vv = np.array([0., 5., 9., 7.])
    assert list(rsg.signal_eval(tt, vv, [1.0, 1.5])) == [5.0, 6.0]


def test_signal_below_first_key_is_an_error():
    with pytest.raises(rsg.RSGError) as e:
        rsg.signal_eval([1.0, 2.0], [0.0, 1.0], [0.5])
    assert e.value.code == -1


def test_create_rejects_bad_node_id_before_touching_the_device():
    x, y, z, cells, _ = cases.structured_mesh(2, 2, 2, 1.0, 1.0, 1.0)
    cells = cells.copy()
    cells[3, 5] = x.shape[0] + 7
    c6 = cases.isotropic_c(1e9, 1e9)
    with pytest.raises(rsg.RSGError) as e:
        rsg.DeviceSolver(x, y, z, cells, cases.Field(cases.FIELD_CONSTANT, np.ones(1)),
                         cases.cij_cell_fields(np.broadcast_to(c6, (cells.shape[0], 6, 6))), 1e-9)
    assert e.value.code == -1 and "invalid node id" in str(e.value)


def test_create_reports_reference_dx_error():
    x, y, z, cells, _ = cases.structured_mesh(1, 1, 1, 1.0, 1.0, 1.0)
    c6 = cases.isotropic_c(1e9, 1e9)
    with pytest.raises(rsg.RSGError) as e:
        rsg.DeviceSolver(np.zeros_like(x), y, z, cells, cases.Field(cases.FIELD_CONSTANT, np.ones(1)),
                         cases.cij_cell_fields(np.broadcast_to(c6, (1, 6, 6))), 1e-9)
    assert e.value.code == -2 and "DX" in str(e.value)


@pytest.mark.parametrize("axis", ["DX", "DY", "DZ"])
def test_create_reports_the_first_bad_node_on_a_large_mesh(axis):
    """Node spacings are computed in parallel chunks; the error still names the first node."""
    x, y, z, cells, _ = cases.structured_mesh(60, 60, 24, 1.0, 1.0, 1.0)  # 90,585 nodes
    assert x.shape[0] > 1 << 16
    xyz = [x, y, z]
    k = ["DX", "DY", "DZ"].index(axis)
    xyz[k] = np.zeros_like(xyz[k])
    c6 = cases.isotropic_c(1e9, 1e9)
    with pytest.raises(rsg.RSGError) as e:
        rsg.DeviceSolver(*xyz, cells, cases.Field(cases.FIELD_CONSTANT, np.ones(1)),
                         [cases.Field(cases.FIELD_CONSTANT, np.array([c6[i, j]])) for i, j in cases.CIJ_PAIRS],
                         1e-9)
    assert e.value.code == -2 and "%s (node 0)" % axis in str(e.value)
