"""The binary job and result files of the C++ driver (pannde_amd/rsg_driver).

* RSGCASE1 (input): node and cell counts, coordinates, hexahedra, density and C11..C66 each with
  its NODE/CELL/CONSTANT type, dt, the model form, the force and moment discs with their signal
  tables, sensor nodes, the step count and an optional initial state.
* RSGOUT01 (output): counts, the final time, V and S read back through ``getStates()`` and the
  sensor velocities of every step.

The same RSGCASE1 layout is what ``rsg_driver -f job.vtu --dump-case FILE`` writes for a VTK job.
"""
from __future__ import annotations

import struct

import numpy as np


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def write_case_file(path, case, nsteps=None, V0=None, S0=None):
    """Writes ``case`` (a ``cases.Case``) as RSGCASE1; ``nsteps`` defaults to ``case.nsteps``."""
    nsteps = case.nsteps if nsteps is None else nsteps
    with open(path, "wb") as f:
        f.write(b"RSGCASE1")
        f.write(struct.pack("<ii", case.n_nodes, case.n_cells))
        for a in (case.x, case.y, case.z):
            f.write(_f64(a).tobytes())
        f.write(_i32(case.cells).tobytes())
        for fld in [case.rho] + list(case.cij):
            f.write(struct.pack("<i", int(fld.type)))
            f.write(_f64(fld.data).tobytes())
        f.write(struct.pack("<di", float(case.dt), int(case.moment_form)))
        for discs in (case.force, case.moment):
            f.write(struct.pack("<i", len(discs)))
            for d in discs:
                f.write(struct.pack("<i5d", int(d.component), *[float(c) for c in d.center], float(d.radius),
                                    float(d.ds)))
                t, v = _f64(d.signal_t), _f64(d.signal_v)
                f.write(struct.pack("<i", t.shape[0]))
                f.write(t.tobytes())
                f.write(v.tobytes())
        s = _i32(case.sensors)
        f.write(struct.pack("<i", s.shape[0]))
        f.write(s.tobytes())
        f.write(struct.pack("<i", int(nsteps)))
        if V0 is None and S0 is None:
            f.write(struct.pack("<i", 0))
        else:
            f.write(struct.pack("<i", 1))
            f.write(_f64(V0).tobytes())
            f.write(_f64(S0).tobytes())


def read_out_file(path):
    """(V [3, nodes], S [6, cells], time, traces [steps, sensors, 3]) of an RSGOUT01 file."""
    with open(path, "rb") as f:
        b = f.read()
    if b[:8] != b"RSGOUT01":
        raise ValueError("%s: not an RSGOUT01 file" % path)
    nn, nc, ns, nsens = struct.unpack_from("<4i", b, 8)
    (t,) = struct.unpack_from("<d", b, 24)
    o = 32
    V = np.frombuffer(b, np.float64, 3 * nn, o).reshape(3, nn).copy()
    o += 24 * nn
    S = np.frombuffer(b, np.float64, 6 * nc, o).reshape(6, nc).copy()
    o += 48 * nc
    tr = np.frombuffer(b, np.float64, ns * nsens * 3, o).reshape(ns, nsens, 3).copy()
    return V, S, t, tr
