"""Python access to the CPU checkers.  TEST INFRASTRUCTURE ONLY.

* ``run_oracle`` calls the C restatement (oracle/cpu_rsg.c -> liboracle_rsg.so);
* ``run_reference`` runs the reference's own solver compiled by ``make -C oracle ref``
  (oracle/_ref/ref_rsg) on a case file, optionally under mpirun.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module;
the product (pannde_amd) never does.
"""
from __future__ import annotations

import ctypes
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_rsg.so")
REF_BIN = os.path.join(HERE, "_ref", "ref_rsg")
MPIRUN = "/opt/conda/bin/mpirun"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class OrcField(ctypes.Structure):
    _fields_ = [("data", _dp), ("type", ctypes.c_int32)]


class OrcDisc(ctypes.Structure):
    _fields_ = [("component", ctypes.c_int32), ("center", ctypes.c_double * 3), ("radius", ctypes.c_double),
                ("ds", ctypes.c_double), ("n_signal", ctypes.c_int32), ("sig_t", _dp), ("sig_v", _dp)]


class OrcProblem(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int32), ("n_cells", ctypes.c_int32), ("x", _dp), ("y", _dp), ("z", _dp),
                ("cells", _ip), ("rho", OrcField), ("c", OrcField * 21), ("dt", ctypes.c_double),
                ("n_force", ctypes.c_int32), ("force", ctypes.POINTER(OrcDisc)),
                ("n_moment", ctypes.c_int32), ("moment", ctypes.POINTER(OrcDisc)),
                ("moment_form", ctypes.c_int32)]


_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_run.argtypes = [ctypes.POINTER(OrcProblem), ctypes.c_int32, _dp, _dp, _dp, ctypes.c_int32, _ip, _dp]
        L.orc_run.restype = ctypes.c_int
        L.orc_dd_probe.argtypes = [ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_dd_probe.restype = ctypes.c_int
        L.orc_build_links.argtypes = [ctypes.c_int32, ctypes.c_int32, _ip, _ip]
        L.orc_field_at_node.argtypes = [OrcField, _ip, ctypes.c_int32]
        L.orc_field_at_node.restype = ctypes.c_double
        L.orc_field_at_cell.argtypes = [OrcField, _ip, ctypes.c_int32]
        L.orc_field_at_cell.restype = ctypes.c_double
        L.orc_signal_sorted.argtypes = [ctypes.c_int32, _dp, _dp, _dp, _dp]
        L.orc_signal_sorted.restype = ctypes.c_int32
        L.orc_signal_eval.argtypes = [ctypes.c_int32, _dp, _dp, ctypes.c_double]
        L.orc_signal_eval.restype = ctypes.c_double
        L.orc_disc_check.argtypes = [ctypes.c_double * 3, ctypes.c_double, ctypes.c_double, ctypes.c_double * 3]
        L.orc_disc_check.restype = ctypes.c_int
        L.orc_schedule.argtypes = [ctypes.c_double, ctypes.c_int32, _dp, ctypes.POINTER(ctypes.c_int64)]
        L.orc_schedule.restype = ctypes.c_int64
        _LIB = L
    return _LIB


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class _Problem:
    def __init__(self, case):
        self.keep = []
        p = OrcProblem()
        self.x, self.y, self.z = _f64(case.x), _f64(case.y), _f64(case.z)
        self.cells = _i32(case.cells).reshape(-1)
        p.n_nodes, p.n_cells = self.x.shape[0], self.cells.shape[0] // 8
        p.x, p.y, p.z = (a.ctypes.data_as(_dp) for a in (self.x, self.y, self.z))
        p.cells = self.cells.ctypes.data_as(_ip)
        p.rho = self.field(case.rho)
        for k in range(21):
            p.c[k] = self.field(case.cij[k])
        p.dt = float(case.dt)
        fl, ml = self.discs(case.force), self.discs(case.moment)
        p.n_force, p.force = len(case.force), fl
        p.n_moment, p.moment = len(case.moment), ml
        p.moment_form = int(case.moment_form)
        self.p = p

    def field(self, f):
        a = _f64(f.data)
        self.keep.append(a)
        return OrcField(a.ctypes.data_as(_dp), int(f.type))

    def discs(self, discs):
        arr = (OrcDisc * max(1, len(discs)))()
        for i, d in enumerate(discs):
            t, v = _f64(d.signal_t), _f64(d.signal_v)
            self.keep += [t, v]
            arr[i].component = int(d.component)
            for k in range(3):
                arr[i].center[k] = float(d.center[k])
            arr[i].radius, arr[i].ds = float(d.radius), float(d.ds)
            arr[i].n_signal, arr[i].sig_t, arr[i].sig_v = t.shape[0], t.ctypes.data_as(_dp), v.ctypes.data_as(_dp)
        self.keep.append(arr)
        return ctypes.cast(arr, ctypes.POINTER(OrcDisc))


def run_oracle(case, nsteps=None, V0=None, S0=None, sensors=None, t0=0.0):
    """Runs the C restatement: returns (V [3,nn], S [6,nc], time, traces [steps,ns,3])."""
    nsteps = case.nsteps if nsteps is None else nsteps
    pr = _Problem(case)
    nn, nc = pr.p.n_nodes, pr.p.n_cells
    V = np.zeros((3, nn)) if V0 is None else _f64(V0).copy()
    S = np.zeros((6, nc)) if S0 is None else _f64(S0).copy()
    sens = _i32(case.sensors if sensors is None else sensors)
    tr = np.zeros((nsteps, sens.shape[0], 3))
    t = ctypes.c_double(t0)
    rc = lib().orc_run(ctypes.byref(pr.p), nsteps, V.ctypes.data_as(_dp), S.ctypes.data_as(_dp), ctypes.byref(t),
                       sens.shape[0], sens.ctypes.data_as(_ip), tr.ctypes.data_as(_dp))
    if rc != 0:
        raise ValueError("oracle: RightHex nodePrep error %s" % ["", "DX", "DY", "DZ"][rc])
    return V, S, t.value, tr


def dd_probe(x, y, z, cells, node_field, cell_field):
    x, y, z, node_field, cell_field = (_f64(a) for a in (x, y, z, node_field, cell_field))
    cells = _i32(cells).reshape(-1)
    nn, nc = x.shape[0], cells.shape[0] // 8
    out = [np.zeros((nc, 3)), np.zeros((nn, 3)), np.zeros((nn, 3)), np.zeros((nc, 3))]
    rc = lib().orc_dd_probe(nn, nc, *(a.ctypes.data_as(_dp) for a in (x, y, z)), cells.ctypes.data_as(_ip),
                            node_field.ctypes.data_as(_dp), cell_field.ctypes.data_as(_dp),
                            *(o.ctypes.data_as(_dp) for o in out))
    if rc != 0:
        raise ValueError("oracle: RightHex nodePrep error %s" % ["", "DX", "DY", "DZ"][rc])
    return tuple(out)  # dd_cc [nc,3], dd_nc [nn,3], node spacing [nn,3], cell spacing [nc,3]


def build_links(n_nodes, cells):
    cells = _i32(cells).reshape(-1)
    out = np.empty(n_nodes * 8, dtype=np.int32)
    lib().orc_build_links(n_nodes, cells.shape[0] // 8, cells.ctypes.data_as(_ip), out.ctypes.data_as(_ip))
    return out.reshape(n_nodes, 8)


def field_at_node(f, node_cells, n):
    a = _f64(f.data)
    nc = _i32(node_cells).reshape(-1)
    return lib().orc_field_at_node(OrcField(a.ctypes.data_as(_dp), f.type), nc.ctypes.data_as(_ip), n)


def field_at_cell(f, cells, c):
    a = _f64(f.data)
    cc = _i32(cells).reshape(-1)
    return lib().orc_field_at_cell(OrcField(a.ctypes.data_as(_dp), f.type), cc.ctypes.data_as(_ip), c)


def signal_eval(t, v, at):
    t, v = _f64(t), _f64(v)
    ts, vs = np.empty_like(t), np.empty_like(v)
    m = lib().orc_signal_sorted(t.shape[0], t.ctypes.data_as(_dp), v.ctypes.data_as(_dp), ts.ctypes.data_as(_dp),
                                vs.ctypes.data_as(_dp))
    return np.array([lib().orc_signal_eval(m, ts.ctypes.data_as(_dp), vs.ctypes.data_as(_dp), float(q))
                     for q in np.atleast_1d(at)])


def disc_check(center, radius, ds, p):
    return bool(lib().orc_disc_check((ctypes.c_double * 3)(*center), radius, ds, (ctypes.c_double * 3)(*p)))


def schedule(dt, write_times):
    w = _f64(write_times)
    out = np.zeros(w.shape[0], dtype=np.int64)
    lib().orc_schedule(dt, w.shape[0], w.ctypes.data_as(_dp), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return out


# ---------------------------------------------------------------- the compiled reference

def have_reference():
    return os.path.exists(REF_BIN)


# The driver's job and result formats (RSGCASE1 / RSGOUT01), which oracle/ref_harness.cpp reads
# and writes too.
from pannde_amd.case_file import read_out_file, write_case_file  # noqa: E402


def run_reference(case, nsteps=None, ranks=1, warmup=0, V0=None, S0=None, timeout=3600, workdir=None, repeat=1):
    """Runs oracle/_ref/ref_rsg (the reference HostRSGTDUT / HostUTModel) on ``case``.
    Returns (V, S, time, traces, timing) with timing = dict(seconds, setup_seconds, ranks,
    windows).  repeat > 1 times the nsteps window that many times after the warm-up
    (seconds = the median window; the outputs are then not those of warmup + nsteps)."""
    if not have_reference():
        raise FileNotFoundError(REF_BIN)
    d = tempfile.mkdtemp(prefix="rsgref_", dir=workdir)
    try:
        cp, op = os.path.join(d, "case.bin"), os.path.join(d, "out.bin")
        write_case_file(cp, case, nsteps, V0, S0)
        cmd = [REF_BIN, cp, op, "--warmup", str(warmup), "--repeat", str(int(repeat))]
        if ranks > 1:
            cmd = [MPIRUN, "-n", str(ranks)] + cmd
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=timeout, env=env)
        if r.returncode != 0:
            raise RuntimeError("ref_rsg failed (%d):\n%s" % (r.returncode, r.stdout[-4000:]))
        timing = {}
        for line in r.stdout.splitlines():
            if line.startswith("REF_TIMING"):
                for kv in line.split()[1:]:
                    k, v = kv.split("=")
                    timing[k] = [float(x) for x in v.split(",")] if k == "windows" else float(v)
        V, S, t, tr = read_out_file(op)
        return V, S, t, tr, timing
    finally:
        shutil.rmtree(d, ignore_errors=True)
