"""ctypes binding of libpannde_rsg.so (include/pannde_rsg.h) and a Python mirror of the
reference's Model interface for the RSG step.

``DeviceSolver`` plays the role HostRSGTDUT / HostUTModel play in the reference
(modules/HostSolver/include/HostRSGTDUT.hpp:87-169, HostUTModel.hpp:84-193): it is built
from a domain (mesh, density, C11..C66), a dt and a list of sources, ``solve(nsteps)``
advances and returns the simulation time, and ``get_states()`` returns the live fields
under the reference's state names (Vx, Vy, Vz, Sxx, Syy, Szz, Syz, Sxz, Sxy).  All the
arithmetic runs in the HIP library; this module only marshals arrays.  There is no CPU
fallback: if the library is missing or the device fails, the call raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Sequence

import numpy as np

from . import cases as _cases

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PANNDE_RSG_LIB") or os.path.join(HERE, "libpannde_rsg.so")

RSG_OK = 0
RSG_FP64, RSG_FP32 = 0, 1
RSG_FORCE, RSG_MOMENT = 0, 1
RSG_BATCH_GROUP_BYTES, RSG_BATCH_XCD_RUN, RSG_BATCH_LANES = 1, 2, 3
STATE_V = ("Vx", "Vy", "Vz")
STATE_S = ("Sxx", "Syy", "Szz", "Syz", "Sxz", "Sxy")

# Every entry point of include/pannde_rsg.h.
EXPORTS = ("rsg_create", "rsg_set_channels", "rsg_set_sensors", "rsg_step", "rsg_get_state",
           "rsg_set_state", "rsg_get_traces", "rsg_synchronize", "rsg_get_time", "rsg_energy", "rsg_set_timing",
           "rsg_get_kernel_times", "rsg_device_bytes", "rsg_destroy", "rsg_last_error",
           "rsg_disc_support", "rsg_signal_eval", "rsg_create_ex", "rsg_get_tuning", "rsg_batch_create",
           "rsg_batch_add", "rsg_batch_remove", "rsg_batch_run", "rsg_batch_synchronize", "rsg_batch_slot_status",
           "rsg_batch_destroy", "rsg_batch_set_option", "rsg_energy_noalloc")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class MeshDesc(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int32), ("n_cells", ctypes.c_int32), ("x", _dp), ("y", _dp),
                ("z", _dp), ("cell_nodes", _ip)]


class FieldDesc(ctypes.Structure):
    _fields_ = [("data", _dp), ("type", ctypes.c_int32)]


class ModelDesc(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("density", FieldDesc), ("cij", FieldDesc * 21),
                ("moment_form", ctypes.c_int32)]


class ChannelDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("component", ctypes.c_int32), ("n_support", ctypes.c_int32),
                ("support", _ip), ("n_signal", ctypes.c_int32), ("signal_t", _dp), ("signal_v", _dp)]


class TuningDesc(ctypes.Structure):
    _fields_ = [("block_velocity", ctypes.c_int32), ("block_stress", ctypes.c_int32),
                ("per_thread_velocity", ctypes.c_int32), ("per_thread_stress", ctypes.c_int32),
                ("lds_velocity", ctypes.c_int32), ("lds_stress", ctypes.c_int32), ("tile", ctypes.c_int32 * 3),
                ("c_layout", ctypes.c_int32),
                ("xcd_remap", ctypes.c_int32), ("coeff_classes", ctypes.c_int32),
                ("node_classes", ctypes.c_int32), ("cell_classes", ctypes.c_int32),
                ("links_velocity", ctypes.c_int32), ("links_stress", ctypes.c_int32)]


def tuning_desc(tuning):
    """dict -> TuningDesc (missing keys take the library defaults)."""
    t = TuningDesc()
    if tuning:
        for k, v in tuning.items():
            if k == "tile":
                for d in range(3):
                    t.tile[d] = int(v[d])
            else:
                setattr(t, k, int(v))
    return t


class RSGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("rsg error %d: %s" % (code, msg))
        self.code = code


_LIB = None


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is missing: build it with pannde_amd.build.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.rsg_create.argtypes = [ctypes.POINTER(MeshDesc), ctypes.POINTER(ModelDesc), ctypes.c_int32,
                                 ctypes.POINTER(vp)]
        L.rsg_set_channels.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ChannelDesc)]
        L.rsg_set_sensors.argtypes = [vp, ctypes.c_int32, _ip]
        L.rsg_step.argtypes = [vp, ctypes.c_int32, _dp, _dp]
        L.rsg_get_state.argtypes = [vp, _dp, _dp]
        L.rsg_set_state.argtypes = [vp, _dp, _dp]
        L.rsg_get_traces.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, _dp]
        L.rsg_synchronize.argtypes = [vp]
        L.rsg_get_time.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_int64)]
        L.rsg_energy.argtypes = [vp, _dp, _dp]
        L.rsg_energy_noalloc.argtypes = [vp, _dp, _dp]
        L.rsg_set_timing.argtypes = [vp, ctypes.c_int32]
        L.rsg_get_kernel_times.argtypes = [vp, _dp, _dp, ctypes.POINTER(ctypes.c_int64)]
        L.rsg_device_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
        L.rsg_destroy.argtypes = [vp]
        L.rsg_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.rsg_disc_support.argtypes = [ctypes.POINTER(MeshDesc), ctypes.c_int32, _dp, ctypes.c_double,
                                       ctypes.c_double, _ip, _ip]
        L.rsg_signal_eval.argtypes = [ctypes.c_int32, _dp, _dp, ctypes.c_int32, _dp, _dp]
        L.rsg_create_ex.argtypes = [ctypes.POINTER(MeshDesc), ctypes.POINTER(ModelDesc), ctypes.c_int32,
                                    ctypes.POINTER(TuningDesc), ctypes.POINTER(vp)]
        L.rsg_get_tuning.argtypes = [vp, ctypes.POINTER(TuningDesc)]
        L.rsg_batch_create.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.POINTER(vp)]
        L.rsg_batch_add.argtypes = [vp, vp, ctypes.c_int64, _ip]
        L.rsg_batch_remove.argtypes = [vp, ctypes.c_int32]
        L.rsg_batch_run.argtypes = [vp, ctypes.c_int32, _ip]
        L.rsg_batch_synchronize.argtypes = [vp]
        L.rsg_batch_slot_status.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
                                            ctypes.POINTER(ctypes.c_int64)]
        L.rsg_batch_destroy.argtypes = [vp]
        L.rsg_batch_set_option.argtypes = [vp, ctypes.c_int32, ctypes.c_int64]
        for name in EXPORTS:
            getattr(L, name).restype = ctypes.c_int
        _LIB = L
    return _LIB


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().rsg_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def _check(rc):
    if rc != RSG_OK:
        raise RSGError(rc, last_error())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _pd(a):
    return a.ctypes.data_as(_dp)


def _pi(a):
    return a.ctypes.data_as(_ip)


class _Mesh:
    """Keeps the arrays of a MeshDesc alive."""

    def __init__(self, x, y, z, cells):
        self.x, self.y, self.z = _f64(x), _f64(y), _f64(z)
        self.cells = _i32(np.asarray(cells).reshape(-1, 8))
        self.desc = MeshDesc(self.x.shape[0], self.cells.shape[0], _pd(self.x), _pd(self.y), _pd(self.z),
                             _pi(self.cells))


def disc_support(mesh, at, center, radius, ds) -> np.ndarray:
    """HostNormalXDZ::checkAt over cell centres (at=0) or nodes (at=1); ids ascending."""
    m = mesh if isinstance(mesh, _Mesh) else _Mesh(*mesh)
    n = m.desc.n_cells if at == 0 else m.desc.n_nodes
    ids = np.empty(max(1, n), dtype=np.int32)
    cnt = ctypes.c_int32(0)
    c = _f64(center)
    _check(lib().rsg_disc_support(ctypes.byref(m.desc), at, _pd(c), float(radius), float(ds), _pi(ids),
                                  ctypes.byref(cnt)))
    return ids[:cnt.value].copy()


def signal_eval(t, v, at) -> np.ndarray:
    t, v, at = _f64(t), _f64(v), _f64(np.atleast_1d(at))
    out = np.empty_like(at)
    _check(lib().rsg_signal_eval(t.shape[0], _pd(t), _pd(v), at.shape[0], _pd(at), _pd(out)))
    return out


class DeviceSolver:
    """The RSG DeviceSolver on the GPU behind the C ABI.

    ``force`` / ``moment`` are lists of ``cases.Disc`` (HostNormalXDZ sources) in the
    reference's row order; ``precision`` is RSG_FP64 (reference arithmetic) or RSG_FP32."""

    def __init__(self, x, y, z, cells, rho, cij, dt, force=(), moment=(), moment_form=0,
                 sensors=(), precision=RSG_FP64, tuning=None):
        L = lib()
        self._mesh = _Mesh(x, y, z, cells)
        self.n_nodes = self._mesh.desc.n_nodes
        self.n_cells = self._mesh.desc.n_cells
        self.precision = precision
        self._keep = []
        md = ModelDesc()
        md.dt = float(dt)
        md.moment_form = int(moment_form)

        def fd(f):
            arr = _f64(f.data)
            self._keep.append(arr)
            return FieldDesc(_pd(arr), int(f.type))

        md.density = fd(rho)
        for k in range(21):
            md.cij[k] = fd(cij[k])
        h = ctypes.c_void_p()
        td = tuning_desc(tuning)
        _check(L.rsg_create_ex(ctypes.byref(self._mesh.desc), ctypes.byref(md), int(precision), ctypes.byref(td),
                               ctypes.byref(h)))
        self._h = h
        self._keep = []
        self.dt = float(dt)
        self.set_sources(list(force), list(moment))
        self.set_sensors(sensors)

    @classmethod
    def from_case(cls, case: "_cases.Case", precision=RSG_FP64, tuning=None) -> "DeviceSolver":
        return cls(case.x, case.y, case.z, case.cells, case.rho, case.cij, case.dt, case.force, case.moment,
                   case.moment_form, case.sensors, precision, tuning)

    def tuning(self):
        t = TuningDesc()
        _check(lib().rsg_get_tuning(self._h, ctypes.byref(t)))
        return dict(block_velocity=t.block_velocity, block_stress=t.block_stress,
                    per_thread_velocity=t.per_thread_velocity, per_thread_stress=t.per_thread_stress,
                    lds_velocity=t.lds_velocity, lds_stress=t.lds_stress, tile=tuple(t.tile),
                    c_layout=t.c_layout, xcd_remap=t.xcd_remap, coeff_classes=t.coeff_classes,
                    node_classes=t.node_classes, cell_classes=t.cell_classes,
                    links_velocity=t.links_velocity, links_stress=t.links_stress)

    # -- sources / sensors
    def set_sources(self, force=(), moment=()):
        chans, keep = [], []
        for kind, discs, at in ((RSG_FORCE, force, 0), (RSG_MOMENT, moment, 1)):
            for d in discs:
                sup = disc_support(self._mesh, at, d.center, d.radius, d.ds)
                t, v = _f64(d.signal_t), _f64(d.signal_v)
                keep += [sup, t, v]
                chans.append(ChannelDesc(kind, int(d.component), sup.shape[0], _pi(sup), t.shape[0], _pd(t), _pd(v)))
        arr = (ChannelDesc * max(1, len(chans)))(*chans)
        _check(lib().rsg_set_channels(self._h, len(chans), arr))
        self.n_channels = len(chans)

    def set_channels(self, channels):
        """Raw channels: (kind, component, support ids, signal_t or None, signal_v or None)."""
        descs, keep = [], []
        for kind, comp, sup, t, v in channels:
            sup = _i32(sup)
            keep.append(sup)
            if t is None:
                descs.append(ChannelDesc(kind, comp, sup.shape[0], _pi(sup), 0, None, None))
            else:
                t, v = _f64(t), _f64(v)
                keep += [t, v]
                descs.append(ChannelDesc(kind, comp, sup.shape[0], _pi(sup), t.shape[0], _pd(t), _pd(v)))
        arr = (ChannelDesc * max(1, len(descs)))(*descs)
        _check(lib().rsg_set_channels(self._h, len(descs), arr))
        self.n_channels = len(descs)

    def set_sensors(self, nodes):
        self.sensors = _i32(nodes)
        _check(lib().rsg_set_sensors(self._h, self.sensors.shape[0], _pi(self.sensors)))

    # -- stepping (Model::solve)
    def solve(self, nsteps=1, amplitudes=None) -> float:
        t = ctypes.c_double(0)
        if amplitudes is None:
            _check(lib().rsg_step(self._h, int(nsteps), None, ctypes.byref(t)))
        else:
            a = _f64(amplitudes)
            _check(lib().rsg_step(self._h, int(nsteps), _pd(a), ctypes.byref(t)))
        return t.value

    def synchronize(self):
        _check(lib().rsg_synchronize(self._h))

    # -- state (Model::getStates)
    def get_state(self):
        V = np.empty((3, self.n_nodes))
        S = np.empty((6, self.n_cells))
        _check(lib().rsg_get_state(self._h, _pd(V), _pd(S)))
        return V, S

    def get_states(self) -> Dict[str, np.ndarray]:
        V, S = self.get_state()
        out = {n: V[k] for k, n in enumerate(STATE_V)}
        out.update({n: S[k] for k, n in enumerate(STATE_S)})
        return out

    def set_state(self, V=None, S=None):
        v = None if V is None else _f64(V)
        s = None if S is None else _f64(S)
        _check(lib().rsg_set_state(self._h, None if v is None else _pd(v), None if s is None else _pd(s)))

    def traces(self, first=0, n=None) -> np.ndarray:
        t = ctypes.c_double(0)
        steps = ctypes.c_int64(0)
        _check(lib().rsg_get_time(self._h, ctypes.byref(t), ctypes.byref(steps)))
        if n is None:
            n = steps.value - first
        out = np.empty((max(0, n), self.sensors.shape[0], 3))
        _check(lib().rsg_get_traces(self._h, int(first), int(n), _pd(out)))
        return out

    @property
    def time(self) -> float:
        t = ctypes.c_double(0)
        _check(lib().rsg_get_time(self._h, ctypes.byref(t), None))
        return t.value

    def energy(self):
        k, s = ctypes.c_double(0), ctypes.c_double(0)
        _check(lib().rsg_energy(self._h, ctypes.byref(k), ctypes.byref(s)))
        return k.value, s.value

    def set_timing(self, on=True, stride=1):
        """Times the launches of every ``stride``-th step with HIP events (off with on=False);
        kernel_times() returns the totals and the number of timed steps."""
        _check(lib().rsg_set_timing(self._h, int(stride) if on else 0))

    def kernel_times(self):
        v, s, n = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int64(0)
        _check(lib().rsg_get_kernel_times(self._h, ctypes.byref(v), ctypes.byref(s), ctypes.byref(n)))
        return v.value, s.value, n.value

    def device_bytes(self) -> int:
        b = ctypes.c_int64(0)
        _check(lib().rsg_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    def close(self):
        if getattr(self, "_h", None):
            lib().rsg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many DeviceSolver instances advanced together (rsg_batch_*): one launch per kernel
    per step over each group of instances, K-step blocks replayed from a HIP graph.
    group_bytes / xcd_run / lanes set the scheduling options (None keeps the library default)."""

    def __init__(self, precision=RSG_FP64, steps_per_block=50, grid=0, use_graph=True, group_bytes=None,
                 xcd_run=None, lanes=None):
        h = ctypes.c_void_p()
        _check(lib().rsg_batch_create(int(precision), int(steps_per_block), int(grid), 1 if use_graph else 0,
                                      ctypes.byref(h)))
        self._h = h
        self.members = {}
        if group_bytes is not None:
            self.set_option(RSG_BATCH_GROUP_BYTES, group_bytes)
        if xcd_run is not None:
            self.set_option(RSG_BATCH_XCD_RUN, xcd_run)
        if lanes is not None:
            self.set_option(RSG_BATCH_LANES, lanes)

    def set_option(self, option: int, value: int):
        _check(lib().rsg_batch_set_option(self._h, int(option), int(value)))

    def add(self, solver: DeviceSolver, nsteps: int) -> int:
        slot = ctypes.c_int32(0)
        _check(lib().rsg_batch_add(self._h, solver._h, int(nsteps), ctypes.byref(slot)))
        self.members[slot.value] = solver
        return slot.value

    def remove(self, slot: int) -> DeviceSolver:
        _check(lib().rsg_batch_remove(self._h, int(slot)))
        return self.members.pop(slot)

    def run(self, nblocks=1) -> int:
        left = ctypes.c_int32(0)
        _check(lib().rsg_batch_run(self._h, int(nblocks), ctypes.byref(left)))
        return left.value

    def synchronize(self):
        _check(lib().rsg_batch_synchronize(self._h))

    def status(self, slot):
        d, t = ctypes.c_int64(0), ctypes.c_int64(0)
        _check(lib().rsg_batch_slot_status(self._h, int(slot), ctypes.byref(d), ctypes.byref(t)))
        return d.value, t.value

    def close(self):
        if getattr(self, "_h", None):
            lib().rsg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
