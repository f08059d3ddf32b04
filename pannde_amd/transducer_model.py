"""TransducerModel on the DeviceSolver: ``python -m pannde_amd.transducer_model -f job.vtu -o out``.

Follows applications/source/CaseRunners/TransducerModel.cpp on one GPU: reads the job
(mesh, density and C11..C66, transducers ``XD<k>/...``, ``dt``, ``write_times``;
TransducerModel.cpp:160-204), starts from zero (:150-159), records the velocities of the
nodes any transducer's window accepts (:106-130), and for every write time steps while
``time < write_time`` (:63-83), then writes the solution and the velocity table.  Outputs,
named as the reference's VTK gateway names them (VTKWrite.hpp:200-232):

* ``<out>_node_coords.vtt``: node_x, node_y, node_z of the recorded nodes (:58);
* ``<out>_<k>.vtu``: the mesh with Vx, Vy, Vz (point data) and Sxx..Sxy (cell data) (:77);
* ``<out>_velocities_<k>.vtt``: Vx, Vy, Vz of the recorded nodes (:78).

The MPI partition and halo exchange of the reference disappear: the whole mesh lives on
one GPU.  ``--fp32`` selects the mixed-precision form.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import rsg, vtkio


def run(job: str, out: str, precision=rsg.RSG_FP64, fields=True, fmt="appended", log=None):
    log = log or (lambda *a: None)
    case, write_times = vtkio.job_case(job)
    steps = vtkio.schedule_steps(case.dt, write_times)
    log("job %s: %d nodes, %d cells, %d transducers, dt %.6e, %d write times, %d steps"
        % (job, case.n_nodes, case.n_cells, len(case.force), case.dt, len(write_times), steps[-1] if len(steps) else 0))
    sens = case.sensors
    vtkio.write_vtt(out + "_node_coords.vtt", {"node_x": case.x[sens], "node_y": case.y[sens], "node_z": case.z[sens]})
    s = rsg.DeviceSolver.from_case(case, precision)
    pts = np.stack([case.x, case.y, case.z], 1)
    done = 0
    t0 = time.perf_counter()
    try:
        for k, n in enumerate(steps):
            if n > done:
                s.solve(int(n - done))
                done = int(n)
            tr = s.traces(done - 1, 1)[0] if done and sens.shape[0] else np.zeros((sens.shape[0], 3))
            vtkio.write_vtt("%s_velocities_%d.vtt" % (out, k), {"Vx": tr[:, 0], "Vy": tr[:, 1], "Vz": tr[:, 2]})
            if fields:
                st = s.get_states()
                vtkio.write_vtu("%s_%d.vtu" % (out, k), pts, case.cells,
                                {n_: st[n_] for n_ in rsg.STATE_V}, {n_: st[n_] for n_ in rsg.STATE_S},
                                {"time": np.array([s.time])}, fmt=fmt)
            log("t=%.6e saved write index %d (%d steps)" % (s.time, k, done))
    finally:
        s.close()
    return dict(steps=int(done), seconds=time.perf_counter() - t0, write_times=len(write_times))


def main(argv=None):
    ap = argparse.ArgumentParser(description="TransducerModel job on the MI355X DeviceSolver")
    ap.add_argument("-f", dest="job", required=True, help="job .vtu")
    ap.add_argument("-o", dest="out", required=True, help="output base name")
    ap.add_argument("--fp32", action="store_true", help="mixed-precision form")
    ap.add_argument("--no-fields", action="store_true", help="write the velocity tables only")
    ap.add_argument("--ascii", action="store_true", help="ascii solution files")
    a = ap.parse_args(argv)
    r = run(a.job, a.out, rsg.RSG_FP32 if a.fp32 else rsg.RSG_FP64, fields=not a.no_fields,
            fmt="ascii" if a.ascii else "appended", log=lambda *m: print(*m, flush=True))
    print("done: %d steps in %.3f s" % (r["steps"], r["seconds"]), flush=True)


if __name__ == "__main__":
    sys.exit(main())
