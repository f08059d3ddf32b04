"""TransducerModel on the DeviceSolver: ``python -m pannde_amd.transducer_model -f job.vtu -o out``.

Follows applications/source/CaseRunners/TransducerModel.cpp on one GPU: reads the job
(mesh, density and C11..C66, transducers ``XD<k>/...``, ``dt``, ``write_times``;
TransducerModel.cpp:160-204), starts from zero (:150-159), records the velocities of the
nodes any transducer's window accepts (:106-130), and for every write time steps while
``time < write_time`` (:63-83), then writes the solution and the velocity table.  Outputs,
named as the reference's VTK gateway names them (VTKGateway.hpp:218-283, VTKWrite.hpp:100-232):
the solution under the output base, the tables under their own names in the tables
directory (the working directory, as the reference writes them):

* ``node_coords.vtt``: node_x, node_y, node_z of the recorded nodes (TransducerModel.cpp:60);
* ``<out>_<k>.vtu``: the mesh with Vx, Vy, Vz (point data) and Sxx..Sxy (cell data) (:77);
* ``velocities_<k>.vtt``: Vx, Vy, Vz of the recorded nodes (:78).

With ``partitioned`` the files take the layout the reference writes on several processes
(VTKPWrite.hpp:100-137): ``.pvtu`` / ``.pvtt`` summaries, each with one piece
``<name>_0.vtu`` / ``<name>_0.vtt`` beside it, which the reference's post-processing
(py_utils/PVTTReader.py) reads.  The MPI partition and halo exchange of the reference
disappear: the whole mesh lives on one GPU.  ``--fp32`` selects the mixed-precision form.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import rsg, vtkio


def output_names(out: str, n_writes: int, tables_dir: str = "", partitioned: bool = False):
    """(node-coordinate table, [velocity table per write], [solution per write]) file names."""
    t_ext, g_ext = (".pvtt", ".pvtu") if partitioned else (".vtt", ".vtu")
    coords = os.path.join(tables_dir, "node_coords" + t_ext)
    vel = [os.path.join(tables_dir, "velocities_%d%s" % (k, t_ext)) for k in range(n_writes)]
    sol = ["%s_%d%s" % (out, k, g_ext) for k in range(n_writes)]
    return coords, vel, sol


def run(job: str, out: str, precision=rsg.RSG_FP64, fields=True, fmt="appended", log=None, tables_dir: str = "",
        partitioned: bool = False):
    log = log or (lambda *a: None)
    case, write_times = vtkio.job_case(job)
    steps = vtkio.schedule_steps(case.dt, write_times)
    log("job %s: %d nodes, %d cells, %d transducers, dt %.6e, %d write times, %d steps"
        % (job, case.n_nodes, case.n_cells, len(case.force), case.dt, len(write_times), steps[-1] if len(steps) else 0))
    coords_name, vel_names, sol_names = output_names(out, len(write_times), tables_dir, partitioned)
    write_table = vtkio.write_pvtt if partitioned else vtkio.write_vtt
    write_grid = vtkio.write_pvtu if partitioned else vtkio.write_vtu
    sens = case.sensors
    write_table(coords_name, {"node_x": case.x[sens], "node_y": case.y[sens], "node_z": case.z[sens]})
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
            write_table(vel_names[k], {"Vx": tr[:, 0], "Vy": tr[:, 1], "Vz": tr[:, 2]})
            if fields:
                st = s.get_states()
                write_grid(sol_names[k], pts, case.cells, {n_: st[n_] for n_ in rsg.STATE_V},
                           {n_: st[n_] for n_ in rsg.STATE_S}, {"time": np.array([s.time])}, fmt=fmt)
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
    ap.add_argument("--tables-dir", default="", help="directory of the tables (default: the working directory)")
    ap.add_argument("--partitioned", action="store_true",
                    help=".pvtu/.pvtt summaries with one piece each, as the reference writes on several processes")
    a = ap.parse_args(argv)
    r = run(a.job, a.out, rsg.RSG_FP32 if a.fp32 else rsg.RSG_FP64, fields=not a.no_fields,
            fmt="ascii" if a.ascii else "appended", log=lambda *m: print(*m, flush=True), tables_dir=a.tables_dir,
            partitioned=a.partitioned)
    print("done: %d steps in %.3f s" % (r["steps"], r["seconds"]), flush=True)


if __name__ == "__main__":
    sys.exit(main())
