"""WaterColumnModel on the DeviceSolver: ``python -m pannde_amd.water_column_model -f job.vtu -o out``.

Follows applications/source/CaseRunners/WaterColumnModel.cpp on one GPU.  It reads the job:
mesh, density and C11..C66, water columns ``WC<k>/...`` applied as XX, YY, ZZ moment
excitations of the HostUTModel (:127-135, :171-180), ``dt`` and ``write_times``. It starts from
zero (:136-145) and, for every write time, steps while ``time < write_time`` (:96-114). Then it
writes the solution ``<out>_<k>.vtu``, or with ``partitioned`` the reference's multi-process
layout, ``<out>_<k>.pvtu`` with one piece.  The reference writes no tables for this model.
``--fp32`` selects the mixed-precision form.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import rsg, vtkio
from .transducer_model import output_names


def run(job: str, out: str, precision=rsg.RSG_FP64, fmt="appended", log=None, partitioned: bool = False):
    log = log or (lambda *a: None)
    case, write_times = vtkio.water_column_case(job)
    steps = vtkio.schedule_steps(case.dt, write_times)
    log("job %s: %d nodes, %d cells, %d water columns, dt %.6e, %d write times, %d steps"
        % (job, case.n_nodes, case.n_cells, len(case.moment) // 3, case.dt, len(write_times),
           steps[-1] if len(steps) else 0))
    _, _, sol_names = output_names(out, len(write_times), "", partitioned)
    write_grid = vtkio.write_pvtu if partitioned else vtkio.write_vtu
    s = rsg.DeviceSolver.from_case(case, precision)
    pts = np.stack([case.x, case.y, case.z], 1)
    done = 0
    t0 = time.perf_counter()
    try:
        for k, n in enumerate(steps):
            if n > done:
                s.solve(int(n - done))
                done = int(n)
            st = s.get_states()
            write_grid(sol_names[k], pts, case.cells, {n_: st[n_] for n_ in rsg.STATE_V},
                       {n_: st[n_] for n_ in rsg.STATE_S}, {"time": np.array([s.time])}, fmt=fmt)
            log("t=%.6e saved write index %d (%d steps)" % (s.time, k, done))
    finally:
        s.close()
    return dict(steps=int(done), seconds=time.perf_counter() - t0, write_times=len(write_times))


def main(argv=None):
    ap = argparse.ArgumentParser(description="WaterColumnModel job on the MI355X DeviceSolver")
    ap.add_argument("-f", dest="job", required=True, help="job .vtu")
    ap.add_argument("-o", dest="out", required=True, help="output base name")
    ap.add_argument("--fp32", action="store_true", help="mixed-precision form")
    ap.add_argument("--ascii", action="store_true", help="ascii solution files")
    ap.add_argument("--partitioned", action="store_true",
                    help=".pvtu summaries with one piece each, as the reference writes on several processes")
    a = ap.parse_args(argv)
    r = run(a.job, a.out, rsg.RSG_FP32 if a.fp32 else rsg.RSG_FP64, fmt="ascii" if a.ascii else "appended",
            log=lambda *m: print(*m, flush=True), partitioned=a.partitioned)
    print("done: %d steps in %.3f s" % (r["steps"], r["seconds"]), flush=True)


if __name__ == "__main__":
    sys.exit(main())
