"""Job files from the parametric case generators, without VTK:
``python -m pannde_amd.case_maker KIND -o job.vtu [options]``.

The role of the reference's case makers (py_utils/Case_Makers/*.py), which need VTK's Python
module: build a mesh with its fields, sources and timing and write it as a ``.vtu`` job that
``transducer_model`` / ``water_column_model`` (and the reference's own drivers) read.  Kinds:

* ``plate``: the CFRP laminate [0/45/90/-45]2s with seeded through-holes and an optional
  surface crack (``cases.cfrp_plate``);
* ``bracket``: the Ti-6Al-4V L-bracket with a seeded surface crack (``cases.ti_bracket``);
* ``block``: the isotropic aluminium block (``cases.al_block``);
* ``water-column``: the block driven by an XX/YY/ZZ moment column (``cases.water_column_block``),
  written as a WaterColumnModel job.

Write times follow the case makers' recipe
(Bracket_Crack_Transducer_Case_Maker_Parametric.py:124-127): ``t = arange(0, duration, dt)``,
``NT = int(len(t) / write_interval)``, ``write_times = arange(NT + 1) * write_interval * dt``, with
``duration = nsteps * dt``.
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from . import cases, vtkio


def write_times(dt: float, nsteps: int, write_interval: int) -> np.ndarray:
    t = np.arange(0, nsteps * dt, dt)
    nt = int(t.shape[0] / write_interval)
    return np.arange(nt + 1) * write_interval * dt


def make_case(kind: str, seed: int, nsteps: int, **p) -> cases.Case:
    if kind == "plate":
        crack = None
        if p.get("crack_length"):
            lx, ly = p["nx"] * p["dx"], p["ny"] * p["dx"]
            crack = cases.Crack((0.5 * lx, 0.3 * ly, 16 * p["cells_per_ply"] * p["dx"]), p["crack_angle"],
                                p["crack_length"], p["crack_depth"])
        return cases.cfrp_plate(seed=seed, nx=p["nx"], ny=p["ny"], cells_per_ply=p["cells_per_ply"], dx=p["dx"],
                                nsteps=nsteps, n_holes=p["holes"], renumber=p["renumber"], crack=crack)
    if kind == "bracket":
        return cases.ti_bracket(seed=seed, leg=p["leg"], width=p["width"], height=p["height"], dx=p["dx"],
                                nsteps=nsteps, crack_length=p.get("crack_length"), crack_depth=p.get("crack_depth"),
                                crack_angle=p.get("crack_angle"))
    if kind == "block":
        return cases.al_block(seed=seed, nx=p["nx"], ny=p["ny"], nz=p["nz"], dx=p["dx"], nsteps=nsteps)
    if kind == "water-column":
        return cases.water_column_block(seed=seed, nx=p["nx"], ny=p["ny"], nz=p["nz"], dx=p["dx"], nsteps=nsteps)
    raise ValueError("unknown kind %s" % kind)


def make_job(kind: str, out: str, seed: int = 1, nsteps: int = 500, write_interval: int = 50, fmt="appended",
             compress=False, **params):
    """Builds the case and writes its job file; returns (case, write_times)."""
    case = make_case(kind, seed, nsteps, **params)
    wt = write_times(case.dt, nsteps, write_interval)
    if kind == "water-column":
        vtkio.case_to_water_column_job(case, out, wt, fmt=fmt, compress=compress)
    else:
        vtkio.case_to_job(case, out, wt, fmt=fmt, compress=compress)
    return case, wt


def main(argv=None):
    ap = argparse.ArgumentParser(description="Write an RSG job file from a parametric case generator")
    ap.add_argument("kind", choices=["plate", "bracket", "block", "water-column"])
    ap.add_argument("-o", dest="out", required=True, help="job .vtu")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--nsteps", type=int, default=500)
    ap.add_argument("--write-interval", type=int, default=50, help="steps between written solutions")
    ap.add_argument("--dx", type=float, default=None, help="cell size in m (default: the kind's)")
    ap.add_argument("--nx", type=int, default=40)
    ap.add_argument("--ny", type=int, default=40)
    ap.add_argument("--nz", type=int, default=20)
    ap.add_argument("--cells-per-ply", type=int, default=4)
    ap.add_argument("--holes", type=int, default=5)
    ap.add_argument("--renumber", action="store_true", help="plate: seeded permutation then RCM numbering")
    ap.add_argument("--leg", type=int, default=512)
    ap.add_argument("--width", type=int, default=96)
    ap.add_argument("--height", type=int, default=96)
    ap.add_argument("--crack-length", type=float, default=None, help="m")
    ap.add_argument("--crack-depth", type=float, default=None, help="m")
    ap.add_argument("--crack-angle", type=float, default=None, help="degrees")
    ap.add_argument("--compress", action="store_true", help="zlib-compressed arrays")
    a = ap.parse_args(argv)
    dx = a.dx if a.dx is not None else (1.0e-4 if a.kind in ("block", "water-column") else 5.0e-5)
    params = dict(nx=a.nx, ny=a.ny, nz=a.nz, dx=dx, cells_per_ply=a.cells_per_ply, holes=a.holes,
                  renumber=a.renumber, leg=a.leg, width=a.width, height=a.height, crack_length=a.crack_length,
                  crack_depth=a.crack_depth if a.crack_depth is not None else (1.0e-3 if a.kind == "plate" else None),
                  crack_angle=a.crack_angle if a.crack_angle is not None else (0.0 if a.kind == "plate" else None))
    case, wt = make_job(a.kind, a.out, a.seed, a.nsteps, a.write_interval, compress=a.compress, **params)
    print("%s: %s, %d nodes, %d cells, dt %.6e, %d write times" % (a.out, case.name, case.n_nodes, case.n_cells,
                                                                    case.dt, len(wt)), flush=True)


if __name__ == "__main__":
    sys.exit(main())
