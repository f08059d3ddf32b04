import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from pannde_amd import cases  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def golden_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    """Returns (case, ref) with ref = dict(V, S, time, traces) from the reference run."""
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    cij = [cases.Field(int(d["c%02d_type" % k]), d["c%02d" % k]) for k in range(21)]
    discs = {}
    for kind in ("force", "moment"):
        discs[kind] = []
        for i in range(int(d["n_" + kind])):
            p = "%s%d_" % (kind, i)
            discs[kind].append(cases.Disc(int(d[p + "comp"]), tuple(d[p + "center"]), float(d[p + "radius"]),
                                          float(d[p + "ds"]), d[p + "t"], d[p + "v"]))
    case = cases.Case(name=name, x=d["x"], y=d["y"], z=d["z"], cells=d["cells"],
                      rho=cases.Field(int(d["rho_type"]), d["rho"]), cij=cij, dt=float(d["dt"]),
                      nsteps=int(d["nsteps"]), force=discs["force"], moment=discs["moment"],
                      moment_form=int(d["moment_form"]), sensors=d["sensors"])
    ref = dict(V=d["ref_V"], S=d["ref_S"], time=float(d["ref_time"]), traces=d["ref_traces"])
    return case, ref


@pytest.fixture(scope="session")
def golden():
    return {n: load_golden(n) for n in golden_names()}
