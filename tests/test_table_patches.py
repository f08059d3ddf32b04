"""Table builds patch only the keys whose special flag changes (set-up uploads every plain key;
a build rewrites the keys of the entities special before or now).  On the CPU, through the
host stand-ins for the HIP runtime (tests/sanitize/hip_host_stub.cpp, whose STUB_DIGEST file
lists every device buffer's size and content hash when it is freed): a handle whose sources
and sensors are replaced several times ends with the same link-and-key records on the device
as a handle created with the final ones."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r'''
import sys
sys.path.insert(0, sys.argv[1])
from pannde_amd import cases, rsg
mode, prec = sys.argv[2], int(sys.argv[3])
case = cases.water_column_block(seed=9, nx=16, ny=14, nz=7, nsteps=10)
other = cases.al_block(seed=9, nx=16, ny=14, nz=7, nsteps=10, n_sensors=9)
if mode == "live":
    s = rsg.DeviceSolver(case.x, case.y, case.z, case.cells, case.rho, case.cij, case.dt, case.force,
                         case.moment, 1, case.sensors, prec)
    s.set_sources(other.force, case.moment[:1])
    s.set_sensors(other.sensors)
    s.set_sources(other.force, [])
else:
    s = rsg.DeviceSolver(case.x, case.y, case.z, case.cells, case.rho, case.cij, case.dt, other.force, [], 1,
                         other.sensors, prec)
print("sizes", case.n_nodes * 16, case.n_cells * 16)
s.close()
'''


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime_api.h"),
                    reason="needs g++ and the HIP headers")
@pytest.mark.timeout(600)
def test_replaced_sources_leave_the_records_of_a_fresh_handle(tmp_path):
    lib = str(tmp_path / "libstub.so")
    cmd = ["g++", "-x", "c++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", os.path.join(ROOT, "pannde_amd", "csrc", "rsg_host.cpp"),
           os.path.join(ROOT, "tests", "sanitize", "hip_host_stub.cpp"), "-o", lib, "-lpthread"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:]
    for prec in ("0", "1"):
        records = {}
        for mode in ("live", "fresh"):
            out = str(tmp_path / ("digest_%s_%s.txt" % (mode, prec)))
            env = dict(os.environ, PANNDE_RSG_LIB=lib, STUB_DIGEST=out)
            r = subprocess.run([sys.executable, "-c", SCRIPT, ROOT, mode, prec], stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True, timeout=300, env=env)
            assert r.returncode == 0, r.stdout[-4000:]
            sizes = set(r.stdout.split("sizes")[1].split())
            # the 16-byte link-and-key records of the nodes and of the cells, allocated once
            records[mode] = sorted(l for l in open(out).read().splitlines() if l.split()[0] in sizes)
        assert len(records["fresh"]) == 2, records
        assert records["live"] == records["fresh"]
