"""The set-up's threading does not change what it uploads: rsg_create on 1 and on 8 threads
(PANNDE_RSG_SETUP_THREADS) of a mesh large enough for 8 chunks, in FP64 and FP32 and with
per-entity coefficients, gives byte-identical device buffers.  The host code runs against
the host-memory stand-ins for the HIP runtime (tests/sanitize/hip_host_stub.cpp), whose
STUB_DIGEST file lists every device buffer's size and content hash."""
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
case = cases.cfrp_plate(seed=3, nx=104, ny=96, nsteps=2, n_holes=2, renumber=False,
                        crack=cases.Crack((2.2e-3, 2.0e-3, 64 * 5e-5), 30.0, 1.5e-3, 8e-4))
for prec in (rsg.RSG_FP64, rsg.RSG_FP32):
    for tun in (None, dict(coeff_classes=-1)):
        s = rsg.DeviceSolver.from_case(case, prec, tuning=tun)
        s.close()
print("ok", case.n_cells)
'''


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists("/opt/rocm/include/hip/hip_runtime_api.h"),
                    reason="needs g++ and the HIP headers")
@pytest.mark.timeout(600)
def test_setup_uploads_do_not_depend_on_thread_count(tmp_path):
    lib = str(tmp_path / "libstub.so")
    cmd = ["g++", "-x", "c++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", os.path.join(ROOT, "pannde_amd", "csrc", "rsg_host.cpp"),
           os.path.join(ROOT, "tests", "sanitize", "hip_host_stub.cpp"), "-o", lib, "-lpthread"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:]
    digests = []
    for threads in (1, 8):
        out = str(tmp_path / ("digest%d.txt" % threads))
        env = dict(os.environ, PANNDE_RSG_LIB=lib, STUB_DIGEST=out, PANNDE_RSG_SETUP_THREADS=str(threads))
        r = subprocess.run([sys.executable, "-c", SCRIPT, ROOT], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=300, env=env)
        assert r.returncode == 0 and "ok" in r.stdout, r.stdout[-4000:]
        n_cells = int(r.stdout.split()[-1])
        assert n_cells > 8 * (1 << 16)  # every set-up loop runs in 8 chunks on 8 threads
        digests.append(open(out).read())
    assert digests[0] and digests[0] == digests[1]
