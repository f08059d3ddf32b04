"""The C++ drop-in: DeviceSolver::DeviceRSGTDUT / DeviceUTModel (include/DeviceSolver)
implementing PanNDE::Model over the C ABI.

CPU: the header compiles against this repository's interface headers and, where the
reference tree is present, against the reference's own PanNDE interfaces together with a
caller that uses the reference's HostData containers and HostNormalXDZ sources.
GPU: the TransducerModel-style driver (pannde_amd/rsg_driver) runs every golden fixture
through the PanNDE::Model interface and must reproduce the reference's fields, time and
traces bit for bit, both with table-driven sources and with opaque MultiVariate sources
(host-evaluated every step), and when stepping one solve() at a time or in chunks."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT, golden_names, load_golden
from oracle import oracle_rsg as orc

REF = "/root/reference/modules"
DRIVER = os.path.join(ROOT, "pannde_amd", "rsg_driver")


def gxx(args, source=None):
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only"] + args, input=source, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout


def test_header_compiles_against_own_interfaces():
    gxx(["-Wall", "-Werror", "-x", "c++", "-I" + os.path.join(ROOT, "include", "PanNDE"),
         "-I" + os.path.join(ROOT, "include"), "-"],
        source='#include "DeviceSolver/DeviceRSGTDUT.hpp"\n#include "DeviceSolver/HostContainers.hpp"\n')


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree absent")
def test_reference_caller_compiles_with_device_models():
    inc = ["-I" + os.path.join(REF, d, "include") for d in ("PanNDE", "HostData", "Stubs")]
    gxx(["-w"] + inc + ["-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "dropin_reference_caller.cpp")])


def run_driver(case, *flags):
    d = tempfile.mkdtemp(prefix="rsgdrv_")
    cp, op = os.path.join(d, "case.bin"), os.path.join(d, "out.bin")
    orc.write_case_file(cp, case)
    r = subprocess.run([DRIVER, cp, op] + list(flags), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout
    return orc.read_out_file(op)


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [(), ("--opaque",), ("--chunk", "7")])
@pytest.mark.parametrize("name", golden_names())
def test_driver_matches_reference_fixture(name, flags):
    case, ref = load_golden(name)
    V, S, t, tr = run_driver(case, *flags)
    assert t == ref["time"]
    assert np.array_equal(V, ref["V"]) and np.array_equal(S, ref["S"])
    assert np.array_equal(tr, ref["traces"])


@pytest.mark.gpu
def test_driver_fp32_within_tolerance():
    case, ref = load_golden("three_transducers")
    V, S, _, tr = run_driver(case, "--fp32")
    err = np.linalg.norm(tr - ref["traces"]) / np.linalg.norm(ref["traces"])
    assert err < 1e-4, err
