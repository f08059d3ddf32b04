"""The C++ drivers' VTK XML reader (pannde_amd/csrc/vtk_job.hpp) and job semantics, on the CPU.

* The reader is compiled into tests/vtk/vtu_dump.cpp and must return exactly what
  pannde_amd.vtkio.read_vtu returns (points, hexahedra, every PointData/CellData/FieldData
  array) for every encoding, header type and compression the writer produces, and for files
  written by hand from the format's definition: big-endian inline base64, appended base64, a zlib
  block header built field by field (a partial last block, and a full one recorded as 0), voxels.
* ``rsg_driver -f job.vtu --dump-case FILE`` writes the job it read as RSGCASE1 without touching
  the GPU: byte-identical to the Python job reader's case for TransducerModel and WaterColumnModel
  jobs (fields NODE- or CELL-placed, transducers, water columns, dt, the write-time schedule's step
  count, the recorded nodes).
"""
import base64
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

from pannde_amd import case_file, cases, vtkio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "pannde_amd", "rsg_driver")


@pytest.fixture(scope="module")
def dumper(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("vtu_dump") / "vtu_dump")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", os.path.join(ROOT, "tests", "vtk", "vtu_dump.cpp"),
                    "-o", exe, "-lz"], check=True)
    return exe


def cpp_read(exe, path):
    r = subprocess.run([exe, str(path)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise vtkio.VtkError(r.stderr.strip())
    lines = r.stdout.splitlines()
    npts, ncells = int(lines[0].split()[1]), int(lines[1].split()[1])
    pts = np.array([float(t) for t in lines[2].split()[1:]]).reshape(npts, 3)
    cells = np.array([int(t) for t in lines[3].split()[1:]], np.int64).reshape(ncells, 8)
    arrays = {"P": {}, "C": {}, "F": {}}
    for ln in lines[4:]:
        t = ln.split()
        ncomp = int(t[2])
        a = np.array([float(x) for x in t[3:]])
        arrays[t[0]][t[1]] = a.reshape(-1, ncomp) if ncomp > 1 else a
    return pts, cells, arrays


def assert_same(exe, path):
    v = vtkio.read_vtu(str(path))
    pts, cells, arrays = cpp_read(exe, path)
    assert np.array_equal(pts, v.points) and np.array_equal(cells, v.cells)
    for key, py in (("P", v.point_data), ("C", v.cell_data), ("F", v.field_data)):
        assert sorted(arrays[key]) == sorted(py), key
        for name, a in py.items():
            assert np.array_equal(arrays[key][name], np.asarray(a, np.float64)), (key, name)
    return v


def two_hexes():
    x, y, z, cells, _ = cases.structured_mesh(2, 1, 1, 0.5, 1.0, 2.0)
    return np.stack([x, y, z], 1), cells


@pytest.mark.parametrize("fmt", ["ascii", "binary", "appended"])
@pytest.mark.parametrize("header", ["UInt32", "UInt64"])
@pytest.mark.parametrize("compress", [False, True])
def test_reader_equals_python_reader(dumper, tmp_path, fmt, header, compress):
    if fmt == "ascii" and compress:
        pytest.skip("ascii data is never compressed")
    rng = np.random.default_rng(5)
    pts, cells = two_hexes()
    n = pts.shape[0]
    pdata = {"density": rng.standard_normal(n), "f32": rng.standard_normal(n).astype(np.float32),
             "vec": rng.standard_normal((n, 3)), "ids": np.arange(n, dtype=np.int32) - 5,
             "i16": (np.arange(n) * -300).astype(np.int16)}
    cdata = {"C11": rng.standard_normal(2) * 1e11, "u8": np.array([7, 250], np.uint8), "i64": np.array([-2 ** 40, 3]),
             "u16": np.array([65535, 1], np.uint16), "i8": np.array([-128, 127], np.int8)}
    fdata = {"dt": np.array([1.2345e-9]), "write_times": np.linspace(0, 1e-6, 7), "XD0/Radius": np.array([2e-3])}
    p = tmp_path / "a.vtu"
    vtkio.write_vtu(str(p), pts, cells, pdata, cdata, fdata, fmt=fmt, header=header, compress=compress)
    assert_same(dumper, p)


def test_several_compressed_blocks(dumper, tmp_path):
    pts, cells = two_hexes()
    big = np.random.default_rng(1).standard_normal(20000)  # 160 kB: five 32 kB blocks
    p = tmp_path / "b.vtu"
    vtkio.write_vtu(str(p), pts, cells, field_data={"big": big}, fmt="appended", compress=True)
    assert np.array_equal(assert_same(dumper, p).field_data["big"], big)


def test_hand_written_files(dumper, tmp_path):
    from test_vtkio import ASCII_JOB, replace_array

    p = tmp_path / "h.vtu"
    p.write_text(ASCII_JOB)
    assert_same(dumper, p)
    # big-endian inline binary: base64 of [UInt32 byte count][data]
    payload = struct.pack(">I", 16) + struct.pack(">2d", 100.0, 200.0)
    el = '<DataArray type="Float64" Name="density" format="binary">%s</DataArray>' % base64.b64encode(payload).decode()
    xml = replace_array(ASCII_JOB, 'Name="density" format="ascii">100', el)
    p.write_text(xml.replace('byte_order="LittleEndian"', 'byte_order="BigEndian"'))
    assert assert_same(dumper, p).cell_data["density"].tolist() == [100.0, 200.0]
    # a voxel: vertices 2 and 3 (and 6 and 7) swapped relative to the hexahedron
    xml = ASCII_JOB.replace('<DataArray type="Int64" Name="connectivity" format="ascii">0 1 4 3 6 7 10 9',
                            '<DataArray type="Int64" Name="connectivity" format="ascii">0 1 3 4 6 7 9 10')
    p.write_text(xml.replace(">12 12<", ">11 12<"))
    assert assert_same(dumper, p).cells[0].tolist() == [0, 1, 4, 3, 6, 7, 10, 9]
    # other cell types are an error in both readers
    p.write_text(ASCII_JOB.replace(">12 12<", ">12 10<"))
    with pytest.raises(vtkio.VtkError, match="hexahedral"):
        cpp_read(dumper, p)


@pytest.mark.parametrize("full_last", [False, True])
def test_zlib_header_from_the_format(dumper, tmp_path, full_last):
    # [number of blocks, block size, size of a partial last block (0: the last block is full),
    #  compressed sizes], UInt64 here, base64 of the header and of the blocks as separate runs
    from test_vtkio import ASCII_JOB, replace_array

    data = np.arange(12 if full_last else 10, dtype=np.float64) * 0.5
    raw = data.tobytes()
    blocks = [raw[:48], raw[48:]]
    comp = [zlib.compress(b) for b in blocks]
    head = struct.pack("<5Q", 2, 48, 0 if full_last else len(blocks[1]), len(comp[0]), len(comp[1]))
    text = (base64.b64encode(head) + base64.b64encode(b"".join(comp))).decode()
    xml = ASCII_JOB.replace('version="0.1" byte_order="LittleEndian"',
                            'version="1.0" byte_order="LittleEndian" header_type="UInt64" '
                            'compressor="vtkZLibDataCompressor"')
    xml = replace_array(xml, 'Name="dt"', '<DataArray type="Float64" Name="dt" format="binary">%s</DataArray>' % text)
    p = tmp_path / "z.vtu"
    p.write_text(xml)
    assert np.array_equal(assert_same(dumper, p).field_data["dt"], data)


def test_appended_base64(dumper, tmp_path):
    # appended base64: offsets count characters of the encoded section; each array's header and
    # data are separate base64 runs
    from test_vtkio import ASCII_JOB, replace_array

    a = np.array([3.5, -1.25])
    b = np.arange(12, dtype=np.float32) + 0.5
    enc = []
    for arr in (a, b):
        d = arr.tobytes()
        enc.append(base64.b64encode(struct.pack("<I", len(d))) + base64.b64encode(d))
    xml = replace_array(ASCII_JOB, 'Name="C11" format="ascii">',
                        '<DataArray type="Float64" Name="C11" format="appended" offset="0"/>')
    xml = replace_array(xml, 'Name="density" format="ascii">1 2',
                        '<DataArray type="Float32" Name="density" format="appended" offset="%d"/>' % len(enc[0]))
    xml = xml.replace("</VTKFile>", '<AppendedData encoding="base64">\n_%s\n</AppendedData>\n</VTKFile>'
                      % b"".join(enc).decode())
    p = tmp_path / "ab.vtu"
    p.write_text(xml)
    v = assert_same(dumper, p)
    assert np.array_equal(v.cell_data["C11"], a) and np.array_equal(v.point_data["density"], b)


def dump_job(tmp_path, job, water=False):
    out = tmp_path / "cpp.bin"
    cmd = [DRIVER, "-f", str(job), "--dump-case", str(out)] + (["--water-column"] if water else [])
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    return out.read_bytes()


def py_job(tmp_path, job, water=False):
    case, _ = (vtkio.water_column_case if water else vtkio.job_case)(str(job))
    p = tmp_path / "py.bin"
    case_file.write_case_file(str(p), case)
    return p.read_bytes(), case


@pytest.mark.parametrize("fmt,compress", [("appended", True), ("binary", False), ("ascii", False)])
def test_transducer_job_equals_python_job(tmp_path, fmt, compress):
    case = cases.cfrp_plate(seed=3, nx=14, ny=12, cells_per_ply=1, dx=2e-4, nsteps=5, n_holes=1, renumber=True)
    case.force = list(case.force) + [cases.Disc(2, (1.1e-3, 0.9e-3, case.force[0].center[2]), 6e-4,
                                                case.force[0].ds, case.force[0].signal_t, -case.force[0].signal_v)]
    job = tmp_path / "job.vtu"
    vtkio.case_to_job(case, str(job), np.array([0.0, 3.3, 9.0]) * case.dt, fmt=fmt, compress=compress)
    ref, got_case = py_job(tmp_path, job)
    assert got_case.sensors.shape[0] > 0 and got_case.nsteps == 9 and len(got_case.force) == 2
    assert dump_job(tmp_path, job) == ref


def test_node_placed_fields_are_read_from_point_data(tmp_path):
    case = cases.al_block(seed=2, nx=6, ny=5, nz=4, nsteps=3, n_sensors=1)
    rng = np.random.default_rng(0)
    case.rho = cases.Field(cases.FIELD_NODE, 2700.0 * (1 + 0.1 * rng.random(case.n_nodes)))
    case.cij = [cases.Field(cases.FIELD_NODE, np.full(case.n_nodes, float(f.data[0])) * (1 + 0.05 * rng.random(
        case.n_nodes))) if k % 4 == 0 else f for k, f in enumerate(case.cij)]
    job = tmp_path / "job.vtu"
    vtkio.case_to_job(case, str(job), np.array([2.5]) * case.dt)
    ref, got_case = py_job(tmp_path, job)
    assert got_case.rho.type == cases.FIELD_NODE and got_case.cij[4].type == cases.FIELD_NODE
    assert dump_job(tmp_path, job) == ref


def test_water_column_job_equals_python_job(tmp_path):
    from test_vtkio import two_column_case

    case = two_column_case(nsteps=12)
    job = tmp_path / "wc.vtu"
    vtkio.case_to_water_column_job(case, str(job), np.array([0.0, 4.5, 11.0]) * case.dt, fmt="binary", compress=True)
    ref, got_case = py_job(tmp_path, job, water=True)
    assert got_case.moment_form == 1 and len(got_case.moment) == 6
    assert dump_job(tmp_path, job, water=True) == ref


@pytest.mark.parametrize("kind", ["plate", "bracket", "water-column"])
def test_case_maker_jobs(tmp_path, kind):
    from pannde_amd import case_maker

    p = tmp_path / ("%s.vtu" % kind)
    argv = [kind, "-o", str(p), "--seed", "7", "--nsteps", "21", "--write-interval", "10"]
    argv += {"plate": ["--nx", "16", "--ny", "14", "--cells-per-ply", "1", "--holes", "1", "--crack-length", "4e-4",
                       "--crack-depth", "3e-4", "--crack-angle", "30"],
             "bracket": ["--leg", "16", "--width", "6", "--height", "5", "--crack-length", "3e-4",
                         "--crack-depth", "1e-4", "--crack-angle", "60"],
             "water-column": ["--nx", "7", "--ny", "6", "--nz", "5"]}[kind]
    case_maker.main(argv)
    water = kind == "water-column"
    ref, _ = py_job(tmp_path, p, water=water)
    assert dump_job(tmp_path, p, water=water) == ref


def test_job_errors(tmp_path):
    case = cases.al_block(seed=2, nx=5, ny=4, nz=3, nsteps=3, n_sensors=1)
    job = tmp_path / "job.vtu"
    vtkio.case_to_job(case, str(job), np.array([1.0]) * case.dt, fmt="ascii")
    text = job.read_text()
    bad = tmp_path / "bad.vtu"
    bad.write_text(text.replace('Name="C44"', 'Name="C4_4"'))
    r = subprocess.run([DRIVER, "-f", str(bad), "--dump-case", str(tmp_path / "x.bin")], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 1 and "no field named C44" in r.stdout
    r = subprocess.run([DRIVER, "-f", str(job), "--water-column", "--dump-case", str(tmp_path / "x.bin")],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 1 and "NWaterColumns" in r.stdout
    r = subprocess.run([DRIVER, "-f", str(job)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 1 and "-o out" in r.stdout


def test_corrupt_sizes_fail_cleanly(dumper, tmp_path):
    # sizes in the file that exceed the data (a huge byte count, a zlib block claiming far more
    # than its compressed bytes can hold) are errors, not allocations
    from test_vtkio import ASCII_JOB, replace_array

    head = struct.pack("<Q", 1 << 40)
    el = '<DataArray type="Float64" Name="dt" format="binary">%s</DataArray>' % base64.b64encode(
        head + struct.pack("<d", 1.0)).decode()
    xml = ASCII_JOB.replace('version="0.1" byte_order="LittleEndian"',
                            'version="1.0" byte_order="LittleEndian" header_type="UInt64"')
    p = tmp_path / "huge.vtu"
    p.write_text(replace_array(xml, 'Name="dt"', el))
    with pytest.raises(vtkio.VtkError, match="ends early"):
        cpp_read(dumper, p)
    comp = zlib.compress(struct.pack("<2d", 1.0, 2.0))
    head = struct.pack("<4Q", 1, 1 << 40, 1 << 40, len(comp))
    el = '<DataArray type="Float64" Name="dt" format="binary">%s</DataArray>' % (
        base64.b64encode(head) + base64.b64encode(comp)).decode()
    xml = xml.replace('header_type="UInt64"', 'header_type="UInt64" compressor="vtkZLibDataCompressor"')
    p.write_text(replace_array(xml, 'Name="dt"', el))
    with pytest.raises(vtkio.VtkError, match="not valid"):
        cpp_read(dumper, p)
