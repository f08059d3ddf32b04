"""VTK XML reader/writer (no VTK in this image): round trips over every encoding, header
type and compression, plus files written by hand from the format's definition (ascii,
big-endian inline base64, a zlib block header built field by field, a voxel cell), and the
job conversion (fields, transducers, timing, sensors) against the oracle's schedule."""
import base64
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np
import pytest

from oracle import oracle_rsg as orc
from pannde_amd import cases, vtkio


def two_hexes():
    x, y, z, cells, _ = cases.structured_mesh(2, 1, 1, 0.5, 1.0, 2.0)
    return np.stack([x, y, z], 1), cells


@pytest.mark.parametrize("fmt", ["ascii", "binary", "appended"])
@pytest.mark.parametrize("header", ["UInt32", "UInt64"])
@pytest.mark.parametrize("compress", [False, True])
def test_round_trip(tmp_path, fmt, header, compress):
    if fmt == "ascii" and compress:
        pytest.skip("ascii data is never compressed")
    rng = np.random.default_rng(3)
    pts, cells = two_hexes()
    pdata = {"density": rng.standard_normal(pts.shape[0]), "f32": rng.standard_normal(pts.shape[0]).astype(np.float32),
             "vec": rng.standard_normal((pts.shape[0], 3)), "ids": np.arange(pts.shape[0], dtype=np.int32) - 5}
    cdata = {"C11": rng.standard_normal(2) * 1e11, "u8": np.array([7, 250], np.uint8), "i64": np.array([-2 ** 40, 3])}
    fdata = {"dt": np.array([1.2345e-9]), "write_times": np.linspace(0, 1e-6, 7), "XD0/Radius": np.array([2e-3])}
    p = tmp_path / "a.vtu"
    vtkio.write_vtu(str(p), pts, cells, pdata, cdata, fdata, fmt=fmt, header=header, compress=compress)
    v = vtkio.read_vtu(str(p))
    assert np.array_equal(v.points, pts) and np.array_equal(v.cells, cells)
    for src, dst in ((pdata, v.point_data), (cdata, v.cell_data), (fdata, v.field_data)):
        for k, a in src.items():
            assert dst[k].dtype == np.asarray(a).dtype, k
            assert np.array_equal(dst[k], a), k
    assert v.value("dt") == 1.2345e-9
    with pytest.raises(vtkio.VtkError, match="scalar"):
        v.value("write_times")


def test_large_compressed_arrays_use_several_blocks(tmp_path):
    pts, cells = two_hexes()
    big = np.random.default_rng(1).standard_normal(20000)  # 160 kB: five 32 kB blocks
    p = tmp_path / "b.vtu"
    vtkio.write_vtu(str(p), pts, cells, field_data={"big": big}, fmt="appended", compress=True)
    assert np.array_equal(vtkio.read_vtu(str(p)).field_data["big"], big)


ASCII_JOB = """<?xml version="1.0"?>
<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">
  <UnstructuredGrid>
    <FieldData>
      <DataArray type="Float64" Name="dt" NumberOfTuples="1" format="ascii">2.5e-08</DataArray>
    </FieldData>
    <Piece NumberOfPoints="12" NumberOfCells="2">
      <PointData>
        <DataArray type="Float32" Name="density" format="ascii">1 2 3 4 5 6 7 8 9 10 11 12</DataArray>
      </PointData>
      <CellData>
        <DataArray type="Float64" Name="density" format="ascii">100 200</DataArray>
        <DataArray type="Float64" Name="C11" format="ascii">1.5e11 2.5e11</DataArray>
      </CellData>
      <Points>
        <DataArray type="Float64" NumberOfComponents="3" format="ascii">
          0 0 0  1 0 0  2 0 0  0 1 0  1 1 0  2 1 0
          0 0 1  1 0 1  2 0 1  0 1 1  1 1 1  2 1 1
        </DataArray>
      </Points>
      <Cells>
        <DataArray type="Int64" Name="connectivity" format="ascii">0 1 4 3 6 7 10 9  1 2 5 4 7 8 11 10</DataArray>
        <DataArray type="Int64" Name="offsets" format="ascii">8 16</DataArray>
        <DataArray type="UInt8" Name="types" format="ascii">12 12</DataArray>
      </Cells>
    </Piece>
  </UnstructuredGrid>
</VTKFile>
"""


def test_hand_written_ascii_file(tmp_path):
    p = tmp_path / "h.vtu"
    p.write_text(ASCII_JOB)
    v = vtkio.read_vtu(str(p))
    assert v.points.shape == (12, 3) and v.points[5].tolist() == [2, 1, 0]
    assert v.cells.tolist() == [[0, 1, 4, 3, 6, 7, 10, 9], [1, 2, 5, 4, 7, 8, 11, 10]]
    assert v.value("dt") == 2.5e-08
    f = v.field("density")  # PointData comes first, as the reference reads fields
    assert f.type == cases.FIELD_NODE and f.data.tolist() == list(range(1, 13))
    c = v.field("C11")
    assert c.type == cases.FIELD_CELL and c.data.tolist() == [1.5e11, 2.5e11]


def replace_array(xml, name_attr, new_element):
    start = xml.index(name_attr)
    start = xml.rindex("<DataArray", 0, start)
    end = xml.index("</DataArray>", start) + len("</DataArray>")
    return xml[:start] + new_element + xml[end:]


def test_big_endian_inline_binary(tmp_path):
    # inline binary = base64 of [UInt32 byte count][data], here big-endian as the attribute says
    vals = [100.0, 200.0]
    payload = struct.pack(">I", 16) + struct.pack(">2d", *vals)
    el = '<DataArray type="Float64" Name="density" format="binary">%s</DataArray>' % base64.b64encode(
        payload).decode()
    xml = replace_array(ASCII_JOB, 'Name="density" format="ascii">100', el)
    xml = xml.replace('byte_order="LittleEndian"', 'byte_order="BigEndian"')
    p = tmp_path / "be.vtu"
    p.write_text(xml)
    assert vtkio.read_vtu(str(p)).cell_data["density"].tolist() == vals


def test_zlib_block_header_from_the_format(tmp_path):
    # compressed inline: base64(header) + base64(blocks); header (UInt64 here) =
    # [number of blocks, uncompressed block size, size of the last block, compressed sizes]
    data = np.arange(10, dtype=np.float64) * 0.5
    raw = data.tobytes()
    blocks = [raw[:48], raw[48:]]
    comp = [zlib.compress(b) for b in blocks]
    head = struct.pack("<5Q", 2, 48, len(blocks[1]), len(comp[0]), len(comp[1]))
    text = (base64.b64encode(head) + base64.b64encode(b"".join(comp))).decode()
    xml = ASCII_JOB.replace('version="0.1" byte_order="LittleEndian"',
                            'version="1.0" byte_order="LittleEndian" header_type="UInt64" '
                            'compressor="vtkZLibDataCompressor"')
    xml = replace_array(xml, 'Name="dt"', '<DataArray type="Float64" Name="dt" format="binary">%s</DataArray>' % text)
    p = tmp_path / "z.vtu"
    p.write_text(xml)
    assert np.array_equal(vtkio.read_vtu(str(p)).field_data["dt"], data)


def test_voxel_cells_are_reordered(tmp_path):
    xml = ASCII_JOB.replace('<DataArray type="Int64" Name="connectivity" format="ascii">0 1 4 3 6 7 10 9',
                            '<DataArray type="Int64" Name="connectivity" format="ascii">0 1 3 4 6 7 9 10')
    xml = xml.replace(">12 12<", ">11 12<")
    p = tmp_path / "v.vtu"
    p.write_text(xml)
    assert vtkio.read_vtu(str(p)).cells[0].tolist() == [0, 1, 4, 3, 6, 7, 10, 9]


def test_other_cell_types_are_rejected(tmp_path):
    p = tmp_path / "t.vtu"
    p.write_text(ASCII_JOB.replace(">12 12<", ">12 10<"))
    with pytest.raises(vtkio.VtkError, match="hexahedral"):
        vtkio.read_vtu(str(p))


@pytest.mark.parametrize("fmt", ["appended", "binary"])
def test_job_round_trip(tmp_path, fmt):
    case = cases.al_block(seed=8, nx=9, ny=8, nz=5, nsteps=40, n_sensors=3)
    wt = np.arange(5) * 9.5 * case.dt
    p = tmp_path / "job.vtu"
    vtkio.case_to_job(case, str(p), wt, fmt=fmt, compress=(fmt == "binary"))
    got, wt2 = vtkio.job_case(str(p))
    assert np.array_equal(wt2, wt) and got.dt == case.dt
    assert np.array_equal(got.cells, case.cells) and np.array_equal(got.x, case.x) and np.array_equal(got.z, case.z)
    assert np.array_equal(got.rho.data, case.rho.data)
    for a, b in zip(got.cij, case.cij):
        assert np.array_equal(a.data, np.broadcast_to(b.data, a.data.shape))
    d0, d1 = got.force[0], case.force[0]
    assert tuple(d0.center) == tuple(d1.center) and d0.radius == d1.radius and d0.ds == d1.ds
    assert np.array_equal(d0.signal_t, d1.signal_t) and np.array_equal(d0.signal_v, d1.signal_v)
    # sensors: every node a transducer's window accepts (HostNormalXDZ::checkAt at nodes)
    acc = [orc.disc_check(d1.center, d1.radius, d1.ds, (case.x[i], case.y[i], case.z[i]))
           for i in range(case.n_nodes)]
    assert got.sensors.shape[0] > 0 and np.array_equal(got.sensors, np.nonzero(acc)[0])
    # the step count follows the driver's FP64 time accumulation (oracle restatement)
    assert got.nsteps == np.cumsum(orc.schedule(case.dt, wt))[-1]


def test_schedule_matches_the_oracle():
    rng = np.random.default_rng(4)
    for _ in range(20):
        dt = float(rng.uniform(1e-9, 1e-8))
        wt = np.cumsum(rng.uniform(0, 40, size=6)) * dt
        assert np.array_equal(vtkio.schedule_steps(dt, wt), np.cumsum(orc.schedule(dt, wt)))


def test_partitioned_files(tmp_path):
    """.pvtu/.pvtt summaries name their pieces as VTK's parallel writers do and read back whole;
    the table summary has the layout py_utils/PVTTReader.py parses (PTable/Piece Source)."""
    rng = np.random.default_rng(5)
    pts, cells = two_hexes()
    pdata = {"Vx": rng.standard_normal(pts.shape[0])}
    cdata = {"Sxx": rng.standard_normal(2), "Syy": rng.standard_normal(2)}
    p = tmp_path / "out_3.pvtu"
    vtkio.write_pvtu(str(p), pts, cells, pdata, cdata, {"time": np.array([1.5e-6])})
    assert vtkio.piece_name(str(p), 0) == str(tmp_path / "out_3_0.vtu") and (tmp_path / "out_3_0.vtu").exists()
    root = ET.parse(str(p)).getroot()
    assert root.get("type") == "PUnstructuredGrid"
    assert [e.get("Name") for e in root.find("PUnstructuredGrid").find("PCellData")] == ["Sxx", "Syy"]
    v = vtkio.read_pvtu(str(p))
    assert np.array_equal(v.points, pts) and np.array_equal(v.cells, cells)
    assert np.array_equal(v.point_data["Vx"], pdata["Vx"]) and np.array_equal(v.cell_data["Syy"], cdata["Syy"])
    assert v.field_data["time"][0] == 1.5e-6
    cols = {"Vx": rng.standard_normal(4), "Vy": rng.standard_normal(4), "Vz": rng.standard_normal(4)}
    t = tmp_path / "velocities_3.pvtt"
    vtkio.write_pvtt(str(t), cols)
    pieces = [e.attrib["Source"] for e in ET.parse(str(t)).findall("PTable/Piece")]  # PVTTReader's lookup
    assert pieces == ["velocities_3_0.vtt"]
    assert all(np.array_equal(vtkio.read_vtt(str(tmp_path / pieces[0]))[k], v_) for k, v_ in cols.items())
    back = vtkio.read_pvtt(str(t))
    assert all(np.array_equal(back[k], v_) for k, v_ in cols.items())
    with pytest.raises(vtkio.VtkError, match="pvtt"):
        vtkio.write_pvtt(str(tmp_path / "x.vtt"), cols)


def test_pvtu_pieces_append_in_order(tmp_path):
    """Two pieces written by hand: points appended, the second piece's cells renumbered."""
    pts, cells = two_hexes()
    a, b = tmp_path / "g_0.vtu", tmp_path / "g_1.vtu"
    vtkio.write_vtu(str(a), pts, cells, {"Vx": np.arange(pts.shape[0], dtype=np.float64)}, fmt="ascii")
    vtkio.write_vtu(str(b), pts + 10.0, cells[:1], {"Vx": -np.arange(pts.shape[0], dtype=np.float64)}, fmt="binary")
    (tmp_path / "g.pvtu").write_text(
        '<?xml version="1.0"?><VTKFile type="PUnstructuredGrid" version="1.0"><PUnstructuredGrid GhostLevel="0">'
        '<PPointData><PDataArray type="Float64" Name="Vx"/></PPointData>'
        '<PPoints><PDataArray type="Float64" NumberOfComponents="3"/></PPoints>'
        '<Piece Source="g_0.vtu"/><Piece Source="g_1.vtu"/></PUnstructuredGrid></VTKFile>')
    v = vtkio.read_pvtu(str(tmp_path / "g.pvtu"))
    n = pts.shape[0]
    assert v.points.shape == (2 * n, 3) and np.array_equal(v.points[n:], pts + 10.0)
    assert np.array_equal(v.cells, np.concatenate([cells, cells[:1] + n]))
    assert np.array_equal(v.point_data["Vx"], np.concatenate([np.arange(n), -np.arange(n)]).astype(np.float64))


def two_column_case(seed=12, nsteps=30):
    """A water-column block with a second column of its own centre, radius and signals."""
    case = cases.water_column_block(seed=seed, nx=10, ny=9, nz=5, nsteps=nsteps)
    d0 = case.moment[0]
    t = d0.signal_t
    c1 = (0.3 * 10 * 1.0e-4, 0.35 * 9 * 1.0e-4, d0.center[2])
    case.moment = list(case.moment) + [cases.Disc(k, c1, 0.6 * d0.radius, d0.ds, t, s * d0.signal_v[::-1])
                                       for k, s in enumerate((0.7, -1.3, 0.4))]
    return case


def test_water_column_job_round_trip(tmp_path):
    case = two_column_case()
    wt = np.arange(4) * 8.5 * case.dt
    p = tmp_path / "wc.vtu"
    vtkio.case_to_water_column_job(case, str(p), wt)
    got, wt2 = vtkio.water_column_case(str(p))
    assert np.array_equal(wt2, wt) and got.dt == case.dt and got.moment_form == 1 and not got.force
    assert np.array_equal(got.cells, case.cells) and np.array_equal(got.z, case.z)
    assert len(got.moment) == len(case.moment) == 6
    for a, b in zip(got.moment, case.moment):
        assert a.component == b.component and tuple(a.center) == tuple(b.center) and a.radius == b.radius
        assert a.ds == b.ds and np.array_equal(a.signal_t, b.signal_t) and np.array_equal(a.signal_v, b.signal_v)
    assert got.sensors.shape[0] == 0
    bad = cases.water_column_block(seed=3)
    bad.moment = bad.moment[:2]
    with pytest.raises(vtkio.VtkError, match="triples"):
        vtkio.case_to_water_column_job(bad, str(tmp_path / "bad.vtu"), wt)


@pytest.mark.skipif(not orc.have_reference(), reason="oracle/_ref/ref_rsg not built (reference tree absent)")
def test_two_water_columns_oracle_matches_reference():
    """Moment terms of two columns on the same components are summed column by column for
    each cell vertex, as HostUTModel::evalMoments does: the oracle equals the reference."""
    case = two_column_case(nsteps=20)
    V, S, t, tr = orc.run_oracle(case)
    V2, S2, t2, tr2, _ = orc.run_reference(case, ranks=2)
    assert np.abs(S).max() > 0
    assert t == t2 and np.array_equal(V, V2) and np.array_equal(S, S2)


@pytest.mark.parametrize("kind", ["plate", "bracket", "block", "water-column"])
def test_case_maker_jobs(tmp_path, kind):
    """Job files from the case generators read back as the generator's case, with the case
    makers' write-time recipe."""
    from pannde_amd import case_maker

    p = tmp_path / ("%s.vtu" % kind)
    argv = [kind, "-o", str(p), "--seed", "7", "--nsteps", "37", "--write-interval", "10"]
    argv += {"plate": ["--nx", "24", "--ny", "20", "--cells-per-ply", "1", "--holes", "1", "--crack-length", "4e-4",
                       "--crack-depth", "3e-4", "--crack-angle", "30"],
             "bracket": ["--leg", "24", "--width", "8", "--height", "6", "--crack-length", "3e-4",
                         "--crack-depth", "1e-4", "--crack-angle", "60"],
             "block": ["--nx", "7", "--ny", "6", "--nz", "5"],
             "water-column": ["--nx", "7", "--ny", "6", "--nz", "5"]}[kind]
    case_maker.main(argv)
    read = vtkio.water_column_case if kind == "water-column" else vtkio.job_case
    got, wt = read(str(p))
    dt = got.dt
    assert np.array_equal(wt, np.arange(int(np.arange(0, 37 * dt, dt).shape[0] / 10) + 1) * 10 * dt)
    params = dict(nx=24 if kind == "plate" else 7, ny=20 if kind == "plate" else 6, nz=5,
                  dx=1.0e-4 if kind in ("block", "water-column") else 5.0e-5, cells_per_ply=1, holes=1,
                  renumber=False, leg=24, width=8, height=6)
    if kind == "plate":
        params.update(crack_length=4e-4, crack_depth=3e-4, crack_angle=30.0)
    if kind == "bracket":
        params.update(crack_length=3e-4, crack_depth=1e-4, crack_angle=60.0)
    ref = case_maker.make_case(kind, 7, 37, **params)
    assert got.dt == ref.dt and np.array_equal(got.cells, ref.cells) and np.array_equal(got.x, ref.x)
    for a, b in zip(got.cij, ref.cij):
        assert np.array_equal(a.data, np.broadcast_to(b.data, a.data.shape))
    src = got.moment if kind == "water-column" else got.force
    assert len(src) == len(ref.moment if kind == "water-column" else ref.force) > 0
    if kind in ("plate", "bracket"):
        assert ref.meta["n_dup"] > 0  # the crack opened
