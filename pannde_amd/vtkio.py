"""VTK XML I/O for RSG jobs and results, without VTK.

Reads the unstructured-grid job files the reference's drivers consume
(modules/VTKIO/include/internal/VTKRead.hpp:115-297, through VTK's XML reader) and writes the
solution and table files they produce (VTKWrite.hpp:200-232: ``<base>_<k>.vtu`` grids,
``<name>.vtt`` / ``<name>_<k>.vtt`` tables; on several processes the partitioned ``.pvtu`` /
``.pvtt`` summaries with ``<base>_<piece>`` pieces, VTKPWrite.hpp:100-137).  Supported, as VTK's
XML format defines them:

* ``ascii``, inline ``binary`` (base64) and ``appended`` data (raw or base64);
* ``header_type`` UInt32 or UInt64, ``byte_order`` LittleEndian or BigEndian;
* ``compressor="vtkZLibDataCompressor"`` (block header: number of blocks, block size, size of a
  partial last block or 0 when the last block is full, compressed sizes);
* every numeric DataArray type; cells of type 12 (hexahedron) and 11 (voxel, reordered to the
  hexahedron vertex order).

Job semantics follow the reference: a named value is a one-tuple FieldData array
(VTKRead.hpp:115-120), arrays come from FieldData (:216-225), a field is looked up in PointData
first, then CellData (:151-157); transducers use the ``XD<k>/...`` names and timing the ``dt`` /
``write_times`` names of StandardNames.hpp:186-244.
"""
from __future__ import annotations

import base64
import os
import re
import xml.etree.ElementTree as ET
import zlib
from typing import Dict, Optional

import numpy as np

from . import cases

_TYPES = {"Int8": "i1", "UInt8": "u1", "Int16": "i2", "UInt16": "u2", "Int32": "i4", "UInt32": "u4",
          "Int64": "i8", "UInt64": "u8", "Float32": "f4", "Float64": "f8"}
_NAMES = {np.dtype(v).newbyteorder("<").str: k for k, v in _TYPES.items()}
VTK_VOXEL, VTK_HEXAHEDRON = 11, 12
_VOXEL_TO_HEX = [0, 1, 3, 2, 4, 5, 7, 6]


class VtkError(ValueError):
    pass


def _b64len(nbytes: int) -> int:
    return 4 * ((nbytes + 2) // 3)


class _Decoder:
    """Decodes DataArray payloads of one file (its header type, byte order, compressor)."""

    def __init__(self, root, appended: Optional[bytes], appended_encoding: str):
        self.order = "<" if root.get("byte_order", "LittleEndian") == "LittleEndian" else ">"
        ht = root.get("header_type", "UInt32")
        if ht not in ("UInt32", "UInt64"):
            raise VtkError("header_type %s" % ht)
        self.hdt = np.dtype(self.order + _TYPES[ht])
        comp = root.get("compressor", "")
        if comp not in ("", "vtkZLibDataCompressor"):
            raise VtkError("compressor %s is not supported" % comp)
        self.zlib = comp == "vtkZLibDataCompressor"
        self.appended = appended
        self.appended_b64 = appended_encoding == "base64"

    def _payload(self, data: bytes, b64: bool) -> bytes:
        """Binary payload (header + data) starting at data[0]; data is base64 text when b64."""
        hb = self.hdt.itemsize
        if not self.zlib:
            if not b64:
                n = int(np.frombuffer(data[:hb], self.hdt)[0])
                return data[hb:hb + n]
            # one base64 stream (header and data), or the header encoded on its own
            text = data.strip()
            raw = base64.b64decode(text[:_b64len(hb)] if len(text) >= _b64len(hb) else text)
            n = int(np.frombuffer(raw[:hb], self.hdt)[0])
            try:
                whole = base64.b64decode(text[:_b64len(hb + n)])
                if len(whole) >= hb + n and int(np.frombuffer(whole[:hb], self.hdt)[0]) == n:
                    return whole[hb:hb + n]
            except ValueError:
                pass
            return base64.b64decode(text[_b64len(hb):])[:n]
        # compressed: [nblocks, block size, last block size, compressed sizes...], then blocks
        if b64:
            text = data.strip()
            head = base64.b64decode(text[:_b64len(3 * hb)])
            nb = int(np.frombuffer(head[:hb], self.hdt)[0])
            head = base64.b64decode(text[:_b64len((3 + nb) * hb)])
            sizes = np.frombuffer(head[:(3 + nb) * hb], self.hdt).astype(np.int64)
            body = base64.b64decode(text[_b64len((3 + nb) * hb):])
        else:
            nb = int(np.frombuffer(data[:hb], self.hdt)[0])
            sizes = np.frombuffer(data[:(3 + nb) * hb], self.hdt).astype(np.int64)
            body = data[(3 + nb) * hb:]
        out, pos = [], 0
        for c in sizes[3:3 + nb]:
            out.append(zlib.decompress(body[pos:pos + int(c)]))
            pos += int(c)
        blob = b"".join(out)
        # the third header word is the size of a partial last block, 0 when the last block is full
        last = int(sizes[2]) or int(sizes[1])
        expect = (nb - 1) * int(sizes[1]) + last if nb else 0
        if len(blob) != expect:
            raise VtkError("compressed block sizes do not match the header")
        return blob

    def array(self, el) -> np.ndarray:
        t = el.get("type")
        if t not in _TYPES:
            raise VtkError("DataArray type %s" % t)
        dt = np.dtype(self.order + _TYPES[t])
        fmt = el.get("format", "ascii")
        if fmt == "ascii":
            txt = (el.text or "").split()
            a = np.array(txt, dtype=np.float64 if dt.kind == "f" else np.int64).astype(dt.newbyteorder("="))
        elif fmt == "binary":
            a = np.frombuffer(self._payload((el.text or "").encode(), True), dt)
        elif fmt == "appended":
            if self.appended is None:
                raise VtkError("appended DataArray without an AppendedData section")
            a = np.frombuffer(self._payload(self.appended[int(el.get("offset", "0")):], self.appended_b64), dt)
        else:
            raise VtkError("DataArray format %s" % fmt)
        nc = int(el.get("NumberOfComponents", "1"))
        a = a.astype(dt.newbyteorder("="))
        return a.reshape(-1, nc) if nc > 1 else a


class VtuFile:
    """Contents of one .vtu file: points [n,3], hexahedra [nc,8], and the named arrays of
    PointData, CellData and FieldData (1-D for one component, [tuples, components] otherwise)."""

    def __init__(self, points, cells, point_data, cell_data, field_data):
        self.points = points
        self.cells = cells
        self.point_data = point_data
        self.cell_data = cell_data
        self.field_data = field_data

    def value(self, name: str) -> float:
        """A one-tuple FieldData array (VTKRead::getValue)."""
        a = self.array(name)
        if a.size != 1:
            raise VtkError("%s: not a scalar quantity" % name)
        return float(a[0])

    def array(self, name: str) -> np.ndarray:
        """A FieldData array as FP64 (VTKRead::getArray)."""
        if name not in self.field_data:
            raise VtkError("no FieldData array named %s" % name)
        return np.asarray(self.field_data[name], dtype=np.float64).reshape(-1)

    def field(self, name: str) -> cases.Field:
        """A node or cell field, PointData first (VTKRead::getField)."""
        if name in self.point_data:
            return cases.Field(cases.FIELD_NODE, np.asarray(self.point_data[name], np.float64).reshape(-1))
        if name in self.cell_data:
            return cases.Field(cases.FIELD_CELL, np.asarray(self.cell_data[name], np.float64).reshape(-1))
        raise VtkError("no field named %s" % name)


def read_vtu(path: str) -> VtuFile:
    with open(path, "rb") as f:
        raw = f.read()
    appended, enc = None, "raw"
    k = raw.find(b"<AppendedData")
    if k >= 0:
        close = raw.find(b">", k)
        m = re.search(rb'encoding="(\w+)"', raw[k:close])
        enc = m.group(1).decode() if m else "raw"
        start = raw.find(b"_", close) + 1
        end = raw.rfind(b"</AppendedData>")
        appended = raw[start:end]
        xml = raw[:k] + b"</VTKFile>"
    else:
        xml = raw
    root = ET.fromstring(xml)
    if root.tag != "VTKFile" or root.get("type") != "UnstructuredGrid":
        raise VtkError("not a VTK UnstructuredGrid file")
    dec = _Decoder(root, appended, enc)
    grid = root.find("UnstructuredGrid")
    piece = grid.find("Piece")
    if len(grid.findall("Piece")) != 1:
        raise VtkError("exactly one Piece is supported")

    def arrays(node):
        out = {}
        if node is not None:
            for el in node.findall("DataArray"):
                out[el.get("Name")] = dec.array(el)
        return out

    field_data = arrays(grid.find("FieldData"))
    pts = piece.find("Points").find("DataArray")
    points = np.asarray(dec.array(pts), np.float64).reshape(-1, 3)
    cel = {el.get("Name"): el for el in piece.find("Cells").findall("DataArray")}
    conn = np.asarray(dec.array(cel["connectivity"]), np.int64).reshape(-1)
    offs = np.asarray(dec.array(cel["offsets"]), np.int64).reshape(-1)
    types = np.asarray(dec.array(cel["types"]), np.int64).reshape(-1)
    n_cells = int(piece.get("NumberOfCells"))
    if offs.shape[0] != n_cells or types.shape[0] != n_cells:
        raise VtkError("cell arrays do not match NumberOfCells")
    if n_cells and (np.any(np.diff(np.concatenate([[0], offs])) != 8) or not np.all(np.isin(types, (11, 12)))):
        raise VtkError("only hexahedral (type 12) or voxel (type 11) cells are supported")
    cells = conn.reshape(n_cells, 8).astype(np.int32)
    vox = types == VTK_VOXEL
    if vox.any():
        cells[vox] = cells[vox][:, _VOXEL_TO_HEX]
    if points.shape[0] != int(piece.get("NumberOfPoints")):
        raise VtkError("Points do not match NumberOfPoints")
    return VtuFile(points, cells, arrays(piece.find("PointData")), arrays(piece.find("CellData")), field_data)


def _encode(a: np.ndarray, fmt: str, header: str, compress: bool, block: int = 1 << 15):
    """(attributes, inline text or appended bytes) of one array."""
    a = np.ascontiguousarray(a)
    le = a.astype(a.dtype.newbyteorder("<"))
    if fmt == "ascii":
        if a.dtype.kind == "f":
            text = " ".join(repr(float(v)) for v in a.reshape(-1))
        else:
            text = " ".join(str(int(v)) for v in a.reshape(-1))
        return text
    hdt = np.dtype("<u4" if header == "UInt32" else "<u8")
    data = le.tobytes()
    if compress:
        chunks = [data[i:i + block] for i in range(0, len(data), block)] or [b""]
        comp = [zlib.compress(c) for c in chunks]
        last = len(chunks[-1]) if data else 0
        head = np.array([len(chunks), block, last] + [len(c) for c in comp], hdt).tobytes()
        body = b"".join(comp)
        if fmt == "binary":
            return (base64.b64encode(head) + base64.b64encode(body)).decode()
        return head + body
    head = np.array([len(data)], hdt).tobytes()
    if fmt == "binary":
        return base64.b64encode(head + data).decode()
    return head + data


def write_vtu(path: str, points, cells, point_data=None, cell_data=None, field_data=None, fmt="appended",
              header="UInt64", compress=False):
    """Writes one UnstructuredGrid piece of hexahedra.  fmt: ascii, binary or appended (raw)."""
    if fmt not in ("ascii", "binary", "appended"):
        raise VtkError("format %s" % fmt)
    points = np.asarray(points, np.float64).reshape(-1, 3)
    cells = np.asarray(cells, np.int64).reshape(-1, 8)
    nc = cells.shape[0]
    blobs = []
    offset = [0]

    def da(name, a, extra=""):
        a = np.asarray(a)
        if a.dtype == np.bool_:
            a = a.astype(np.uint8)
        t = _NAMES[a.dtype.newbyteorder("<").str]
        ncomp = a.shape[1] if a.ndim == 2 else 1
        attrs = 'type="%s" Name="%s" NumberOfComponents="%d" format="%s"%s' % (t, name, ncomp, fmt, extra)
        enc = _encode(a, fmt, header, compress)
        if fmt == "appended":
            s = '<DataArray %s offset="%d"/>' % (attrs, offset[0])
            blobs.append(enc)
            offset[0] += len(enc)
            return s
        return "<DataArray %s>%s</DataArray>" % (attrs, enc)

    def section(tag, d):
        if not d:
            return ""
        return "<%s>%s</%s>" % (tag, "".join(da(k, v) for k, v in d.items()), tag)

    fd = ""
    if field_data:
        fd = "<FieldData>%s</FieldData>" % "".join(
            da(k, np.atleast_1d(np.asarray(v)), ' NumberOfTuples="%d"' % np.atleast_1d(np.asarray(v)).shape[0])
            for k, v in field_data.items())
    body = ('<UnstructuredGrid>%s<Piece NumberOfPoints="%d" NumberOfCells="%d">' % (fd, points.shape[0], nc)
            + section("PointData", point_data) + section("CellData", cell_data)
            + "<Points>%s</Points>" % da("Points", points)
            + "<Cells>%s%s%s</Cells>" % (da("connectivity", cells.reshape(-1)),
                                         da("offsets", np.arange(8, 8 * nc + 1, 8, dtype=np.int64)),
                                         da("types", np.full(nc, VTK_HEXAHEDRON, np.uint8)))
            + "</Piece></UnstructuredGrid>")
    attrs = 'type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="%s"' % header
    if compress and fmt != "ascii":
        attrs += ' compressor="vtkZLibDataCompressor"'
    with open(path, "wb") as f:
        f.write(('<?xml version="1.0"?>\n<VTKFile %s>\n' % attrs).encode())
        f.write(body.encode())
        if fmt == "appended":
            f.write(b'\n<AppendedData encoding="raw">\n_')
            for b in blobs:
                f.write(b)
            f.write(b"\n</AppendedData>")
        f.write(b"\n</VTKFile>\n")


def write_vtt(path: str, columns: Dict[str, np.ndarray]):
    """A vtkTable file (RowData columns), ascii, as the reference's table writer produces
    for the sensor coordinates and velocities (TransducerModel.cpp:58,78)."""
    n = len(next(iter(columns.values()))) if columns else 0
    rows = "".join('<DataArray type="Float64" Name="%s" NumberOfComponents="1" format="ascii">%s</DataArray>'
                   % (k, " ".join(repr(float(x)) for x in np.asarray(v, np.float64))) for k, v in columns.items())
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="Table" version="1.0" byte_order="LittleEndian">\n'
                '<Table><Piece NumberOfCols="%d" NumberOfRows="%d"><RowData>%s</RowData></Piece></Table>\n'
                "</VTKFile>\n" % (len(columns), n, rows))


def read_vtt(path: str) -> Dict[str, np.ndarray]:
    root = ET.parse(path).getroot()
    dec = _Decoder(root, None, "raw")
    rd = root.find("Table").find("Piece").find("RowData")
    return {el.get("Name"): dec.array(el) for el in rd.findall("DataArray")}


# ---------------------------------------------------------------- partitioned (parallel) files

def piece_name(path: str, piece: int) -> str:
    """The file name VTK's parallel XML writers give piece ``piece`` of ``path``:
    ``<base>_<piece>.<serial extension>`` beside the summary file."""
    base, ext = os.path.splitext(path)
    return "%s_%d.%s" % (base, piece, {".pvtu": "vtu", ".pvtt": "vtt"}[ext])


def _p_arrays(tag: str, d: Dict[str, np.ndarray]) -> str:
    if not d:
        return ""
    out = []
    for k, v in d.items():
        a = np.asarray(v)
        ncomp = a.shape[1] if a.ndim == 2 else 1
        out.append('<PDataArray type="%s" Name="%s" NumberOfComponents="%d"/>'
                   % (_NAMES[a.dtype.newbyteorder("<").str], k, ncomp))
    return "<%s>%s</%s>" % (tag, "".join(out), tag)


def write_pvtu(path: str, points, cells, point_data=None, cell_data=None, field_data=None, fmt="appended",
               header="UInt64", compress=False):
    """A one-piece partitioned grid: the ``.pvtu`` summary (the arrays' names and types and
    the piece's file) and the piece ``<base>_0.vtu``, as the reference writes solutions when
    it runs on several processes (VTKPWrite.hpp:100-137 through vtkXMLPUnstructuredGridWriter)."""
    if not path.endswith(".pvtu"):
        raise VtkError("%s: a partitioned grid file ends in .pvtu" % path)
    piece = piece_name(path, 0)
    write_vtu(piece, points, cells, point_data, cell_data, field_data, fmt=fmt, header=header, compress=compress)
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="PUnstructuredGrid" version="1.0" byte_order="LittleEndian" '
                'header_type="%s">\n<PUnstructuredGrid GhostLevel="0">%s%s%s'
                '<PPoints><PDataArray type="Float64" Name="Points" NumberOfComponents="3"/></PPoints>'
                '<Piece Source="%s"/></PUnstructuredGrid>\n</VTKFile>\n'
                % (header, _p_arrays("PFieldData", field_data), _p_arrays("PPointData", point_data),
                   _p_arrays("PCellData", cell_data), os.path.basename(piece)))


def write_pvtt(path: str, columns: Dict[str, np.ndarray]):
    """A one-piece partitioned table: ``.pvtt`` summary plus the piece ``<base>_0.vtt``, the
    layout the reference's post-processing reads (py_utils/PVTTReader.py: ``PTable/Piece``
    ``Source`` attributes, each piece a vtkTable)."""
    if not path.endswith(".pvtt"):
        raise VtkError("%s: a partitioned table file ends in .pvtt" % path)
    piece = piece_name(path, 0)
    write_vtt(piece, columns)
    cols = {k: np.asarray(v, np.float64) for k, v in columns.items()}
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n<VTKFile type="PTable" version="1.0" byte_order="LittleEndian">\n'
                '<PTable>%s<Piece Source="%s"/></PTable>\n</VTKFile>\n'
                % (_p_arrays("PRowData", cols), os.path.basename(piece)))


def _pieces(path: str, kind: str):
    root = ET.parse(path).getroot()
    if root.tag != "VTKFile" or root.get("type") != kind:
        raise VtkError("%s: not a VTK %s file" % (path, kind))
    here = os.path.dirname(os.path.abspath(path))
    return [os.path.join(here, el.get("Source")) for el in root.find(kind).findall("Piece")]


def read_pvtt(path: str) -> Dict[str, np.ndarray]:
    """The rows of every piece of a ``.pvtt`` table, in piece order (as PVTTReader
    concatenates them); piece paths are taken relative to the summary file."""
    parts = [read_vtt(p) for p in _pieces(path, "PTable")]
    if not parts:
        return {}
    return {k: np.concatenate([np.asarray(t[k]).reshape(-1) for t in parts]) for k in parts[0]}


def read_pvtu(path: str) -> VtuFile:
    """The pieces of a ``.pvtu`` grid appended in piece order (points of later pieces after
    those of earlier ones, their cells renumbered accordingly; points shared by two pieces
    appear once per piece, as VTK's reader returns them); FieldData from the first piece."""
    parts = [read_vtu(p) for p in _pieces(path, "PUnstructuredGrid")]
    if not parts:
        raise VtkError("%s has no pieces" % path)
    base = np.cumsum([0] + [q.points.shape[0] for q in parts[:-1]])
    cells = np.concatenate([q.cells.astype(np.int64) + b for q, b in zip(parts, base)]).astype(np.int32)

    def cat(key):
        return {k: np.concatenate([getattr(q, key)[k] for q in parts]) for k in getattr(parts[0], key)}

    return VtuFile(np.concatenate([q.points for q in parts]), cells, cat("point_data"), cat("cell_data"),
                   parts[0].field_data)


# ---------------------------------------------------------------- jobs

ELASTIC = ["C%d%d" % (i + 1, j + 1) for i in range(6) for j in range(i, 6)]


def _job_domain(v: VtuFile):
    """Mesh, density, C11..C66 and the window depth (z-size of cell 0, Boss.hpp:530-536 and
    :639-645) of a job file."""
    x, y, z = (np.ascontiguousarray(v.points[:, d]) for d in range(3))
    cells = v.cells
    if cells.shape[0] == 0:
        raise VtkError("the job has no cells")
    ds = abs(z[cells[0, 0]] - z[cells[0, 6]])
    return x, y, z, cells, v.field("density"), [v.field(n) for n in ELASTIC], ds


def job_case(path: str, name: Optional[str] = None) -> "tuple[cases.Case, np.ndarray]":
    """A TransducerModel job (TransducerModel.cpp:160-204): mesh, density and C11..C66
    fields, transducers XD<k>/{XCenter,YCenter,ZCenter,Radius,SignalTimes,SignalValues} driving
    the z force with the z-size of cell 0 as window depth (Boss.hpp:502-536), dt and
    write_times.  Sensors are the nodes any transducer's window accepts
    (TransducerModel.cpp:106-130).  Returns (case, write_times)."""
    v = read_vtu(path)
    x, y, z, cells, rho, cij, ds = _job_domain(v)
    discs = []
    nxd = int(v.value("NTransducers"))
    for k in range(nxd):
        p = "XD%d/" % k
        center = (v.value(p + "XCenter"), v.value(p + "YCenter"), v.value(p + "ZCenter"))
        discs.append(cases.Disc(2, center, v.value(p + "Radius"), ds, v.array(p + "SignalTimes"),
                                v.array(p + "SignalValues")))
    dt = v.value("dt")
    write_times = v.array("write_times")
    sens = np.zeros(x.shape[0], dtype=bool)
    for d in discs:  # HostNormalXDZ::checkAt at every node
        dist = np.sqrt((x - d.center[0]) ** 2 + (y - d.center[1]) ** 2 + (z - d.center[2]) ** 2)
        sens |= (np.abs(z - d.center[2]) <= 0.75 * d.ds) & (dist < d.radius) & (dist >= 0)
    steps = schedule_steps(dt, write_times)
    case = cases.Case(name=name or os.path.splitext(os.path.basename(path))[0], x=x, y=y, z=z, cells=cells,
                      rho=rho, cij=cij, dt=dt, nsteps=int(steps[-1]) if len(steps) else 0, force=discs,
                      sensors=np.nonzero(sens)[0].astype(np.int32))
    return case, write_times


WATER_COLUMN_SIGNALS = ("SignalSigmaXX", "SignalSigmaYY", "SignalSigmaZZ")


def water_column_case(path: str, name: Optional[str] = None) -> "tuple[cases.Case, np.ndarray]":
    """A WaterColumnModel job (WaterColumnModel.cpp:85-180): mesh, density and C11..C66,
    water columns WC<k>/{XCenter,YCenter,ZCenter,Radius,SignalTimes,SignalSigmaXX,
    SignalSigmaYY,SignalSigmaZZ} (StandardNames.hpp:262-290), each an XX, YY, ZZ moment
    excitation of the HostUTModel form with the z-size of cell 0 as window depth
    (Boss.hpp:555-650), dt and write_times.  The discs are listed column by column, XX, YY,
    ZZ within a column, which is the order the reference sums them in for each stress
    component (HostUTModel.hpp:296-311).  No sensors: the reference records none.  Returns
    (case, write_times)."""
    v = read_vtu(path)
    x, y, z, cells, rho, cij, ds = _job_domain(v)
    discs = []
    for k in range(int(v.value("NWaterColumns"))):
        p = "WC%d/" % k
        center = (v.value(p + "XCenter"), v.value(p + "YCenter"), v.value(p + "ZCenter"))
        t = v.array(p + "SignalTimes")
        for comp, sig in enumerate(WATER_COLUMN_SIGNALS):
            discs.append(cases.Disc(comp, center, v.value(p + "Radius"), ds, t, v.array(p + sig)))
    dt = v.value("dt")
    write_times = v.array("write_times")
    steps = schedule_steps(dt, write_times)
    case = cases.Case(name=name or os.path.splitext(os.path.basename(path))[0], x=x, y=y, z=z, cells=cells,
                      rho=rho, cij=cij, dt=dt, nsteps=int(steps[-1]) if len(steps) else 0, force=[],
                      sensors=np.zeros(0, np.int32))
    case.moment, case.moment_form = discs, 1
    return case, write_times


def schedule_steps(dt: float, write_times) -> np.ndarray:
    """Cumulative step counts at each write time: ``while(time < t_k) time = solve()`` with
    time accumulated by FP64 addition (TransducerModel.cpp:63-83)."""
    out, t, n = [], 0.0, 0
    for tw in np.asarray(write_times, np.float64):
        while t < tw:
            t += dt
            n += 1
        out.append(n)
    return np.array(out, dtype=np.int64)


def _domain_arrays(case: cases.Case):
    pdata, cdata = {}, {}
    fields = [("density", case.rho)] + list(zip(ELASTIC, case.cij))
    for name, f in fields:
        if f.type == cases.FIELD_NODE:
            pdata[name] = f.data
        elif f.type == cases.FIELD_CELL:
            cdata[name] = f.data
        else:
            cdata[name] = np.full(case.n_cells, float(f.data[0]))
    return pdata, cdata


def case_to_job(case: cases.Case, path: str, write_times, fmt="appended", header="UInt64", compress=False):
    """Writes a Case with z-force transducers as a TransducerModel job file."""
    pdata, cdata = _domain_arrays(case)
    fd = {"NTransducers": np.array([float(len(case.force))]), "dt": np.array([case.dt]),
          "write_times": np.asarray(write_times, np.float64)}
    for k, d in enumerate(case.force):
        if d.component != 2:
            raise VtkError("job files carry z-force transducers only")
        p = "XD%d/" % k
        fd[p + "XCenter"], fd[p + "YCenter"], fd[p + "ZCenter"] = (np.array([float(c)]) for c in d.center)
        fd[p + "Radius"] = np.array([float(d.radius)])
        fd[p + "SignalTimes"] = np.asarray(d.signal_t, np.float64)
        fd[p + "SignalValues"] = np.asarray(d.signal_v, np.float64)
    write_vtu(path, np.stack([case.x, case.y, case.z], 1), case.cells, pdata, cdata, fd, fmt=fmt, header=header,
              compress=compress)


def case_to_water_column_job(case: cases.Case, path: str, write_times, fmt="appended", header="UInt64",
                             compress=False):
    """Writes a Case of the HostUTModel form whose moment discs come in XX, YY, ZZ triples (one
    per water column, sharing centre, radius and signal times) as a WaterColumnModel job."""
    m = list(case.moment)
    if case.moment_form != 1 or case.force or len(m) % 3:
        raise VtkError("a water-column job needs XX, YY, ZZ moment triples of the HostUTModel form")
    pdata, cdata = _domain_arrays(case)
    fd = {"NWaterColumns": np.array([float(len(m) // 3)]), "dt": np.array([case.dt]),
          "write_times": np.asarray(write_times, np.float64)}
    for k in range(len(m) // 3):
        col = m[3 * k:3 * k + 3]
        if [d.component for d in col] != [0, 1, 2] or any(
                tuple(d.center) != tuple(col[0].center) or d.radius != col[0].radius
                or not np.array_equal(d.signal_t, col[0].signal_t) for d in col):
            raise VtkError("water column %d: discs must be XX, YY, ZZ with one centre, radius and time table" % k)
        p = "WC%d/" % k
        fd[p + "XCenter"], fd[p + "YCenter"], fd[p + "ZCenter"] = (np.array([float(c)]) for c in col[0].center)
        fd[p + "Radius"] = np.array([float(col[0].radius)])
        fd[p + "SignalTimes"] = np.asarray(col[0].signal_t, np.float64)
        for d, sig in zip(col, WATER_COLUMN_SIGNALS):
            fd[p + sig] = np.asarray(d.signal_v, np.float64)
    write_vtu(path, np.stack([case.x, case.y, case.z], 1), case.cells, pdata, cdata, fd, fmt=fmt, header=header,
              compress=compress)
