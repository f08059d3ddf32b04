// vtk_job.hpp -- VTK XML job files for the C++ drivers, without VTK.
//
// Reads the UnstructuredGrid job files the reference's drivers consume
// (modules/VTKIO/include/internal/VTKRead.hpp:115-297, through VTK's XML reader) and writes
// the solution and table files they produce (VTKWrite.hpp:200-232; on several processes the
// partitioned summaries with one piece each, VTKPWrite.hpp:100-137).  Read as VTK's XML format
// defines it:
//   * ascii, inline binary (base64) and appended data (raw or base64);
//   * header_type UInt32 or UInt64, byte_order LittleEndian or BigEndian;
//   * compressor vtkZLibDataCompressor (header: number of blocks, block size, size of a partial
//     last block or 0 when the last block is full, compressed sizes; then the blocks);
//   * every numeric DataArray type; hexahedra (type 12) and voxels (type 11, reordered to the
//     hexahedron vertex order).
// Job semantics: a named value is a one-tuple FieldData array (VTKRead.hpp:115-120), arrays come
// from FieldData (:216-225), a field is looked up in PointData first, then CellData (:151-157).
#pragma once

#include <zlib.h>

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace vtkjob {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ------------------------------------------------------------------------------ XML

struct Element {
  std::string tag, text;
  std::vector<std::pair<std::string, std::string>> attr;
  std::vector<Element> kids;

  const std::string* find_attr(const std::string& k) const {
    for (auto& a : attr)
      if (a.first == k) return &a.second;
    return nullptr;
  }
  std::string get(const std::string& k, const std::string& dflt) const {
    const std::string* v = find_attr(k);
    return v ? *v : dflt;
  }
  const Element* child(const std::string& t) const {
    for (auto& e : kids)
      if (e.tag == t) return &e;
    return nullptr;
  }
  std::vector<const Element*> children(const std::string& t) const {
    std::vector<const Element*> out;
    for (auto& e : kids)
      if (e.tag == t) out.push_back(&e);
    return out;
  }
};

// A small non-validating XML reader: elements, attributes, text, comments, CDATA, processing
// instructions and the five predefined entities (plus numeric character references).
class XmlReader {
 public:
  explicit XmlReader(const std::string& s) : s_(s) {}
  Element document() {
    skip_misc();
    Element root = element(0);
    return root;
  }

 private:
  bool at(const char* p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
  void skip_ws() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) i_++;
  }
  void skip_past(const char* end) {
    const size_t k = s_.find(end, i_);
    if (k == std::string::npos) throw Error("XML: unterminated construct");
    i_ = k + std::strlen(end);
  }
  void skip_misc() {
    for (;;) {
      skip_ws();
      if (at("<?"))
        skip_past("?>");
      else if (at("<!--"))
        skip_past("-->");
      else if (at("<!"))
        skip_past(">");
      else
        return;
    }
  }
  std::string name() {
    const size_t b = i_;
    while (i_ < s_.size() && !std::isspace((unsigned char)s_[i_]) && s_[i_] != '>' && s_[i_] != '/' &&
           s_[i_] != '=')
      i_++;
    if (b == i_) throw Error("XML: malformed name");
    return s_.substr(b, i_ - b);
  }
  static std::string unescape(const std::string& in) {
    if (in.find('&') == std::string::npos) return in;
    std::string out;
    out.reserve(in.size());
    for (size_t k = 0; k < in.size();) {
      if (in[k] != '&') {
        out += in[k++];
        continue;
      }
      const size_t e = in.find(';', k);
      if (e == std::string::npos) throw Error("XML: unterminated entity");
      const std::string ent = in.substr(k + 1, e - k - 1);
      if (ent == "lt") out += '<';
      else if (ent == "gt") out += '>';
      else if (ent == "amp") out += '&';
      else if (ent == "quot") out += '"';
      else if (ent == "apos") out += '\'';
      else if (!ent.empty() && ent[0] == '#') {
        const long c = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                                                          : std::strtol(ent.c_str() + 1, nullptr, 10);
        if (c <= 0 || c > 127) throw Error("XML: character reference outside ASCII");
        out += (char)c;
      } else {
        throw Error("XML: unknown entity &" + ent + ";");
      }
      k = e + 1;
    }
    return out;
  }
  Element element(int depth) {
    if (depth > 64) throw Error("XML: nesting too deep");
    if (!at("<")) throw Error("XML: expected an element");
    i_++;
    Element e;
    e.tag = name();
    for (;;) {
      skip_ws();
      if (i_ >= s_.size()) throw Error("XML: unterminated element <" + e.tag + ">");
      if (at("/>")) {
        i_ += 2;
        return e;
      }
      if (s_[i_] == '>') {
        i_++;
        break;
      }
      std::string k = name();
      skip_ws();
      if (i_ >= s_.size() || s_[i_] != '=') throw Error("XML: malformed attribute in <" + e.tag + ">");
      i_++;
      skip_ws();
      const char q = i_ < s_.size() ? s_[i_] : 0;
      if (q != '"' && q != '\'') throw Error("XML: unquoted attribute in <" + e.tag + ">");
      const size_t end = s_.find(q, i_ + 1);
      if (end == std::string::npos) throw Error("XML: unterminated attribute in <" + e.tag + ">");
      e.attr.emplace_back(k, unescape(s_.substr(i_ + 1, end - i_ - 1)));
      i_ = end + 1;
    }
    for (;;) {
      if (i_ >= s_.size()) throw Error("XML: unterminated element <" + e.tag + ">");
      if (at("</")) {
        i_ += 2;
        const std::string n = name();
        skip_ws();
        if (n != e.tag || i_ >= s_.size() || s_[i_] != '>') throw Error("XML: mismatched closing tag </" + n + ">");
        i_++;
        return e;
      }
      if (at("<!--")) {
        skip_past("-->");
      } else if (at("<![CDATA[")) {
        const size_t k = s_.find("]]>", i_);
        if (k == std::string::npos) throw Error("XML: unterminated CDATA");
        e.text += s_.substr(i_ + 9, k - i_ - 9);
        i_ = k + 3;
      } else if (at("<?")) {
        skip_past("?>");
      } else if (s_[i_] == '<') {
        e.kids.push_back(element(depth + 1));
      } else {
        size_t k = s_.find('<', i_);
        if (k == std::string::npos) k = s_.size();
        e.text += unescape(s_.substr(i_, k - i_));
        i_ = k;
      }
    }
  }

  const std::string& s_;
  size_t i_ = 0;
};

// ------------------------------------------------------------------------------ payloads

// Bytes from raw memory or from a base64 text.  Base64 is read quantum by quantum and a padded
// quantum ends one encoded run, not the text: VTK encodes a binary array's header and its data
// as separate runs, and both read as one byte stream here.
class ByteSource {
 public:
  static ByteSource raw(const unsigned char* p, size_t n) {
    ByteSource s;
    s.raw_ = p;
    s.len_ = n;
    return s;
  }
  static ByteSource base64(const char* p, size_t n) {
    ByteSource s;
    s.b64_ = p;
    s.len_ = n;
    return s;
  }
  void read(size_t n, std::vector<unsigned char>& out) {
    out.clear();
    // sizes come from the file: check them against what is left before allocating
    if (n > (len_ - pos_) / (raw_ ? 1 : 4) * (raw_ ? 1 : 3) + (raw_ ? 0 : 3)) throw Error("binary data ends early");
    out.reserve(n);
    if (raw_) {
      out.insert(out.end(), raw_ + pos_, raw_ + pos_ + n);
      pos_ += n;
      return;
    }
    while (out.size() < n) {
      if (pend_pos_ < pend_n_) {
        out.push_back(pend_[pend_pos_++]);
        continue;
      }
      int v[4], got = 0, pad = 0;
      while (got < 4) {
        if (pos_ >= len_) throw Error("base64 data ends early");
        const char c = b64_[pos_++];
        if (std::isspace((unsigned char)c)) continue;
        if (c == '=') {
          v[got++] = 0;
          pad++;
          continue;
        }
        const int d = digit(c);
        if (d < 0 || pad) throw Error("invalid base64 data");
        v[got++] = d;
      }
      if (pad > 2) throw Error("invalid base64 padding");
      const uint32_t w = ((uint32_t)v[0] << 18) | ((uint32_t)v[1] << 12) | ((uint32_t)v[2] << 6) | (uint32_t)v[3];
      pend_[0] = (unsigned char)(w >> 16);
      pend_[1] = (unsigned char)(w >> 8);
      pend_[2] = (unsigned char)w;
      pend_n_ = 3 - pad;
      pend_pos_ = 0;
    }
  }

 private:
  static int digit(char c) {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+') return 62;
    if (c == '/') return 63;
    return -1;
  }
  const unsigned char* raw_ = nullptr;
  const char* b64_ = nullptr;
  size_t len_ = 0, pos_ = 0;
  unsigned char pend_[3] = {0, 0, 0};
  int pend_n_ = 0, pend_pos_ = 0;
};

struct Array {
  std::vector<double> v;  // every type converted to FP64 (integers are exact below 2^53)
  int ncomp = 1;
};

// DataArray payloads of one file: its byte order, header type and compressor.
class Decoder {
 public:
  Decoder(const Element& root, const unsigned char* app, size_t app_len, bool app_b64)
      : app_(app), app_len_(app_len), app_b64_(app_b64) {
    big_ = root.get("byte_order", "LittleEndian") == "BigEndian";
    const std::string ht = root.get("header_type", "UInt32");
    if (ht != "UInt32" && ht != "UInt64") throw Error("header_type " + ht + " is not supported");
    hb_ = ht == "UInt64" ? 8 : 4;
    const std::string comp = root.get("compressor", "");
    if (!comp.empty() && comp != "vtkZLibDataCompressor") throw Error("compressor " + comp + " is not supported");
    zlib_ = !comp.empty();
  }

  Array array(const Element& el) const {
    const Type t = type_of(el.get("type", ""));
    const int es = t.size;
    Array a;
    a.ncomp = std::atoi(el.get("NumberOfComponents", "1").c_str());
    if (a.ncomp < 1) throw Error("DataArray " + el.get("Name", "") + ": NumberOfComponents < 1");
    const std::string fmt = el.get("format", "ascii");
    if (fmt == "ascii") {
      const char* p = el.text.c_str();
      for (;;) {
        while (*p && std::isspace((unsigned char)*p)) p++;
        if (!*p) break;
        char* end = nullptr;
        const double x = std::strtod(p, &end);
        if (end == p) throw Error("DataArray " + el.get("Name", "") + ": malformed ascii value");
        a.v.push_back(x);
        p = end;
      }
      return a;
    }
    std::vector<unsigned char> bytes;
    if (fmt == "binary") {
      ByteSource src = ByteSource::base64(el.text.data(), el.text.size());
      payload(src, bytes);
    } else if (fmt == "appended") {
      if (!app_) throw Error("appended DataArray without an AppendedData section");
      const size_t off = (size_t)std::strtoull(el.get("offset", "0").c_str(), nullptr, 10);
      if (off > app_len_) throw Error("appended DataArray offset past the data");
      ByteSource src = app_b64_ ? ByteSource::base64((const char*)app_ + off, app_len_ - off)
                                : ByteSource::raw(app_ + off, app_len_ - off);
      payload(src, bytes);
    } else {
      throw Error("DataArray format " + fmt + " is not supported");
    }
    if (bytes.size() % es) throw Error("DataArray " + el.get("Name", "") + ": size is not a whole number of values");
    const size_t n = bytes.size() / es;
    a.v.resize(n);
    for (size_t i = 0; i < n; i++) a.v[i] = value(t, bytes.data() + i * es);
    return a;
  }

 private:
  struct Type {
    int size;
    char kind;  // 'f' float, 'i' signed, 'u' unsigned
  };
  static Type type_of(const std::string& t) {
    if (t == "Int8") return {1, 'i'};
    if (t == "UInt8") return {1, 'u'};
    if (t == "Int16") return {2, 'i'};
    if (t == "UInt16") return {2, 'u'};
    if (t == "Int32") return {4, 'i'};
    if (t == "UInt32") return {4, 'u'};
    if (t == "Int64") return {8, 'i'};
    if (t == "UInt64") return {8, 'u'};
    if (t == "Float32") return {4, 'f'};
    if (t == "Float64") return {8, 'f'};
    throw Error("DataArray type " + t + " is not supported");
  }
  // one value of type t at p in the file's byte order (this host is little-endian)
  double value(Type t, const unsigned char* p) const {
    unsigned char b[8];
    const int es = t.size;
    for (int k = 0; k < es; k++) b[k] = big_ ? p[es - 1 - k] : p[k];
    if (t.kind == 'f') {
      if (es == 8) {
        double x;
        std::memcpy(&x, b, 8);
        return x;
      }
      float x;
      std::memcpy(&x, b, 4);
      return x;
    }
    uint64_t u = 0;
    for (int k = es - 1; k >= 0; k--) u = (u << 8) | b[k];
    if (t.kind == 'i' && es < 8 && ((u >> (8 * es - 1)) & 1)) u |= ~0ull << (8 * es);  // sign-extend
    return t.kind == 'i' ? (double)(int64_t)u : (double)u;
  }
  uint64_t header(const unsigned char* p) const {
    uint64_t u = 0;
    for (int k = 0; k < hb_; k++) u |= (uint64_t)(big_ ? p[hb_ - 1 - k] : p[k]) << (8 * k);
    return u;
  }
  // The data bytes of one array: [size] data, or [blocks, block size, last size, compressed
  // sizes] and the zlib blocks.
  void payload(ByteSource& src, std::vector<unsigned char>& out) const {
    std::vector<unsigned char> h;
    if (!zlib_) {
      src.read(hb_, h);
      src.read((size_t)header(h.data()), out);
      return;
    }
    src.read(3 * (size_t)hb_, h);
    const uint64_t nb = header(h.data()), bs = header(h.data() + hb_), last = header(h.data() + 2 * hb_);
    if (nb > (1ull << 32) || (bs == 0 && nb)) throw Error("compressed header is not valid");
    std::vector<unsigned char> sizes, comp;
    src.read((size_t)nb * hb_, sizes);
    out.clear();
    for (uint64_t b = 0; b < nb; b++) {
      src.read((size_t)header(sizes.data() + b * hb_), comp);
      const uint64_t want = (b + 1 == nb && last) ? last : bs;
      // a deflate stream expands at most about 1032:1; a larger claim is a corrupt header
      if (want > (uint64_t)comp.size() * 1032 + 64) throw Error("compressed block size is not valid");
      const size_t at = out.size();
      out.resize(at + want);
      uLongf got = (uLongf)want;
      if (uncompress(out.data() + at, &got, comp.data(), (uLong)comp.size()) != Z_OK || got != want)
        throw Error("compressed block does not match the header");
    }
  }

  const unsigned char* app_;
  size_t app_len_;
  bool app_b64_, big_ = false, zlib_ = false;
  int hb_ = 4;
};

// ------------------------------------------------------------------------------ grids

struct Vtu {
  std::vector<double> x, y, z;
  std::vector<int32_t> cells;  // [cells][8], hexahedron vertex order
  std::map<std::string, Array> point_data, cell_data, field_data;
  int32_t npts = 0, ncells = 0;

  // a one-tuple FieldData array (VTKRead::getValue)
  double value(const std::string& name) const {
    const Array& a = array_ref(name);
    if (a.v.size() != 1) throw Error(name + ": not a scalar quantity");
    return a.v[0];
  }
  // a FieldData array (VTKRead::getArray)
  const std::vector<double>& array(const std::string& name) const { return array_ref(name).v; }
  // a node (0) or cell (1) field, PointData first (VTKRead::getField)
  int field(const std::string& name, std::vector<double>& out) const {
    auto p = point_data.find(name);
    if (p != point_data.end()) {
      out = p->second.v;
      return 0;
    }
    auto c = cell_data.find(name);
    if (c != cell_data.end()) {
      out = c->second.v;
      return 1;
    }
    throw Error("no field named " + name);
  }

 private:
  const Array& array_ref(const std::string& name) const {
    auto it = field_data.find(name);
    if (it == field_data.end()) throw Error("no FieldData array named " + name);
    return it->second;
  }
};

inline std::string slurp(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw Error("cannot open " + path);
  std::string s;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
  std::fclose(f);
  return s;
}

inline Vtu read_vtu(const std::string& path) {
  const std::string raw = slurp(path);
  std::string xml;
  const unsigned char* app = nullptr;
  size_t app_len = 0;
  bool app_b64 = false;
  const size_t k = raw.find("<AppendedData");
  if (k != std::string::npos) {
    const size_t close = raw.find('>', k);
    if (close == std::string::npos) throw Error(path + ": malformed AppendedData");
    const std::string tag = raw.substr(k, close - k);
    const size_t e = tag.find("encoding=");
    if (e != std::string::npos && e + 10 < tag.size()) {
      const char q = tag[e + 9];
      const size_t v1 = tag.find(q, e + 10);
      app_b64 = v1 != std::string::npos && tag.substr(e + 10, v1 - e - 10) == "base64";
    }
    const size_t start = raw.find('_', close);
    const size_t end = raw.rfind("</AppendedData>");
    if (start == std::string::npos || end == std::string::npos || end < start)
      throw Error(path + ": malformed AppendedData");
    app = reinterpret_cast<const unsigned char*>(raw.data()) + start + 1;
    app_len = end - start - 1;
    xml = raw.substr(0, k) + "</VTKFile>";
  } else {
    xml = raw;
  }
  const Element root = XmlReader(xml).document();
  if (root.tag != "VTKFile" || root.get("type", "") != "UnstructuredGrid")
    throw Error(path + ": not a VTK UnstructuredGrid file");
  const Decoder dec(root, app, app_len, app_b64);
  const Element* grid = root.child("UnstructuredGrid");
  if (!grid) throw Error(path + ": no UnstructuredGrid element");
  const auto pieces = grid->children("Piece");
  if (pieces.size() != 1) throw Error(path + ": exactly one Piece is supported");
  const Element& piece = *pieces[0];

  Vtu v;
  auto arrays = [&](const Element* sec, std::map<std::string, Array>& out) {
    if (!sec) return;
    for (const Element* el : sec->children("DataArray")) out[el->get("Name", "")] = dec.array(*el);
  };
  arrays(grid->child("FieldData"), v.field_data);
  arrays(piece.child("PointData"), v.point_data);
  arrays(piece.child("CellData"), v.cell_data);

  const long long npts = std::atoll(piece.get("NumberOfPoints", "0").c_str());
  const long long ncells = std::atoll(piece.get("NumberOfCells", "0").c_str());
  if (npts < 0 || ncells < 0 || npts > INT32_MAX || ncells > INT32_MAX / 8)
    throw Error(path + ": point or cell count out of range");
  v.npts = (int32_t)npts;
  v.ncells = (int32_t)ncells;
  const Element* pts_sec = piece.child("Points");
  const Element* pts_el = pts_sec ? pts_sec->child("DataArray") : nullptr;
  if (!pts_el) throw Error(path + ": no Points array");
  const Array pts = dec.array(*pts_el);
  if (pts.v.size() != (size_t)npts * 3) throw Error(path + ": Points do not match NumberOfPoints");
  v.x.resize(npts);
  v.y.resize(npts);
  v.z.resize(npts);
  for (long long i = 0; i < npts; i++) {
    v.x[i] = pts.v[3 * i];
    v.y[i] = pts.v[3 * i + 1];
    v.z[i] = pts.v[3 * i + 2];
  }
  const Element* cel = piece.child("Cells");
  if (!cel) throw Error(path + ": no Cells element");
  std::map<std::string, const Element*> by_name;
  for (const Element* el : cel->children("DataArray")) by_name[el->get("Name", "")] = el;
  for (const char* n : {"connectivity", "offsets", "types"})
    if (!by_name.count(n)) throw Error(path + ": no cell array " + n);
  const Array conn = dec.array(*by_name["connectivity"]);
  const Array offs = dec.array(*by_name["offsets"]);
  const Array types = dec.array(*by_name["types"]);
  if (offs.v.size() != (size_t)ncells || types.v.size() != (size_t)ncells)
    throw Error(path + ": cell arrays do not match NumberOfCells");
  if (conn.v.size() != (size_t)ncells * 8)
    throw Error(path + ": only hexahedral (type 12) or voxel (type 11) cells are supported");
  static const int kVoxelToHex[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  v.cells.resize((size_t)ncells * 8);
  for (long long c = 0; c < ncells; c++) {
    const double prev = c ? offs.v[c - 1] : 0.0;
    const int type = (int)types.v[c];
    if (offs.v[c] - prev != 8 || (type != 11 && type != 12))
      throw Error(path + ": only hexahedral (type 12) or voxel (type 11) cells are supported");
    for (int k8 = 0; k8 < 8; k8++) {
      const double id = conn.v[8 * c + (type == 11 ? kVoxelToHex[k8] : k8)];
      if (id < 0 || id >= (double)npts) throw Error(path + ": cell vertex out of range");
      v.cells[8 * c + k8] = (int32_t)id;
    }
  }
  return v;
}

// ------------------------------------------------------------------------------ writers

struct Column {
  std::string name;
  std::vector<double> v;
};

inline FILE* open_out(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw Error("cannot write " + path);
  return f;
}

inline void put_ascii(FILE* f, const double* v, size_t n) {
  for (size_t i = 0; i < n; i++) std::fprintf(f, i ? " %.17g" : "%.17g", v[i]);
}

// A vtkTable file of FP64 columns (RowData), ascii, as the reference's table writer produces
// for the sensor coordinates and velocities (TransducerModel.cpp:58,78).
inline void write_vtt(const std::string& path, const std::vector<Column>& cols) {
  FILE* f = open_out(path);
  const size_t rows = cols.empty() ? 0 : cols[0].v.size();
  std::fprintf(f,
               "<?xml version=\"1.0\"?>\n<VTKFile type=\"Table\" version=\"1.0\" byte_order=\"LittleEndian\">\n"
               "<Table><Piece NumberOfCols=\"%zu\" NumberOfRows=\"%zu\"><RowData>",
               cols.size(), rows);
  for (auto& c : cols) {
    std::fprintf(f, "<DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"1\" format=\"ascii\">",
                 c.name.c_str());
    put_ascii(f, c.v.data(), c.v.size());
    std::fprintf(f, "</DataArray>");
  }
  std::fprintf(f, "</RowData></Piece></Table>\n</VTKFile>\n");
  std::fclose(f);
}

// One UnstructuredGrid piece of hexahedra: FP64 point and cell data, a FieldData time, and the
// points and cells; appended raw data with UInt64 headers, or ascii.
struct GridOut {
  const std::vector<double>*x, *y, *z;
  const std::vector<int32_t>* cells;
  std::vector<Column> point_data, cell_data;
  double time = 0.0;
};

inline void write_vtu(const std::string& path, const GridOut& g, bool ascii) {
  const size_t nn = g.x->size(), nc = g.cells->size() / 8;
  std::vector<std::pair<std::vector<unsigned char>, size_t>> blobs;  // appended payloads
  size_t offset = 0;
  std::string xml;
  char buf[256];
  auto data_array = [&](const char* type, const std::string& name, int ncomp, const void* data, size_t n, int es,
                        const char* extra) {
    std::snprintf(buf, sizeof buf, "<DataArray type=\"%s\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"%s\"%s",
                  type, name.c_str(), ncomp, ascii ? "ascii" : "appended", extra);
    xml += buf;
    if (ascii) {
      xml += ">";
      std::string txt;
      for (size_t i = 0; i < n; i++) {
        if (i) txt += ' ';
        if (es == 8 && type[0] == 'F') {
          std::snprintf(buf, sizeof buf, "%.17g", static_cast<const double*>(data)[i]);
        } else if (es == 8) {
          std::snprintf(buf, sizeof buf, "%lld", (long long)static_cast<const int64_t*>(data)[i]);
        } else {
          std::snprintf(buf, sizeof buf, "%d", (int)static_cast<const uint8_t*>(data)[i]);
        }
        txt += buf;
      }
      xml += txt + "</DataArray>";
      return;
    }
    std::snprintf(buf, sizeof buf, " offset=\"%zu\"/>", offset);
    xml += buf;
    std::vector<unsigned char> b(8 + n * es);
    const uint64_t len = (uint64_t)n * es;
    std::memcpy(b.data(), &len, 8);  // little-endian host
    if (n) std::memcpy(b.data() + 8, data, n * es);
    offset += b.size();
    blobs.emplace_back(std::move(b), 0);
  };
  xml += "<UnstructuredGrid><FieldData>";
  data_array("Float64", "time", 1, &g.time, 1, 8, " NumberOfTuples=\"1\"");
  std::snprintf(buf, sizeof buf, "</FieldData><Piece NumberOfPoints=\"%zu\" NumberOfCells=\"%zu\">", nn, nc);
  xml += buf;
  if (!g.point_data.empty()) {
    xml += "<PointData>";
    for (auto& c : g.point_data) data_array("Float64", c.name, 1, c.v.data(), c.v.size(), 8, "");
    xml += "</PointData>";
  }
  if (!g.cell_data.empty()) {
    xml += "<CellData>";
    for (auto& c : g.cell_data) data_array("Float64", c.name, 1, c.v.data(), c.v.size(), 8, "");
    xml += "</CellData>";
  }
  std::vector<double> pts(3 * nn);
  for (size_t i = 0; i < nn; i++) {
    pts[3 * i] = (*g.x)[i];
    pts[3 * i + 1] = (*g.y)[i];
    pts[3 * i + 2] = (*g.z)[i];
  }
  xml += "<Points>";
  data_array("Float64", "Points", 3, pts.data(), pts.size(), 8, "");
  xml += "</Points><Cells>";
  std::vector<int64_t> conn(g.cells->begin(), g.cells->end()), offs(nc);
  for (size_t c = 0; c < nc; c++) offs[c] = 8 * (int64_t)(c + 1);
  std::vector<uint8_t> types(nc, 12);
  data_array("Int64", "connectivity", 1, conn.data(), conn.size(), 8, "");
  data_array("Int64", "offsets", 1, offs.data(), offs.size(), 8, "");
  data_array("UInt8", "types", 1, types.data(), types.size(), 1, "");
  xml += "</Cells></Piece></UnstructuredGrid>";

  FILE* f = open_out(path);
  std::fprintf(f,
               "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" "
               "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n");
  std::fwrite(xml.data(), 1, xml.size(), f);
  if (!ascii) {
    std::fprintf(f, "\n<AppendedData encoding=\"raw\">\n_");
    for (auto& b : blobs) std::fwrite(b.first.data(), 1, b.first.size(), f);
    std::fprintf(f, "\n</AppendedData>");
  }
  std::fprintf(f, "\n</VTKFile>\n");
  std::fclose(f);
}

// The file name VTK's parallel XML writers give piece k of `path`: <base>_<k>.<serial ext>.
inline std::string piece_name(const std::string& path, int k) {
  const size_t dot = path.rfind('.');
  const std::string base = path.substr(0, dot), ext = path.substr(dot);
  return base + "_" + std::to_string(k) + (ext == ".pvtu" ? ".vtu" : ".vtt");
}
inline std::string base_name(const std::string& path) {
  const size_t s = path.rfind('/');
  return s == std::string::npos ? path : path.substr(s + 1);
}

// One-piece partitioned files, the layout the reference writes on several processes
// (VTKPWrite.hpp:100-137): the summary names the arrays and the piece beside it.
inline void write_pvtt(const std::string& path, const std::vector<Column>& cols) {
  const std::string piece = piece_name(path, 0);
  write_vtt(piece, cols);
  FILE* f = open_out(path);
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"PTable\" version=\"1.0\" byte_order=\"LittleEndian\">\n"
                  "<PTable><PRowData>");
  for (auto& c : cols)
    std::fprintf(f, "<PDataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"1\"/>", c.name.c_str());
  std::fprintf(f, "</PRowData><Piece Source=\"%s\"/></PTable>\n</VTKFile>\n", base_name(piece).c_str());
  std::fclose(f);
}

inline void write_pvtu(const std::string& path, const GridOut& g, bool ascii) {
  const std::string piece = piece_name(path, 0);
  write_vtu(piece, g, ascii);
  FILE* f = open_out(path);
  std::fprintf(f,
               "<?xml version=\"1.0\"?>\n<VTKFile type=\"PUnstructuredGrid\" version=\"1.0\" "
               "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n<PUnstructuredGrid GhostLevel=\"0\">"
               "<PFieldData><PDataArray type=\"Float64\" Name=\"time\" NumberOfComponents=\"1\"/></PFieldData>");
  std::fprintf(f, "<PPointData>");
  for (auto& c : g.point_data)
    std::fprintf(f, "<PDataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"1\"/>", c.name.c_str());
  std::fprintf(f, "</PPointData><PCellData>");
  for (auto& c : g.cell_data)
    std::fprintf(f, "<PDataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"1\"/>", c.name.c_str());
  std::fprintf(f,
               "</PCellData><PPoints><PDataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\"/>"
               "</PPoints><Piece Source=\"%s\"/></PUnstructuredGrid>\n</VTKFile>\n",
               base_name(piece).c_str());
  std::fclose(f);
}

}  // namespace vtkjob
