// HostContainers.hpp -- minimal host-side implementations of the PanNDE container
// interfaces (Array, Mesh, Field, FieldBundle) for programs that drive the DeviceSolver
// without the reference's HostData module.  Semantics follow HostData (HexMesh.hpp,
// HostField.hpp): connected-cell slots by HexMesh::buildLinks' rule (slot = vertex ^ 6),
// CELL/CONSTANT fields read at nodes as the sum over present cells divided by 8, NODE
// fields read at cells as the mean of the 8 vertices.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "Array.hpp"
#include "Field.hpp"
#include "Mesh.hpp"

namespace DeviceSolver {

template <typename T>
class VectorArray : public PanNDE::Array<T> {
 public:
  VectorArray() = default;
  explicit VectorArray(std::vector<T> v) : v_(std::move(v)) {}
  T& at(int64_t i) override { return v_.at((size_t)i); }
  int64_t size() override { return (int64_t)v_.size(); }
  T* data() override { return v_.data(); }
  void resize(int64_t n) override { v_.resize((size_t)n); }
  void copy(PanNDE::Array<T>* o) override { v_.assign(o->data(), o->data() + o->size()); }
  void copy(std::shared_ptr<PanNDE::Array<T>> o) override { copy(o.get()); }
  void copy(PanNDE::CArray<T>& o) override { v_.assign(o.data, o.data + o.length); }
  const PanNDE::CArray<T> getCArray() override { return PanNDE::CArray<T>{v_.data(), (int64_t)v_.size()}; }

 private:
  std::vector<T> v_;
};

// Single-partition hexahedral mesh.
class HexMesh : public PanNDE::Mesh {
 public:
  HexMesh(std::vector<double> x, std::vector<double> y, std::vector<double> z, std::vector<int32_t> cells)
      : x_(std::move(x)), y_(std::move(y)), z_(std::move(z)), cells_(std::move(cells)) {
    const int32_t nn = (int32_t)x_.size(), nc = (int32_t)(cells_.size() / 8);
    for (int32_t v : cells_)
      if (v < 0 || v >= nn) throw std::out_of_range("invalid node id: " + std::to_string(v));
    links_.assign((size_t)nn * 8, -1);
    for (int32_t k = 0; k < nc; k++)
      for (int v = 0; v < 8; v++) links_[(size_t)cells_[(size_t)k * 8 + v] * 8 + (v ^ 6)] = k;
  }
  int32_t nodeCount() override { return (int32_t)x_.size(); }
  int32_t cellCount() override { return (int32_t)(cells_.size() / 8); }
  int32_t partitionId() override { return 0; }
  void nodeCoordinate(int32_t n, double c[3]) override {
    c[0] = x_.at(n);
    c[1] = y_.at(n);
    c[2] = z_.at(n);
  }
  std::shared_ptr<PanNDE::Array<double>> nodeCoordinate(int32_t n) override {
    return std::make_shared<VectorArray<double>>(std::vector<double>{x_.at(n), y_.at(n), z_.at(n)});
  }
  void cell(int32_t c, int32_t* nodes) override {
    for (int k = 0; k < 8; k++) nodes[k] = cells_.at((size_t)c * 8 + k);
  }
  std::shared_ptr<PanNDE::Array<int32_t>> cell(int32_t c) override {
    return std::make_shared<VectorArray<int32_t>>(
        std::vector<int32_t>(cells_.begin() + (size_t)c * 8, cells_.begin() + (size_t)c * 8 + 8));
  }
  void connectedCells(int32_t n, int32_t* cells) override {
    for (int k = 0; k < 8; k++) cells[k] = links_.at((size_t)n * 8 + k);
  }
  std::shared_ptr<PanNDE::Array<int32_t>> connectedCells(int32_t n) override {
    return std::make_shared<VectorArray<int32_t>>(
        std::vector<int32_t>(links_.begin() + (size_t)n * 8, links_.begin() + (size_t)n * 8 + 8));
  }
  int64_t globalNodeId(int32_t n) override { return n; }
  int32_t nodeHomePartition(int32_t) override { return 0; }
  int64_t globalCellId(int32_t c) override { return c; }
  int32_t cellHomePartition(int32_t) override { return 0; }
  void copy(PanNDE::Mesh*) override { throw std::runtime_error("HexMesh::copy is not supported"); }
  void copy(std::shared_ptr<PanNDE::Mesh> o) override { copy(o.get()); }

 private:
  std::vector<double> x_, y_, z_;
  std::vector<int32_t> cells_, links_;
};

class SimpleField : public PanNDE::Field {
 public:
  SimpleField(std::shared_ptr<PanNDE::Mesh> mesh, FieldType tp) : mesh_(std::move(mesh)), type_(tp) {
    v_.assign(tp == NODE ? mesh_->nodeCount() : tp == CELL ? mesh_->cellCount() : 1, 0.0);
  }
  FieldType type() override { return type_; }
  std::shared_ptr<PanNDE::Mesh> mesh() override { return mesh_; }
  double* data() override { return v_.data(); }
  double& at(int i) override { return type_ == CONSTANT ? v_.at(0) : v_.at((size_t)i); }
  double atCell(int c) override {
    if (type_ != NODE) return at(c);
    int32_t box[8];
    mesh_->cell(c, box);
    double v = 0.0;
    for (int k = 0; k < 8; k++) v += (-1 == box[k]) ? 0. : at(box[k]);
    return v / 8.0;
  }
  double atNode(int n) override {
    if (type_ == NODE) return at(n);
    int32_t box[8];
    mesh_->connectedCells(n, box);
    double v = 0.0;
    for (int k = 0; k < 8; k++) v += (-1 == box[k]) ? 0. : at(box[k]);
    return v / 8.0;
  }
  int32_t size() override { return (int32_t)v_.size(); }
  void mapFrom(std::shared_ptr<PanNDE::Field> o) override { mapFrom(o.get()); }
  void mapFrom(PanNDE::Field* o) override {
    for (int32_t k = 0; k < size(); k++) at(k) = type_ == NODE ? o->atNode(k) : o->atCell(k);
  }

 private:
  std::shared_ptr<PanNDE::Mesh> mesh_;
  FieldType type_;
  std::vector<double> v_;
};

class SimpleFieldBundle : public PanNDE::FieldBundle {
 public:
  explicit SimpleFieldBundle(std::shared_ptr<PanNDE::Mesh> mesh) : mesh_(std::move(mesh)) {}
  std::shared_ptr<PanNDE::Mesh>& mesh() override { return mesh_; }
  std::shared_ptr<PanNDE::Field> field(std::string k) override { return fields_.at(k); }
  std::string fieldName(int i) override { return names_.at((size_t)i); }
  int fieldCount() override { return (int)names_.size(); }
  void emplaceField(std::string k, std::shared_ptr<PanNDE::Field> f) override {
    if (!fields_.emplace(k, f).second) throw std::runtime_error(k + " already in use.");
    names_.push_back(k);
  }
  void emplaceField(std::string k, PanNDE::Field::FieldType tp) override {
    emplaceField(k, std::make_shared<SimpleField>(mesh_, tp));
  }

 private:
  std::shared_ptr<PanNDE::Mesh> mesh_;
  std::map<std::string, std::shared_ptr<PanNDE::Field>> fields_;
  std::vector<std::string> names_;
};

}  // namespace DeviceSolver
