// PanNDE::Mesh / MeshFactory: hexahedral mesh interface (reference
// modules/PanNDE/include/Mesh.hpp:74-370; same names and signatures, written for this
// build).  Cells list 8 node ids in CGNS order [000,100,110,010,001,101,111,011];
// connectedCells(n) lists the cell in each octant slot around node n, -1 for none.
#pragma once

#include <cstdint>
#include <memory>

#include "Array.hpp"

namespace PanNDE {

class Mesh {
 public:
  struct Node {
    std::shared_ptr<PanNDE::Array<double>> coord;
    int64_t global_id;
    int32_t owner_partition;
  };
  struct Cell {
    std::shared_ptr<PanNDE::Array<int32_t>> nodes;
    int64_t global_id;
    int32_t owner_partition;
  };
  virtual int32_t nodeCount() = 0;
  virtual int32_t cellCount() = 0;
  virtual int32_t partitionId() = 0;
  virtual void nodeCoordinate(int32_t node_id, double coord[3]) = 0;
  virtual std::shared_ptr<PanNDE::Array<double>> nodeCoordinate(int32_t node_id) = 0;
  virtual void cell(int32_t cell_id, int32_t* nodes) = 0;
  virtual std::shared_ptr<PanNDE::Array<int32_t>> cell(int32_t cell_id) = 0;
  virtual void connectedCells(int32_t node_id, int32_t* cells) = 0;
  virtual std::shared_ptr<PanNDE::Array<int32_t>> connectedCells(int32_t node_id) = 0;
  virtual int64_t globalNodeId(int32_t node_id) = 0;
  virtual int32_t nodeHomePartition(int32_t node_id) = 0;
  virtual int64_t globalCellId(int32_t cell_id) = 0;
  virtual int32_t cellHomePartition(int32_t cell_id) = 0;
  virtual void copy(PanNDE::Mesh* other) = 0;
  virtual void copy(std::shared_ptr<PanNDE::Mesh> other) = 0;
  virtual ~Mesh() {}
};

class MeshFactory {
 public:
  virtual PanNDE::Mesh::Node makeNode(double coord[3], int64_t global_id, int32_t owner = 0) = 0;
  virtual PanNDE::Mesh::Cell makeCell(int32_t* nodes, int32_t Nnodes, int64_t global_id,
                                      int32_t owner = 0) = 0;
  virtual std::shared_ptr<PanNDE::Mesh> makeManagedMesh(PanNDE::Mesh::Node* nodes, int64_t Nnodes,
                                                        PanNDE::Mesh::Cell* cells, int64_t Ncells,
                                                        int32_t partition = 0) = 0;
  virtual std::shared_ptr<PanNDE::Mesh> makeManagedMesh(
      std::shared_ptr<PanNDE::Array<PanNDE::Mesh::Node>> nodes,
      std::shared_ptr<PanNDE::Array<PanNDE::Mesh::Cell>> cells, int32_t partition = 0) = 0;
  virtual PanNDE::Mesh* newMesh(PanNDE::Mesh::Node* nodes, int64_t Nnodes, PanNDE::Mesh::Cell* cells,
                                int64_t Ncells, int32_t partition = 0) = 0;
  virtual PanNDE::Mesh* newMesh(std::shared_ptr<PanNDE::Array<PanNDE::Mesh::Node>> nodes,
                                std::shared_ptr<PanNDE::Array<PanNDE::Mesh::Cell>> cells,
                                int32_t partition = 0) = 0;
  virtual void deleteMesh(PanNDE::Mesh* mesh) = 0;
  virtual std::shared_ptr<PanNDE::Array<std::shared_ptr<PanNDE::Mesh>>> makeManagedMeshArray() = 0;
  virtual ~MeshFactory() {}
};

}  // namespace PanNDE
