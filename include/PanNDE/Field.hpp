// PanNDE::Field / FieldBundle / FieldFactory (reference modules/PanNDE/include/Field.hpp:73-349;
// same names and signatures; the code is this library's own).  A NODE field is read at a cell as
// the mean of its 8 vertices; a CELL or CONSTANT field is read at a node as the sum over
// the present connected cells divided by 8.
#pragma once

#include <cstdint>
#include <memory>
#include <string>

#include "Mesh.hpp"

namespace PanNDE {

class Field {
 public:
  enum FieldType { NODE = 0, CELL = 1, CONSTANT = 2 };
  virtual FieldType type() = 0;
  virtual std::shared_ptr<PanNDE::Mesh> mesh() = 0;
  virtual double* data() = 0;
  virtual double& at(int index) = 0;
  virtual double atCell(int cidx) = 0;
  virtual double atNode(int nidx) = 0;
  virtual int32_t size() = 0;
  virtual void mapFrom(std::shared_ptr<PanNDE::Field> other) = 0;
  virtual void mapFrom(PanNDE::Field* other) = 0;
  virtual ~Field() {}
};

class FieldBundle {
 public:
  virtual std::shared_ptr<PanNDE::Mesh>& mesh() = 0;
  virtual std::shared_ptr<PanNDE::Field> field(std::string keyname) = 0;
  virtual std::string fieldName(int idx) = 0;
  virtual int fieldCount() = 0;
  virtual void emplaceField(std::string keyname, std::shared_ptr<PanNDE::Field> field) = 0;
  virtual void emplaceField(std::string keyname, PanNDE::Field::FieldType type) = 0;
  virtual ~FieldBundle() {}
};

class FieldFactory {
 public:
  virtual std::shared_ptr<PanNDE::Field> makeManagedField(std::shared_ptr<PanNDE::Mesh> mesh,
                                                          PanNDE::Field::FieldType tp) = 0;
  virtual std::shared_ptr<PanNDE::Field> makeManagedField(PanNDE::Mesh* mesh, PanNDE::Field::FieldType tp) = 0;
  virtual PanNDE::Field* newField(std::shared_ptr<PanNDE::Mesh> mesh, PanNDE::Field::FieldType tp) = 0;
  virtual PanNDE::Field* newField(PanNDE::Mesh* mesh, PanNDE::Field::FieldType tp) = 0;
  virtual void deleteField(PanNDE::Field* field) = 0;
  virtual std::shared_ptr<PanNDE::Array<std::shared_ptr<PanNDE::Field>>> makeManagedFieldArray() = 0;
  virtual std::shared_ptr<PanNDE::FieldBundle> makeEmptyManagedFieldBundle() = 0;
  virtual ~FieldFactory() {}
};

}  // namespace PanNDE
