// PanNDE::Array / DataBundle / ArrayFactory: the container interface the RSG path reads
// node coordinates, cell vertex lists and signal tables through.  Same names and
// signatures as the reference's modules/PanNDE/include/Array.hpp:67-310, so code written
// against one compiles against the other; the code is this library's own.
#pragma once

#include <cstdint>
#include <memory>
#include <string>

namespace PanNDE {

template <typename T>
struct CArray {
  T* data = nullptr;
  int64_t length = 0;
};

template <typename T>
class Array {
 public:
  virtual T& at(int64_t idx) = 0;
  virtual int64_t size() = 0;
  virtual T* data() = 0;
  virtual void resize(int64_t newsize) = 0;
  virtual void copy(PanNDE::Array<T>* other) = 0;
  virtual void copy(std::shared_ptr<PanNDE::Array<T>> other) = 0;
  virtual void copy(CArray<T>& other) = 0;
  virtual const CArray<T> getCArray() = 0;
  virtual ~Array() {}
};

template <typename T>
class DataBundle {
 public:
  virtual T scalar(std::string keyname) = 0;
  virtual std::shared_ptr<PanNDE::Array<T>> array(std::string keyname) = 0;
  virtual std::string scalarName(int idx) = 0;
  virtual std::string arrayName(int idx) = 0;
  virtual int scalarCount() = 0;
  virtual int arrayCount() = 0;
  virtual void emplaceScalar(std::string keyname, T value) = 0;
  virtual void emplaceArray(std::string keyname, std::shared_ptr<PanNDE::Array<T>> array) = 0;
  virtual ~DataBundle() {}
};

template <typename T>
class ArrayFactory {
 public:
  virtual std::shared_ptr<PanNDE::Array<T>> makeManagedArray() = 0;
  virtual PanNDE::Array<T>* newArray() = 0;
  virtual void deleteArray(PanNDE::Array<T>* array) = 0;
  virtual std::shared_ptr<PanNDE::DataBundle<T>> makeManagedDataBundle() = 0;
  virtual ~ArrayFactory() {}
};

}  // namespace PanNDE
