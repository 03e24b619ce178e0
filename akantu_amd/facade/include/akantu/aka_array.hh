/**
 * aka_array.hh -- Array<T>: contiguous (size x nb_component) row-major storage,
 * the layout of the reference's Array (src/common/aka_array.hh:139-224,
 * aka_array_tmpl.hh:440-450).
 */
#pragma once

#include <algorithm>
#include <string>
#include <vector>

#include "aka_common.hh"

namespace akantu {

template <typename T>
class Array {
  // One element per cell; a plain vector<bool> would not hand out bool&.
  struct Cell {
    T v;
  };
  static_assert(sizeof(Cell) == sizeof(T), "cell adds padding");

 public:
  Array() = default;
  explicit Array(UInt size, UInt nb_component = 1, const T& value = T(), const ID& id = "")
      : size_(size), nb_component_(nb_component), values_(size_t(size) * nb_component, Cell{value}),
        id_(id) {}
  Array(UInt size, UInt nb_component, const ID& id)
      : size_(size), nb_component_(nb_component), values_(size_t(size) * nb_component), id_(id) {}

  UInt size() const { return size_; }
  UInt getNbComponent() const { return nb_component_; }
  const ID& getID() const { return id_; }
  T* storage() { return reinterpret_cast<T*>(values_.data()); }
  const T* storage() const { return reinterpret_cast<const T*>(values_.data()); }

  T& operator()(UInt i, UInt j = 0) { return values_[size_t(i) * nb_component_ + j].v; }
  const T& operator()(UInt i, UInt j = 0) const { return values_[size_t(i) * nb_component_ + j].v; }
  T& operator[](size_t k) { return values_[k].v; }
  const T& operator[](size_t k) const { return values_[k].v; }

  void resize(UInt size) {
    size_ = size;
    values_.resize(size_t(size) * nb_component_);
  }
  void resize(UInt size, UInt nb_component) {
    size_ = size;
    nb_component_ = nb_component;
    values_.resize(size_t(size) * nb_component_);
  }
  void set(const T& v) { std::fill(values_.begin(), values_.end(), Cell{v}); }
  void clear() { set(T()); }
  void copy(const Array& other) {
    size_ = other.size_;
    nb_component_ = other.nb_component_;
    values_ = other.values_;
  }

  T* begin() { return storage(); }
  T* end() { return storage() + values_.size(); }
  const T* begin() const { return storage(); }
  const T* end() const { return storage() + values_.size(); }

 private:
  UInt size_ = 0;
  UInt nb_component_ = 1;
  std::vector<Cell> values_;
  ID id_;
};

}  // namespace akantu
