// Test harness: contiguous [n1][n2] arrays (row pointers into one block).
#pragma once
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include "pointers.h"
namespace LAMMPS_NS {
class Memory {
 public:
  template <typename T>
  T **grow(T **&array, int n1, int n2, const char *) {
    T *data = static_cast<T *>(std::calloc(static_cast<size_t>(n1) * n2 + 1, sizeof(T)));
    T **rows = static_cast<T **>(std::malloc(sizeof(T *) * (n1 > 0 ? n1 : 1)));
    for (int i = 0; i < n1; ++i) rows[i] = data + static_cast<size_t>(i) * n2;
    if (array) {
      // keep the old rows (same column count), as LAMMPS's grow does
      std::memcpy(data, array[0], sizeof(T) * static_cast<size_t>(old_rows(array)) * n2);
      destroy(array);
    }
    if (n1 == 0) rows[0] = data;
    rows_[rows] = n1;
    array = rows;
    return array;
  }
  template <typename T>
  void destroy(T **&array) {
    if (!array) return;
    rows_.erase(array);
    std::free(array[0]);
    std::free(array);
    array = nullptr;
  }

 private:
  template <typename T>
  int old_rows(T **a) {
    auto it = rows_.find(a);
    return it == rows_.end() ? 0 : it->second;
  }
  struct Hash {
    size_t operator()(const void *p) const { return reinterpret_cast<size_t>(p); }
  };
  std::unordered_map<const void *, int, Hash> rows_;
};
}  // namespace LAMMPS_NS
