// Test harness: errors end the run with the message (LAMMPS aborts instead).
#pragma once
#include <stdexcept>
#include <string>
#include "pointers.h"
#define FLERR __FILE__, __LINE__
namespace LAMMPS_NS {
class Error {
 public:
  [[noreturn]] void all(const std::string &file, int line, const std::string &str) {
    throw std::runtime_error(file + ":" + std::to_string(line) + ": " + str);
  }
  [[noreturn]] void one(const std::string &file, int line, const std::string &str) {
    all(file, line, str);
  }
};
}  // namespace LAMMPS_NS
