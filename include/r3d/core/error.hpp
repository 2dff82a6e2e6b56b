// Exceptions of the r3d API (reference: src/r3d/core/error.hpp:12-25 and
// src/r3d/core/cancel.hpp:10-13), without <format>.
#pragma once

#include <stdexcept>
#include <string>

namespace r3d {

class Error : public std::runtime_error {
 public:
  using std::runtime_error::runtime_error;
};

class Cancelled : public Error {
 public:
  Cancelled() : Error("cancelled") {}
};

[[noreturn]] inline void throw_error(const std::string& message) { throw Error(message); }

}  // namespace r3d
