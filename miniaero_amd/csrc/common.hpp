// common.hpp -- error reporting shared by the host side of libminiaero_amd.
#pragma once
#include <string>

#include "../../include/miniaero.h"

namespace ma {

// Thread-local message behind ma_last_error().  Defined once, in mesh.cpp:
// the library mixes g++- and hipcc-compiled objects, and an inline function's
// static would not reliably be shared between them.
std::string& last_error();

inline ma_status fail(ma_status code, const std::string& msg) {
  last_error() = msg;
  return code;
}

}  // namespace ma
