// common.hpp -- error reporting shared by the host side of libminiaero_amd.
#pragma once
#include <string>

#include "../../include/miniaero.h"

namespace ma {

// Thread-local message behind ma_last_error().
inline std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}

inline ma_status fail(ma_status code, const std::string& msg) {
  last_error() = msg;
  return code;
}

}  // namespace ma
