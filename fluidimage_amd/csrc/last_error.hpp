// Thread-local message of the last failed C-ABI call, shared by every
// translation unit of libpiv_hip (one instance: inline function static).
#pragma once

#include <string>

namespace fia {

inline std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}

}  // namespace fia
