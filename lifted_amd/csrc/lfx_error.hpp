// Host-side error reporting of lifted-hip.
#pragma once
#include <string>

#include "../../include/lifted_hip.h"

namespace lfx {

struct Error {
  int code;
  std::string msg;
};
[[noreturn]] inline void fail(int code, const std::string& m) { throw Error{code, m}; }

}  // namespace lfx
