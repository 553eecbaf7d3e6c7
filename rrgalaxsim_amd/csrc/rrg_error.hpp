// Error type used inside librrg.so; never crosses the C ABI.
#pragma once

#include <stdexcept>
#include <string>

#include "rrg.h"

namespace rrg {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

}  // namespace rrg
