// params.json / ini readers shared by librrg.so and the rrgsim CLI.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

#include "rrg.h"
#include "rrg_error.hpp"

namespace rrg {
namespace io {

void params_default(rrg_params* p);
void params_finalize(rrg_params* p);
void params_parse_json(const std::string& path, rrg_params* p, rrg_ini_params* ini);
void ini_read(const rrg_ini_params* ini, const char* base_dir, int64_t* n, double* pos, double* vel, double* mass,
              double* soft2);

}  // namespace io
}  // namespace rrg
