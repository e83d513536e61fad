// common.hpp - host-side helpers shared by the library's translation units.
#pragma once
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/erm.h"

namespace erm {

// Thread-local message of the last failure that had no context to hold it.
void set_global_error(const char* fmt, ...);
const char* global_error();

std::string vformat(const char* fmt, va_list ap);
std::string format(const char* fmt, ...);

// getpar-compatible parsing (getpar.cpp)
int parse_par(int argc, const char* const* argv, erm_params* p, std::string* err);
void params_default(erm_params* p);

// synthetic models / raw I/O (model.cpp)
int make_model3d(int kind, int nx, int ny, int nz, float dx, float dy, float dz, unsigned long long seed,
                 float* vp, float* vs, float* rho);
int write_raw(const char* path, const float* data, size_t n);
int read_raw(const char* path, float* data, size_t n);

}  // namespace erm
