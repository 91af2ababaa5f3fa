// common.h -- shared definitions of the host side (the drop-in classes).
//
// `real` is double, as in the reference's default build (src/common.h:22-26).
// Device errors end the process with a message, as the reference's CHECK
// macro does (src/common.h:31-42); input errors are thrown as InputError and
// turned into "message + exit(1)" by the CLI (the reference prints and exits
// at the point of failure, e.g. src/model.cu:128-310).
#pragma once

#include <stdexcept>
#include <string>

typedef double real;

namespace lsqt_host {

struct InputError : std::runtime_error {
  explicit InputError(const std::string& msg) : std::runtime_error(msg) {}
};

// Aborts the process on a non-zero status from the device library.
void check_status(int status, const char* call, const char* file, int line);

}  // namespace lsqt_host

#define LSQT_CHECK(call) ::lsqt_host::check_status((call), #call, __FILE__, __LINE__)
