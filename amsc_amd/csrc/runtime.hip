// runtime.hip -- error state and stream selection of the C ABI.
#include "common.hpp"
#include "amsc_mpm.h"

#include <string>

namespace amsc {

namespace {
thread_local std::string g_error;
hipStream_t g_stream = nullptr;
}  // namespace

void set_error(const std::string &msg) { g_error = msg; }

int fail(int code, const std::string &msg) {
  g_error = msg;
  return code;
}

hipStream_t current_stream() { return g_stream; }

const char *last_error_cstr() { return g_error.c_str(); }
void set_stream(hipStream_t s) { g_stream = s; }

}  // namespace amsc

extern "C" {

const char *amsc_last_error(void) { return amsc::last_error_cstr(); }

int amsc_set_stream(void *stream) {
  amsc::set_stream((hipStream_t)stream);
  return AMSC_OK;
}

int amsc_synchronize(void) {
  AMSC_HIP(hipStreamSynchronize(amsc::current_stream()));
  return AMSC_OK;
}

}  // extern "C"
