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

int amsc_device_alloc(size_t bytes, void **ptr) {
  if (!ptr) return amsc::fail(AMSC_ERR_INVALID, "NULL output pointer");
  *ptr = nullptr;
  if (hipMalloc(ptr, bytes > 0 ? bytes : 1) != hipSuccess)
    return amsc::fail(AMSC_ERR_NOMEM, "device allocation failed");
  return AMSC_OK;
}

int amsc_device_free(void *ptr) {
  if (ptr) AMSC_HIP(hipFree(ptr));
  return AMSC_OK;
}

int amsc_copy_to_device(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return AMSC_OK;
  AMSC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return AMSC_OK;
}

int amsc_copy_on_device(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return AMSC_OK;
  AMSC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                          amsc::current_stream()));
  return AMSC_OK;
}

int amsc_copy_to_host(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return AMSC_OK;
  AMSC_HIP(hipDeviceSynchronize());
  AMSC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return AMSC_OK;
}

}  // extern "C"
