// Thread-local "last error" text for context-free C-ABI calls.
#ifndef DMM_HOST_ERROR_H_
#define DMM_HOST_ERROR_H_

#include <cstdarg>
#include <cstdio>
#include <string>

namespace dmm {

std::string& thread_error();

inline int fail(std::string* slot, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (slot) *slot = buf;
  thread_error() = buf;
  return code;
}

}  // namespace dmm

#endif  // DMM_HOST_ERROR_H_
