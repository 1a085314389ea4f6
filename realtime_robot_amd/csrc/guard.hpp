// No C++ exception crosses the C ABI: every entry point's body runs inside
// RTR_GUARD_BEGIN / RTR_GUARD_END, which turn a host exception into a status
// code and the thread's error string.
#pragma once
#include <new>

#include "../../include/rtr.h"

namespace rtr {
void set_error(const char* fmt, ...);
}

// return the error code of a failing RTR_* call
#define RTR_TRY(x)                   \
  do {                               \
    int _rc = (x);                   \
    if (_rc != RTR_OK) return _rc;   \
  } while (0)

#define RTR_GUARD_BEGIN try {
#define RTR_GUARD_END                                 \
  }                                                   \
  catch (const std::bad_alloc&) {                     \
    rtr::set_error("host allocation failed");         \
    return RTR_ERR_CAPACITY;                          \
  }                                                   \
  catch (...) {                                       \
    rtr::set_error("unexpected host exception");      \
    return RTR_ERR_DEVICE;                            \
  }
