// Exception barrier of the C-ABI.  Every int-returning extern "C" entry point
// wraps its body in ALPSG_TRY ... ALPSG_CATCH, so that a C++ exception from
// inside the library (std::bad_alloc from a host container, for instance)
// becomes a status code and the message of alpsg_last_error() instead of
// reaching the caller's frames, which std::terminate would end.
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "../../include/alpsg.h"

int alpsg_internal_fail(int code, const std::string &msg);

#define ALPSG_TRY try {
#define ALPSG_CATCH                                                                            \
  }                                                                                            \
  catch (const std::bad_alloc &) {                                                             \
    return alpsg_internal_fail(ALPSG_E_ARG, "host memory exhausted (request too large)");      \
  }                                                                                            \
  catch (const std::length_error &) {                                                          \
    return alpsg_internal_fail(ALPSG_E_ARG, "host container size exceeded (request too large)"); \
  }                                                                                            \
  catch (const std::exception &e_) {                                                           \
    return alpsg_internal_fail(ALPSG_E_DEVICE, std::string("internal error: ") + e_.what());  \
  }                                                                                            \
  catch (...) {                                                                                \
    return alpsg_internal_fail(ALPSG_E_DEVICE, "internal error");                              \
  }
