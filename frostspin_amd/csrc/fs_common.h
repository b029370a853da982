// Internal helpers shared by the frostspin_amd translation units.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "frostspin_amd.h"

namespace fs {

// Thread-local last-error storage behind fs_last_error().
void set_error(const char *fmt, ...);
void clear_error();

// Run `body`, converting any C++ exception into an FS_ERR_* status so nothing
// escapes across the C ABI.
template <class F> int guarded(F &&body) {
  try {
    clear_error();
    return body();
  } catch (const std::bad_alloc &) {
    set_error("host out of memory");
    return FS_ERR_NOMEM;
  } catch (const std::exception &e) {
    set_error("%s", e.what());
    return FS_ERR_INVALID;
  } catch (...) {
    set_error("unknown C++ exception");
    return FS_ERR_INVALID;
  }
}

} // namespace fs

#define FS_CHECK_ARG(cond, ...)                                                \
  do {                                                                         \
    if (!(cond)) {                                                             \
      fs::set_error(__VA_ARGS__);                                              \
      return FS_ERR_INVALID;                                                   \
    }                                                                          \
  } while (0)
