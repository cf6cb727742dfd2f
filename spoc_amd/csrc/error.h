// Error reporting shared by the C-ABI's translation units: report() sets
// the message that spoc_last_error() returns and passes the code through.
#pragma once

#include <exception>
#include <new>
#include <system_error>

#include "spoc_abi.h"

namespace spoc {
int report(int code, const char *fmt, ...);

// Runs the body of a C-ABI entry point.  No C++ exception may cross the C
// boundary (it would end the caller's process): host memory that cannot be
// had, or a host thread the system will not start, becomes SPOC_ENOMEM, any
// other exception SPOC_EINVAL, each with its message.
template <class F>
int guarded(F &&body) {
  try {
    return body();
  } catch (const std::bad_alloc &) {
    return report(SPOC_ENOMEM, "host memory exhausted");
  } catch (const std::system_error &ex) {
    return report(SPOC_ENOMEM, "host resources exhausted: %s", ex.what());
  } catch (const std::exception &ex) {
    return report(SPOC_EINVAL, "%s", ex.what());
  } catch (...) {
    return report(SPOC_EINVAL, "unexpected exception");
  }
}
}  // namespace spoc
