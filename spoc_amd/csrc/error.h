// Error reporting shared by the C-ABI's translation units: report() sets
// the message that spoc_last_error() returns and passes the code through.
#pragma once

namespace spoc {
int report(int code, const char *fmt, ...);
}  // namespace spoc
