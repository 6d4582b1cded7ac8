// Error channel shared by the C-ABI translation units: records the message
// returned by stride_last_error() on the calling thread.
#pragma once

#include <string>

namespace stride {
int set_error(int code, const std::string& msg);
}  // namespace stride
