// status.hpp -- error text of the C ABI, shared by its translation units.
#pragma once

#include <string>

namespace bine {

// Store msg as the calling thread's bine_last_error() text; returns code.
int record_error(int code, const std::string& msg);

}  // namespace bine
