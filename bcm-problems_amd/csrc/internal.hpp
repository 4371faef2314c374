// Internal helpers shared by the C ABI translation units.
#pragma once

#include <string>

namespace bcm {
// Stores the message returned by bcm_last_error() on this thread.
void set_last_error(const std::string& msg);
}  // namespace bcm
