// Thread-local last-error string behind dphd_last_error().
#pragma once

#include <string>

namespace dphd {
void set_error(const std::string& message);
}  // namespace dphd
