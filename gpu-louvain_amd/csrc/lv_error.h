// Last-error message shared by all C-ABI entry points of the library.
#pragma once

#include <string>

namespace lv {
void set_error(const std::string &msg);
}  // namespace lv
