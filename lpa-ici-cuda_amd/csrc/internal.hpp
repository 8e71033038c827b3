// Internal helpers shared by the C ABI translation units.
#ifndef LPA_ICI_INTERNAL_HPP
#define LPA_ICI_INTERNAL_HPP

#include <string>

#include "lpa_ici.h"

namespace lpa {

// Record `msg` as this thread's lpa_last_error() text and return `st`.
lpa_status fail(lpa_status st, const std::string &msg);
// Clear this thread's error text.
void clear_error();

}  // namespace lpa

#endif
