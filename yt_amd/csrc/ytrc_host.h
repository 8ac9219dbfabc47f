// ytrc_host.h -- host-side helpers shared between the library's translation units.
#pragma once
#include <string>

namespace ytrc {

// Records `msg` as the calling thread's last error (ytrc_last_error) and
// returns `code`.
int set_error(int code, const std::string &msg);

}  // namespace ytrc
