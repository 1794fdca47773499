// Internal declarations shared by the host set-up code and the HIP engine.
#pragma once
#include <string>
#include <array>

namespace dsmc {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string& msg);

}  // namespace dsmc
