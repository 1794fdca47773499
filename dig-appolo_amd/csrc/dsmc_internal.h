// Internal declarations shared by the host set-up code and the HIP engine.
#pragma once
#include <string>
#include <array>
#include "dsmc_engine.h"

namespace dsmc {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string& msg);

// Checks launch settings (fields at DSMC_TUNE_KEEP are not checked); on a bad
// one records "<who>: ..." and returns DSMC_ERR_ARG.
int check_tuning(const dsmc_tuning* t, const char* who);

}  // namespace dsmc
