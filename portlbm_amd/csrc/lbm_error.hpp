// Error reporting shared by the translation units of the C ABI: records the
// message returned by lbm_last_error() for the calling thread.
#pragma once

#include <string>

int lbm_set_error(int code, const std::string &msg);
