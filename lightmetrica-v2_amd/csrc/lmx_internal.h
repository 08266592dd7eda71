// lmx_internal.h -- helpers shared by the translation units of liblmx.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

namespace lmx {

// Records `msg` as this thread's last error and returns `code`.
int set_error(int code, const std::string& msg);
// Records a HIP runtime failure and returns LMX_EHIP.
int hip_error(hipError_t e, const char* what);

}  // namespace lmx
