// lmx_internal.h -- helpers shared by the translation units of liblmx.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include "lmx.h"
#include "wbvh.h"

namespace lmx {

// Records `msg` as this thread's last error and returns `code`.
int set_error(int code, const std::string& msg);
// Records a HIP runtime failure and returns LMX_EHIP.
int hip_error(hipError_t e, const char* what);

// Enqueues closest-hit traversal of rays[0, n) on stream s, where n is *n_dev
// when n_dev is non-null (a count written by an earlier kernel; n_max bounds
// it and sizes the grid), else n.  With anyhit, a ray stops at the first
// triangle it hits (occlusion queries).
// With cursors_set the launch's cursors were already set on the stream, by a
// kernel that ran trace_cursors_set (trace_cursors.h) on the counter block
// and grid trace_shape gives for n_max.
int launch_trace(lmx_accel* a, const WRay* rays, uint32_t n, const uint32_t* n_dev, uint32_t n_max,
                 WHit* hits, hipStream_t s, bool anyhit = false, bool cursors_set = false);
// The counter block and workgroup count of a trace over at most n_max rays.
int trace_shape(lmx_accel* a, uint32_t n_max, uint32_t** counters, uint32_t* grid);

}  // namespace lmx
