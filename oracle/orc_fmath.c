/* orc_fmath.c -- TEST INFRASTRUCTURE ONLY: the host build of lmx_fmath.h
 * (the FP32 sine, cosine, tangent, exponential, logarithm and power the
 * kernels and the render oracle share), exported so that tests can pin it
 * against double precision and against the same functions built for the GPU
 * (lmx_fmath_eval in liblmx). */
#include <stddef.h>

#include "orc.h"
#include "lmx_fmath.h"

void orc_fmath_eval(int kind, const float *x, const float *y, size_t n, float *out)
{
    for (size_t i = 0; i < n; i++) {
        switch (kind) {
        case ORC_FM_SIN: out[i] = lmx_sinf(x[i]); break;
        case ORC_FM_COS: out[i] = lmx_cosf(x[i]); break;
        case ORC_FM_TAN: out[i] = lmx_tanf(x[i]); break;
        case ORC_FM_EXP: out[i] = lmx_expf(x[i]); break;
        case ORC_FM_LOG: out[i] = x[i] > 0.0f ? lmx_logf_pos(x[i]) : lmx_u2f(0x7fc00000u); break;
        default: out[i] = lmx_powf(x[i], y[i]); break;
        }
    }
}
