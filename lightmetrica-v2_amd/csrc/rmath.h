// rmath.h -- FP32 vector math, shading frames, BSDF sampling and Philox for
// the wavefront path samplers.  Formulas and operation order follow the
// reference where it defines them (cited per function); square roots and
// divisions are IEEE-rounded (the reference's SSE Normalize uses rsqrtps,
// include/lightmetrica/math.h:1872-1875, which is not reproducible).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "lmx_fmath.h"

#define RD __device__ __forceinline__

namespace lmx {

#define LMX_PI_F 3.14159265358979323846f
#define LMX_INV_PI_F 0.318309886183790671538f     // (float)(1/pi), Math::InvPi<float>()
#define LMX_EPS_ISECT 1e-4f                        // Math::EpsIsect(), math.h:1668
#define LMX_FLT_MAX 3.402823466e+38f               // Math::Inf(), math.h:1663

struct f3 { float x, y, z; };
__host__ __device__ __forceinline__ f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
RD f3 add(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
RD f3 sub(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
RD f3 mul(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
RD f3 neg(f3 a) { return mk3(-a.x, -a.y, -a.z); }
// _mm_dp_ps order: x and y products summed first (math.h:1728-1731)
RD float dot(f3 a, f3 b) { const float xy = a.x * b.x + a.y * b.y; return xy + a.z * b.z; }
RD f3 cross(f3 a, f3 b)   // math.h:1784-1787
{
    return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
RD f3 normalize(f3 a)
{
    const float l = sqrtf(dot(a, a));
    return mk3(a.x / l, a.y / l, a.z / l);
}
RD f3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

// OrthonormalBasis (math.h:2355-2359): c from the larger of |a.x|, |a.y|,
// b = c x a; the shading frame is ToWorld = Mat3(b, c, a)
// (surfacegeometry.h:58-59).
RD void onb(f3 a, f3& b, f3& c)
{
    c = fabsf(a.x) > fabsf(a.y) ? normalize(mk3(a.z, 0.0f, -a.x)) : normalize(mk3(0.0f, a.z, -a.y));
    b = cross(c, a);
}

struct Frame { f3 p, gn, sn, u, v; };
RD f3 to_world(const Frame& f, f3 l) { return add(add(mul(f.u, l.x), mul(f.v, l.y)), mul(f.sn, l.z)); }
RD f3 to_local(const Frame& f, f3 w) { return mk3(dot(f.u, w), dot(f.v, w), dot(f.sn, w)); }

// Sampler::UniformConcentricDiskSample / CosineSampleHemisphere
// (include/lightmetrica/sampler.h:44-67)
RD void concentric_disk(float u0, float u1, float& ox, float& oy)
{
    const float vx = 2.0f * u0 - 1.0f, vy = 2.0f * u1 - 1.0f;
    if (vx == 0.0f && vy == 0.0f) { ox = 0.0f; oy = 0.0f; return; }
    const float pi4 = LMX_PI_F * 0.25f;
    float r, theta;
    if (vx > -vy) {
        if (vx > vy) { r = vx; theta = pi4 * vy / vx; }
        else { r = vy; theta = pi4 * (2.0f - vx / vy); }
    } else {
        if (vx < vy) { r = -vx; theta = pi4 * (4.0f + vy / vx); }
        else { r = -vy; theta = pi4 * (6.0f - vx / vy); }
    }
    float st, ct;
    lmx_sincosf(theta, &st, &ct);
    ox = r * ct;
    oy = r * st;
}

RD f3 cosine_hemisphere(float u0, float u1)
{
    float x, y;
    concentric_disk(u0, u1, x, y);
    return mk3(x, y, sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y)));
}

// Sampler::UniformSampleTriangle (sampler.h:97-101)
RD void uniform_triangle(float u0, float u1, float& b0, float& b1)
{
    const float s = sqrtf(fmaxf(0.0f, u0));
    b0 = 1.0f - s;
    b1 = u1 * s;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC 2011).  The uniform of
// (seed, sample, slot) is word slot%4 of the block with counter
// (sample lo, sample hi, slot/4, LMX_PHILOX_TAG) under key (seed lo, seed hi),
// top 24 bits scaled to [0, 1).

#define LMX_PHILOX_TAG 0x6c6d7870u

RD uint4 philox4x32_10(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k.x += 0x9E3779B9u; k.y += 0xBB67AE85u; }
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    }
    return c;
}

RD float u24(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

}  // namespace lmx
