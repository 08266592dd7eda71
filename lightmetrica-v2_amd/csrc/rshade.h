// rshade.h -- device-side scene, sampling and shading routines shared by
// the wavefront path samplers (renderer::pt in lmx_render.hip and the
// bidirectional sampler in lmx_bdpt.hip), and the lmx_scene object.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>

#include "lmx.h"
#include "rmath.h"
#include "wbvh.h"

namespace lmx {

struct RMat {
    int32_t type;
    float R[3];
    float exponent;
    float Le[3];
    float roughness;
    float eta[3], k[3];
    float reserved;
};
static_assert(sizeof(RMat) == 64, "RMat layout");

// One light primitive: its triangles [first, first + count) carry an area
// CDF at cdf[cdf_off .. cdf_off + count] (Distribution1D, dist.h:37-111).
struct RLight {
    uint32_t first, count, cdf_off;
    float inv_area;
};

struct RScene {
    const uint32_t* tris;
    const float* verts;
    const float* normals;      // may be null
    const uint32_t* tri_mat;
    const RMat* mats;
    const int32_t* tri_light;  // light primitive of each triangle, -1 if none
    const RLight* lights;
    const float* light_cdf;
    uint32_t nlight;
    f3 cam_pos, cam_vx, cam_vy, cam_vz;
    float tan_fov, aspect;
    int32_t width, height;
};

// Uniform of (sample, slot) from the table or from Philox.
// Device-side control block.
struct Ctl {
    uint32_t n_trace;          // rays in the queue traced next
    uint32_t n_next;           // continuing paths appended by shade
    uint32_t issued;           // camera samples started
    uint32_t iterations;
    unsigned long long rays;   // extension rays traced so far
    uint32_t n_shadow;         // next-event shadow rays of the iteration
    uint32_t stop;             // time mode: no new camera samples once set
    unsigned long long shadow_rays;
    uint32_t bin[4];           // material-binned shading: paths per class this iteration
    uint32_t ray_out_n;        // rays copied to RParams::ray_out so far
    uint32_t film_scaled;      // the film was rescaled on the device at the render's end
};

// Shading classes of material-binned shading: paths that end at their hit
// (a miss, an emitter without a BSDF) and the three BSDFs.
#define SHADE_CLASSES 4

struct Queue {
    float4* ray;               // 2 per path: (o, tmin), (d, tmax)
    float4* thr;               // throughput rgb
    uint4* meta;               // sample, segment, pixel, flags
};

#define META_DEAD 1u

struct GuideLaunch;            // the learned sampler's state for a shading launch

// Next-event estimation: shadow rays of one iteration (at most one per path
// shaded plus one per camera path started)
struct NeeQ {
    float4* ray;               // 2 per shadow ray
    float4* c;                 // weighted contribution rgb, 0
    uint4* meta;               // sample, pixel, camera-vertex flag, raster x bits
    float* ry;                 // raster y of camera-vertex connections
    float4* hits;
};
#define NEE_NONE 0
#define NEE_DIRECT 1           // renderer::ptdirect: light sampling only
#define NEE_MIS 2              // renderer::ptmis: light and BSDF sampling, balance heuristic

// Parameters of the unidirectional wavefront sampler (plain and guided).
struct RParams {
    int32_t max_nv, min_nv, rr_depth;
    float rr_prob;
    uint2 key;
    uint32_t total;
    int32_t raster_mode;
    const float* table;
    uint32_t table_width;
    float* sample_out;         // 8 floats per sample, or null
    float* film;               // W x H x 3 accumulation
    const GuideLaunch* guide;  // null: BSDF sampling only
    int32_t rr_mode;           // LMX_RR_*
    int32_t nee;               // NEE_*
    uint32_t seg_stride;       // uniform slots per vertex (4; 8 with next-event estimation)
    uint32_t so_stride;        // floats per sample_out record (8; 16 with next-event estimation)
    NeeQ nq;
    const uint32_t* perm = nullptr;   // material-binned shading: class c's paths at perm[c cap ...], or null
    uint32_t perm_cap = 0;
    float4* ray_out = nullptr;        // measurement hook: the first ray_out_cap rays traced
    uint32_t ray_out_cap = 0;
};

// Russian-roulette continuation probability of a path with throughput thr
RD float rr_continue(const RParams& p, f3 thr)
{
    return p.rr_mode == LMX_RR_THROUGHPUT ? fminf(1.0f, fmaxf(thr.x, fmaxf(thr.y, thr.z))) : p.rr_prob;
}

template <class P>
struct Uniforms {
    const P& p;
    uint32_t sample;
    RD float one(uint32_t slot) const
    {
        if (p.table) return p.table[(size_t)sample * p.table_width + slot];
        const uint4 o = philox4x32_10(make_uint4(sample, 0u, slot >> 2, LMX_PHILOX_TAG), p.key);
        const uint32_t w = (slot & 3) == 0 ? o.x : (slot & 3) == 1 ? o.y : (slot & 3) == 2 ? o.z : o.w;
        return u24(w);
    }
    // the four slots of one aligned block (slot % 4 == 0)
    RD float4 block(uint32_t slot) const
    {
        if (p.table) {
            const float* t = p.table + (size_t)sample * p.table_width + slot;
            return make_float4(t[0], t[1], t[2], t[3]);
        }
        const uint4 o = philox4x32_10(make_uint4(sample, 0u, slot >> 2, LMX_PHILOX_TAG), p.key);
        return make_float4(u24(o.x), u24(o.y), u24(o.z), u24(o.w));
    }
};

// Pinhole sensor: raster -> direction (sensor_pinhole.cpp:79-90) and its
// inverse with the behind/outside rejection (:165-185).
RD f3 camera_dir(const RScene& s, float rx, float ry)
{
    const float x = 2.0f * rx - 1.0f, y = 2.0f * ry - 1.0f;
    const f3 we = normalize(mk3(s.aspect * s.tan_fov * x, s.tan_fov * y, -1.0f));
    return add(add(mul(s.cam_vx, we.x), mul(s.cam_vy, we.y)), mul(s.cam_vz, we.z));
}

RD bool camera_raster(const RScene& s, f3 wo, float& rx, float& ry)
{
    const f3 we = mk3(dot(s.cam_vx, wo), dot(s.cam_vy, wo), dot(s.cam_vz, wo));
    if (we.z >= 0.0f) return false;
    rx = (-we.x / we.z / s.tan_fov / s.aspect + 1.0f) * 0.5f;
    ry = (-we.y / we.z / s.tan_fov + 1.0f) * 0.5f;
    return !(rx < 0.0f || rx > 1.0f || ry < 0.0f || ry > 1.0f);
}

// Film_HDR::Splat pixel (film_hdr.cpp:218-223)
RD uint32_t raster_pixel(const RScene& s, float rx, float ry)
{
    int px = (int)(rx * (float)s.width), py = (int)(ry * (float)s.height);
    px = px < 0 ? 0 : (px > s.width - 1 ? s.width - 1 : px);
    py = py < 0 ? 0 : (py > s.height - 1 ? s.height - 1 : py);
    return (uint32_t)py * (uint32_t)s.width + (uint32_t)px;
}

// IntersectionUtils::CreateTriangleIntersection (intersectionutils.h:59-135)
RD void surface_at(const RScene& s, f3 o, f3 d, float t, float b0, float b1, uint32_t id, Frame& f)
{
    const uint32_t i0 = s.tris[3 * id], i1 = s.tris[3 * id + 1], i2 = s.tris[3 * id + 2];
    f.p = add(o, mul(d, t));
    const f3 p1 = ld3(s.verts + 3 * i0), p2 = ld3(s.verts + 3 * i1), p3 = ld3(s.verts + 3 * i2);
    f.gn = normalize(cross(sub(p2, p1), sub(p3, p1)));
    if (s.normals) {
        const f3 n1 = ld3(s.normals + 3 * i0), n2 = ld3(s.normals + 3 * i1), n3 = ld3(s.normals + 3 * i2);
        const float w0 = 1.0f - b0 - b1;
        f.sn = normalize(add(add(mul(n1, w0), mul(n2, b0)), mul(n3, b1)));
        if (isnan(f.sn.x) || isnan(f.sn.y) || isnan(f.sn.z)) f.sn = f.gn;
    } else {
        f.sn = f.gn;
    }
    onb(f.sn, f.u, f.v);
}

RD f3 reflect_sn(const Frame& f, f3 wi) { return sub(mul(f.sn, 2.0f * dot(wi, f.sn)), wi); }

// BSDF queries split as in the reference (renderer_pt.cpp:129-160): sample a
// direction, then pdf (projected solid angle) and f.  Lambertian:
// bsdf_diffuse.cpp:69-117; normalized Phong: f = R (n+2)/(2 pi) cos^n(a),
// lobe sampling cos(a) = u0^(1/(n+1)) about the mirror direction.
// BSDF_CookTorrance with the GGX distribution (bsdf_cooktorrance.cpp:36,
// 184-226, 256-292), in the shading frame.  The shadowing term uses
// |wo.H| for both directions as the reference writes it (:281-282); for the
// half vector wi.H = wo.H, so only rounding can tell the two apart.
RD float ggx_local_tan(f3 v)
{
    const float t = 1.0f - v.z * v.z;
    return t <= 0.0f ? 0.0f : sqrtf(t) / v.z;
}
RD float ggx_D(float r, f3 H)
{
    const float cosH = H.z;
    const float tanH = ggx_local_tan(H);
    if (cosH <= 0.0f) return 0.0f;
    const float t1 = r * r;
    const float t = r * r + tanH * tanH;
    const float t2 = LMX_PI_F * cosH * cosH * cosH * cosH * t * t;
    return t1 / t2;
}
RD f3 ggx_sample_H(float r, float u0, float u1)
{
    const float eps = 1e-7f;                                   // Math::Eps()
    const float a = (1.0f - eps) * u0 + eps;                   // ToOpenClosed
    const float b = (1.0f - 2.0f * eps) * u1 + eps;            // ToOpenOpen
    const float den = sqrtf(1.0f - (1.0f - r * r) * a);
    const float cosT = sqrtf(1.0f - a) / den;
    const float sinT = r * (sqrtf(a) / den);
    const float phi = LMX_PI_F * (2.0f * b - 1.0f);
    float sp, cp;
    lmx_sincosf(phi, &sp, &cp);
    return mk3(sinT * cp, sinT * sp, cosT);
}
RD float ggx_G(f3 li, f3 lo, f3 H)
{
    const float nH = H.z, no = lo.z, ni = li.z;
    const float woH = fabsf(dot(lo, H)), wiH = fabsf(dot(lo, H));
    return fminf(1.0f, fminf(2.0f * nH * no / woH, 2.0f * nH * ni / wiH));
}
RD float ggx_F(float eta, float k, float c)
{
    const float tmp = (eta * eta + k * k) * (c * c);
    const float rParl2 = (tmp - eta * (2.0f * c) + 1.0f) / (tmp + eta * (2.0f * c) + 1.0f);
    const float tmpF = eta * eta + k * k;
    const float rPerp2 = (tmpF - eta * (2.0f * c) + c * c) / (tmpF + eta * (2.0f * c) + c * c);
    return (rParl2 + rPerp2) * 0.5f;
}
RD f3 ggx_f(const RMat& m, f3 li, f3 lo)
{
    const f3 H = normalize(add(li, lo));
    const float D = ggx_D(m.roughness, H);
    const float G = ggx_G(li, lo, H);
    const float c = dot(li, H);
    f3 out;
    out.x = m.R[0] * D * G * ggx_F(m.eta[0], m.k[0], c) / (4.0f * li.z) / lo.z;
    out.y = m.R[1] * D * G * ggx_F(m.eta[1], m.k[1], c) / (4.0f * li.z) / lo.z;
    out.z = m.R[2] * D * G * ggx_F(m.eta[2], m.k[2], c) / (4.0f * li.z) / lo.z;
    return out;
}

RD bool bsdf_sample(const RMat& m, const Frame& f, f3 wi, float u0, float u1, f3& wo)
{
    const f3 li = to_local(f, wi);
    if (li.z <= 0.0f) return false;
    if (m.type == LMX_MAT_DIFFUSE) {
        wo = to_world(f, cosine_hemisphere(u0, u1));
        return true;
    }
    if (m.type == LMX_MAT_PHONG) {
        const f3 r = reflect_sn(f, wi);
        f3 ru, rv;
        onb(r, ru, rv);
        const float ca = lmx_powf(u0, 1.0f / (m.exponent + 1.0f));
        const float sa = sqrtf(fmaxf(0.0f, 1.0f - ca * ca));
        const float phi = 2.0f * LMX_PI_F * u1;
        float sp, cp;
        lmx_sincosf(phi, &sp, &cp);
        const f3 l = mk3(sa * cp, sa * sp, ca);
        wo = add(add(mul(ru, l.x), mul(rv, l.y)), mul(r, l.z));
        return true;
    }
    if (m.type == LMX_MAT_GGX) {
        const f3 H = ggx_sample_H(m.roughness, u0, u1);
        const f3 ni = neg(li);
        const f3 lo = sub(ni, mul(H, 2.0f * dot(ni, H)));
        if (lo.z <= 0.0f) return false;
        wo = to_world(f, lo);
        return true;
    }
    return false;
}

RD float bsdf_pdf(const RMat& m, const Frame& f, f3 wi, f3 wo)
{
    const f3 li = to_local(f, wi), lo = to_local(f, wo);
    if (li.z <= 0.0f || lo.z <= 0.0f) return 0.0f;
    if (m.type == LMX_MAT_DIFFUSE) return LMX_INV_PI_F;
    if (m.type == LMX_MAT_PHONG) {
        const float ca = fmaxf(0.0f, dot(reflect_sn(f, wi), wo));
        const float pdf_sa = (m.exponent + 1.0f) * (0.5f * LMX_INV_PI_F) * lmx_powf(ca, m.exponent);
        return pdf_sa / lo.z;
    }
    if (m.type == LMX_MAT_GGX) {
        const f3 H = normalize(add(li, lo));
        const float D = ggx_D(m.roughness, H);
        return D * H.z / (4.0f * dot(lo, H)) / lo.z;
    }
    return 0.0f;
}

// BSDFUtils::ShadingNormalCorrection for transport EL (bsdfutils.h:55-66)
RD float shading_correction(const Frame& f, f3 wi, f3 wo)
{
    const f3 li = to_local(f, wi), lo = to_local(f, wo);
    const float wiDotNg = dot(wi, f.gn), woDotNg = dot(wo, f.gn);
    if (wiDotNg * li.z <= 0.0f || woDotNg * lo.z <= 0.0f) return 0.0f;
    return 1.0f;
}

RD f3 bsdf_eval(const RMat& m, const Frame& f, f3 wi, f3 wo)
{
    const f3 li = to_local(f, wi), lo = to_local(f, wo);
    if (li.z <= 0.0f || lo.z <= 0.0f) return mk3(0.0f, 0.0f, 0.0f);
    const float corr = shading_correction(f, wi, wo);
    if (m.type == LMX_MAT_DIFFUSE)
        return mk3(m.R[0] * LMX_INV_PI_F * corr, m.R[1] * LMX_INV_PI_F * corr, m.R[2] * LMX_INV_PI_F * corr);
    if (m.type == LMX_MAT_PHONG) {
        const float ca = fmaxf(0.0f, dot(reflect_sn(f, wi), wo));
        const float k = (m.exponent + 2.0f) * (0.5f * LMX_INV_PI_F) * lmx_powf(ca, m.exponent);
        return mk3(m.R[0] * k * corr, m.R[1] * k * corr, m.R[2] * k * corr);
    }
    if (m.type == LMX_MAT_GGX) {
        const f3 v = ggx_f(m, li, lo);
        return mk3(v.x * corr, v.y * corr, v.z * corr);
    }
    return mk3(0.0f, 0.0f, 0.0f);
}

// BSDF f for either transport direction: BSDFUtils::ShadingNormalCorrection
// (bsdfutils.h:55-66) is 1 for EL and wiDotNs woDotNg / (woDotNs wiDotNg)
// for LE.
RD f3 bsdf_eval_t(const RMat& m, const Frame& f, f3 wi, f3 wo, bool le)
{
    const f3 li = to_local(f, wi), lo = to_local(f, wo);
    if (li.z <= 0.0f || lo.z <= 0.0f) return mk3(0.0f, 0.0f, 0.0f);
    const float wiDotNg = dot(wi, f.gn), woDotNg = dot(wo, f.gn);
    const float wiDotNs = li.z, woDotNs = lo.z;
    float corr;
    if (wiDotNg * wiDotNs <= 0.0f || woDotNg * woDotNs <= 0.0f) corr = 0.0f;
    else if (le) corr = wiDotNs * woDotNg / (woDotNs * wiDotNg);
    else corr = 1.0f;
    if (m.type == LMX_MAT_DIFFUSE)
        return mk3(m.R[0] * LMX_INV_PI_F * corr, m.R[1] * LMX_INV_PI_F * corr, m.R[2] * LMX_INV_PI_F * corr);
    if (m.type == LMX_MAT_PHONG) {
        const float ca = fmaxf(0.0f, dot(reflect_sn(f, wi), wo));
        const float k = (m.exponent + 2.0f) * (0.5f * LMX_INV_PI_F) * lmx_powf(ca, m.exponent);
        return mk3(m.R[0] * k * corr, m.R[1] * k * corr, m.R[2] * k * corr);
    }
    if (m.type == LMX_MAT_GGX) {
        const f3 v = ggx_f(m, li, lo);
        return mk3(v.x * corr, v.y * corr, v.z * corr);
    }
    return mk3(0.0f, 0.0f, 0.0f);
}

RD bool is_scatterer(const RMat& m)
{
    return m.type == LMX_MAT_DIFFUSE || m.type == LMX_MAT_PHONG || m.type == LMX_MAT_GGX;
}

// Pinhole importance W_e(wo) = 1 / (A cos^3) with A = 4 tan^2(fov/2) aspect,
// zero outside the film (sensor_pinhole.cpp:137-154).  It is also the
// sensor's direction pdf in projected solid angle.
RD float importance(const RScene& s, f3 wo)
{
    float rx, ry;
    if (!camera_raster(s, wo, rx, ry)) return 0.0f;
    const float cos_t = -dot(s.cam_vz, wo);
    const float inv = 1.0f / cos_t;
    const float A = s.tan_fov * s.tan_fov * s.aspect * 4.0f;
    return inv * inv * inv / A;
}

RD f3 dir_to(f3 from, f3 to) { return normalize(sub(to, from)); }

// RenderUtils::GeometryTerm (renderutils.h:46-56)
RD float gterm(f3 pa, f3 na, bool dega, f3 pb, f3 nb, bool degb)
{
    f3 d = sub(pb, pa);
    const float l2 = dot(d, d);
    const float l = sqrtf(l2);
    d = mk3(d.x / l, d.y / l, d.z / l);
    float t = 1.0f;
    if (!dega) t *= fabsf(dot(na, d));
    if (!degb) t *= fabsf(dot(nb, neg(d)));
    return t / l2;
}
RD float pos_pdf_light(const RScene& s, int l) { return s.lights[l].inv_area * (1.0f / (float)s.nlight); }

// Area-light emission: Le one-sided (light_area.cpp:105-110); direction pdf
// 1/pi in projected solid angle when above the surface (:84-89).
RD bool emits_towards(const Frame& f, f3 wo) { return to_local(f, wo).z > 0.0f; }

// Light point: Distribution1D::SampleReuse + UniformSampleTriangle
// (triangleutils.h:73-120)
RD void sample_light_point(const RScene& s, int l, float u0, float u1, f3& p, f3& gn, uint32_t& tri)
{
    const RLight L = s.lights[l];
    const float* cdf = s.light_cdf + L.cdf_off;
    uint32_t lo = 0, hi = L.count + 1;          // upper_bound over cdf[0..count]
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] <= u0) lo = mid + 1; else hi = mid;
    }
    int i = (int)lo - 1;
    i = i < 0 ? 0 : (i > (int)L.count - 1 ? (int)L.count - 1 : i);
    const float u2 = (u0 - cdf[i]) / (cdf[i + 1] - cdf[i]);
    float b0, b1;
    uniform_triangle(u2, u1, b0, b1);
    tri = L.first + (uint32_t)i;
    const uint32_t i0 = s.tris[3 * tri], i1 = s.tris[3 * tri + 1], i2 = s.tris[3 * tri + 2];
    const f3 p1 = ld3(s.verts + 3 * i0), p2 = ld3(s.verts + 3 * i1), p3 = ld3(s.verts + 3 * i2);
    p = add(add(mul(p1, 1.0f - b0 - b1), mul(p2, b0)), mul(p3, b1));
    gn = normalize(cross(sub(p2, p1), sub(p3, p1)));
}

}  // namespace lmx

struct lmx_scene;
// The unidirectional wavefront loop (lmx_render.hip): generation, trace,
// shade (guided when rp.guide is set), graph-replayed until every sample is
// done.  `cap` paths in flight.
// film_n > 0: once every sample is issued and every path is done, the
// film's first film_n floats are scaled by film_w on the device (sets
// Ctl::film_scaled; the readback after it ends the loop).
int wavefront_render(lmx_scene* sc, const lmx::RParams& rp, uint32_t cap, hipStream_t st, uint64_t* rays,
                     uint32_t* iters, double render_time = 0.0, uint64_t* samples = nullptr, size_t film_n = 0,
                     float film_w = 0.0f);
// The guided shading step of one iteration (lmx_guide.hip).
int launch_guided_shade(lmx_scene* sc, const lmx::RParams& rp, lmx::Queue in, lmx::Queue out, hipStream_t st);

namespace lmx {

// Reserves `count` consecutive queue slots per thread of a workgroup of NW
// waves with one atomic per workgroup: exclusive prefix sum over each wave,
// the wave totals combined in LDS, their sum added to the counter by thread 0.
// Every thread of the workgroup must call it; `lds` holds NW + 1 words.
template <int NW>
RD uint32_t block_reserve(uint32_t* counter, uint32_t count, uint32_t* lds)
{
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = count;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off);
        if (lane >= (uint32_t)off) incl += v;
    }
    if (lane == 63u) lds[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < NW; w++) { const uint32_t c = lds[w]; lds[w] = tot; tot += c; }
        lds[NW] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint32_t r = lds[NW] + lds[wave] + incl - count;
    __syncthreads();
    return r;
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // n records the size only once the allocation succeeded, so a reuse
    // check on n never hands a null buffer to a kernel after a failure; a
    // failed hipMalloc's error is also taken off the runtime's last-error
    // slot, which the caller's next launch check would otherwise report
    hipError_t alloc(size_t count)
    {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = 0;
        T* q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(1, count) * sizeof(T));
        if (e != hipSuccess) { (void)hipGetLastError(); return e; }
        p = q;
        n = count;
        return e;
    }
};

}  // namespace lmx

using lmx::Ctl;
using lmx::DevBuf;
using lmx::RLight;
using lmx::RMat;
using lmx::RScene;

// Control block of the bidirectional sampler (one batch of samples at a time).
struct BCtl {
    uint32_t n_trace;          // subpath rays queued for the next trace
    uint32_t n_next;           // continuations appended by the shade step
    uint32_t n_shadow;         // shadow-queue slots of the batch (live and dead)
    uint32_t batch_first;      // first sample of the current batch
    unsigned long long rays;   // extension rays traced
    unsigned long long shadow_rays;   // live connection (shadow) rays traced
    uint32_t pass_first;       // first sample index of the current pass (photons or eye samples)
    float radius;              // vertex merging: merge radius of the pass
    float cell;                // vertex merging: grid cell edge (2 radius)
    float c_merge;             // vertex merging: N (pi radius^2)^2
    unsigned long long merges; // vertex merging: merged paths evaluated
    uint32_t n_mq;             // vertex merging: eye vertices queued for merging in the batch
    uint32_t pad;
    unsigned long long dropped;   // vertex merging: merges whose contribution was not finite (reported)
    uint32_t cap_n[2];         // measurement hook: extension and shadow rays copied out so far
    uint32_t conn_next;        // connection rounds: first sample of the batch not yet connected
    uint32_t conn_lo, conn_hi; // connection rounds: the samples [lo, hi) of the current round
    uint32_t conn_base;        // connection rounds: slot offset of sample lo
};

// The scene object behind the C ABI: accel, shading data on the device and
// the wavefront buffers of the samplers (reused across renders).
struct lmx_scene {
    lmx_accel* accel = nullptr;
    DevBuf<uint32_t> tris, tri_mat;
    DevBuf<int32_t> tri_light;      // light primitive of each triangle, -1 if none
    DevBuf<float> verts, normals;
    DevBuf<RMat> mats;
    DevBuf<RLight> lights;
    DevBuf<float> light_cdf;
    RScene rs{};
    int width = 0, height = 0;
    float bounds[6] = {0, 0, 0, 0, 0, 0};   // scene AABB: min xyz, max xyz
    int cus = 0;
    // wavefront buffers, reused across renders of the same capacity
    uint32_t cap = 0;
    DevBuf<float4> ray[2], thr[2];
    DevBuf<uint4> meta[2];
    DevBuf<float4> hits;
    // next-event shadow queue (2 x cap entries), allocated on first use
    DevBuf<float4> nee_ray, nee_c, nee_hits;
    DevBuf<uint4> nee_meta;
    DevBuf<float> nee_ry;
    uint32_t nee_cap = 0;
    DevBuf<uint32_t> perm;              // material-binned shading, SHADE_CLASSES x cap
    DevBuf<Ctl> ctl;
    Ctl* h_ctl = nullptr;      // pinned
    hipGraphExec_t wf_exec = nullptr;   // the wavefront iterations' graph, updated in place per render
    hipStream_t cap_stream = nullptr;   // records that graph and runs nothing
    uint32_t tuned_cap = 0;
    // bidirectional sampler buffers (sized for bd_batch samples, bd_maxv vertices)
    uint32_t bd_batch = 0, bd_maxv = 0, bd_conn = 0;   // bd_conn: shadow-queue slots per sample and round
    uint32_t bd_nconn = 0;              // shadow-queue slots of one connection round
    uint32_t bd_tuned_batch = 0, bd_tuned_key = 0;      // autotuned batch for one (variant, max_num_vertices)
    uint64_t bd_tuned_limit = 0;        // largest candidate the tuning run allowed (render size rounded up)
    DevBuf<float4> bd_vert[2];          // per subpath vertex: 4 float4 (see lmx_bdpt.hip)
    DevBuf<float> bd_V;                 // per light subpath vertex (bdpt variant)
    DevBuf<uint32_t> bd_nsub[2];
    DevBuf<uint32_t> bd_epix;
    DevBuf<float2> bd_eraster;
    DevBuf<float4> bd_ray[2], bd_thr[2];
    DevBuf<uint4> bd_meta[2];
    DevBuf<float4> bd_hits;
    DevBuf<float4> bd_sray, bd_sc, bd_saux, bd_shits;
    DevBuf<uint4> bd_smeta;
    DevBuf<uint32_t> bd_off, bd_part;   // per-sample connection counts -> slot offsets, and scan partials
    DevBuf<uint32_t> bd_dmask;          // bdpt: per-sample mask of the eye vertices with a direct strategy
    DevBuf<BCtl> bd_ctl;
    BCtl* bd_hctl = nullptr;      // pinned
    // vertex merging: photon subpaths and the hash grid over their vertices
    uint32_t vc_np = 0, vc_maxv = 0, vc_H = 0;
    DevBuf<float4> vc_vert;             // photon subpath vertices, as bd_vert[1]
    DevBuf<uint32_t> vc_nsub;
    DevBuf<float4> vc_items;            // position, photon * maxv + vertex
    DevBuf<uint32_t> vc_start, vc_cursor, vc_part;
    DevBuf<uint32_t> vc_mq;             // merge queue: eye sample * 32 + vertex, bd_batch x (bd_maxv - 1)
    DevBuf<float4> vc_ray[2], vc_thr[2], vc_hits;   // photon-phase ray queues (vc_np rays each)
    DevBuf<uint4> vc_meta[2];
    DevBuf<float4> ppm_st;              // renderer::ppm measurement-point state, 2 per point
    ~lmx_scene()
    {
        if (wf_exec) (void)hipGraphExecDestroy(wf_exec);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        if (h_ctl) (void)hipHostFree(h_ctl);
        if (bd_hctl) (void)hipHostFree(bd_hctl);
        lmx_accel_free(accel);
    }
};
