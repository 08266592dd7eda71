/*
 * orc_render.c -- CPU oracle of the path samplers: the reference's
 * renderer::pt per-sample loop (src/liblightmetrica/renderer/renderer_pt.cpp:
 * 64-231) with its pinhole sensor, area lights, Lambertian BSDF and film
 * splat, plus the normalized Phong BSDF the benchmark scenes need (the
 * reference has none; closed forms pin it, see orc.h).
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h).
 *
 * Uniform numbers come either from a caller table or from Philox4x32-10
 * (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as easy as 1, 2,
 * 3", SC 2011), addressed by (seed, sample, slot).  The slot layout per
 * sample is fixed (ORC_PT_* below) so that a table and the generator feed the
 * same consumption order: the reference's own order of two Next2D() calls in
 * one argument list is unspecified (renderer_pt.cpp:84,136); here it is left
 * to right.
 *
 * Sines, cosines, the tangent and powers come from lmx_fmath.h, the FP32
 * functions the GPU kernels are compiled with (the header is pinned against
 * double precision by tests/test_fmath.py), so that oracle and kernels round
 * every operation of a sample alike; the vendor libraries (glibc here, the
 * ROCm device library there) may differ by an ulp.  The reference's own
 * sampler mappings, which tests pin against vectors from its headers, stay
 * in orc_isect.c with the C library's functions.
 */
#include "orc.h"
#include "lmx_fmath.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define PI_F ((float)3.141592653589793238462643383279502884e+00)
#define INV_PI_F ((float)(1.0 / 3.141592653589793238462643383279502884e+00))
#define EPS_ISECT 1e-4f              /* Math::EpsIsect(), math.h:1668 */

/* ---------------------------------------------------------------------- */
/* Philox4x32-10 */

static inline void mulhilo(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo)
{
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo(0xCD9E8D57u, c2, &hi1, &lo1);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* uniform in [0,1) from the top 24 bits */
static inline float u24(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

float orc_uniform(uint64_t seed, uint64_t sample, uint32_t slot)
{
    const uint32_t ctr[4] = {(uint32_t)sample, (uint32_t)(sample >> 32), slot >> 2, ORC_PHILOX_TAG};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    return u24(o[slot & 3]);
}

/* ---------------------------------------------------------------------- */
/* FP32 vector helpers (dot sums x,y first, then z) */

typedef struct { float x, y, z; } v3;
static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 vneg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float vdot(v3 a, v3 b) { const float xy = a.x * b.x + a.y * b.y; return xy + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b)
{
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
/* Math::Normalize with an exact square root (the reference's SSE build uses
 * rsqrtps, math.h:1872-1875, whose error is CPU dependent) */
static inline v3 vnorm(v3 a)
{
    const float l = sqrtf(vdot(a, a));
    return mk(a.x / l, a.y / l, a.z / l);
}
static inline v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }

/* Shading frame: OrthonormalBasis (math.h:2355-2359) and
 * ToWorld = Mat3(dpdu, dpdv, sn) (surfacegeometry.h:58-59) */
typedef struct { v3 p, gn, sn, u, v; } frame;

static inline void onb(v3 a, v3 *b, v3 *c)
{
    *c = fabsf(a.x) > fabsf(a.y) ? vnorm(mk(a.z, 0.0f, -a.x)) : vnorm(mk(0.0f, a.z, -a.y));
    *b = vcross(*c, a);
}
static inline v3 to_world(const frame *f, v3 l)
{
    return vadd(vadd(vmul(f->u, l.x), vmul(f->v, l.y)), vmul(f->sn, l.z));
}
static inline v3 to_local(const frame *f, v3 w)
{
    return mk(vdot(f->u, w), vdot(f->v, w), vdot(f->sn, w));
}

/* ---------------------------------------------------------------------- */
/* scene */

struct orc_scene {
    orc_scene_desc d;
    orc_qbvh *q;
    /* per light primitive: area CDF over its triangles (dist.h) */
    float **cdf;
    float *inv_area;
};

orc_scene *orc_scene_new(const orc_scene_desc *d)
{
    orc_scene *s = (orc_scene *)calloc(1, sizeof(orc_scene));
    s->d = *d;
    s->q = orc_qbvh_build(d->verts, d->tris, d->ntri);
    s->cdf = (float **)calloc(d->nlight ? d->nlight : 1, sizeof(float *));
    s->inv_area = (float *)calloc(d->nlight ? d->nlight : 1, sizeof(float));
    for (uint32_t l = 0; l < d->nlight; l++) {
        const uint32_t first = d->light_prims[2 * l], n = d->light_prims[2 * l + 1];
        /* TriangleUtils::CreateTriangleAreaDist, triangleutils.h:47-70 */
        float *cdf = (float *)calloc(n + 1, sizeof(float));
        float sum = 0.0f;
        cdf[0] = 0.0f;
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t *t = &d->tris[3 * (first + i)];
            const v3 p1 = ld3(&d->verts[3 * t[0]]), p2 = ld3(&d->verts[3 * t[1]]), p3 = ld3(&d->verts[3 * t[2]]);
            const v3 c = vcross(vsub(p2, p1), vsub(p3, p1));
            const float area = sqrtf(vdot(c, c)) * 0.5f;
            cdf[i + 1] = cdf[i] + area;          /* Distribution1D::Add, dist.h:44-48 */
            sum += area;
        }
        const float inv = 1.0f / cdf[n];           /* Distribution1D::Normalize, dist.h:52-60 */
        for (uint32_t i = 0; i <= n; i++) cdf[i] *= inv;
        s->cdf[l] = cdf;
        s->inv_area[l] = 1.0f / sum;
    }
    return s;
}

void orc_scene_free(orc_scene *s)
{
    if (!s) return;
    orc_qbvh_free(s->q);
    for (uint32_t l = 0; l < s->d.nlight; l++) free(s->cdf[l]);
    free(s->cdf);
    free(s->inv_area);
    free(s);
}

/* Scene3::Intersect: closest hit on [1e-4, FLT_MAX] (scene3.cpp:458-478) */
static int scene_intersect(const orc_scene *s, v3 o, v3 d, orc_hit *h)
{
    orc_ray r;
    r.o[0] = o.x; r.o[1] = o.y; r.o[2] = o.z; r.tmin = EPS_ISECT;
    r.d[0] = d.x; r.d[1] = d.y; r.d[2] = d.z; r.tmax = FLT_MAX;
    orc_qbvh_intersect_ex(s->q, &r, 1, h, ORC_TEST_WATERTIGHT, ORC_TIE_MIN_ID, 1);
    return h->id != 0xffffffffu;
}

/* IntersectionUtils::CreateTriangleIntersection (intersectionutils.h:59-135):
 * p = o + d t, gn from the face, sn interpolated from vertex normals (NaN ->
 * gn) or gn, then the shading frame. */
static void surface_at(const orc_scene *s, v3 o, v3 d, const orc_hit *h, frame *f)
{
    const uint32_t *t = &s->d.tris[3 * h->id];
    f->p = vadd(o, vmul(d, h->t));
    const v3 p1 = ld3(&s->d.verts[3 * t[0]]), p2 = ld3(&s->d.verts[3 * t[1]]), p3 = ld3(&s->d.verts[3 * t[2]]);
    f->gn = vnorm(vcross(vsub(p2, p1), vsub(p3, p1)));
    if (s->d.normals) {
        const v3 n1 = ld3(&s->d.normals[3 * t[0]]), n2 = ld3(&s->d.normals[3 * t[1]]), n3 = ld3(&s->d.normals[3 * t[2]]);
        const float w0 = 1.0f - h->b0 - h->b1;
        f->sn = vnorm(vadd(vadd(vmul(n1, w0), vmul(n2, h->b0)), vmul(n3, h->b1)));
        if (isnan(f->sn.x) || isnan(f->sn.y) || isnan(f->sn.z)) f->sn = f->gn;
    } else {
        f->sn = f->gn;
    }
    onb(f->sn, &f->u, &f->v);
}

/* ---------------------------------------------------------------------- */
/* BSDFs.  Queries follow the reference's split: sample a direction, then
 * evaluate pdf (projected solid angle) and f separately
 * (renderer_pt.cpp:129-160). */

/* Sampler::UniformConcentricDiskSample and CosineSampleHemisphere
 * (sampler.h:44-76), with the sine and cosine of lmx_fmath.h */
static void cos_hemisphere(const float u[2], float out[3])
{
    const float vx = 2.0f * u[0] - 1.0f, vy = 2.0f * u[1] - 1.0f;
    float x = 0.0f, y = 0.0f;
    if (!(vx == 0.0f && vy == 0.0f)) {
        const float pi4 = PI_F * 0.25f;
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
        x = r * ct;
        y = r * st;
    }
    out[0] = x; out[1] = y;
    out[2] = sqrtf(fmaxf(0.0f, 1.0f - x * x - y * y));
}

/* BSDFUtils::ShadingNormalCorrection for transport EL (bsdfutils.h:55-66) */
static inline float shading_correction(const frame *f, v3 wi, v3 wo)
{
    const v3 li = to_local(f, wi), lo = to_local(f, wo);
    const float wiDotNg = vdot(wi, f->gn), woDotNg = vdot(wo, f->gn);
    const float wiDotNs = li.z, woDotNs = lo.z;
    if (wiDotNg * wiDotNs <= 0.0f || woDotNg * woDotNs <= 0.0f) return 0.0f;
    return 1.0f;
}

/* Reflection of wi about the shading normal, and the Phong lobe frame */
static inline v3 reflect_sn(const frame *f, v3 wi)
{
    return vsub(vmul(f->sn, 2.0f * vdot(wi, f->sn)), wi);
}

/* BSDF_CookTorrance, GGX distribution (bsdf_cooktorrance.cpp:36,184-226,
 * 256-292), in the shading frame; the shadowing term takes |wo.H| for both
 * directions as written there (:281-282). */
static float ggx_tan(v3 v)
{
    const float t = 1.0f - v.z * v.z;                          /* Math::LocalTan, math.h:1903-1907 */
    return t <= 0.0f ? 0.0f : sqrtf(t) / v.z;
}
static float ggx_D(float r, v3 H)
{
    const float cosH = H.z, tanH = ggx_tan(H);
    if (cosH <= 0.0f) return 0.0f;
    const float t1 = r * r;
    const float t = r * r + tanH * tanH;
    const float t2 = PI_F * cosH * cosH * cosH * cosH * t * t;
    return t1 / t2;
}
static v3 ggx_sample_H(float r, const float u[2])
{
    const float eps = 1e-7f;
    const float a = (1.0f - eps) * u[0] + eps, b = (1.0f - 2.0f * eps) * u[1] + eps;
    const float den = sqrtf(1.0f - (1.0f - r * r) * a);
    const float cosT = sqrtf(1.0f - a) / den, sinT = r * (sqrtf(a) / den);
    const float phi = PI_F * (2.0f * b - 1.0f);
    float sp, cp;
    lmx_sincosf(phi, &sp, &cp);
    return mk(sinT * cp, sinT * sp, cosT);
}
static float ggx_G(v3 li, v3 lo, v3 H)
{
    const float nH = H.z, no = lo.z, ni = li.z;
    const float woH = fabsf(vdot(lo, H)), wiH = fabsf(vdot(lo, H));
    return fminf(1.0f, fminf(2.0f * nH * no / woH, 2.0f * nH * ni / wiH));
}
static float ggx_F(float eta, float k, float c)
{
    const float tmp = (eta * eta + k * k) * (c * c);
    const float rParl2 = (tmp - eta * (2.0f * c) + 1.0f) / (tmp + eta * (2.0f * c) + 1.0f);
    const float tmpF = eta * eta + k * k;
    const float rPerp2 = (tmpF - eta * (2.0f * c) + c * c) / (tmpF + eta * (2.0f * c) + c * c);
    return (rParl2 + rPerp2) * 0.5f;
}
static void ggx_f(const orc_material *m, v3 li, v3 lo, float out[3])
{
    const v3 H = vnorm(vadd(li, lo));
    const float D = ggx_D(m->roughness, H), G = ggx_G(li, lo, H), c = vdot(li, H);
    for (int k = 0; k < 3; k++) out[k] = m->R[k] * D * G * ggx_F(m->eta[k], m->k[k], c) / (4.0f * li.z) / lo.z;
}

/* Samples wo; returns 0 when the BSDF gives no direction (the reference's
 * SampleDirection returns without writing wo, bsdf_diffuse.cpp:69-79). */
static int bsdf_sample(const orc_material *m, const frame *f, v3 wi, const float u[2], v3 *wo)
{
    const v3 li = to_local(f, wi);
    if (li.z <= 0.0f) return 0;
    if (m->type == ORC_MAT_DIFFUSE) {
        float l[3];
        cos_hemisphere(u, l);
        *wo = to_world(f, mk(l[0], l[1], l[2]));
        return 1;
    }
    if (m->type == ORC_MAT_PHONG) {
        /* cos(alpha) = u0^(1/(n+1)), phi = 2 pi u1 about the mirror direction */
        const v3 r = reflect_sn(f, wi);
        v3 ru, rv;
        onb(r, &ru, &rv);
        const float ca = lmx_powf(u[0], 1.0f / (m->exponent + 1.0f));
        const float sa = sqrtf(fmaxf(0.0f, 1.0f - ca * ca));
        const float phi = 2.0f * PI_F * u[1];
        float sp, cp;
        lmx_sincosf(phi, &sp, &cp);
        const v3 l = mk(sa * cp, sa * sp, ca);
        *wo = vadd(vadd(vmul(ru, l.x), vmul(rv, l.y)), vmul(r, l.z));
        return 1;
    }
    if (m->type == ORC_MAT_GGX) {
        const v3 H = ggx_sample_H(m->roughness, u);
        const v3 ni = vneg(li);
        const v3 lo = vsub(ni, vmul(H, 2.0f * vdot(ni, H)));
        if (lo.z <= 0.0f) return 0;
        *wo = to_world(f, lo);
        return 1;
    }
    return 0;
}

/* pdf in projected solid angle */
static float bsdf_pdf(const orc_material *m, const frame *f, v3 wi, v3 wo)
{
    const v3 li = to_local(f, wi), lo = to_local(f, wo);
    if (li.z <= 0.0f || lo.z <= 0.0f) return 0.0f;
    if (m->type == ORC_MAT_DIFFUSE) return INV_PI_F;           /* sampler.h:70-76 */
    if (m->type == ORC_MAT_PHONG) {
        const float ca = fmaxf(0.0f, vdot(reflect_sn(f, wi), wo));
        const float pdf_sa = (m->exponent + 1.0f) * (0.5f * INV_PI_F) * lmx_powf(ca, m->exponent);
        return pdf_sa / lo.z;
    }
    if (m->type == ORC_MAT_GGX) {
        const v3 H = vnorm(vadd(li, lo));
        return ggx_D(m->roughness, H) * H.z / (4.0f * vdot(lo, H)) / lo.z;
    }
    return 0.0f;
}

static void bsdf_eval(const orc_material *m, const frame *f, v3 wi, v3 wo, float out[3])
{
    out[0] = out[1] = out[2] = 0.0f;
    const v3 li = to_local(f, wi), lo = to_local(f, wo);
    if (li.z <= 0.0f || lo.z <= 0.0f) return;
    const float corr = shading_correction(f, wi, wo);
    if (m->type == ORC_MAT_DIFFUSE) {                           /* bsdf_diffuse.cpp:93-103 */
        for (int c = 0; c < 3; c++) out[c] = m->R[c] * INV_PI_F * corr;
    } else if (m->type == ORC_MAT_PHONG) {
        const float ca = fmaxf(0.0f, vdot(reflect_sn(f, wi), wo));
        const float k = (m->exponent + 2.0f) * (0.5f * INV_PI_F) * lmx_powf(ca, m->exponent);
        for (int c = 0; c < 3; c++) out[c] = m->R[c] * k * corr;
    } else if (m->type == ORC_MAT_GGX) {
        ggx_f(m, li, lo, out);
        for (int c = 0; c < 3; c++) out[c] *= corr;
    }
}

/* ---------------------------------------------------------------------- */
/* pinhole sensor (sensor_pinhole.cpp:79-98,165-185) */

typedef struct {
    v3 pos, vx, vy, vz;
    float tan_fov, aspect;
} camera;

static void camera_init(const orc_scene_desc *d, camera *c)
{
    c->pos = ld3(d->cam_pos); c->vx = ld3(d->cam_vx); c->vy = ld3(d->cam_vy); c->vz = ld3(d->cam_vz);
    const float fov = d->fov_deg * PI_F / 180.0f;                 /* Math::Radians, math.h:1674 */
    c->tan_fov = lmx_tanf(fov * 0.5f);
    c->aspect = (float)d->width / (float)d->height;
}

static v3 camera_dir(const camera *c, float rx, float ry)
{
    const float x = 2.0f * rx - 1.0f, y = 2.0f * ry - 1.0f;
    const v3 we = vnorm(mk(c->aspect * c->tan_fov * x, c->tan_fov * y, -1.0f));
    return vadd(vadd(vmul(c->vx, we.x), vmul(c->vy, we.y)), vmul(c->vz, we.z));
}

static int camera_raster(const camera *c, v3 wo, float *rx, float *ry)
{
    const v3 we = mk(vdot(c->vx, wo), vdot(c->vy, wo), vdot(c->vz, wo));
    if (we.z >= 0.0f) return 0;
    *rx = (-we.x / we.z / c->tan_fov / c->aspect + 1.0f) * 0.5f;
    *ry = (-we.y / we.z / c->tan_fov + 1.0f) * 0.5f;
    if (*rx < 0.0f || *rx > 1.0f || *ry < 0.0f || *ry > 1.0f) return 0;
    return 1;
}

/* ---------------------------------------------------------------------- */
/* renderer::pt, one sample */

typedef struct {
    const float *table;      /* row 0 belongs to sample `first` */
    uint32_t width;
    uint64_t seed;
    uint64_t first;
} usrc;

static inline float U(const usrc *s, uint64_t sample, uint32_t slot)
{
    if (s->table) return s->table[(sample - s->first) * s->width + slot];
    return orc_uniform(s->seed, sample, slot);
}

/* Writes out[0..1] = raster position (or pixel centre in stratified mode),
 * out[2..4] = summed contribution C of the sample, out[5] = number of splats. */
static void pt_sample(const orc_scene *s, const camera *cam, const orc_pt_params *pp,
                      const usrc *us, uint64_t sample, float out[6])
{
    out[0] = out[1] = out[2] = out[3] = out[4] = out[5] = 0.0f;
    /* sensor selection consumes slot 0 (one sensor, pdf 1: scene3.cpp:506-539) */
    float rx, ry;
    v3 wo;
    const float ux = U(us, sample, ORC_PT_RASTER), uy = U(us, sample, ORC_PT_RASTER + 1);
    if (pp->raster_mode == ORC_RASTER_STRATIFIED) {
        const uint64_t npix = (uint64_t)s->d.width * (uint64_t)s->d.height;
        const uint32_t pix = (uint32_t)(sample % npix);
        const uint32_t px = pix % (uint32_t)s->d.width, py = pix / (uint32_t)s->d.width;
        rx = ((float)px + ux) / (float)s->d.width;
        ry = ((float)py + uy) / (float)s->d.height;
        wo = camera_dir(cam, rx, ry);
        out[0] = ((float)px + 0.5f) / (float)s->d.width;
        out[1] = ((float)py + 0.5f) / (float)s->d.height;
    } else {
        wo = camera_dir(cam, ux, uy);
        if (!camera_raster(cam, wo, &rx, &ry)) return;           /* renderer_pt.cpp:94-99 */
        out[0] = rx; out[1] = ry;
    }
    /* throughput = We / pdfPE / pdfE = 1, and fs / pdfD = 1 for the sensor
     * vertex (both are the importance, sensor_pinhole.cpp:105-133) */
    float thr[3] = {1.0f, 1.0f, 1.0f};
    v3 o = cam->pos, wi = mk(0, 0, 0);
    frame fr;
    const orc_material *mat = NULL;
    int num_vertices = 1;
    for (int seg = 0;; seg++) {
        if (pp->max_num_vertices != -1 && num_vertices >= pp->max_num_vertices) break;
        const uint32_t base = ORC_PT_SEG0 + 4u * (uint32_t)seg;
        if (seg > 0) {
            if (!mat || (mat->type != ORC_MAT_DIFFUSE && mat->type != ORC_MAT_PHONG && mat->type != ORC_MAT_GGX)) break;
            const float u2[2] = {U(us, sample, base), U(us, sample, base + 1)};
            if (!bsdf_sample(mat, &fr, wi, u2, &wo)) break;
            const float pdf = bsdf_pdf(mat, &fr, wi, wo);
            float fs[3];
            bsdf_eval(mat, &fr, wi, wo, fs);
            if (fs[0] == 0.0f && fs[1] == 0.0f && fs[2] == 0.0f) break;   /* :148-151 */
            for (int c = 0; c < 3; c++) thr[c] *= fs[c] / pdf;
            o = fr.p;
        }
        orc_hit h;
        if (!scene_intersect(s, o, wo, &h)) break;                  /* :168-175 */
        surface_at(s, o, wo, &h, &fr);
        mat = &s->d.mats[s->d.tri_mat[h.id]];
        if (mat->Le[0] > 0.0f || mat->Le[1] > 0.0f || mat->Le[2] > 0.0f) {   /* :183-194 */
            if (num_vertices + 1 >= pp->min_num_vertices) {
                /* Light_Area::EvaluateDirection: Le if LocalCos(-d) > 0 */
                const v3 lw = to_local(&fr, vneg(wo));
                if (lw.z > 0.0f) {
                    for (int c = 0; c < 3; c++) out[2 + c] += thr[c] * mat->Le[c];
                    out[5] += 1.0f;
                }
            }
        }
        /* Russian roulette (:207-215), from depth rr_depth on */
        if (seg + 1 >= pp->rr_depth) {
            /* rr_mode 1: throughput roulette, an extension without a reference counterpart */
            float q = pp->rr_prob;
            if (pp->rr_mode == 1) {
                q = fmaxf(thr[0], fmaxf(thr[1], thr[2]));
                q = q < 1.0f ? q : 1.0f;
            }
            if (!(q > 0.0f) || U(us, sample, base + 3) > q) break;
            for (int c = 0; c < 3; c++) thr[c] /= q;
        }
        wi = vneg(wo);
        num_vertices++;
    }
}

typedef struct {
    const orc_scene *s; const orc_pt_params *pp; usrc us; camera cam;
    uint64_t first; float *out;
} pt_ctx;

typedef struct { pt_ctx *c; uint64_t b, e; } pt_job;

static void *pt_run(void *p)
{
    pt_job *j = (pt_job *)p;
    for (uint64_t i = j->b; i < j->e; i++)
        pt_sample(j->c->s, &j->c->cam, j->c->pp, &j->c->us, j->c->first + i, &j->c->out[6 * i]);
    return NULL;
}

void orc_pt_samples(const orc_scene *s, const orc_pt_params *pp, const float *table, uint32_t table_width,
                    uint64_t first_sample, uint64_t n, float *out, int nthreads)
{
    pt_ctx c;
    c.s = s; c.pp = pp; c.us.table = table; c.us.width = table_width;
    c.us.seed = ((uint64_t)pp->seed_hi << 32) | pp->seed_lo;
    c.us.first = first_sample;
    camera_init(&s->d, &c.cam);
    c.first = first_sample; c.out = out;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    pt_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t].c = &c;
        jobs[t].b = n * (uint64_t)t / (uint64_t)nthreads;
        jobs[t].e = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, pt_run, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

/* Film_HDR::Splat pixel (film_hdr.cpp:218-223): clamp(int(r W)), row 0 at the
 * bottom. */
void orc_film_splat(const float *samples, uint64_t n, int32_t width, int32_t height, double *film)
{
    for (uint64_t i = 0; i < n; i++) {
        const float *sm = &samples[6 * i];
        if (sm[5] == 0.0f) continue;
        int px = (int)(sm[0] * (float)width), py = (int)(sm[1] * (float)height);
        px = px < 0 ? 0 : (px > width - 1 ? width - 1 : px);
        py = py < 0 ? 0 : (py > height - 1 ? height - 1 : py);
        double *d = &film[3 * ((size_t)py * (size_t)width + (size_t)px)];
        d[0] += sm[2]; d[1] += sm[3]; d[2] += sm[4];
    }
}

/* Test hook: one BSDF query on a flat frame with normal n (gn = sn = n):
 * samples wo from u, returns 1 and writes wo, the projected-solid-angle pdf
 * and f; returns 0 when the BSDF gives no direction. */
int orc_bsdf_query(const orc_material *m, const float n[3], const float wi[3], const float u[2],
                   float wo[3], float *pdf, float f[3])
{
    frame fr;
    fr.p = mk(0, 0, 0);
    fr.gn = fr.sn = vnorm(ld3(n));
    onb(fr.sn, &fr.u, &fr.v);
    v3 w;
    if (!bsdf_sample(m, &fr, ld3(wi), u, &w)) return 0;
    wo[0] = w.x; wo[1] = w.y; wo[2] = w.z;
    *pdf = bsdf_pdf(m, &fr, ld3(wi), w);
    bsdf_eval(m, &fr, ld3(wi), w, f);
    return 1;
}

/* pdf and f for a given pair of directions on the same flat frame */
void orc_bsdf_eval(const orc_material *m, const float n[3], const float wi[3], const float wo[3],
                   float *pdf, float f[3])
{
    frame fr;
    fr.p = mk(0, 0, 0);
    fr.gn = fr.sn = vnorm(ld3(n));
    onb(fr.sn, &fr.u, &fr.v);
    *pdf = bsdf_pdf(m, &fr, ld3(wi), ld3(wo));
    bsdf_eval(m, &fr, ld3(wi), ld3(wo), f);
}

/* ====================================================================== */
/* renderer::vcm in bdpt mode (renderer_vcm.cpp:51-341,662-713) with
 * SubpathSampler::TraceSubpath_ (subpathsampler.cpp:42-203): an eye and a
 * light subpath per sample, every (s,t) connection with s+t <= max vertices,
 * contribution f w / p with the power heuristic over all strategies that can
 * generate the path.  Restated literally: each connection builds the full
 * vertex list and evaluates EvaluateF / EvaluatePathPDF for every strategy
 * (path pdfs in double, so long paths neither overflow nor underflow).
 *
 * One deliberate difference: the reference samples an area light's position
 * and its emission direction from the same 2-D number
 * (light_area.cpp:72-80, SamplePositionAndDirection(uED, uEP) at
 * subpathsampler.cpp:86), which correlates them and biases every strategy
 * with s >= 2.  Here the position takes uEP and the direction uED.
 *
 * Uniform slots per sample: an eye block then a light block, each
 * 8 + 4 (max_vertices - 1) wide.  Within a block: 1-2 position (uEP),
 * 3 emitter selection (uEP_C), 4-5 direction (uED), 6 component (uED_C);
 * direction sampling at subpath vertex j >= 1 uses 8 + 4 (j-1) + 0..2. */

enum { BV_CAMERA = 1, BV_LIGHT = 2, BV_SURFACE = 3 };

typedef struct {
    int kind;                 /* BV_* role of this vertex */
    int degenerate;           /* the pinhole's position has no surface */
    frame fr;                 /* p, gn, sn, u, v */
    const orc_material *mat;  /* surface material (BV_SURFACE, or a light surface) */
    uint32_t tri;             /* triangle id (surfaces / light points) */
    int light;                /* light primitive of the triangle, -1 if none */
    float alpha[3];           /* subpath throughput up to this vertex */
} bvert;

typedef struct {
    const orc_scene *s;
    camera cam;
    int *tri_light;           /* per triangle: light primitive index or -1 */
    float *tri_cdf_base;      /* unused placeholder */
} bctx;

static float importance(const camera *c, v3 wo)
{
    float rx, ry;
    if (!camera_raster(c, wo, &rx, &ry)) return 0.0f;
    const v3 we = mk(vdot(c->vx, wo), vdot(c->vy, wo), vdot(c->vz, wo));
    const float cos_t = -we.z;
    const float inv = 1.0f / cos_t;
    const float A = c->tan_fov * c->tan_fov * c->aspect * 4.0f;
    return inv * inv * inv / A;                     /* sensor_pinhole.cpp:137-154 */
}

/* BSDF f with the transport direction's shading-normal correction
 * (bsdfutils.h:55-66): EL -> 1, LE -> wiDotNs woDotNg / (woDotNs wiDotNg) */
static void bsdf_eval_dir(const orc_material *m, const frame *f, v3 wi, v3 wo, int le, float out[3])
{
    out[0] = out[1] = out[2] = 0.0f;
    const v3 li = to_local(f, wi), lo = to_local(f, wo);
    if (li.z <= 0.0f || lo.z <= 0.0f) return;
    const float wiDotNg = vdot(wi, f->gn), woDotNg = vdot(wo, f->gn);
    const float wiDotNs = li.z, woDotNs = lo.z;
    float corr;
    if (wiDotNg * wiDotNs <= 0.0f || woDotNg * woDotNs <= 0.0f) corr = 0.0f;
    else if (le) corr = wiDotNs * woDotNg / (woDotNs * wiDotNg);
    else corr = 1.0f;
    if (m->type == ORC_MAT_DIFFUSE) {
        for (int c = 0; c < 3; c++) out[c] = m->R[c] * INV_PI_F * corr;
    } else if (m->type == ORC_MAT_PHONG) {
        const float ca = fmaxf(0.0f, vdot(reflect_sn(f, wi), wo));
        const float k = (m->exponent + 2.0f) * (0.5f * INV_PI_F) * lmx_powf(ca, m->exponent);
        for (int c = 0; c < 3; c++) out[c] = m->R[c] * k * corr;
    } else if (m->type == ORC_MAT_GGX) {
        ggx_f(m, li, lo, out);
        for (int c = 0; c < 3; c++) out[c] *= corr;
    }
}

/* Primitive::EvaluateDirection for a vertex in its role */
static void v_eval(const bctx *b, const bvert *v, v3 wi, v3 wo, int le, float out[3])
{
    out[0] = out[1] = out[2] = 0.0f;
    if (v->kind == BV_CAMERA) {
        const float w = importance(&b->cam, wo);
        out[0] = out[1] = out[2] = w;
    } else if (v->kind == BV_LIGHT) {                 /* light_area.cpp:105-110 */
        if (to_local(&v->fr, wo).z <= 0.0f) return;
        for (int c = 0; c < 3; c++) out[c] = v->mat->Le[c];
    } else if (v->mat && (v->mat->type == ORC_MAT_DIFFUSE || v->mat->type == ORC_MAT_PHONG ||
                          v->mat->type == ORC_MAT_GGX)) {
        bsdf_eval_dir(v->mat, &v->fr, wi, wo, le, out);
    }
}

/* Primitive::EvaluateDirectionPDF, projected solid angle */
static float v_pdf(const bctx *b, const bvert *v, v3 wi, v3 wo)
{
    if (v->kind == BV_CAMERA) return importance(&b->cam, wo);
    if (v->kind == BV_LIGHT) return to_local(&v->fr, wo).z <= 0.0f ? 0.0f : INV_PI_F;
    if (v->mat && (v->mat->type == ORC_MAT_DIFFUSE || v->mat->type == ORC_MAT_PHONG || v->mat->type == ORC_MAT_GGX))
        return bsdf_pdf(v->mat, &v->fr, wi, wo);
    return 0.0f;
}

/* RenderUtils::GeometryTerm (renderutils.h:46-56) */
static float gterm(const bvert *a, const bvert *c)
{
    v3 d = vsub(c->fr.p, a->fr.p);
    const float l2 = vdot(d, d);
    const float l = sqrtf(l2);
    d = mk(d.x / l, d.y / l, d.z / l);
    float t = 1.0f;
    if (!a->degenerate) t *= fabsf(vdot(a->fr.sn, d));
    if (!c->degenerate) t *= fabsf(vdot(c->fr.sn, vneg(d)));
    return t / l2;
}

static v3 dir_to(const bvert *from, const bvert *to) { return vnorm(vsub(to->fr.p, from->fr.p)); }

/* EvaluatePositionGivenDirectionPDF x EvaluateEmitterPDF of an endpoint */
static float pos_pdf(const bctx *b, const bvert *v)
{
    if (v->kind == BV_CAMERA) return 1.0f;
    if (v->light < 0) return 0.0f;
    return b->s->inv_area[v->light] * (1.0f / (float)b->s->d.nlight);
}

/* Area light sample (TriangleUtils::SampleTriangleMesh, triangleutils.h:73-120) */
static void sample_light_point(const bctx *b, int l, const float u[2], bvert *v)
{
    const orc_scene *s = b->s;
    const uint32_t first = s->d.light_prims[2 * l], n = s->d.light_prims[2 * l + 1];
    const float *cdf = s->cdf[l];
    /* Distribution1D::SampleReuse (dist.h:71-77) */
    int lo = 0;
    for (uint32_t k = 1; k <= n; k++)
        if (cdf[k] <= u[0]) lo = (int)k; else break;
    int i = lo > (int)n - 1 ? (int)n - 1 : lo;
    float u2[2] = {(u[0] - cdf[i]) / (cdf[i + 1] - cdf[i]), u[1]};
    float bb[2];
    orc_sample_uniform_triangle(u2, bb);
    const uint32_t tri = first + (uint32_t)i;
    const uint32_t *t = &s->d.tris[3 * tri];
    const v3 p1 = ld3(&s->d.verts[3 * t[0]]), p2 = ld3(&s->d.verts[3 * t[1]]), p3 = ld3(&s->d.verts[3 * t[2]]);
    v->fr.p = vadd(vadd(vmul(p1, 1.0f - bb[0] - bb[1]), vmul(p2, bb[0])), vmul(p3, bb[1]));
    v->fr.gn = vnorm(vcross(vsub(p2, p1), vsub(p3, p1)));
    v->fr.sn = v->fr.gn;
    onb(v->fr.sn, &v->fr.u, &v->fr.v);
    v->kind = BV_LIGHT; v->degenerate = 0; v->tri = tri; v->light = l;
    v->mat = &s->d.mats[s->d.tri_mat[tri]];
}

/* TraceSubpath_ for one direction; returns the number of vertices */
static int trace_subpath(const bctx *b, const orc_bdpt_params *pp, const usrc *us, uint64_t sample,
                         uint32_t block, int light_dir, bvert *vs)
{
    const orc_scene *s = b->s;
    const int maxv = pp->max_num_vertices;
    int nv = 0;
    v3 init_wo = mk(0, 0, 0);
    for (int step = 0; maxv == -1 || step < maxv; step++) {
        if (step == 0) {
            bvert v;
            memset(&v, 0, sizeof v);
            const float uEP[2] = {U(us, sample, block + 1), U(us, sample, block + 2)};
            const float uEPC = U(us, sample, block + 3);
            const float uED[2] = {U(us, sample, block + 4), U(us, sample, block + 5)};
            if (light_dir) {
                if (s->d.nlight == 0) return 0;
                int l = (int)(uEPC * (float)s->d.nlight);       /* scene3.cpp:510-513 */
                l = l < 0 ? 0 : (l > (int)s->d.nlight - 1 ? (int)s->d.nlight - 1 : l);
                sample_light_point(b, l, uEP, &v);
                float lw[3];
                cos_hemisphere(uED, lw);
                init_wo = to_world(&v.fr, mk(lw[0], lw[1], lw[2]));
                const float a = 1.0f / pos_pdf(b, &v);
                v.alpha[0] = v.alpha[1] = v.alpha[2] = a;
            } else {
                v.kind = BV_CAMERA; v.degenerate = 1; v.light = -1;
                v.fr.p = b->cam.pos;
                if (pp->raster_mode == ORC_RASTER_STRATIFIED) {
                    const uint64_t npix = (uint64_t)s->d.width * (uint64_t)s->d.height;
                    const uint32_t pix = (uint32_t)(sample % npix);
                    const uint32_t px = pix % (uint32_t)s->d.width, py = pix / (uint32_t)s->d.width;
                    init_wo = camera_dir(&b->cam, ((float)px + uED[0]) / (float)s->d.width,
                                         ((float)py + uED[1]) / (float)s->d.height);
                } else {
                    init_wo = camera_dir(&b->cam, uED[0], uED[1]);
                }
                v.alpha[0] = v.alpha[1] = v.alpha[2] = 1.0f;
            }
            vs[nv++] = v;
            continue;
        }
        bvert *pv = &vs[nv - 1];
        v3 wi, wo;
        if (step == 1) {
            wi = mk(0, 0, 0);
            wo = init_wo;
        } else {
            const bvert *ppv = &vs[nv - 2];
            wi = dir_to(pv, ppv);
            const uint32_t base = block + ORC_PT_SEG0 + 4u * (uint32_t)(step - 2);
            const float uD[2] = {U(us, sample, base), U(us, sample, base + 1)};
            if (pv->kind != BV_SURFACE || !bsdf_sample(pv->mat, &pv->fr, wi, uD, &wo)) break;
        }
        float fs[3];
        v_eval(b, pv, wi, wo, light_dir, fs);
        if (fs[0] == 0.0f && fs[1] == 0.0f && fs[2] == 0.0f) break;
        const float pdf = v_pdf(b, pv, wi, wo);
        orc_hit h;
        if (!scene_intersect(s, pv->fr.p, wo, &h)) break;
        bvert v;
        memset(&v, 0, sizeof v);
        surface_at(s, pv->fr.p, wo, &h, &v.fr);
        v.kind = BV_SURFACE; v.degenerate = 0; v.tri = h.id;
        v.mat = &s->d.mats[s->d.tri_mat[h.id]];
        v.light = b->tri_light[h.id];
        for (int c = 0; c < 3; c++) v.alpha[c] = pv->alpha[c] * (fs[c] / pdf);
        vs[nv++] = v;
    }
    return nv;
}

/* The full path x_0 (light end) .. x_{n-1} (camera end) of strategy s. */
static int connect(const bctx *b, const bvert *L, int nL, const bvert *E, int nE, int s, int t, bvert *x)
{
    (void)nL; (void)nE;
    int n = 0;
    for (int i = 0; i < s; i++) x[n++] = L[i];
    for (int i = t - 1; i >= 0; i--) x[n++] = E[i];
    if (s == 0) {
        if (x[0].light < 0) return 0;                            /* :83-87 */
        x[0].kind = BV_LIGHT;
    } else if (t > 0) {
        /* Scene3::Visible (scene3.h:107-116) */
        const bvert *vL = &L[s - 1], *vE = &E[t - 1];
        v3 d = vsub(vE->fr.p, vL->fr.p);
        const float l = sqrtf(vdot(d, d));
        d = mk(d.x / l, d.y / l, d.z / l);
        orc_ray r;
        r.o[0] = vL->fr.p.x; r.o[1] = vL->fr.p.y; r.o[2] = vL->fr.p.z; r.tmin = EPS_ISECT;
        r.d[0] = d.x; r.d[1] = d.y; r.d[2] = d.z; r.tmax = l * (1.0f - EPS_ISECT);
        orc_hit h;
        orc_qbvh_intersect_ex(b->s->q, &r, 1, &h, ORC_TEST_WATERTIGHT, ORC_TIE_MIN_ID, 1);
        if (h.id != 0xffffffffu) return 0;
    }
    return n;
}

/* VCMPath::EvaluateF (renderer_vcm.cpp:116-200) */
static void eval_f(const bctx *b, const bvert *x, int n, int s, float out[3])
{
    const int t = n - s;
    float fL[3] = {1, 1, 1}, fE[3] = {1, 1, 1}, cst[3] = {0, 0, 0}, tmp[3];
    if (s > 0) {
        /* EvaluatePosition of the light endpoint = 1 */
        for (int i = 0; i < s - 1; i++) {
            const v3 wi = i >= 1 ? dir_to(&x[i], &x[i - 1]) : mk(0, 0, 0);
            const v3 wo = dir_to(&x[i], &x[i + 1]);
            v_eval(b, &x[i], wi, wo, 1, tmp);
            const float G = gterm(&x[i], &x[i + 1]);
            for (int c = 0; c < 3; c++) fL[c] *= tmp[c] * G;
        }
    }
    if (fL[0] == 0.0f && fL[1] == 0.0f && fL[2] == 0.0f) { out[0] = out[1] = out[2] = 0.0f; return; }
    if (t > 0) {
        for (int i = n - 1; i > s; i--) {
            const v3 wi = i < n - 1 ? dir_to(&x[i], &x[i + 1]) : mk(0, 0, 0);
            const v3 wo = dir_to(&x[i], &x[i - 1]);
            v_eval(b, &x[i], wi, wo, 0, tmp);
            const float G = gterm(&x[i], &x[i - 1]);
            for (int c = 0; c < 3; c++) fE[c] *= tmp[c] * G;
        }
    }
    if (fE[0] == 0.0f && fE[1] == 0.0f && fE[2] == 0.0f) { out[0] = out[1] = out[2] = 0.0f; return; }
    if (s == 0) {
        v_eval(b, &x[0], mk(0, 0, 0), dir_to(&x[0], &x[1]), 0, cst);
    } else {
        float fsL[3], fsE[3];
        const bvert *vL = &x[s - 1], *vE = &x[s];
        v_eval(b, vL, s >= 2 ? dir_to(vL, &x[s - 2]) : mk(0, 0, 0), dir_to(vL, vE), 1, fsL);
        v_eval(b, vE, s + 1 < n ? dir_to(vE, &x[s + 1]) : mk(0, 0, 0), dir_to(vE, vL), 0, fsE);
        const float G = gterm(vL, vE);
        for (int c = 0; c < 3; c++) cst[c] = fsL[c] * G * fsE[c];
    }
    for (int c = 0; c < 3; c++) out[c] = fL[c] * cst[c] * fE[c];
}

/* VCMPath::EvaluatePathPDF (renderer_vcm.cpp:202-256), vertex connection only */
static double path_pdf(const bctx *b, const bvert *x, int n, int s)
{
    const int t = n - s;
    if (s == 0 && t > 0) { if (x[0].kind == BV_CAMERA) return 0.0; }        /* delta position */
    else if (s > 0 && t == 0) { if (x[n - 1].kind == BV_CAMERA) return 0.0; }
    double pdf = 1.0;
    if (s > 0) {
        bvert x0 = x[0];
        if (x0.kind != BV_LIGHT) { if (x0.light < 0) return 0.0; x0.kind = BV_LIGHT; }
        pdf *= (double)pos_pdf(b, &x0);
        for (int i = 0; i < s - 1; i++) {
            const bvert *vi = i == 0 ? &x0 : &x[i];
            const v3 wi = i >= 1 ? dir_to(&x[i], &x[i - 1]) : mk(0, 0, 0);
            pdf *= (double)v_pdf(b, vi, wi, dir_to(&x[i], &x[i + 1])) * (double)gterm(&x[i], &x[i + 1]);
        }
    }
    if (t > 0) {
        pdf *= (double)pos_pdf(b, &x[n - 1]);
        for (int i = n - 1; i >= s + 1; i--) {
            const v3 wi = i + 1 < n ? dir_to(&x[i], &x[i + 1]) : mk(0, 0, 0);
            pdf *= (double)v_pdf(b, &x[i], wi, dir_to(&x[i], &x[i - 1])) * (double)gterm(&x[i], &x[i - 1]);
        }
    }
    return pdf;
}

static void bdpt_sample(const bctx *b, const orc_bdpt_params *pp, const usrc *us, uint64_t sample, float *out)
{
    const int maxv = pp->max_num_vertices;
    const int stride = (maxv + 1) * (maxv + 1);
    memset(out, 0, sizeof(float) * 8 * (size_t)stride);
    bvert E[64], L[64], x[128];
    const uint32_t wsub = ORC_PT_SEG0 + 4u * (uint32_t)(maxv - 1);
    const int nE = trace_subpath(b, pp, us, sample, 0, 0, E);           /* eye first, :671-672 */
    const int nL = trace_subpath(b, pp, us, sample, wsub, 1, L);
    const int npix = b->s->d.width * b->s->d.height;
    for (int t = 1; t <= nE; t++) {
        int minS = 2 - t; if (minS < 0) minS = 0;
        if (pp->min_num_vertices - t > minS) minS = pp->min_num_vertices - t;
        int maxS = maxv - t; if (nL < maxS) maxS = nL;
        for (int s = minS; s <= maxS; s++) {
            const int n = connect(b, L, nL, E, nE, s, t, x);
            if (n == 0) continue;
            float f[3];
            eval_f(b, x, n, s, f);
            if (f[0] == 0.0f && f[1] == 0.0f && f[2] == 0.0f) continue;
            const double ps = path_pdf(b, x, n, s);
            if (ps == 0.0) continue;
            double invw = 0.0;
            for (int k = 0; k <= n; k++) {
                const double pk = path_pdf(b, x, n, k);
                if (pk > 0.0) { const double r = pk / ps; invw += r * r; }
            }
            const double w = 1.0 / invw;
            /* raster position of the path's last edge (renderer_vcm.cpp:343-350) */
            float rx = 0, ry = 0;
            if (t == 1) {
                if (!camera_raster(&b->cam, dir_to(&x[n - 1], &x[n - 2]), &rx, &ry)) continue;
            } else if (pp->raster_mode == ORC_RASTER_STRATIFIED) {
                const uint32_t pix = (uint32_t)(sample % (uint64_t)npix);
                rx = ((float)(pix % (uint32_t)b->s->d.width) + 0.5f) / (float)b->s->d.width;
                ry = ((float)(pix / (uint32_t)b->s->d.width) + 0.5f) / (float)b->s->d.height;
            } else {
                if (!camera_raster(&b->cam, dir_to(&x[n - 1], &x[n - 2]), &rx, &ry)) continue;
            }
            float *o = &out[8 * (size_t)(t * (maxv + 1) + s)];
            o[0] = rx; o[1] = ry;
            for (int c = 0; c < 3; c++) o[2 + c] = (float)((double)f[c] * w / ps);
            o[5] = 1.0f; o[6] = (float)w;
        }
    }
}

typedef struct { bctx *b; const orc_bdpt_params *pp; usrc us; uint64_t first, n; float *out; int stride; } bd_job_ctx;
typedef struct { bd_job_ctx *c; uint64_t b, e; } bd_job;

static void *bd_run(void *p)
{
    bd_job *j = (bd_job *)p;
    for (uint64_t i = j->b; i < j->e; i++)
        bdpt_sample(j->c->b, j->c->pp, &j->c->us, j->c->first + i, &j->c->out[(size_t)8 * j->c->stride * i]);
    return NULL;
}

void orc_bdpt_samples(const orc_scene *s, const orc_bdpt_params *pp, const float *table, uint32_t table_width,
                      uint64_t first_sample, uint64_t n, float *out, int nthreads)
{
    bctx b;
    b.s = s;
    camera_init(&s->d, &b.cam);
    b.tri_light = (int *)malloc(sizeof(int) * (s->d.ntri ? s->d.ntri : 1));
    for (uint32_t i = 0; i < s->d.ntri; i++) b.tri_light[i] = -1;
    for (uint32_t l = 0; l < s->d.nlight; l++)
        for (uint32_t k = 0; k < s->d.light_prims[2 * l + 1]; k++) b.tri_light[s->d.light_prims[2 * l] + k] = (int)l;
    b.tri_cdf_base = NULL;
    bd_job_ctx c;
    c.b = &b; c.pp = pp; c.us.table = table; c.us.width = table_width;
    c.us.seed = ((uint64_t)pp->seed_hi << 32) | pp->seed_lo;
    c.us.first = first_sample;
    c.first = first_sample; c.n = n; c.out = out;
    c.stride = (pp->max_num_vertices + 1) * (pp->max_num_vertices + 1);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    bd_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t].c = &c;
        jobs[t].b = n * (uint64_t)t / (uint64_t)nthreads;
        jobs[t].e = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, bd_run, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(b.tri_light);
}

/* ---------------------------------------------------------------------------
 * renderer::vcm in modes vcm and bdpm (renderer_vcm.cpp:102-114 MergeSubpaths,
 * :116-200 EvaluateF, :202-256 EvaluatePathPDF, :258-340 the weights,
 * :580-770 one pass), restated literally for ONE pass with merge radius r:
 * N photon subpaths (light subpaths of the photon uniform source) give the
 * merge candidates, every light vertex j >= 1 of them; each eye sample
 * traces its eye and light subpaths as in bdpt mode, makes every connection
 * (mode vcm) and, at every eye vertex t-1 >= 1, merges with every photon
 * vertex j closer than r: the path is photon vertices 0..j-1 followed by eye
 * vertices t-1..0 (the photon vertex itself is replaced by the eye vertex),
 * its pdf carries pi r^2, and the weight is the power heuristic over every
 * connection and merge strategy with merge terms counted N times
 * (EvaluateMISWeight_VCM) or over the merges alone (EvaluateMISWeight_BDPM).
 * The candidates come from a uniform grid (the reference's kd-tree returns
 * the same set).  Uniform slots as orc_bdpt_samples; photon subpath i takes
 * the light block of sample i of the photon source. */

/* EvaluateF with merging (:116-200) */
static void vcm_f(const bctx *b, const bvert *x, int n, int s, int merge, float out[3])
{
    const int t = n - s;
    float fL[3] = {1, 1, 1}, fE[3] = {1, 1, 1}, cst[3] = {0, 0, 0}, tmp[3];
    out[0] = out[1] = out[2] = 0.0f;
    if (s > 0)
        for (int i = 0; i < (merge ? s : s - 1); i++) {
            const v3 wi = i >= 1 ? dir_to(&x[i], &x[i - 1]) : mk(0, 0, 0);
            v_eval(b, &x[i], wi, dir_to(&x[i], &x[i + 1]), 1, tmp);
            const float G = gterm(&x[i], &x[i + 1]);
            for (int c = 0; c < 3; c++) fL[c] *= tmp[c] * G;
        }
    if (fL[0] == 0.0f && fL[1] == 0.0f && fL[2] == 0.0f) return;
    if (t > 0)
        for (int i = n - 1; i > s; i--) {
            const v3 wi = i < n - 1 ? dir_to(&x[i], &x[i + 1]) : mk(0, 0, 0);
            v_eval(b, &x[i], wi, dir_to(&x[i], &x[i - 1]), 0, tmp);
            const float G = gterm(&x[i], &x[i - 1]);
            for (int c = 0; c < 3; c++) fE[c] *= tmp[c] * G;
        }
    if (fE[0] == 0.0f && fE[1] == 0.0f && fE[2] == 0.0f) return;
    if (!merge) {
        eval_f(b, x, n, s, out);
        return;
    }
    /* the merged vertex scatters from the eye side towards the light side */
    v_eval(b, &x[s], dir_to(&x[s], &x[s + 1]), dir_to(&x[s], &x[s - 1]), 0, cst);
    for (int c = 0; c < 3; c++) out[c] = fL[c] * cst[c] * fE[c];
}

/* EvaluatePathPDF with merging (:202-256), product area measure in double */
static double vcm_pdf(const bctx *b, const bvert *x, int n, int s, int merge, float r)
{
    const int t = n - s;
    if (!merge) return path_pdf(b, x, n, s);
    if (s == 0 || t == 0) return 0.0;
    if (x[s].kind == BV_CAMERA) return 0.0;             /* delta position */
    double pdf = 1.0;
    bvert x0 = x[0];
    if (x0.kind != BV_LIGHT) { if (x0.light < 0) return 0.0; x0.kind = BV_LIGHT; }
    pdf *= (double)pos_pdf(b, &x0);
    for (int i = 0; i < s; i++) {
        const bvert *vi = i == 0 ? &x0 : &x[i];
        const v3 wi = i >= 1 ? dir_to(&x[i], &x[i - 1]) : mk(0, 0, 0);
        pdf *= (double)v_pdf(b, vi, wi, dir_to(&x[i], &x[i + 1])) * (double)gterm(&x[i], &x[i + 1]);
    }
    pdf *= (double)pos_pdf(b, &x[n - 1]);
    for (int i = n - 1; i >= s + 1; i--) {
        const v3 wi = i + 1 < n ? dir_to(&x[i], &x[i + 1]) : mk(0, 0, 0);
        pdf *= (double)v_pdf(b, &x[i], wi, dir_to(&x[i], &x[i - 1])) * (double)gterm(&x[i], &x[i - 1]);
    }
    return pdf * (double)(PI_F * r * r);
}

/* EvaluateMISWeight_VCM / _BDPM (:258-340) */
static double vcm_weight(const bctx *b, const bvert *x, int n, int s0, int merge0, float r, uint32_t np, int bdpm)
{
    const double ps = vcm_pdf(b, x, n, s0, merge0, r);
    double inv = 0.0;
    for (int s = 0; s <= n; s++)
        for (int ty = bdpm ? 1 : 0; ty < 2; ty++) {
            const double pi = vcm_pdf(b, x, n, s, ty, r);
            if (pi > 0.0) { const double q = pi / ps; inv += q * q * (ty ? (double)np : 1.0); }
        }
    return 1.0 / inv;
}

typedef struct {
    const bctx *b;
    const orc_vcm_params *pp;
    usrc us;
    const bvert *ph;          /* photon vertices, np x maxv */
    const int *ph_n;          /* vertices per photon subpath */
    /* grid over the photon vertices j >= 1 (cell size r) */
    float lo[3], cell;
    int dim[3];
    int *cstart, *items;      /* items: photon * maxv + j */
    float *out;
    int stride;
    uint64_t first, n;
} vcm_ctx;

static void vcm_sample(const vcm_ctx *c, uint64_t sample, float *out)
{
    const bctx *b = c->b;
    const orc_vcm_params *pp = c->pp;
    const int maxv = pp->max_num_vertices;
    const int M1 = maxv + 1;
    const int bdpm = pp->mode == ORC_VCM_BDPM;
    const float r = pp->radius;
    memset(out, 0, sizeof(float) * 8 * (size_t)c->stride);
    orc_bdpt_params bp = {pp->max_num_vertices, pp->min_num_vertices, pp->seed_lo, pp->seed_hi, pp->raster_mode};
    bvert E[64], L[64], x[128];
    const uint32_t wsub = ORC_PT_SEG0 + 4u * (uint32_t)(maxv - 1);
    const int nE = trace_subpath(b, &bp, &c->us, sample, 0, 0, E);
    const int nL = trace_subpath(b, &bp, &c->us, sample, wsub, 1, L);
    const int npix = b->s->d.width * b->s->d.height;
    for (int t = 1; t <= nE; t++) {
        float rx = 0.0f, ry = 0.0f;
        if (pp->raster_mode == ORC_RASTER_STRATIFIED) {
            const uint32_t pix = (uint32_t)(sample % (uint64_t)npix);
            rx = ((float)(pix % (uint32_t)b->s->d.width) + 0.5f) / (float)b->s->d.width;
            ry = ((float)(pix / (uint32_t)b->s->d.width) + 0.5f) / (float)b->s->d.height;
        }
        /* vertex connection (mode vcm) */
        if (!bdpm) {
            int minS = 2 - t; if (minS < 0) minS = 0;
            if (pp->min_num_vertices - t > minS) minS = pp->min_num_vertices - t;
            int maxS = maxv - t; if (nL < maxS) maxS = nL;
            for (int s = minS; s <= maxS; s++) {
                const int n = connect(b, L, nL, E, nE, s, t, x);
                if (n == 0) continue;
                float f[3];
                eval_f(b, x, n, s, f);
                if (f[0] == 0.0f && f[1] == 0.0f && f[2] == 0.0f) continue;
                const double ps = vcm_pdf(b, x, n, s, 0, r);
                if (ps == 0.0) continue;
                const double w = vcm_weight(b, x, n, s, 0, r, pp->num_photons, 0);
                float px = rx, py = ry;
                if (t == 1 || pp->raster_mode != ORC_RASTER_STRATIFIED)
                    if (!camera_raster(&b->cam, dir_to(&x[n - 1], &x[n - 2]), &px, &py)) continue;
                float *o = &out[8 * (size_t)(2 * (t * M1 + s))];
                o[0] = px; o[1] = py;
                for (int k = 0; k < 3; k++) o[2 + k] = (float)((double)f[k] * w / ps);
                o[5] = 1.0f; o[6] = (float)w;
            }
        }
        /* vertex merging at eye vertex t-1 (not the pinhole) */
        const bvert *vE = &E[t - 1];
        if (vE->kind == BV_CAMERA) continue;
        float px = rx, py = ry;
        if (pp->raster_mode != ORC_RASTER_STRATIFIED)
            if (!camera_raster(&b->cam, dir_to(&E[0], &E[1]), &px, &py)) continue;
        int lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            const float pa = a == 0 ? vE->fr.p.x : (a == 1 ? vE->fr.p.y : vE->fr.p.z);
            lo[a] = (int)floorf((pa - r - c->lo[a]) / c->cell);
            hi[a] = (int)floorf((pa + r - c->lo[a]) / c->cell);
            if (lo[a] < 0) lo[a] = 0;
            if (hi[a] > c->dim[a] - 1) hi[a] = c->dim[a] - 1;
        }
        for (int gz = lo[2]; gz <= hi[2]; gz++)
            for (int gy = lo[1]; gy <= hi[1]; gy++)
                for (int gx = lo[0]; gx <= hi[0]; gx++) {
                    const int cidx = (gz * c->dim[1] + gy) * c->dim[0] + gx;
                    for (int k = c->cstart[cidx]; k < c->cstart[cidx + 1]; k++) {
                        const int it = c->items[k];
                        const int pi = it / maxv, vi = it % maxv;
                        const bvert *pv = &c->ph[(size_t)pi * maxv + vi];
                        const v3 d = vsub(pv->fr.p, vE->fr.p);
                        if (!(vdot(d, d) < r * r)) continue;
                        const int n = vi + t;
                        if (n < pp->min_num_vertices || maxv < n) continue;
                        /* MergeSubpaths(L, E, vi, t) */
                        const bvert *Lp = &c->ph[(size_t)pi * maxv];
                        int m = 0;
                        for (int i = 0; i < vi; i++) x[m++] = Lp[i];
                        for (int i = t - 1; i >= 0; i--) x[m++] = E[i];
                        float f[3];
                        vcm_f(b, x, n, vi, 1, f);
                        if (f[0] == 0.0f && f[1] == 0.0f && f[2] == 0.0f) continue;
                        const double ps = vcm_pdf(b, x, n, vi, 1, r);
                        if (ps == 0.0) continue;
                        const double w = vcm_weight(b, x, n, vi, 1, r, pp->num_photons, bdpm);
                        float *o = &out[8 * (size_t)(2 * (t * M1 + vi + 1) + 1)];
                        o[0] = px; o[1] = py;
                        for (int q = 0; q < 3; q++) o[2 + q] += (float)((double)f[q] * w / ps);
                        o[5] += 1.0f; o[6] += (float)w;
                    }
                }
    }
}

typedef struct { vcm_ctx *c; uint64_t b, e; } vcm_job;

/* photon subpaths [b, e) */
typedef struct {
    const bctx *b;
    const orc_bdpt_params *bp;
    const usrc *pu;
    bvert *ph;
    int *ph_n;
    int maxv;
    uint32_t wsub;
    uint32_t lo, hi;
} vcm_pjob;

static void *vcm_photons(void *p)
{
    vcm_pjob *j = (vcm_pjob *)p;
    bvert tmp[64];
    for (uint32_t i = j->lo; i < j->hi; i++) {
        j->ph_n[i] = trace_subpath(j->b, j->bp, j->pu, i, j->wsub, 1, tmp);
        for (int k = 0; k < j->ph_n[i]; k++) j->ph[(size_t)i * j->maxv + k] = tmp[k];
    }
    return NULL;
}

static void *vcm_run(void *p)
{
    vcm_job *j = (vcm_job *)p;
    for (uint64_t i = j->b; i < j->e; i++)
        vcm_sample(j->c, j->c->first + i, &j->c->out[(size_t)8 * j->c->stride * i]);
    return NULL;
}

void orc_vcm_samples(const orc_scene *s, const orc_vcm_params *pp, const float *table, uint32_t table_width,
                     const float *ptable, uint32_t ptable_width, uint64_t first_sample, uint64_t n, float *out,
                     int nthreads)
{
    bctx b;
    b.s = s;
    camera_init(&s->d, &b.cam);
    b.tri_light = (int *)malloc(sizeof(int) * (s->d.ntri ? s->d.ntri : 1));
    for (uint32_t i = 0; i < s->d.ntri; i++) b.tri_light[i] = -1;
    for (uint32_t l = 0; l < s->d.nlight; l++)
        for (uint32_t k = 0; k < s->d.light_prims[2 * l + 1]; k++) b.tri_light[s->d.light_prims[2 * l] + k] = (int)l;
    b.tri_cdf_base = NULL;
    const int maxv = pp->max_num_vertices;
    /* photon subpaths */
    const uint32_t np = pp->num_photons;
    bvert *ph = (bvert *)calloc((size_t)np * maxv + 1, sizeof(bvert));
    int *ph_n = (int *)calloc(np + 1, sizeof(int));
    usrc pu;
    pu.table = ptable; pu.width = ptable_width;
    pu.seed = ((uint64_t)pp->pseed_hi << 32) | pp->pseed_lo;
    pu.first = 0;
    orc_bdpt_params bp = {pp->max_num_vertices, pp->min_num_vertices, pp->pseed_lo, pp->pseed_hi, pp->raster_mode};
    const uint32_t wsub = ORC_PT_SEG0 + 4u * (uint32_t)(maxv - 1);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    {
        pthread_t pth[256];
        vcm_pjob pj[256];
        for (int t = 0; t < nthreads; t++) {
            pj[t].b = &b; pj[t].bp = &bp; pj[t].pu = &pu; pj[t].ph = ph; pj[t].ph_n = ph_n;
            pj[t].maxv = maxv; pj[t].wsub = wsub;
            pj[t].lo = (uint32_t)((uint64_t)np * (uint64_t)t / (uint64_t)nthreads);
            pj[t].hi = (uint32_t)((uint64_t)np * (uint64_t)(t + 1) / (uint64_t)nthreads);
            pthread_create(&pth[t], NULL, vcm_photons, &pj[t]);
        }
        for (int t = 0; t < nthreads; t++) pthread_join(pth[t], NULL);
    }
    float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
    int count = 0;
    for (uint32_t i = 0; i < np; i++)
        for (int j = 1; j < ph_n[i]; j++) {
            const bvert *v = &ph[(size_t)i * maxv + j];
            const float q[3] = {v->fr.p.x, v->fr.p.y, v->fr.p.z};
            for (int a = 0; a < 3; a++) { if (q[a] < lo[a]) lo[a] = q[a]; if (q[a] > hi[a]) hi[a] = q[a]; }
            count++;
        }
    vcm_ctx c;
    c.b = &b; c.pp = pp; c.ph = ph; c.ph_n = ph_n;
    c.cell = pp->radius;
    long long cells = 1;
    for (int a = 0; a < 3; a++) {
        if (count == 0) { lo[a] = 0.0f; hi[a] = 0.0f; }
        c.lo[a] = lo[a];
        c.dim[a] = (int)floorf((hi[a] - lo[a]) / c.cell) + 1;
        if (c.dim[a] > 512) { c.dim[a] = 512; }
        cells *= c.dim[a];
    }
    /* cells beyond 512 per axis: widen the cell (the set found is the same) */
    float span = 0.0f;
    for (int a = 0; a < 3; a++) if (hi[a] - lo[a] > span) span = hi[a] - lo[a];
    if (span / c.cell > 511.0f) {
        c.cell = span / 511.0f;
        cells = 1;
        for (int a = 0; a < 3; a++) { c.dim[a] = (int)floorf((hi[a] - lo[a]) / c.cell) + 1; cells *= c.dim[a]; }
    }
    c.cstart = (int *)calloc((size_t)cells + 1, sizeof(int));
    c.items = (int *)malloc(sizeof(int) * (size_t)(count + 1));
    int *cellof = (int *)malloc(sizeof(int) * (size_t)((size_t)np * maxv + 1));
    for (uint32_t i = 0; i < np; i++)
        for (int j = 1; j < ph_n[i]; j++) {
            const bvert *v = &ph[(size_t)i * maxv + j];
            const float q[3] = {v->fr.p.x, v->fr.p.y, v->fr.p.z};
            int g[3];
            for (int a = 0; a < 3; a++) {
                g[a] = (int)floorf((q[a] - c.lo[a]) / c.cell);
                if (g[a] < 0) g[a] = 0;
                if (g[a] > c.dim[a] - 1) g[a] = c.dim[a] - 1;
            }
            const int ci = (g[2] * c.dim[1] + g[1]) * c.dim[0] + g[0];
            cellof[(size_t)i * maxv + j] = ci;
            c.cstart[ci + 1]++;
        }
    for (long long k = 0; k < cells; k++) c.cstart[k + 1] += c.cstart[k];
    int *fill = (int *)calloc((size_t)cells, sizeof(int));
    for (uint32_t i = 0; i < np; i++)
        for (int j = 1; j < ph_n[i]; j++) {
            const int ci = cellof[(size_t)i * maxv + j];
            c.items[c.cstart[ci] + fill[ci]++] = (int)i * maxv + j;
        }
    free(fill);
    free(cellof);
    c.us.table = table; c.us.width = table_width;
    c.us.seed = ((uint64_t)pp->seed_hi << 32) | pp->seed_lo;
    c.us.first = first_sample;
    c.out = out;
    c.stride = 2 * (maxv + 1) * (maxv + 1);
    c.first = first_sample; c.n = n;
    pthread_t th[256];
    vcm_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t].c = &c;
        jobs[t].b = n * (uint64_t)t / (uint64_t)nthreads;
        jobs[t].e = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, vcm_run, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(c.cstart); free(c.items); free(ph); free(ph_n); free(b.tri_light);
}

/* ---------------------------------------------------------------------------
 * renderer::bdpt (renderer_bdpt.cpp:50-1002), restated literally: per sample
 * an eye and a light subpath (Subpath::Sample, :80-220) whose vertices hold a
 * BSDF-sampled vertex `sv` and a `direct` vertex sampled on an emitter of the
 * other kind from the previous vertex; Russian roulette with probability 0.5
 * after every vertex but the endpoint; every strategy (s, t, direct) of every
 * length n (:885-950) built by Path::Connect (:233-333), weighed by the power
 * heuristic or the simple 1/#strategies weight over EvaluatePDF (:649-699,
 * :704-784) and divided by SelectionPDF (:348-368).
 *
 * Deliberate differences, as for renderer::vcm: the area light's position and
 * emission direction take independent numbers; a strategy whose own pdf is
 * zero contributes nothing.  max_num_vertices = -1 (unbounded, the
 * reference's default) keeps at most 64 vertices per subpath here.
 *
 * Uniform slots per sample: an eye block then a light block, each
 * 8 + 8 (M - 1) wide (M = max_num_vertices, 32 when unbounded).  Within a
 * block: 1-2 endpoint position, 3 emitter selection, 4-5 endpoint direction;
 * step j >= 1 (vertex j) owns 8 + 8 (j - 1) .. +7: 0-1 direction sampled at
 * vertex j - 1 (j >= 2), 3 roulette after vertex j, 4 emitter selection and
 * 5-6 position of the direct vertex sampled from vertex j - 1. */

#define RB_CAP 64

typedef struct { bvert sv, direct; int has_sv, has_direct; } rsub;

/* Scene3::Visible (scene3.h:107-116) */
static int visible(const bctx *b, v3 p1, v3 p2)
{
    v3 d = vsub(p2, p1);
    const float l = sqrtf(vdot(d, d));
    d = mk(d.x / l, d.y / l, d.z / l);
    orc_ray r;
    r.o[0] = p1.x; r.o[1] = p1.y; r.o[2] = p1.z; r.tmin = EPS_ISECT;
    r.d[0] = d.x; r.d[1] = d.y; r.d[2] = d.z; r.tmax = l * (1.0f - EPS_ISECT);
    orc_hit h;
    orc_qbvh_intersect_ex(b->s->q, &r, 1, &h, ORC_TEST_WATERTIGHT, ORC_TIE_MIN_ID, 1);
    return h.id == 0xffffffffu;
}

/* Emitter::EvaluatePosition (light_area.cpp:112-115, sensor_pinhole.cpp:120-123) */
static float ep_position(const bvert *v, int eval_delta)
{
    if (v->kind == BV_CAMERA) return eval_delta ? 0.0f : 1.0f;
    return 1.0f;
}

/* EvaluatePositionGivenDirectionPDF = EvaluatePositionGivenPreviousPositionPDF
 * for both emitters (light_area.cpp:92-100, sensor_pinhole.cpp:105-113) */
static float ep_pos_pdf(const bctx *b, const bvert *v, int eval_delta)
{
    if (v->kind == BV_CAMERA) return eval_delta ? 0.0f : 1.0f;
    return v->light < 0 ? 0.0f : b->s->inv_area[v->light];
}

/* Scene3::EvaluateEmitterPDF (scene3.cpp:524-540) */
static float emitter_pdf(const bctx *b, const bvert *v)
{
    return v->kind == BV_CAMERA ? 1.0f : 1.0f / (float)b->s->d.nlight;
}

static int rb_subpath(const bctx *b, const orc_bdptref_params *pp, const usrc *us, uint64_t sample, uint32_t block,
                      int light_dir, rsub *vs)
{
    const orc_scene *s = b->s;
    const int maxv = pp->max_num_vertices;
    int nv = 0;
    v3 init_wo = mk(0, 0, 0);
    for (int step = 0; (maxv == -1 || step < maxv) && step < RB_CAP; step++) {
        if (step == 0) {
            /* an emitter, a point on it and the initial direction (:89-107) */
            rsub e;
            memset(&e, 0, sizeof e);
            bvert *v = &e.sv;
            const float uP[2] = {U(us, sample, block + 1), U(us, sample, block + 2)};
            const float uC = U(us, sample, block + 3);
            const float uD[2] = {U(us, sample, block + 4), U(us, sample, block + 5)};
            if (light_dir) {
                if (s->d.nlight == 0) return 0;
                int l = (int)(uC * (float)s->d.nlight);
                l = l < 0 ? 0 : (l > (int)s->d.nlight - 1 ? (int)s->d.nlight - 1 : l);
                sample_light_point(b, l, uP, v);
                float lw[3];
                cos_hemisphere(uD, lw);
                init_wo = to_world(&v->fr, mk(lw[0], lw[1], lw[2]));
            } else {
                v->kind = BV_CAMERA; v->degenerate = 1; v->light = -1;
                v->fr.p = b->cam.pos;
                if (pp->raster_mode == ORC_RASTER_STRATIFIED) {
                    const uint64_t npix = (uint64_t)s->d.width * (uint64_t)s->d.height;
                    const uint32_t pix = (uint32_t)(sample % npix);
                    const uint32_t px = pix % (uint32_t)s->d.width, py = pix / (uint32_t)s->d.width;
                    init_wo = camera_dir(&b->cam, ((float)px + uD[0]) / (float)s->d.width,
                                         ((float)py + uD[1]) / (float)s->d.height);
                } else {
                    init_wo = camera_dir(&b->cam, uD[0], uD[1]);
                }
            }
            e.has_sv = 1;
            vs[nv++] = e;
            continue;
        }
        const uint32_t sb = block + ORC_PT_SEG0 + 8u * (uint32_t)(step - 1);
        const bvert *pv = &vs[nv - 1].sv;
        const bvert *ppv = nv > 1 ? &vs[nv - 2].sv : NULL;
        rsub nx;
        memset(&nx, 0, sizeof nx);
        /* the vertex sampled with the BSDF (:117-160) */
        {
            const v3 wi = ppv ? dir_to(pv, ppv) : mk(0, 0, 0);
            v3 wo = init_wo;
            int ok = 1;
            if (step > 1) {
                const float uD[2] = {U(us, sample, sb), U(us, sample, sb + 1)};
                ok = pv->kind == BV_SURFACE && bsdf_sample(pv->mat, &pv->fr, wi, uD, &wo);
            }
            if (ok) {
                float fs[3];
                v_eval(b, pv, wi, wo, light_dir, fs);
                orc_hit h;
                if (!(fs[0] == 0.0f && fs[1] == 0.0f && fs[2] == 0.0f) && scene_intersect(s, pv->fr.p, wo, &h)) {
                    bvert *v = &nx.sv;
                    surface_at(s, pv->fr.p, wo, &h, &v->fr);
                    v->kind = BV_SURFACE; v->degenerate = 0; v->tri = h.id;
                    v->mat = &s->d.mats[s->d.tri_mat[h.id]];
                    v->light = b->tri_light[h.id];
                    nx.has_sv = 1;
                }
            }
        }
        /* the direct vertex: an emitter of the other kind, sampled from the
         * previous vertex, if visible from it (:166-190) */
        {
            bvert *v = &nx.direct;
            int have = 0;
            if (light_dir) {
                v->kind = BV_CAMERA; v->degenerate = 1; v->light = -1;
                v->fr.p = b->cam.pos;                    /* sensor_pinhole.cpp:73-77 */
                have = 1;
            } else if (s->d.nlight > 0) {
                int l = (int)(U(us, sample, sb + 4) * (float)s->d.nlight);
                l = l < 0 ? 0 : (l > (int)s->d.nlight - 1 ? (int)s->d.nlight - 1 : l);
                const float uP[2] = {U(us, sample, sb + 5), U(us, sample, sb + 6)};
                sample_light_point(b, l, uP, v);
                have = 1;
            }
            nx.has_direct = have && visible(b, pv->fr.p, v->fr.p);
        }
        if (nx.has_sv || nx.has_direct) vs[nv++] = nx;
        if (!nx.has_sv) break;
        if (U(us, sample, sb + 3) > 0.5f) break;           /* :205-211 */
    }
    return nv;
}

/* Path::Connect (:233-333); returns the number of vertices, 0 if the
 * strategy does not apply to these subpaths */
static int rb_connect(const bctx *b, const rsub *L, const rsub *E, int s, int t, int direct, bvert *x)
{
    int n = 0;
    if (s == 0 && t > 0) {
        if (!direct) {
            if (!E[t - 1].has_sv || E[t - 1].sv.light < 0) return 0;
            for (int i = t - 1; i >= 0; i--) x[n++] = E[i].sv;
        } else {
            if (!E[t - 1].has_direct) return 0;
            x[n++] = E[t - 1].direct;
            for (int i = t - 2; i >= 0; i--) x[n++] = E[i].sv;
        }
        x[0].kind = BV_LIGHT;
    } else if (s > 0 && t == 0) {
        if (!direct) {
            /* the light subpath would have to hit the sensor: a pinhole is
             * not part of the scene's surfaces */
            return 0;
        }
        if (!L[s - 1].has_direct) return 0;
        for (int i = 0; i < s - 1; i++) x[n++] = L[i].sv;
        x[n++] = L[s - 1].direct;
        x[n - 1].kind = BV_CAMERA;
    } else {
        if (!L[s - 1].has_sv || !E[t - 1].has_sv) return 0;
        if (!visible(b, L[s - 1].sv.fr.p, E[t - 1].sv.fr.p)) return 0;
        for (int i = 0; i < s; i++) x[n++] = L[i].sv;
        for (int i = t - 1; i >= 0; i--) x[n++] = E[i].sv;
    }
    return n;
}

/* Path::EvaluateCst (:378-412) */
static void rb_cst(const bctx *b, const bvert *x, int n, int s, float out[3])
{
    const int t = n - s;
    out[0] = out[1] = out[2] = 0.0f;
    float e[3];
    if (s == 0 && t > 0) {
        const float P = ep_position(&x[0], 1);
        v_eval(b, &x[0], mk(0, 0, 0), dir_to(&x[0], &x[1]), 0, e);
        for (int c = 0; c < 3; c++) out[c] = P * e[c];
    } else if (s > 0 && t == 0) {
        const float P = ep_position(&x[n - 1], 1);
        v_eval(b, &x[n - 1], mk(0, 0, 0), dir_to(&x[n - 1], &x[n - 2]), 1, e);
        for (int c = 0; c < 3; c++) out[c] = P * e[c];
    } else {
        const bvert *vL = &x[s - 1], *vE = &x[s];
        float fsL[3], fsE[3];
        v_eval(b, vL, s >= 2 ? dir_to(vL, &x[s - 2]) : mk(0, 0, 0), dir_to(vL, vE), 1, fsL);
        v_eval(b, vE, s + 1 < n ? dir_to(vE, &x[s + 1]) : mk(0, 0, 0), dir_to(vE, vL), 0, fsE);
        const float G = gterm(vL, vE);
        for (int c = 0; c < 3; c++) out[c] = fsL[c] * G * fsE[c];
    }
}

static int black(const float v[3]) { return v[0] == 0.0f && v[1] == 0.0f && v[2] == 0.0f; }

/* Path::EvaluateUnweightContribution (:489-597): alpha_L c_st alpha_E */
static void rb_unweighted(const bctx *b, const bvert *x, int n, int s, int direct, float out[3])
{
    const int t = n - s;
    float aL[3] = {1, 1, 1}, aE[3] = {1, 1, 1}, fs[3];
    out[0] = out[1] = out[2] = 0.0f;
    if (s > 0) {
        const float a = ep_position(&x[0], 0) / ep_pos_pdf(b, &x[0], 0) / emitter_pdf(b, &x[0]);
        aL[0] = aL[1] = aL[2] = a;
        for (int i = 0; i < s - 1; i++) {
            const v3 wi = i >= 1 ? dir_to(&x[i], &x[i - 1]) : mk(0, 0, 0);
            const v3 wo = dir_to(&x[i], &x[i + 1]);
            v_eval(b, &x[i], wi, wo, 1, fs);
            if (black(fs)) return;
            const float d = (t == 0 && i == s - 2 && direct)
                ? ep_pos_pdf(b, &x[i + 1], 0) / gterm(&x[i + 1], &x[i]) * emitter_pdf(b, &x[i + 1])
                : v_pdf(b, &x[i], wi, wo);
            for (int c = 0; c < 3; c++) aL[c] *= fs[c] / d;
        }
    }
    if (black(aL)) return;
    if (t > 0) {
        const float a = ep_position(&x[n - 1], 0) / ep_pos_pdf(b, &x[n - 1], 0) / emitter_pdf(b, &x[n - 1]);
        aE[0] = aE[1] = aE[2] = a;
        for (int i = n - 1; i > s; i--) {
            const v3 wi = i < n - 1 ? dir_to(&x[i], &x[i + 1]) : mk(0, 0, 0);
            const v3 wo = dir_to(&x[i], &x[i - 1]);
            v_eval(b, &x[i], wi, wo, 0, fs);
            if (black(fs)) return;
            const float d = (s == 0 && i == 1 && direct)
                ? ep_pos_pdf(b, &x[i - 1], 0) / gterm(&x[i - 1], &x[i]) * emitter_pdf(b, &x[i - 1])
                : v_pdf(b, &x[i], wi, wo);
            for (int c = 0; c < 3; c++) aE[c] *= fs[c] / d;
        }
    }
    if (black(aE)) return;
    float cst[3];
    rb_cst(b, x, n, s, cst);
    if (black(cst)) return;
    for (int c = 0; c < 3; c++) out[c] = aL[c] * cst[c] * aE[c];
}

/* Path::Samplable + Path::EvaluatePDF (:599-699), product area measure; in
 * double, so long paths neither overflow nor underflow.  No BSDF here is a
 * delta distribution; the pinhole's position is one. */
static double rb_pdf(const bctx *b, const bvert *x, int n, int s, int direct)
{
    const int t = n - s;
    if (s > 0 && t > 0 && direct) return 0.0;
    if (s > 0 && t == 0 && x[n - 1].kind == BV_CAMERA) return 0.0;
    if (s == 0 && t > 0 && x[0].kind == BV_CAMERA) return 0.0;
    double pdf = 1.0;
    if (s > 0) {
        pdf *= (double)ep_pos_pdf(b, &x[0], 0) * (double)emitter_pdf(b, &x[0]);
        for (int i = 0; i < s - 1; i++) {
            if (t == 0 && i == s - 2 && direct) {
                pdf *= (double)ep_pos_pdf(b, &x[i + 1], 0) * (double)emitter_pdf(b, &x[i + 1]);
            } else {
                const v3 wi = i >= 1 ? dir_to(&x[i], &x[i - 1]) : mk(0, 0, 0);
                pdf *= (double)v_pdf(b, &x[i], wi, dir_to(&x[i], &x[i + 1])) * (double)gterm(&x[i], &x[i + 1]);
            }
        }
    }
    if (t > 0) {
        pdf *= (double)ep_pos_pdf(b, &x[n - 1], 0) * (double)emitter_pdf(b, &x[n - 1]);
        for (int i = n - 1; i >= s + 1; i--) {
            if (s == 0 && i == 1 && direct) {
                pdf *= (double)ep_pos_pdf(b, &x[0], 0) * (double)emitter_pdf(b, &x[0]);
            } else {
                const v3 wi = i + 1 < n ? dir_to(&x[i], &x[i + 1]) : mk(0, 0, 0);
                pdf *= (double)v_pdf(b, &x[i], wi, dir_to(&x[i], &x[i - 1])) * (double)gterm(&x[i], &x[i - 1]);
            }
        }
    }
    return pdf;
}

/* MISWeight_PowerHeuristics / MISWeight_Simple (:704-784) */
static double rb_mis(const bctx *b, const bvert *x, int n, int s0, int d0, int simple)
{
    const double ps = rb_pdf(b, x, n, s0, d0);
    double inv = 0.0;
    int nonzero = 0;
    for (int s = 0; s <= n; s++)
        for (int d = 0; d < 2; d++) {
            if (s > 0 && n - s > 0 && d) continue;
            const double pi = rb_pdf(b, x, n, s, d);
            if (pi > 0.0) { const double r = pi / ps; inv += r * r; nonzero++; }
        }
    return simple ? 1.0 / (double)nonzero : 1.0 / inv;
}

static void bdptref_sample(const bctx *b, const orc_bdptref_params *pp, const usrc *us, uint64_t sample, float *out)
{
    const int M = pp->max_num_vertices == -1 ? 32 : pp->max_num_vertices;
    const int stride = 2 * (M + 1) * (M + 1);
    memset(out, 0, sizeof(float) * 8 * (size_t)stride);
    rsub L[RB_CAP], E[RB_CAP];
    bvert x[2 * RB_CAP];
    const uint32_t wsub = ORC_PT_SEG0 + 8u * (uint32_t)(M - 1);
    const int nL = rb_subpath(b, pp, us, sample, wsub, 1, L);      /* light first (:866-867) */
    const int nE = rb_subpath(b, pp, us, sample, 0, 0, E);
    const int npix = b->s->d.width * b->s->d.height;
    for (int n = 2; n <= nE + nL; n++) {
        if (pp->max_num_vertices != -1 && (n > pp->max_num_vertices || n < pp->min_num_vertices)) continue;
        const int minS = n - nE > 0 ? n - nE : 0;
        const int maxS = nL < n ? nL : n;
        for (int s = minS; s <= maxS; s++)
            for (int d = 0; d < 2; d++) {
                const int t = n - s;
                if (s > 0 && t > 0 && d) continue;
                if (s > M || t > M) continue;                    /* beyond the record layout */
                const int nv = rb_connect(b, L, E, s, t, d, x);
                if (nv == 0) continue;
                float cs[3];
                rb_unweighted(b, x, nv, s, d, cs);
                if (black(cs)) continue;
                if (rb_pdf(b, x, nv, s, d) == 0.0) continue;
                const double w = rb_mis(b, x, nv, s, d, pp->mis);
                /* Path::SelectionPDF (:348-368) */
                float sel = 1.0f;
                for (int i = 1; i < s - 1; i++) sel *= 0.5f;
                for (int i = 1; i < t - 1; i++) sel *= 0.5f;
                float C[3];
                for (int c = 0; c < 3; c++) C[c] = (float)((double)cs[c] * w) / sel;
                if (black(C)) continue;
                /* Path::RasterPosition (:370-376); the eye subpath's own first
                 * edge gives the sample's pixel in stratified mode */
                float rx = 0.0f, ry = 0.0f;
                const int eye_edge = t >= 2 && !(s == 0 && d == 1 && t == 2);
                if (eye_edge && pp->raster_mode == ORC_RASTER_STRATIFIED) {
                    const uint32_t pix = (uint32_t)(sample % (uint64_t)npix);
                    rx = ((float)(pix % (uint32_t)b->s->d.width) + 0.5f) / (float)b->s->d.width;
                    ry = ((float)(pix / (uint32_t)b->s->d.width) + 0.5f) / (float)b->s->d.height;
                } else if (!camera_raster(&b->cam, dir_to(&x[nv - 1], &x[nv - 2]), &rx, &ry)) {
                    continue;
                }
                float *o = &out[8 * (size_t)(2 * (t * (M + 1) + s) + d)];
                o[0] = rx; o[1] = ry;
                for (int c = 0; c < 3; c++) o[2 + c] = C[c];
                o[5] = 1.0f; o[6] = (float)w;
            }
    }
}

typedef struct { bctx *b; const orc_bdptref_params *pp; usrc us; float *out; int stride; } rb_job_ctx;
typedef struct { rb_job_ctx *c; uint64_t b, e; } rb_job;

static void *rb_run(void *p)
{
    rb_job *j = (rb_job *)p;
    for (uint64_t i = j->b; i < j->e; i++)
        bdptref_sample(j->c->b, j->c->pp, &j->c->us, j->c->us.first + i, &j->c->out[(size_t)8 * j->c->stride * i]);
    return NULL;
}

void orc_bdptref_samples(const orc_scene *s, const orc_bdptref_params *pp, const float *table, uint32_t table_width,
                         uint64_t first_sample, uint64_t n, float *out, int nthreads)
{
    bctx b;
    b.s = s;
    camera_init(&s->d, &b.cam);
    b.tri_light = (int *)malloc(sizeof(int) * (s->d.ntri ? s->d.ntri : 1));
    for (uint32_t i = 0; i < s->d.ntri; i++) b.tri_light[i] = -1;
    for (uint32_t l = 0; l < s->d.nlight; l++)
        for (uint32_t k = 0; k < s->d.light_prims[2 * l + 1]; k++) b.tri_light[s->d.light_prims[2 * l] + k] = (int)l;
    b.tri_cdf_base = NULL;
    const int M = pp->max_num_vertices == -1 ? 32 : pp->max_num_vertices;
    rb_job_ctx c;
    c.b = &b; c.pp = pp; c.us.table = table; c.us.width = table_width;
    c.us.seed = ((uint64_t)pp->seed_hi << 32) | pp->seed_lo;
    c.us.first = first_sample;
    c.out = out;
    c.stride = 2 * (M + 1) * (M + 1);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    rb_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t].c = &c;
        jobs[t].b = n * (uint64_t)t / (uint64_t)nthreads;
        jobs[t].e = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, rb_run, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(b.tri_light);
}

/* ---------------------------------------------------------------------------
 * renderer::ptdirect (mode 1, renderer_ptdirect.cpp:72-297) and
 * renderer::ptmis (mode 2, renderer_ptmis.cpp:69-326): renderer::pt with
 * next-event estimation at every vertex, the camera's included.  Uniform
 * block of vertex v at 8 + 8 v: direction 0-1, roulette after the hit of
 * segment v 3, light selection 4, light position 5-6.  out[16]: raster x, y,
 * C rgb, count (connections and hits splatted at the sample's pixel), 0, 0,
 * raster x, y, C rgb, count of the camera-vertex connection, 0, 0. */
static void nee_sample(const bctx *b, const orc_pt_params *pp, int mode, const usrc *us, uint64_t sample,
                       float out[16])
{
    const orc_scene *s = b->s;
    const camera *cam = &b->cam;
    memset(out, 0, 16 * sizeof(float));
    float rx, ry;
    v3 wo;
    const float ux = U(us, sample, ORC_PT_RASTER), uy = U(us, sample, ORC_PT_RASTER + 1);
    if (pp->raster_mode == ORC_RASTER_STRATIFIED) {
        const uint64_t npix = (uint64_t)s->d.width * (uint64_t)s->d.height;
        const uint32_t pix = (uint32_t)(sample % npix);
        const uint32_t px = pix % (uint32_t)s->d.width, py = pix / (uint32_t)s->d.width;
        rx = ((float)px + ux) / (float)s->d.width;
        ry = ((float)py + uy) / (float)s->d.height;
        wo = camera_dir(cam, rx, ry);
        out[0] = ((float)px + 0.5f) / (float)s->d.width;
        out[1] = ((float)py + 0.5f) / (float)s->d.height;
    } else {
        wo = camera_dir(cam, ux, uy);
        if (!camera_raster(cam, wo, &rx, &ry)) return;           /* renderer_ptmis.cpp:95-103 */
        out[0] = rx; out[1] = ry;
    }
    float thr[3] = {1.0f, 1.0f, 1.0f};
    bvert cur;
    memset(&cur, 0, sizeof cur);
    cur.kind = BV_CAMERA; cur.degenerate = 1; cur.fr.p = cam->pos; cur.light = -1;
    v3 wi = mk(0, 0, 0);
    int nv = 1;
    for (int v = 0;; v++) {
        if (pp->max_num_vertices != -1 && nv >= pp->max_num_vertices) break;
        const uint32_t base = ORC_PT_SEG0 + 8u * (uint32_t)v;
        const int scat = cur.kind == BV_CAMERA ||
                         (cur.mat && (cur.mat->type == ORC_MAT_DIFFUSE || cur.mat->type == ORC_MAT_PHONG ||
                                      cur.mat->type == ORC_MAT_GGX));
        if (!scat) break;
        /* direct light sampling (ptmis :130-195, ptdirect :109-150) */
        if (s->d.nlight > 0 && (mode == 1 || nv + 1 >= pp->min_num_vertices)) {
            const float uL = U(us, sample, base + 4);
            const float up[2] = {U(us, sample, base + 5), U(us, sample, base + 6)};
            int l = (int)(uL * (float)s->d.nlight);
            l = l < 0 ? 0 : (l > (int)s->d.nlight - 1 ? (int)s->d.nlight - 1 : l);
            bvert L;
            memset(&L, 0, sizeof L);
            sample_light_point(b, l, up, &L);
            const v3 ppL = dir_to(&cur, &L);
            float fsE[3], pB;
            if (cur.kind == BV_CAMERA) {
                const float W = importance(cam, ppL);
                fsE[0] = fsE[1] = fsE[2] = W;
                pB = W;
            } else {
                bsdf_eval_dir(cur.mat, &cur.fr, wi, ppL, 0, fsE);
                pB = bsdf_pdf(cur.mat, &cur.fr, wi, ppL);
            }
            const float G = gterm(&cur, &L);
            const float inv_pl = (float)s->d.nlight / s->inv_area[l];
            float C[3];
            const int emits = vdot(L.fr.gn, vneg(ppL)) > 0.0f;
            for (int c = 0; c < 3; c++) C[c] = emits ? thr[c] * fsE[c] * G * L.mat->Le[c] * inv_pl : 0.0f;
            /* black contributions are dropped before the weight (:171) */
            if (mode == 2 && !(C[0] == 0.0f && C[1] == 0.0f && C[2] == 0.0f)) {
                const float pL = s->inv_area[l] / G * (1.0f / (float)s->d.nlight);
                const float w = pL / (pL + pB);
                for (int c = 0; c < 3; c++) C[c] *= w;
            }
            if (!(C[0] == 0.0f && C[1] == 0.0f && C[2] == 0.0f)) {
                float r2x = 0.0f, r2y = 0.0f;
                int ok = 1;
                if (cur.kind == BV_CAMERA) ok = camera_raster(cam, ppL, &r2x, &r2y);
                /* Scene3::Visible (scene3.h:107-116) */
                v3 d = vsub(L.fr.p, cur.fr.p);
                const float len = sqrtf(vdot(d, d));
                d = mk(d.x / len, d.y / len, d.z / len);
                orc_ray r;
                r.o[0] = cur.fr.p.x; r.o[1] = cur.fr.p.y; r.o[2] = cur.fr.p.z; r.tmin = EPS_ISECT;
                r.d[0] = d.x; r.d[1] = d.y; r.d[2] = d.z; r.tmax = len * (1.0f - EPS_ISECT);
                orc_hit h;
                orc_qbvh_intersect_ex(s->q, &r, 1, &h, ORC_TEST_WATERTIGHT, ORC_TIE_MIN_ID, 1);
                if (ok && h.id == 0xffffffffu) {
                    if (cur.kind == BV_CAMERA) {
                        out[8] = r2x; out[9] = r2y;
                        for (int c = 0; c < 3; c++) out[10 + c] += C[c];
                        out[13] += 1.0f;
                    } else {
                        for (int c = 0; c < 3; c++) out[2 + c] += C[c];
                        out[5] += 1.0f;
                    }
                }
            }
        }
        /* next direction; qB = its density (projected solid angle) times the
         * cosine at this vertex (1 at the pinhole) */
        float qB;
        if (cur.kind == BV_CAMERA) {
            qB = importance(cam, wo);
        } else {
            const float u2[2] = {U(us, sample, base), U(us, sample, base + 1)};
            if (!bsdf_sample(cur.mat, &cur.fr, wi, u2, &wo)) break;
            const float pdf = bsdf_pdf(cur.mat, &cur.fr, wi, wo);
            float fs[3];
            bsdf_eval(cur.mat, &cur.fr, wi, wo, fs);
            if (fs[0] == 0.0f && fs[1] == 0.0f && fs[2] == 0.0f) break;
            for (int c = 0; c < 3; c++) thr[c] *= fs[c] / pdf;
            qB = pdf * fabsf(to_local(&cur.fr, wo).z);
        }
        orc_hit h;
        if (!scene_intersect(s, cur.fr.p, wo, &h)) break;
        frame fr;
        surface_at(s, cur.fr.p, wo, &h, &fr);
        const orc_material *mat = &s->d.mats[s->d.tri_mat[h.id]];
        /* light hit (ptmis :229-257; ptdirect ignores it) */
        if (mode == 2 && (mat->Le[0] > 0.0f || mat->Le[1] > 0.0f || mat->Le[2] > 0.0f) &&
            nv + 1 >= pp->min_num_vertices) {
            const float cl = to_local(&fr, vneg(wo)).z;
            if (cl > 0.0f) {
                const int l = b->tri_light[h.id];
                const float pL = s->inv_area[l] * h.t * h.t / cl * (1.0f / (float)s->d.nlight);
                const float w = qB / (qB + pL);
                for (int c = 0; c < 3; c++) out[2 + c] += thr[c] * mat->Le[c] * w;
                out[5] += 1.0f;
            }
        }
        /* Russian roulette (ptmis :261-269), from depth rr_depth on */
        if (v + 1 >= pp->rr_depth) {
            float q = pp->rr_prob;
            if (pp->rr_mode == 1) {
                q = fmaxf(thr[0], fmaxf(thr[1], thr[2]));
                q = q < 1.0f ? q : 1.0f;
            }
            if (!(q > 0.0f) || U(us, sample, base + 3) > q) break;
            for (int c = 0; c < 3; c++) thr[c] /= q;
        }
        wi = vneg(wo);
        cur.kind = BV_SURFACE; cur.degenerate = 0; cur.fr = fr; cur.mat = mat; cur.tri = h.id;
        cur.light = b->tri_light[h.id];
        nv++;
    }
}

typedef struct { bctx *b; const orc_pt_params *pp; int mode; usrc us; uint64_t first; float *out; } nee_job_ctx;
typedef struct { nee_job_ctx *c; uint64_t b, e; } nee_job;

static void *nee_run(void *p)
{
    nee_job *j = (nee_job *)p;
    for (uint64_t i = j->b; i < j->e; i++)
        nee_sample(j->c->b, j->c->pp, j->c->mode, &j->c->us, j->c->first + i, &j->c->out[(size_t)16 * i]);
    return NULL;
}

void orc_ptnee_samples(const orc_scene *s, const orc_pt_params *pp, int mode, const float *table,
                       uint32_t table_width, uint64_t first_sample, uint64_t n, float *out, int nthreads)
{
    bctx b;
    b.s = s;
    camera_init(&s->d, &b.cam);
    b.tri_light = (int *)malloc(sizeof(int) * (s->d.ntri ? s->d.ntri : 1));
    for (uint32_t i = 0; i < s->d.ntri; i++) b.tri_light[i] = -1;
    for (uint32_t l = 0; l < s->d.nlight; l++)
        for (uint32_t k = 0; k < s->d.light_prims[2 * l + 1]; k++) b.tri_light[s->d.light_prims[2 * l] + k] = (int)l;
    b.tri_cdf_base = NULL;
    nee_job_ctx c;
    c.b = &b; c.pp = pp; c.mode = mode;
    c.us.table = table; c.us.width = table_width;
    c.us.seed = ((uint64_t)pp->seed_hi << 32) | pp->seed_lo;
    c.us.first = first_sample;
    c.first = first_sample; c.out = out;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    nee_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t].c = &c;
        jobs[t].b = n * (uint64_t)t / (uint64_t)nthreads;
        jobs[t].e = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, nee_run, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(b.tri_light);
}

/* ---------------------------------------------------------------------------
 * renderer::pm (renderer_pm.cpp:61-220) with photonmap::kdtree / naive
 * (photonmap_kdtree.cpp:123-165, photonmap_naive.cpp:45-56: every photon
 * with |p - x|^2 < r^2), restated literally through TraceSubpath_
 * (subpathsampler.cpp:42-203):
 *  - photon subpath i: an emitter point and direction; the callback of
 *    every vertex stores a photon at a diffuse or glossy vertex (position,
 *    throughput at the vertex, direction to the previous vertex, vertex count)
 *    and then plays roulette 0.5 (the endpoint included), the survivor's
 *    throughput doubled.  Uniform row of the photon: 1-2 position, 3 emitter
 *    selection, 4-5 direction, 6 roulette after the endpoint; vertex j >= 1
 *    owns ORC_PT_SEG0 + 8 (j - 1) .. +7: 0-1 the direction sampled at vertex
 *    j - 1 (j >= 2), 3 the roulette after vertex j;
 *  - eye sample: the eye subpath of renderer::vcm's layout without roulette;
 *    at every vertex past the camera an emitter adds throughput x Le, and the
 *    first (finalgather 0) or second (finalgather 1) diffuse or glossy vertex
 *    gathers the photons within r with the kernel 3/pi (1 - d^2/r^2)^2 over
 *    r^2 N, times the BSDF towards the photon's incoming direction and the
 *    photon's throughput, and ends the path.
 * The photon set is found by a uniform grid (the same set as the kd-tree).
 * max_num_vertices -1 is unbounded (subpaths cut at 32 vertices) and then
 * gathers every photon (the reference compares the path length with -1 and
 * so drops every photon). */
#define PM_CAP 32

typedef struct {
    v3 p, wi;
    float thr[3];
    int nv;
} pm_photon;

static int pm_trace_photon(const bctx *b, int maxv, const usrc *us, uint64_t i, bvert *vs)
{
    const orc_scene *s = b->s;
    int nv = 0;
    v3 init_wo = mk(0, 0, 0);
    float run[3] = {0.0f, 0.0f, 0.0f};
    for (int step = 0; (maxv == -1 || step < maxv) && step < PM_CAP; step++) {
        if (step == 0) {
            if (s->d.nlight == 0) return 0;
            bvert v;
            memset(&v, 0, sizeof v);
            const float uEP[2] = {U(us, i, 1), U(us, i, 2)};
            const float uEPC = U(us, i, 3);
            const float uED[2] = {U(us, i, 4), U(us, i, 5)};
            int l = (int)(uEPC * (float)s->d.nlight);
            l = l < 0 ? 0 : (l > (int)s->d.nlight - 1 ? (int)s->d.nlight - 1 : l);
            sample_light_point(b, l, uEP, &v);
            float lw[3];
            cos_hemisphere(uED, lw);
            init_wo = to_world(&v.fr, mk(lw[0], lw[1], lw[2]));
            const float a = 1.0f / pos_pdf(b, &v);
            v.alpha[0] = v.alpha[1] = v.alpha[2] = a;
            vs[nv++] = v;
            if (U(us, i, 6) > 0.5f) break;                   /* renderer_pm.cpp:111-120 */
            for (int c = 0; c < 3; c++) run[c] = v.alpha[c] / 0.5f;
            continue;
        }
        const uint32_t sb = ORC_PT_SEG0 + 8u * (uint32_t)(step - 1);
        bvert *pv = &vs[nv - 1];
        v3 wi, wo;
        if (step == 1) {
            wi = mk(0, 0, 0);
            wo = init_wo;
        } else {
            wi = dir_to(pv, &vs[nv - 2]);
            const float uD[2] = {U(us, i, sb), U(us, i, sb + 1)};
            if (pv->kind != BV_SURFACE || !bsdf_sample(pv->mat, &pv->fr, wi, uD, &wo)) break;
        }
        float fs[3];
        v_eval(b, pv, wi, wo, 1, fs);
        if (fs[0] == 0.0f && fs[1] == 0.0f && fs[2] == 0.0f) break;
        const float pdf = v_pdf(b, pv, wi, wo);
        orc_hit h;
        if (!scene_intersect(s, pv->fr.p, wo, &h)) break;
        bvert v;
        memset(&v, 0, sizeof v);
        surface_at(s, pv->fr.p, wo, &h, &v.fr);
        v.kind = BV_SURFACE; v.degenerate = 0; v.tri = h.id;
        v.mat = &s->d.mats[s->d.tri_mat[h.id]];
        v.light = b->tri_light[h.id];
        for (int c = 0; c < 3; c++) v.alpha[c] = run[c] * (fs[c] / pdf);
        vs[nv++] = v;
        if (U(us, i, sb + 3) > 0.5f) break;
        for (int c = 0; c < 3; c++) run[c] = v.alpha[c] / 0.5f;
    }
    return nv;
}

static int pm_scatters(const orc_material *m)
{
    return m && (m->type == ORC_MAT_DIFFUSE || m->type == ORC_MAT_PHONG || m->type == ORC_MAT_GGX);
}

typedef struct {
    const bctx *b;
    const orc_pm_params *pp;
    usrc us;
    const pm_photon *ph;
    float lo[3], cell;
    int dim[3];
    int *cstart, *items;
    float *out;
    uint64_t first;
} pm_ctx;

static void pm_sample(const pm_ctx *c, uint64_t sample, float *out)
{
    const bctx *b = c->b;
    const orc_pm_params *pp = c->pp;
    const float r = pp->radius;
    memset(out, 0, sizeof(float) * 16);
    int cap = pp->max_num_vertices == -1 ? PM_CAP : pp->max_num_vertices;
    if (cap > 2 + pp->finalgather) cap = 2 + pp->finalgather;     /* the path ends at the gather vertex */
    orc_bdpt_params bp = {cap, 0, pp->seed_lo, pp->seed_hi, pp->raster_mode};
    bvert E[PM_CAP];
    const int nE = trace_subpath(b, &bp, &c->us, sample, 0, 0, E);
    float rx = 0.0f, ry = 0.0f;
    if (pp->raster_mode == ORC_RASTER_STRATIFIED) {
        const uint32_t npix = (uint32_t)(b->s->d.width * b->s->d.height);
        const uint32_t pix = (uint32_t)(sample % (uint64_t)npix);
        rx = ((float)(pix % (uint32_t)b->s->d.width) + 0.5f) / (float)b->s->d.width;
        ry = ((float)(pix / (uint32_t)b->s->d.width) + 0.5f) / (float)b->s->d.height;
    } else if (nE >= 2) {
        if (!camera_raster(&b->cam, dir_to(&E[0], &E[1]), &rx, &ry)) return;
    }
    out[0] = rx; out[1] = ry;
    int gather_next = !pp->finalgather;
    for (int k = 1; k < nE; k++) {
        const bvert *v = &E[k], *pv = &E[k - 1];
        /* an emitter hit: throughput x Le towards the previous vertex (:158-167) */
        if (v->light >= 0) {
            bvert lv = *v;
            lv.kind = BV_LIGHT;
            float le[3];
            v_eval(b, &lv, mk(0, 0, 0), dir_to(v, pv), 0, le);
            for (int q = 0; q < 3; q++) out[2 + q] += v->alpha[q] * le[q];
            out[5] += 1.0f;
        }
        if (!pm_scatters(v->mat)) continue;
        if (!gather_next) { gather_next = 1; continue; }
        /* density estimation (:169-197) */
        out[10] = (float)(k + 1);
        int lo[3], hi[3];
        const float q3[3] = {v->fr.p.x, v->fr.p.y, v->fr.p.z};
        for (int a = 0; a < 3; a++) {
            lo[a] = (int)floorf((q3[a] - r - c->lo[a]) / c->cell);
            hi[a] = (int)floorf((q3[a] + r - c->lo[a]) / c->cell);
            if (lo[a] < 0) lo[a] = 0;
            if (hi[a] > c->dim[a] - 1) hi[a] = c->dim[a] - 1;
        }
        const v3 wiE = dir_to(v, pv);
        for (int gz = lo[2]; gz <= hi[2]; gz++)
            for (int gy = lo[1]; gy <= hi[1]; gy++)
                for (int gx = lo[0]; gx <= hi[0]; gx++) {
                    const int cidx = (gz * c->dim[1] + gy) * c->dim[0] + gx;
                    for (int it = c->cstart[cidx]; it < c->cstart[cidx + 1]; it++) {
                        const pm_photon *ph = &c->ph[c->items[it]];
                        const v3 d = vsub(ph->p, v->fr.p);
                        const float d2 = vdot(d, d);
                        if (!(d2 < r * r)) continue;
                        if (pp->max_num_vertices != -1 && (k + 1) + ph->nv - 1 > pp->max_num_vertices) continue;
                        const float sk = 1.0f - d2 / r / r;
                        const float kern = 3.0f * INV_PI_F * sk * sk;
                        const float p = kern / (r * r * (float)pp->num_photons);
                        float f[3];
                        v_eval(b, v, wiE, ph->wi, 0, f);
                        for (int q = 0; q < 3; q++) out[6 + q] += v->alpha[q] * p * f[q] * ph->thr[q];
                        out[9] += 1.0f;
                    }
                }
        break;
    }
}

typedef struct { pm_ctx *c; uint64_t b, e; } pm_job;

static void *pm_run(void *p)
{
    pm_job *j = (pm_job *)p;
    for (uint64_t i = j->b; i < j->e; i++) pm_sample(j->c, j->c->first + i, &j->c->out[16 * i]);
    return NULL;
}

/* photon subpaths [lo, hi): their photons, in subpath order, into a growing
 * array of the job's own */
typedef struct {
    const bctx *b;
    int maxv;
    const usrc *pu;
    uint32_t lo, hi;
    pm_photon *ph;
    size_t n, cap;
} pm_pjob;

static void *pm_photons(void *p)
{
    pm_pjob *j = (pm_pjob *)p;
    bvert vs[PM_CAP];
    j->ph = NULL; j->n = j->cap = 0;
    for (uint32_t i = j->lo; i < j->hi; i++) {
        const int nv = pm_trace_photon(j->b, j->maxv, j->pu, i, vs);
        /* the photons: diffuse or glossy vertices past the light (:98-110) */
        for (int k = 1; k < nv; k++) {
            if (!pm_scatters(vs[k].mat)) continue;
            if (j->n == j->cap) {
                j->cap = j->cap ? 2 * j->cap : 4096;
                j->ph = (pm_photon *)realloc(j->ph, sizeof(pm_photon) * j->cap);
            }
            pm_photon *q = &j->ph[j->n++];
            q->p = vs[k].fr.p;
            for (int c = 0; c < 3; c++) q->thr[c] = vs[k].alpha[c];
            q->wi = dir_to(&vs[k], &vs[k - 1]);
            q->nv = k + 1;
        }
    }
    return NULL;
}

void orc_pm_samples(const orc_scene *s, const orc_pm_params *pp, const float *table, uint32_t table_width,
                    const float *ptable, uint32_t ptable_width, uint64_t first_sample, uint64_t n, float *out,
                    int nthreads)
{
    bctx b;
    b.s = s;
    camera_init(&s->d, &b.cam);
    b.tri_light = (int *)malloc(sizeof(int) * (s->d.ntri ? s->d.ntri : 1));
    for (uint32_t i = 0; i < s->d.ntri; i++) b.tri_light[i] = -1;
    for (uint32_t l = 0; l < s->d.nlight; l++)
        for (uint32_t k = 0; k < s->d.light_prims[2 * l + 1]; k++) b.tri_light[s->d.light_prims[2 * l] + k] = (int)l;
    b.tri_cdf_base = NULL;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    const uint32_t np = pp->num_photons;
    usrc pu;
    pu.table = ptable; pu.width = ptable_width;
    pu.seed = ((uint64_t)pp->pseed_hi << 32) | pp->pseed_lo;
    pu.first = 0;
    pm_pjob pj[256];
    {
        pthread_t th[256];
        for (int t = 0; t < nthreads; t++) {
            pj[t].b = &b; pj[t].maxv = pp->max_num_vertices; pj[t].pu = &pu;
            pj[t].lo = (uint32_t)((uint64_t)np * (uint64_t)t / (uint64_t)nthreads);
            pj[t].hi = (uint32_t)((uint64_t)np * (uint64_t)(t + 1) / (uint64_t)nthreads);
            pthread_create(&th[t], NULL, pm_photons, &pj[t]);
        }
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    }
    size_t m = 0;
    for (int t = 0; t < nthreads; t++) m += pj[t].n;
    pm_photon *ph = (pm_photon *)malloc(sizeof(pm_photon) * (m + 1));
    m = 0;
    for (int t = 0; t < nthreads; t++) {
        if (pj[t].n) memcpy(&ph[m], pj[t].ph, sizeof(pm_photon) * pj[t].n);
        m += pj[t].n;
        free(pj[t].ph);
    }
    float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
    for (size_t k = 0; k < m; k++) {
        const float a3[3] = {ph[k].p.x, ph[k].p.y, ph[k].p.z};
        for (int a = 0; a < 3; a++) { if (a3[a] < lo[a]) lo[a] = a3[a]; if (a3[a] > hi[a]) hi[a] = a3[a]; }
    }
    pm_ctx c;
    c.b = &b; c.pp = pp; c.ph = ph;
    c.cell = pp->radius;
    float span = 0.0f;
    for (int a = 0; a < 3; a++) {
        if (m == 0) { lo[a] = 0.0f; hi[a] = 0.0f; }
        c.lo[a] = lo[a];
        if (hi[a] - lo[a] > span) span = hi[a] - lo[a];
    }
    if (span / c.cell > 255.0f) c.cell = span / 255.0f;     /* a coarser grid finds the same set */
    long long cells = 1;
    for (int a = 0; a < 3; a++) { c.dim[a] = (int)floorf((hi[a] - lo[a]) / c.cell) + 1; cells *= c.dim[a]; }
    c.cstart = (int *)calloc((size_t)cells + 1, sizeof(int));
    c.items = (int *)malloc(sizeof(int) * (m + 1));
    int *cellof = (int *)malloc(sizeof(int) * (m + 1));
    for (size_t k = 0; k < m; k++) {
        const float a3[3] = {ph[k].p.x, ph[k].p.y, ph[k].p.z};
        int g[3];
        for (int a = 0; a < 3; a++) {
            g[a] = (int)floorf((a3[a] - c.lo[a]) / c.cell);
            if (g[a] < 0) g[a] = 0;
            if (g[a] > c.dim[a] - 1) g[a] = c.dim[a] - 1;
        }
        cellof[k] = (g[2] * c.dim[1] + g[1]) * c.dim[0] + g[0];
        c.cstart[cellof[k] + 1]++;
    }
    for (long long k = 0; k < cells; k++) c.cstart[k + 1] += c.cstart[k];
    int *fill = (int *)calloc((size_t)cells, sizeof(int));
    for (size_t k = 0; k < m; k++) c.items[c.cstart[cellof[k]] + fill[cellof[k]]++] = (int)k;
    free(fill); free(cellof);
    c.us.table = table; c.us.width = table_width;
    c.us.seed = ((uint64_t)pp->seed_hi << 32) | pp->seed_lo;
    c.us.first = first_sample;
    c.out = out;
    c.first = first_sample;
    pthread_t th[256];
    pm_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t].c = &c;
        jobs[t].b = n * (uint64_t)t / (uint64_t)nthreads;
        jobs[t].e = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, pm_run, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(c.cstart); free(c.items); free(ph); free(b.tri_light);
}

/* ---------------------------------------------------------------------------
 * renderer::ppm (renderer_ppm.cpp:84-320), restated literally:
 *  - measurement points: eye subpath i (renderer::vcm's eye layout, no
 *    roulette) stops at its first diffuse or glossy vertex, which becomes the
 *    point (radius r0, N = 0, tau = 0, throughput, direction to the previous
 *    vertex, vertex count), with the emission there if that surface emits
 *    (a light-only surface is no point and ends the path: directly seen
 *    emitters stay black, :131-168);
 *  - pass q: photon subpaths q Np + i of the photon source, roulette 0.5
 *    after every vertex but the endpoint (renderer::bdpt's light-block
 *    layout: vertex j >= 1 owns ORC_PT_SEG0 + 8 (j - 1) .. +7, 0-1 the
 *    direction sampled at vertex j - 1, 3 the roulette after vertex j), a
 *    photon at every diffuse or glossy vertex past the light (:204-240);
 *    every point gathers the pass's photons within its radius (M of them,
 *    dtau = sum f x photon throughput) and updates ratio = (N + a M) /
 *    (N + M), tau = (tau + dtau) ratio, r = r sqrt(ratio), N = N + a M
 *    (:270-297);
 *  - the film after the last pass: throughput x tau / (pi r^2 P Np) +
 *    emission per point (:300-310).
 * max_num_vertices -1 is unbounded (32-vertex subpaths, every photon). */
typedef struct {
    int valid;
    bvert v;
    v3 wi;
    float thr[3], em[3], tau[3];
    float N, r;
    int nv;
    float rx, ry;
} ppm_point;

static int ppm_trace_photon(const bctx *b, int maxv, const usrc *us, uint64_t i, bvert *vs)
{
    const orc_scene *s = b->s;
    int nv = 0;
    v3 init_wo = mk(0, 0, 0);
    float run[3] = {0.0f, 0.0f, 0.0f};
    for (int step = 0; (maxv == -1 || step < maxv) && step < PM_CAP; step++) {
        if (step == 0) {
            if (s->d.nlight == 0) return 0;
            bvert v;
            memset(&v, 0, sizeof v);
            const float uEP[2] = {U(us, i, 1), U(us, i, 2)};
            const float uEPC = U(us, i, 3);
            const float uED[2] = {U(us, i, 4), U(us, i, 5)};
            int l = (int)(uEPC * (float)s->d.nlight);
            l = l < 0 ? 0 : (l > (int)s->d.nlight - 1 ? (int)s->d.nlight - 1 : l);
            sample_light_point(b, l, uEP, &v);
            float lw[3];
            cos_hemisphere(uED, lw);
            init_wo = to_world(&v.fr, mk(lw[0], lw[1], lw[2]));
            const float a = 1.0f / pos_pdf(b, &v);
            v.alpha[0] = v.alpha[1] = v.alpha[2] = a;
            vs[nv++] = v;
            for (int c = 0; c < 3; c++) run[c] = a;              /* the endpoint skips the roulette (:210-214) */
            continue;
        }
        const uint32_t sb = ORC_PT_SEG0 + 8u * (uint32_t)(step - 1);
        bvert *pv = &vs[nv - 1];
        v3 wi, wo;
        if (step == 1) {
            wi = mk(0, 0, 0);
            wo = init_wo;
        } else {
            wi = dir_to(pv, &vs[nv - 2]);
            const float uD[2] = {U(us, i, sb), U(us, i, sb + 1)};
            if (pv->kind != BV_SURFACE || !bsdf_sample(pv->mat, &pv->fr, wi, uD, &wo)) break;
        }
        float fs[3];
        v_eval(b, pv, wi, wo, 1, fs);
        if (fs[0] == 0.0f && fs[1] == 0.0f && fs[2] == 0.0f) break;
        const float pdf = v_pdf(b, pv, wi, wo);
        orc_hit h;
        if (!scene_intersect(s, pv->fr.p, wo, &h)) break;
        bvert v;
        memset(&v, 0, sizeof v);
        surface_at(s, pv->fr.p, wo, &h, &v.fr);
        v.kind = BV_SURFACE; v.degenerate = 0; v.tri = h.id;
        v.mat = &s->d.mats[s->d.tri_mat[h.id]];
        v.light = b->tri_light[h.id];
        for (int c = 0; c < 3; c++) v.alpha[c] = run[c] * (fs[c] / pdf);
        vs[nv++] = v;
        if (U(us, i, sb + 3) > 0.5f) break;                      /* :227-237 */
        for (int c = 0; c < 3; c++) run[c] = v.alpha[c] / 0.5f;
    }
    return nv;
}

typedef struct {
    const bctx *b;
    int maxv;
    const usrc *pu;
    uint64_t first;
    uint32_t lo, hi;
    pm_photon *ph;
    size_t n, cap;
} ppm_pjob;

static void *ppm_photons(void *p)
{
    ppm_pjob *j = (ppm_pjob *)p;
    bvert vs[PM_CAP];
    j->ph = NULL; j->n = j->cap = 0;
    for (uint32_t i = j->lo; i < j->hi; i++) {
        const int nv = ppm_trace_photon(j->b, j->maxv, j->pu, j->first + i, vs);
        for (int k = 1; k < nv; k++) {
            if (!pm_scatters(vs[k].mat)) continue;
            if (j->n == j->cap) {
                j->cap = j->cap ? 2 * j->cap : 4096;
                j->ph = (pm_photon *)realloc(j->ph, sizeof(pm_photon) * j->cap);
            }
            pm_photon *q = &j->ph[j->n++];
            q->p = vs[k].fr.p;
            for (int c = 0; c < 3; c++) q->thr[c] = vs[k].alpha[c];
            q->wi = dir_to(&vs[k], &vs[k - 1]);
            q->nv = k + 1;
        }
    }
    return NULL;
}

typedef struct {
    const bctx *b;
    const orc_ppm_params *pp;
    ppm_point *mp;
    const pm_photon *ph;
    size_t nph;
    float lo[3], cell;
    int dim[3];
    const int *cstart, *items;
    uint64_t b0, e0;
    int sppm;               /* renderer::sppm: tau gains throughput x dtau (renderer_sppm.cpp:313) */
} ppm_ujob;

static void *ppm_update(void *p)
{
    ppm_ujob *j = (ppm_ujob *)p;
    const orc_ppm_params *pp = j->pp;
    for (uint64_t k = j->b0; k < j->e0; k++) {
        ppm_point *m = &j->mp[k];
        if (!m->valid) continue;
        const float r = m->r;
        float dtau[3] = {0.0f, 0.0f, 0.0f};
        float M = 0.0f;
        int lo[3], hi[3];
        const float q3[3] = {m->v.fr.p.x, m->v.fr.p.y, m->v.fr.p.z};
        for (int a = 0; a < 3; a++) {
            lo[a] = (int)floorf((q3[a] - r - j->lo[a]) / j->cell);
            hi[a] = (int)floorf((q3[a] + r - j->lo[a]) / j->cell);
            if (lo[a] < 0) lo[a] = 0;
            if (hi[a] > j->dim[a] - 1) hi[a] = j->dim[a] - 1;
        }
        for (int gz = lo[2]; gz <= hi[2]; gz++)
            for (int gy = lo[1]; gy <= hi[1]; gy++)
                for (int gx = lo[0]; gx <= hi[0]; gx++) {
                    const int cidx = (gz * j->dim[1] + gy) * j->dim[0] + gx;
                    for (int it = j->cstart[cidx]; it < j->cstart[cidx + 1]; it++) {
                        const pm_photon *ph = &j->ph[j->items[it]];
                        const v3 d = vsub(ph->p, m->v.fr.p);
                        if (!(vdot(d, d) < r * r)) continue;
                        if (pp->max_num_vertices != -1 && m->nv + ph->nv - 1 > pp->max_num_vertices) continue;
                        float f[3];
                        v_eval(j->b, &m->v, m->wi, ph->wi, 0, f);
                        for (int c = 0; c < 3; c++) dtau[c] += f[c] * ph->thr[c];
                        M += 1.0f;
                    }
                }
        if (m->N + M == 0.0f) continue;
        const float ratio = (m->N + pp->alpha * M) / (m->N + M);
        for (int c = 0; c < 3; c++) m->tau[c] = (m->tau[c] + (j->sppm ? m->thr[c] * dtau[c] : dtau[c])) * ratio;
        m->r = m->r * sqrtf(ratio);
        m->N = m->N + pp->alpha * M;
    }
    return NULL;
}

typedef struct { const bctx *b; const orc_ppm_params *pp; usrc us; ppm_point *mp; uint64_t lo, hi; } ppm_ejob;

static void *ppm_points(void *p)
{
    ppm_ejob *j = (ppm_ejob *)p;
    const bctx *b = j->b;
    const orc_ppm_params *pp = j->pp;
    orc_bdpt_params bp = {2, 0, pp->seed_lo, pp->seed_hi, pp->raster_mode};   /* the point is vertex 1 or none */
    for (uint64_t i = j->lo; i < j->hi; i++) {
        ppm_point *m = &j->mp[i];
        memset(m, 0, sizeof *m);
        bvert E[2];
        const int nE = trace_subpath(b, &bp, &j->us, i, 0, 0, E);
        if (nE < 2 || !pm_scatters(E[1].mat)) continue;
        m->valid = 1;
        m->v = E[1];
        m->wi = dir_to(&E[1], &E[0]);
        for (int c = 0; c < 3; c++) m->thr[c] = E[1].alpha[c];
        m->nv = 2;
        m->r = pp->initial_radius;
        if (E[1].light >= 0) {
            bvert lv = E[1];
            lv.kind = BV_LIGHT;
            float le[3];
            v_eval(b, &lv, mk(0, 0, 0), m->wi, 0, le);
            for (int c = 0; c < 3; c++) m->em[c] = m->thr[c] * le[c];
        }
        if (pp->raster_mode == ORC_RASTER_STRATIFIED) {
            const uint32_t npix = (uint32_t)(b->s->d.width * b->s->d.height);
            const uint32_t pix = (uint32_t)(i % (uint64_t)npix);
            m->rx = ((float)(pix % (uint32_t)b->s->d.width) + 0.5f) / (float)b->s->d.width;
            m->ry = ((float)(pix / (uint32_t)b->s->d.width) + 0.5f) / (float)b->s->d.height;
        } else if (!camera_raster(&b->cam, dir_to(&E[0], &E[1]), &m->rx, &m->ry)) {
            m->valid = 0;
        }
    }
    return NULL;
}

/* one photon pass of renderer::ppm / sppm: the pass's photon subpaths
 * (photon rows pass Np + i), a uniform grid of cell >= every radius, and the
 * progressive update of the n points */
static void ppm_pass(const bctx *b, const orc_ppm_params *pp, const usrc *pu, uint32_t pass, ppm_point *mp,
                     uint64_t n, int sppm, int nthreads)
{
    pthread_t th[256];
    const uint32_t np = pp->num_photons;
    {
        ppm_pjob pj[256];
        for (int t = 0; t < nthreads; t++) {
            pj[t].b = b; pj[t].maxv = pp->max_num_vertices; pj[t].pu = pu;
            pj[t].first = (uint64_t)pass * np;
            pj[t].lo = (uint32_t)((uint64_t)np * (uint64_t)t / (uint64_t)nthreads);
            pj[t].hi = (uint32_t)((uint64_t)np * (uint64_t)(t + 1) / (uint64_t)nthreads);
            pthread_create(&th[t], NULL, ppm_photons, &pj[t]);
        }
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
        size_t m = 0;
        for (int t = 0; t < nthreads; t++) m += pj[t].n;
        pm_photon *ph = (pm_photon *)malloc(sizeof(pm_photon) * (m + 1));
        m = 0;
        for (int t = 0; t < nthreads; t++) {
            if (pj[t].n) memcpy(&ph[m], pj[t].ph, sizeof(pm_photon) * pj[t].n);
            m += pj[t].n;
            free(pj[t].ph);
        }
        /* a uniform grid of cell >= the largest radius (the initial one) */
        float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
        for (size_t k = 0; k < m; k++) {
            const float a3[3] = {ph[k].p.x, ph[k].p.y, ph[k].p.z};
            for (int a = 0; a < 3; a++) { if (a3[a] < lo[a]) lo[a] = a3[a]; if (a3[a] > hi[a]) hi[a] = a3[a]; }
        }
        float cell = pp->initial_radius, span = 0.0f;
        for (int a = 0; a < 3; a++) {
            if (m == 0) { lo[a] = 0.0f; hi[a] = 0.0f; }
            if (hi[a] - lo[a] > span) span = hi[a] - lo[a];
        }
        if (span / cell > 255.0f) cell = span / 255.0f;
        int dim[3];
        long long cells = 1;
        for (int a = 0; a < 3; a++) { dim[a] = (int)floorf((hi[a] - lo[a]) / cell) + 1; cells *= dim[a]; }
        int *cstart = (int *)calloc((size_t)cells + 1, sizeof(int));
        int *items = (int *)malloc(sizeof(int) * (m + 1));
        int *cellof = (int *)malloc(sizeof(int) * (m + 1));
        for (size_t k = 0; k < m; k++) {
            const float a3[3] = {ph[k].p.x, ph[k].p.y, ph[k].p.z};
            int g[3];
            for (int a = 0; a < 3; a++) {
                g[a] = (int)floorf((a3[a] - lo[a]) / cell);
                if (g[a] < 0) g[a] = 0;
                if (g[a] > dim[a] - 1) g[a] = dim[a] - 1;
            }
            cellof[k] = (g[2] * dim[1] + g[1]) * dim[0] + g[0];
            cstart[cellof[k] + 1]++;
        }
        for (long long k = 0; k < cells; k++) cstart[k + 1] += cstart[k];
        int *fill = (int *)calloc((size_t)cells, sizeof(int));
        for (size_t k = 0; k < m; k++) items[cstart[cellof[k]] + fill[cellof[k]]++] = (int)k;
        free(fill); free(cellof);
        ppm_ujob uj[256];
        for (int t = 0; t < nthreads; t++) {
            uj[t].b = b; uj[t].pp = pp; uj[t].mp = mp; uj[t].ph = ph; uj[t].nph = m; uj[t].sppm = sppm;
            for (int a = 0; a < 3; a++) { uj[t].lo[a] = lo[a]; uj[t].dim[a] = dim[a]; }
            uj[t].cell = cell; uj[t].cstart = cstart; uj[t].items = items;
            uj[t].b0 = n * (uint64_t)t / (uint64_t)nthreads;
            uj[t].e0 = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
            pthread_create(&th[t], NULL, ppm_update, &uj[t]);
        }
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
        free(cstart); free(items); free(ph);
    }
}

void orc_ppm_points(const orc_scene *s, const orc_ppm_params *pp, const float *table, uint32_t table_width,
                    const float *ptable, uint32_t ptable_width, uint64_t n, float *out, int nthreads)
{
    bctx b;
    b.s = s;
    camera_init(&s->d, &b.cam);
    b.tri_light = (int *)malloc(sizeof(int) * (s->d.ntri ? s->d.ntri : 1));
    for (uint32_t i = 0; i < s->d.ntri; i++) b.tri_light[i] = -1;
    for (uint32_t l = 0; l < s->d.nlight; l++)
        for (uint32_t k = 0; k < s->d.light_prims[2 * l + 1]; k++) b.tri_light[s->d.light_prims[2 * l] + k] = (int)l;
    b.tri_cdf_base = NULL;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    ppm_point *mp = (ppm_point *)calloc(n + 1, sizeof(ppm_point));
    pthread_t th[256];
    {
        ppm_ejob ej[256];
        for (int t = 0; t < nthreads; t++) {
            ej[t].b = &b; ej[t].pp = pp; ej[t].mp = mp;
            ej[t].us.table = table; ej[t].us.width = table_width;
            ej[t].us.seed = ((uint64_t)pp->seed_hi << 32) | pp->seed_lo;
            ej[t].us.first = 0;
            ej[t].lo = n * (uint64_t)t / (uint64_t)nthreads;
            ej[t].hi = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
            pthread_create(&th[t], NULL, ppm_points, &ej[t]);
        }
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    }
    usrc pu;
    pu.table = ptable; pu.width = ptable_width;
    pu.seed = ((uint64_t)pp->pseed_hi << 32) | pp->pseed_lo;
    pu.first = 0;
    const uint32_t np = pp->num_photons;
    for (uint32_t pass = 0; pass < pp->num_passes; pass++) ppm_pass(&b, pp, &pu, pass, mp, n, 0, nthreads);
    const double total = (double)pp->num_passes * (double)np;
    for (uint64_t i = 0; i < n; i++) {
        float *o = &out[16 * i];
        memset(o, 0, sizeof(float) * 16);
        const ppm_point *m = &mp[i];
        if (!m->valid) continue;
        /* :303-307 */
        const float p = 1.0f / (m->r * m->r * PI_F * (float)total);
        o[0] = m->rx; o[1] = m->ry;
        for (int c = 0; c < 3; c++) {
            o[2 + c] = m->thr[c] * p * m->tau[c] + m->em[c];
            o[5 + c] = m->em[c];
            o[8 + c] = m->tau[c];
        }
        o[11] = m->N; o[12] = m->r; o[13] = 1.0f;
    }
    free(mp); free(b.tri_light);
}

/* renderer::sppm (renderer_sppm.cpp:98-378): a measurement point per pixel
 * keeps its radius, N and tau across passes; pass q re-traces pixel i's eye
 * subpath (uniform index q W H + i, renderer::vcm's eye layout with the
 * raster jittered inside the pixel, :159-197) to its first diffuse or
 * glossy vertex, which becomes the point's position for the pass (no such
 * vertex: the point skips the pass), adding its emission to the point's
 * sum; then the pass's photons as renderer::ppm's and the update with tau
 * gaining throughput x dtau; the film after the last pass is, per pixel,
 * tau / (pi r^2 P Np) + emission / P (SetPixel, no rescale, :321-327). */
typedef struct { const bctx *b; const orc_ppm_params *pp; usrc us; ppm_point *mp; uint32_t pass; uint64_t lo, hi; } sppm_ejob;

static void *sppm_points(void *p)
{
    sppm_ejob *j = (sppm_ejob *)p;
    const bctx *b = j->b;
    const orc_ppm_params *pp = j->pp;
    const uint64_t npix = (uint64_t)b->s->d.width * (uint64_t)b->s->d.height;
    orc_bdpt_params bp = {2, 0, pp->seed_lo, pp->seed_hi, ORC_RASTER_STRATIFIED};
    for (uint64_t i = j->lo; i < j->hi; i++) {
        ppm_point *m = &j->mp[i];
        m->valid = 0;
        bvert E[2];
        const int nE = trace_subpath(b, &bp, &j->us, (uint64_t)j->pass * npix + i, 0, 0, E);
        if (nE < 2 || !pm_scatters(E[1].mat)) continue;
        m->valid = 1;
        m->v = E[1];
        m->wi = dir_to(&E[1], &E[0]);
        for (int c = 0; c < 3; c++) m->thr[c] = E[1].alpha[c];
        m->nv = 2;
        if (E[1].light >= 0) {
            bvert lv = E[1];
            lv.kind = BV_LIGHT;
            float le[3];
            v_eval(b, &lv, mk(0, 0, 0), m->wi, 0, le);
            for (int c = 0; c < 3; c++) m->em[c] += m->thr[c] * le[c];
        }
        m->rx += 1.0f;                                   /* passes with a point */
    }
    return NULL;
}

void orc_sppm_pixels(const orc_scene *s, const orc_ppm_params *pp, const float *table, uint32_t table_width,
                     const float *ptable, uint32_t ptable_width, float *out, int nthreads)
{
    bctx b;
    b.s = s;
    camera_init(&s->d, &b.cam);
    b.tri_light = (int *)malloc(sizeof(int) * (s->d.ntri ? s->d.ntri : 1));
    for (uint32_t i = 0; i < s->d.ntri; i++) b.tri_light[i] = -1;
    for (uint32_t l = 0; l < s->d.nlight; l++)
        for (uint32_t k = 0; k < s->d.light_prims[2 * l + 1]; k++) b.tri_light[s->d.light_prims[2 * l] + k] = (int)l;
    b.tri_cdf_base = NULL;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    const uint64_t n = (uint64_t)s->d.width * (uint64_t)s->d.height;
    ppm_point *mp = (ppm_point *)calloc(n + 1, sizeof(ppm_point));
    for (uint64_t i = 0; i < n; i++) mp[i].r = pp->initial_radius;
    usrc pu;
    pu.table = ptable; pu.width = ptable_width;
    pu.seed = ((uint64_t)pp->pseed_hi << 32) | pp->pseed_lo;
    pu.first = 0;
    pthread_t th[256];
    for (uint32_t pass = 0; pass < pp->num_passes; pass++) {
        sppm_ejob ej[256];
        for (int t = 0; t < nthreads; t++) {
            ej[t].b = &b; ej[t].pp = pp; ej[t].mp = mp; ej[t].pass = pass;
            ej[t].us.table = table; ej[t].us.width = table_width;
            ej[t].us.seed = ((uint64_t)pp->seed_hi << 32) | pp->seed_lo;
            ej[t].us.first = 0;
            ej[t].lo = n * (uint64_t)t / (uint64_t)nthreads;
            ej[t].hi = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
            pthread_create(&th[t], NULL, sppm_points, &ej[t]);
        }
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
        ppm_pass(&b, pp, &pu, pass, mp, n, 1, nthreads);
    }
    const double total = (double)pp->num_passes * (double)pp->num_photons;
    for (uint64_t i = 0; i < n; i++) {
        float *o = &out[16 * i];
        memset(o, 0, sizeof(float) * 16);
        const ppm_point *m = &mp[i];
        for (int c = 0; c < 3; c++) {
            o[c] = m->tau[c] / (m->r * m->r * PI_F * (float)total) + m->em[c] / (float)pp->num_passes;
            o[3 + c] = m->em[c];
            o[6 + c] = m->tau[c];
        }
        o[9] = m->N; o[10] = m->r; o[11] = m->rx;
    }
    free(mp); free(b.tri_light);
}

/* ---------------------------------------------------------------------------
 * renderer::ltdirect (renderer_ltdirect.cpp:72-250), restated literally: per
 * sample a light subpath whose vertex j >= 1 is kept when the roulette after
 * reaching it passes (its throughput doubled), and every kept vertex k
 * (endpoint included, while k + 1 < max_num_vertices) connects to the
 * pinhole: C = throughput x f_L(towards the camera) x G x V x We, splatted
 * where the direction projects.  Uniform layout of lmx_render_pm's photons
 * (slot 6: the roulette before vertex 1; vertex j owns ORC_PT_SEG0 +
 * 8 (j - 1) .. +7: 0-1 its direction sampled at vertex j - 1, 3 the roulette
 * before vertex j + 1).  Output per sample: max_num_vertices records of 8
 * floats, record k: raster x, y, C rgb (before the W H / N film scale), 1
 * if it contributes, 0, 0. */
static void lt_sample(const bctx *b, const orc_lt_params *pp, const usrc *us, uint64_t i, float *out, int M)
{
    memset(out, 0, sizeof(float) * 8 * (size_t)M);
    bvert vs[PM_CAP];
    const int maxv = pp->max_num_vertices;
    const int nv = pm_trace_photon(b, maxv, us, i, vs);
    for (int k = 0; k < nv; k++) {
        if (maxv != -1 && k + 1 >= maxv) break;          /* :100-103 */
        const bvert *v = &vs[k];
        bvert cam;
        memset(&cam, 0, sizeof cam);
        cam.kind = BV_CAMERA; cam.degenerate = 1; cam.light = -1;
        cam.fr.p = b->cam.pos;
        const v3 ppE = dir_to(v, &cam);
        const v3 wi = k >= 1 ? dir_to(v, &vs[k - 1]) : mk(0, 0, 0);
        float fsL[3];
        v_eval(b, v, wi, ppE, 1, fsL);
        const float fsE = importance(&b->cam, vneg(ppE));
        const float G = gterm(v, &cam);
        if (fsE == 0.0f || G == 0.0f || (fsL[0] == 0.0f && fsL[1] == 0.0f && fsL[2] == 0.0f)) continue;
        /* Scene3::Visible from the vertex to the camera (scene3.h:107-116) */
        v3 d = vsub(cam.fr.p, v->fr.p);
        const float l = sqrtf(vdot(d, d));
        d = mk(d.x / l, d.y / l, d.z / l);
        orc_ray r;
        r.o[0] = v->fr.p.x; r.o[1] = v->fr.p.y; r.o[2] = v->fr.p.z; r.tmin = EPS_ISECT;
        r.d[0] = d.x; r.d[1] = d.y; r.d[2] = d.z; r.tmax = l * (1.0f - EPS_ISECT);
        orc_hit h;
        orc_qbvh_intersect_ex(b->s->q, &r, 1, &h, ORC_TEST_WATERTIGHT, ORC_TIE_MIN_ID, 1);
        if (h.id != 0xffffffffu) continue;
        float rx, ry;
        if (!camera_raster(&b->cam, vneg(ppE), &rx, &ry)) continue;
        float *o = &out[8 * k];
        o[0] = rx; o[1] = ry;
        for (int c = 0; c < 3; c++) o[2 + c] = v->alpha[c] * fsL[c] * G * fsE;
        o[5] = 1.0f;
    }
}

typedef struct { const bctx *b; const orc_lt_params *pp; usrc us; float *out; int M; uint64_t first, lo, hi; } lt_job;

static void *lt_run(void *p)
{
    lt_job *j = (lt_job *)p;
    for (uint64_t i = j->lo; i < j->hi; i++) lt_sample(j->b, j->pp, &j->us, j->first + i, &j->out[(size_t)8 * j->M * i], j->M);
    return NULL;
}

void orc_ltdirect_samples(const orc_scene *s, const orc_lt_params *pp, const float *table, uint32_t table_width,
                          uint64_t first, uint64_t n, float *out, int nthreads)
{
    bctx b;
    b.s = s;
    camera_init(&s->d, &b.cam);
    b.tri_light = (int *)malloc(sizeof(int) * (s->d.ntri ? s->d.ntri : 1));
    for (uint32_t i = 0; i < s->d.ntri; i++) b.tri_light[i] = -1;
    for (uint32_t l = 0; l < s->d.nlight; l++)
        for (uint32_t k = 0; k < s->d.light_prims[2 * l + 1]; k++) b.tri_light[s->d.light_prims[2 * l] + k] = (int)l;
    b.tri_cdf_base = NULL;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    const int M = pp->max_num_vertices == -1 ? PM_CAP : pp->max_num_vertices;
    pthread_t th[256];
    lt_job jobs[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t].b = &b; jobs[t].pp = pp; jobs[t].out = out; jobs[t].M = M; jobs[t].first = first;
        jobs[t].us.table = table; jobs[t].us.width = table_width;
        jobs[t].us.seed = ((uint64_t)pp->seed_hi << 32) | pp->seed_lo;
        jobs[t].us.first = first;
        jobs[t].lo = n * (uint64_t)t / (uint64_t)nthreads;
        jobs[t].hi = n * (uint64_t)(t + 1) / (uint64_t)nthreads;
        pthread_create(&th[t], NULL, lt_run, &jobs[t]);
    }
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(b.tri_light);
}

/* renderer::raycast (renderer_raycast.cpp:59-111): one ray through each
 * pixel centre; a hit shows |sn . d| times the surface's reflectance R
 * (BSDF::Reflectance2 of diffuse and cook-torrance; 1 for a surface without
 * a BSDF), a miss black.  out: W H x 3 floats, pixel y W + x. */
void orc_raycast(const orc_scene *s, float *out)
{
    camera cam;
    camera_init(&s->d, &cam);
    const int w = (int)s->d.width, h = (int)s->d.height;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float *o = &out[3 * ((size_t)y * w + x)];
            o[0] = o[1] = o[2] = 0.0f;
            const v3 d = camera_dir(&cam, ((float)x + 0.5f) / (float)w, ((float)y + 0.5f) / (float)h);
            orc_hit hh;
            if (!scene_intersect(s, cam.pos, d, &hh)) continue;
            frame fr;
            surface_at(s, cam.pos, d, &hh, &fr);
            const orc_material *m = &s->d.mats[s->d.tri_mat[hh.id]];
            const float c = fabsf(vdot(fr.sn, vneg(d)));
            const int refl = m->type == ORC_MAT_DIFFUSE || m->type == ORC_MAT_PHONG || m->type == ORC_MAT_GGX;
            for (int k = 0; k < 3; k++) o[k] = c * (refl ? m->R[k] : 1.0f);
        }
}
