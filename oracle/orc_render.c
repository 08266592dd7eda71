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
 */
#include "orc.h"

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

/* Sampler::CosineSampleHemisphere (sampler.h:44-76) */
void orc_sample_cosine_hemisphere(const float u[2], float out[3]);

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

/* Samples wo; returns 0 when the BSDF gives no direction (the reference's
 * SampleDirection returns without writing wo, bsdf_diffuse.cpp:69-79). */
static int bsdf_sample(const orc_material *m, const frame *f, v3 wi, const float u[2], v3 *wo)
{
    const v3 li = to_local(f, wi);
    if (li.z <= 0.0f) return 0;
    if (m->type == ORC_MAT_DIFFUSE) {
        float l[3];
        orc_sample_cosine_hemisphere(u, l);
        *wo = to_world(f, mk(l[0], l[1], l[2]));
        return 1;
    }
    if (m->type == ORC_MAT_PHONG) {
        /* cos(alpha) = u0^(1/(n+1)), phi = 2 pi u1 about the mirror direction */
        const v3 r = reflect_sn(f, wi);
        v3 ru, rv;
        onb(r, &ru, &rv);
        const float ca = powf(u[0], 1.0f / (m->exponent + 1.0f));
        const float sa = sqrtf(fmaxf(0.0f, 1.0f - ca * ca));
        const float phi = 2.0f * PI_F * u[1];
        const v3 l = mk(sa * cosf(phi), sa * sinf(phi), ca);
        *wo = vadd(vadd(vmul(ru, l.x), vmul(rv, l.y)), vmul(r, l.z));
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
        const float pdf_sa = (m->exponent + 1.0f) * (0.5f * INV_PI_F) * powf(ca, m->exponent);
        return pdf_sa / lo.z;
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
        const float k = (m->exponent + 2.0f) * (0.5f * INV_PI_F) * powf(ca, m->exponent);
        for (int c = 0; c < 3; c++) out[c] = m->R[c] * k * corr;
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
    c->tan_fov = tanf(fov * 0.5f);
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
    const float *table;
    uint32_t width;
    uint64_t seed;
} usrc;

static inline float U(const usrc *s, uint64_t sample, uint32_t slot)
{
    if (s->table) return s->table[sample * s->width + slot];
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
            if (!mat || (mat->type != ORC_MAT_DIFFUSE && mat->type != ORC_MAT_PHONG)) break;
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
            if (U(us, sample, base + 3) > pp->rr_prob) break;
            for (int c = 0; c < 3; c++) thr[c] /= pp->rr_prob;
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
