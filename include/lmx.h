/*
 * lmx.h -- C ABI of the MI355X-native Lightmetrica ray-tracing hot path.
 *
 * Plain pointers and sizes only.  Every call returns 0 on success or a
 * negative LMX_E* code; lmx_last_error() then describes the failure (the
 * string belongs to the calling thread and stays valid until its next call).
 *
 * Device pointers are HIP device addresses on the current device; `stream`
 * is a hipStream_t (NULL = the default stream).  Calls on one lmx_accel must
 * not overlap in time on different streams: the traversal's work counters
 * live in the accel.
 */
#ifndef LMX_H
#define LMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LMX_OK 0
#define LMX_EINVAL (-1)     /* bad argument */
#define LMX_ENOMEM (-2)     /* host or device allocation failed */
#define LMX_EHIP (-3)       /* HIP runtime error (no device, launch failure) */
#define LMX_EBUILD (-4)     /* BVH construction failed */
#define LMX_EIO (-5)        /* file could not be read or written */

/* Triangle test used by traversal. */
#define LMX_ISECT_TRIACCEL 0     /* Wald projection test, reference semantics */
#define LMX_ISECT_WATERTIGHT 1   /* watertight test (default) */

typedef struct lmx_accel lmx_accel;

/* Ray: origin, tmin, direction, tmax (32 B).  Scene3::Intersect uses
 * tmin = 1e-4 and tmax = FLT_MAX (src/liblightmetrica/scene3.cpp:461). */
typedef struct lmx_ray {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} lmx_ray;

/* Closest hit: t, b0 = weight of the triangle's 2nd vertex, b1 = weight of
 * its 3rd vertex (the reference's barycentric convention,
 * include/lightmetrica/intersectionutils.h:96,108), id = global triangle index
 * in enumeration order, LMX_MISS when nothing was hit (t = b0 = b1 = 0). */
typedef struct lmx_hit {
    float t, b0, b1;
    uint32_t id;
} lmx_hit;

#define LMX_MISS 0xffffffffu

typedef struct lmx_build_opts {
    int32_t isect;       /* LMX_ISECT_* */
    int32_t nthreads;    /* host build threads; <= 0 picks the hardware count */
} lmx_build_opts;

typedef struct lmx_accel_info {
    uint32_t num_nodes;      /* 80-byte wide nodes */
    uint32_t num_tris;
    uint32_t max_depth;      /* wide levels, root = 1 */
    int32_t isect;
    uint64_t node_bytes;
    uint64_t tri_bytes;
    uint32_t stack_overflow; /* nonzero once a GPU traversal of this accel ran out of stack
                                (never for a tree lmx_accel_build accepts); synchronises */
    uint32_t reserved;
} lmx_accel_info;

/* Builds the compressed 8-wide BVH on the host over world-space triangles
 * (verts: nvert x 3 floats; tris: ntri x 3 vertex indices).  Replaces
 * Accel::Build(const Scene*) (include/lightmetrica/accel.h:79,
 * src/liblightmetrica/accel/accel_qbvh.cpp:152-389); the caller enumerates
 * triangles in primitive order x face order with the primitive transform
 * already applied (accel_qbvh.cpp:164-189).  Needs no GPU; device memory is
 * allocated on the first lmx_accel_intersect. */
int lmx_accel_build(const float *verts, uint32_t nvert, const uint32_t *tris, uint32_t ntri,
                    const lmx_build_opts *opts, lmx_accel **out);

void lmx_accel_free(lmx_accel *accel);

int lmx_accel_info_get(const lmx_accel *accel, lmx_accel_info *info);

/* Uploads the BVH to the current device (idempotent). */
int lmx_accel_upload(lmx_accel *accel);

/* Batched closest hit on the GPU: rays and hits are device arrays of n
 * records.  Replaces the per-ray Accel3::Intersect(scene, ray, isect, minT,
 * maxT) (include/lightmetrica/accel3.h:68) for batches; each ray carries its
 * own [tmin, tmax].  Enqueued on `stream`, asynchronous. */
int lmx_accel_intersect(lmx_accel *accel, const lmx_ray *rays, uint32_t n, lmx_hit *hits,
                        void *stream);

/* Batched occlusion (the any-hit form behind Scene3::Visible,
 * include/lightmetrica/scene3.h:107-116): hits[i].id is LMX_MISS when ray i
 * is unoccluded on [tmin, tmax], otherwise the id of some triangle it hits
 * there (not necessarily the closest; t, b0, b1 belong to that triangle). */
int lmx_accel_occluded(lmx_accel *accel, const lmx_ray *rays, uint32_t n, lmx_hit *hits, void *stream);

/* lmx_accel_intersect without any device allocation, for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  No counterpart in the reference.  Same arguments, same
 * hits, same kernel.  The accel must already be resident on the current
 * device (lmx_accel_upload), else LMX_EINVAL.  The launch's scratch (region
 * cursors and ray count, 2,304 bytes for any n; n up to 2^31 - 1, else
 * LMX_EINVAL) is one static device array per GPU, held until the process
 * ends and shared by every call on that GPU: calls that overlap on one GPU
 * (different streams, or different threads) race and must be serialised by
 * the caller.  Enqueued on `stream`, asynchronous. */
int lmx_accel_intersect_noalloc(lmx_accel *accel, const lmx_ray *rays, uint32_t n, lmx_hit *hits,
                                void *stream);

/* The same query on the host (the scalar Accel3::Intersect path), walking
 * the same wide BVH in the same order as the kernel.  Host arrays. */
int lmx_accel_intersect_host(const lmx_accel *accel, const lmx_ray *rays, uint32_t n,
                             lmx_hit *hits, int32_t nthreads);

/* Counting walker: total wide-node visits and triangle tests the traversal
 * kernel performs for these rays (host arrays).  Feeds the kernel's
 * algorithmic byte count. */
int lmx_accel_count(const lmx_accel *accel, const lmx_ray *rays, uint32_t n,
                    uint64_t *node_visits, uint64_t *tri_tests, int32_t nthreads);
/* The same counts for the occlusion query (lmx_accel_occluded): the walk
 * stops after the node in which it found a triangle in range. */
int lmx_accel_count_occluded(const lmx_accel *accel, const lmx_ray *rays, uint32_t n,
                             uint64_t *node_visits, uint64_t *tri_tests, int32_t nthreads);

/* ------------------------------------------------------------------------ */
/* Path samplers: scene and renderers (Renderer::Render, renderer.h:81)       */

#define LMX_MAT_NULL 0       /* bsdf::null: no scattering (light-only surfaces) */
#define LMX_MAT_DIFFUSE 1    /* bsdf::diffuse (bsdf_diffuse.cpp) */
#define LMX_MAT_PHONG 2      /* normalized Phong (no counterpart in the reference) */
#define LMX_MAT_GGX 3        /* bsdf::cooktorrance (bsdf_cooktorrance.cpp, GGX distribution, conductor Fresnel) */

#define LMX_RASTER_RANDOM 0      /* raster position from the sample's uniforms, as the reference */
#define LMX_RASTER_STRATIFIED 1  /* sample i -> pixel i mod (W H), jittered inside it */

typedef struct lmx_material {
    int32_t type;        /* LMX_MAT_* */
    float R[3];          /* reflectance */
    float exponent;      /* Phong exponent */
    float Le[3];         /* emitted radiance; > 0 makes every triangle of the material an area light */
    float roughness;     /* GGX roughness (bsdf_cooktorrance.cpp: default 0.1) */
    float eta[3], k[3];  /* conductor index (default 0.14, 0.129, 0.1585; k 4.58625, 3.348125, 2.329375) */
    float reserved;
} lmx_material;

/* Pinhole sensor (sensor_pinhole.cpp:44-202): position, camera basis (the
 * sensor looks along -vz), vertical field of view in degrees, film size. */
typedef struct lmx_camera {
    float pos[3], vx[3], vy[3], vz[3];
    float fov_deg;
    int32_t width, height;
} lmx_camera;

/* World-space scene (host arrays, copied).  Triangles in enumeration order;
 * light_prims lists each light primitive as (first triangle, count). */
typedef struct lmx_scene_desc {
    const float *verts; uint32_t nvert;
    const uint32_t *tris; uint32_t ntri;
    const float *normals;            /* per vertex, or NULL (geometric normals) */
    const uint32_t *tri_mat;         /* material index per triangle */
    const lmx_material *mats; uint32_t nmat;
    const uint32_t *light_prims; uint32_t nlight;
    lmx_camera camera;
} lmx_scene_desc;

typedef struct lmx_scene lmx_scene;

/* Per-sample uniform slots (shared by the generator and uniform tables):
 * slot 0 selects the sensor, 1-2 place the raster position, 3-7 unused;
 * path segment j (j = 0 is the camera ray) owns slots 8+4j .. 8+4j+3:
 * direction 2-D sample, component sample, Russian roulette after the
 * segment's hit.  Table width >= 8 + 4 (max_num_vertices - 1). */
#define LMX_PT_SLOT_RASTER 1
#define LMX_PT_SLOT_SEG0 8
/* Russian roulette of the unidirectional samplers: continue with rr_prob
 * (renderer_pt.cpp:207-215), or -- an extension without a reference
 * counterpart -- with min(1, largest throughput channel), which keeps the
 * throughput of surviving paths from growing */
#define LMX_RR_FIXED 0
#define LMX_RR_THROUGHPUT 1
/* Multiple-importance weights of renderer::bdpt (misweight::powerheuristics,
 * misweight::simple; renderer_bdpt.cpp:704-787) */
#define LMX_MIS_POWER 0
#define LMX_MIS_SIMPLE 1

typedef struct lmx_render_params {
    int32_t max_num_vertices;    /* -1: unbounded (renderer_pt.cpp:59) */
    int32_t min_num_vertices;    /* (renderer_pt.cpp:60) */
    int32_t rr_depth;            /* roulette after the hit of segment j when j + 1 >= rr_depth (1 = reference) */
    float rr_prob;               /* continuation probability (0.5 in the reference) */
    uint64_t seed;
    uint64_t num_samples;        /* total camera samples = spp x W x H */
    int32_t raster_mode;         /* LMX_RASTER_* */
    uint32_t queue_capacity;     /* paths in flight; 0 = autotuned */
    const float *uniform_table;  /* device, num_samples x table_width, or NULL (Philox) */
    uint32_t table_width;
    float *sample_out;           /* device, num_samples x 8 floats (raster x, y, C rgb, #splats, 0, 0), or NULL */
    int32_t rr_mode;             /* LMX_RR_FIXED (reference) or LMX_RR_THROUGHPUT */
    int32_t mis;                 /* renderer::bdpt: LMX_MIS_POWER (reference default) or LMX_MIS_SIMPLE */
    /* Time mode of the reference's scheduler (render_time, scheduler.cpp:57,
     * 108,191-199): > 0 renders for this many seconds instead of num_samples
     * -- new camera samples (the unidirectional samplers) or batches (the
     * bidirectional ones) start until the deadline, every started sample
     * completes, and the film is scaled by W H / (samples taken), reported
     * in lmx_render_stats.samples.  The clock starts after the one-time
     * autotuning of the scene.  <= 0: sample-count mode.  Not with
     * uniform_table or sample_out.  lmx_render_pt, _ptdirect, _ptmis,
     * _vcm_bdpt and _bdpt. */
    double render_time;
    /* lmx_render_pt / _ptdirect / _ptmis: LMX_SHADE_QUEUE shades the traced
     * paths in queue order; LMX_SHADE_BY_MATERIAL first bins them by the
     * BSDF of the surface they hit (one pass over the hits), so each wave of
     * the shading kernel runs one material's code.  Results are the same
     * either way up to the order of the film's float additions. */
    int32_t shade_order;
    /* Measurement hook: when ray_out is a device array of ray_out_cap x 8
     * floats, the first ray_out_cap extension rays traced (lmx_ray records,
     * in launch order) are copied there (lmx_render_pt / _ptdirect / _ptmis),
     * or the first ray_out_cap / 2 subpath extension rays to the first half
     * and the first ray_out_cap / 2 shadow rays to the second
     * (lmx_render_vcm_bdpt / _bdpt), so that the counting walkers
     * (lmx_accel_count, lmx_accel_count_occluded) can price the render's
     * traversal in algorithmic bytes.  NULL: off. */
    uint32_t ray_out_cap;
    float *ray_out;
} lmx_render_params;

#define LMX_SHADE_QUEUE 0
#define LMX_SHADE_BY_MATERIAL 1

typedef struct lmx_render_stats {
    uint64_t rays;               /* extension rays traced */
    uint32_t iterations;         /* wavefront iterations */
    uint32_t queue_capacity;     /* capacity used */
    double ms;                   /* wall time of the render call */
    uint64_t samples;            /* camera samples taken (num_samples unless render_time > 0) */
    uint64_t dropped;            /* lmx_render_vcm: merges dropped because their contribution was not
                                    finite in FP32 (reported so that a test can require none) */
} lmx_render_stats;

/* Builds the accel (watertight, 8-wide) and uploads shading data to the
 * current device. */
int lmx_scene_create(const lmx_scene_desc *desc, lmx_scene **out);
void lmx_scene_free(lmx_scene *scene);

/* renderer::pt (renderer_pt.cpp:64-242) as a wavefront pipeline: film_rgb is
 * a device array of W x H x 3 floats (row 0 = bottom), overwritten with the
 * image (sum of splats x W H / num_samples, Film::Rescale,
 * scheduler.cpp:281-288).  Synchronous on `stream`; stats may be NULL. */
int lmx_render_pt(lmx_scene *scene, const lmx_render_params *params, float *film_rgb, void *stream,
                  lmx_render_stats *stats);

/* renderer::ptdirect (renderer_ptdirect.cpp:72-297) and renderer::ptmis
 * (renderer_ptmis.cpp:69-326): renderer::pt plus next-event estimation at
 * every vertex, the camera's included (a light by uniform selection, a point
 * by area, one shadow ray); ptmis weighs it and the BSDF-sampled light hits
 * by the balance heuristic, ptdirect counts light sampling only.  Same
 * parameters as lmx_render_pt (ptdirect ignores min_num_vertices as the
 * reference does); uniform blocks of 8 slots per vertex v at 8 + 8 v:
 * direction 0-1, unused 2, roulette after the hit of segment v 3, light
 * selection 4, light position 5-6.  sample_out records have 16 floats: the
 * 8 of lmx_render_pt, then raster x, y, C rgb, count of the camera-vertex
 * connection (splatted where the light point projects). */
int lmx_render_ptdirect(lmx_scene *scene, const lmx_render_params *params, float *film_rgb, void *stream,
                        lmx_render_stats *stats);
int lmx_render_ptmis(lmx_scene *scene, const lmx_render_params *params, float *film_rgb, void *stream,
                     lmx_render_stats *stats);

/* renderer::vcm with mode: bdpt (renderer_vcm.cpp:662-713): an eye and a
 * light subpath of at most max_num_vertices vertices per sample, every
 * (s, t) connection with s + t <= max_num_vertices, power-heuristic MIS over
 * all strategies.  Uses max_num_vertices (2..32), min_num_vertices, seed,
 * num_samples, raster_mode; queue_capacity is the batch size in samples
 * (0 = autotuned per scene and max_num_vertices: the fastest of 2^19..2^24
 * that fits half of free device memory).  Uniform table width >= 2 (8 + 4 (max_num_vertices - 1)): an
 * eye block then a light block, each laid out as for lmx_render_pt except
 * 1-2 emitter position, 3 emitter selection, 4-5 emitter direction.
 * sample_out, when given, holds (max_num_vertices + 1)^2 records of 8 floats
 * per sample at index t (max_num_vertices + 1) + s: raster x, y, C rgb
 * (weighted, before the W H / N film scale), 1, MIS weight, 0.  The light
 * position and emission direction use independent numbers (the reference
 * draws both from one, light_area.cpp:72-80, which biases s >= 2). */
int lmx_render_vcm_bdpt(lmx_scene *scene, const lmx_render_params *params, float *film_rgb, void *stream,
                        lmx_render_stats *stats);

/* renderer::bdpt (renderer_bdpt.cpp:814-1000): an eye and a light subpath per
 * sample with Russian roulette 0.5 after every vertex but the endpoint, each
 * subpath vertex also sampling a "direct" vertex on an emitter of the other
 * kind; every strategy (s, t) plus the direct ones (s = 0: a light point
 * sampled from eye vertex t - 2) of every length n, weighed by `mis` over
 * all of them and divided by the roulette's selection probability.  The
 * light subpath's direct vertices (points on a pinhole sensor) carry nothing
 * and are not sampled.  max_num_vertices: -1 (unbounded, the reference's
 * default; subpaths are cut at 32 vertices, reached with probability 2^-30)
 * or 2..32; min_num_vertices applies only when max_num_vertices is given
 * (renderer_bdpt.cpp:887).  queue_capacity as for lmx_render_vcm_bdpt.
 * Uniform table width >= 2 (8 + 8 (M - 1)), M = max_num_vertices or 32: an
 * eye block then a light block; within a block 1-2 endpoint position,
 * 3 emitter selection, 4-5 endpoint direction, and step j >= 1 (subpath
 * vertex j) owns 8 + 8 (j - 1) .. +7: 0-1 direction sampled at vertex j - 1
 * (j >= 2), 3 roulette after vertex j, 4 light selection and 5-6 light
 * position of the direct vertex sampled from eye vertex j - 1.  sample_out,
 * when given, holds 2 (M + 1)^2 records of 8 floats per sample at index
 * 2 (t (M + 1) + s) + direct, laid out as lmx_render_vcm_bdpt's. */
int lmx_render_bdpt(lmx_scene *scene, const lmx_render_params *params, float *film_rgb, void *stream,
                    lmx_render_stats *stats);

/* renderer::vcm in modes vcm and bdpm (renderer_vcm.cpp:584-790, parameters
 * :558-577): num_iteration_pass passes; pass q traces
 * num_photon_trace_samples photon (light) subpaths, gathers their vertices
 * past the light into a hash grid, then params->num_samples eye samples
 * (num_eye_trace_samples), each with (mode vcm) every connection of
 * lmx_render_vcm_bdpt and, at every eye vertex past the pinhole, a merge with
 * every photon vertex within radius r_q (r_0 = initial_radius,
 * r_q = sqrt((alpha + q) / (1 + q)) r_{q-1}); power-heuristic MIS over
 * connections and merges (merges counted num_photon_trace_samples times),
 * merges alone in mode bdpm.  film_rgb is the mean of the passes' images.
 * Eye sample i of pass q uses uniform index q num_samples + i of `seed`,
 * photon i index q num_photon_trace_samples + i of photon_seed, each laid
 * out as lmx_render_vcm_bdpt's sample (the photon uses the light block).
 * With uniform tables, photon_table (num_photon_trace_samples rows of
 * photon_table_width >= 2 (8 + 4 (max_num_vertices - 1)) floats) or
 * sample_out, num_iteration_pass must be 1.  sample_out holds
 * 2 (M + 1)^2 records of 8 floats per eye sample at
 * 2 (t (M + 1) + s) + merge; a merge of eye vertex t - 1 with photon
 * vertex j is recorded at s = j + 1 (its path is that of the connection
 * (j, t)), summed over the photons: raster x, y, C rgb, count, sum of
 * weights, 0.  The render zeroes sample_out first. */
#define LMX_VCM_MODE_VCM 0
#define LMX_VCM_MODE_BDPM 2
typedef struct lmx_vcm_params {
    int32_t mode;                        /* LMX_VCM_MODE_VCM (reference default) or LMX_VCM_MODE_BDPM */
    uint32_t num_iteration_pass;         /* reference default 100 */
    uint32_t num_photon_trace_samples;   /* 1 .. 2^26 (reference default 10000) */
    float initial_radius;                /* > 0 (reference default 0.1) */
    float alpha;                         /* radius reduction, > 0 (reference default 0.7) */
    uint64_t photon_seed;
    const float *photon_table;           /* device, or NULL (Philox with photon_seed) */
    uint32_t photon_table_width;
} lmx_vcm_params;
int lmx_render_vcm(lmx_scene *scene, const lmx_render_params *params, const lmx_vcm_params *vcm, float *film_rgb,
                   void *stream, lmx_render_stats *stats);

/* renderer::pm (renderer_pm.cpp:61-220) with photonmap::kdtree / naive
 * (photonmap_kdtree.cpp:123-165: every photon within `radius`):
 * num_photon_trace_samples photon (light) subpaths with roulette 0.5 after
 * every vertex, the endpoint included, store a photon (position, throughput,
 * direction to the previous vertex, vertex count) at every diffuse or glossy
 * vertex; then params->num_samples eye samples, traced without roulette:
 * every emitter vertex adds throughput x Le, and the first (finalgather 0)
 * or second (finalgather 1, the reference default) diffuse or glossy vertex
 * adds the density estimate -- kernel 3/pi (1 - d^2/r^2)^2 / (r^2 N) x BSDF
 * towards the photon's direction x photon throughput over the photons whose
 * merged path has at most max_num_vertices vertices -- and ends the path.
 * max_num_vertices: -1 (unbounded, the reference default; subpaths cut at 32
 * vertices, every photon gathered) or 2..32.  Film scaled by W H / N.
 * Uniforms: eye sample i index i of `seed` in renderer::vcm's eye layout
 * (table width >= 8 + 4 (min(M, 2 + finalgather) - 1)); photon i index i of
 * photon_seed: 1-2 emitter position, 3 emitter selection, 4-5 direction,
 * 6 roulette after the endpoint, vertex j >= 1 owns 8 + 8 (j - 1) .. +7: 0-1
 * the direction sampled at vertex j - 1, 3 the roulette after vertex j (table
 * width >= 8 + 8 (M - 1), M = max_num_vertices or 32).  sample_out: 16 floats
 * per eye sample: raster x, y, emission C rgb, emitter vertices, density C
 * rgb, photons gathered, vertex count at the gather vertex (0: none), 0 x 5
 * (before the film scale). */
typedef struct lmx_pm_params {
    uint32_t num_photon_trace_samples;   /* 1 .. 2^26 (reference default 100000) */
    int32_t finalgather;                 /* 0 or 1 (reference default 1) */
    float radius;                        /* > 0 (reference default 0.01) */
    uint64_t photon_seed;
    const float *photon_table;           /* device, or NULL (Philox with photon_seed) */
    uint32_t photon_table_width;
} lmx_pm_params;
int lmx_render_pm(lmx_scene *scene, const lmx_render_params *params, const lmx_pm_params *pm, float *film_rgb,
                  void *stream, lmx_render_stats *stats);

/* renderer::ppm (renderer_ppm.cpp:84-320): params->num_samples (1 .. 2^24)
 * measurement points, each the first vertex of an eye sample when it is
 * diffuse or glossy (other surfaces end the path, so an emitter without a
 * BSDF seen directly stays black, as in the reference); then
 * num_iteration_pass passes of num_photon_trace_samples photon subpaths with
 * roulette 0.5 after every vertex but the endpoint, each point gathering the
 * pass's photons within its radius and updating N, tau and the radius
 * progressively (alpha); the film is throughput x tau / (pi r^2 x all
 * photons) + the point's emission, scaled by W H / num_samples.
 * max_num_vertices -1 (unbounded: subpaths cut at 32 vertices, every photon
 * gathered) or 2..32.  Uniforms: point i index i of `seed` (eye endpoint
 * slots 4-5; table width >= 8); photon i of pass q index q N + i of
 * photon_seed in lmx_render_pm's photon layout without slot 6 (table rows
 * num_iteration_pass x N, width >= 8 + 8 (M - 1)).  sample_out: 16 floats per
 * point: raster x, y, C rgb (before the film scale), emission rgb, tau rgb,
 * N, radius, 1 if the sample made a point, 0 x 2. */
typedef struct lmx_ppm_params {
    uint32_t num_iteration_pass;         /* reference default 1000 */
    uint32_t num_photon_trace_samples;   /* per pass, 1 .. 2^26 (reference default 100) */
    float initial_radius;                /* > 0 (reference default 0.1) */
    float alpha;                         /* (0, 1] (reference default 0.7) */
    uint64_t photon_seed;
    const float *photon_table;           /* device, or NULL (Philox with photon_seed) */
    uint32_t photon_table_width;
} lmx_ppm_params;
int lmx_render_ppm(lmx_scene *scene, const lmx_render_params *params, const lmx_ppm_params *ppm, float *film_rgb,
                   void *stream, lmx_render_stats *stats);

/* renderer::sppm (renderer_sppm.cpp:98-378): one measurement point per
 * pixel with its own radius, N and tau; each of num_iteration_pass passes
 * re-traces every pixel's eye sample (raster jittered inside the pixel) to
 * its first diffuse or glossy vertex -- the point's position for that pass,
 * whose emission adds to the pixel's sum -- then traces
 * num_photon_trace_samples photon subpaths as lmx_render_ppm does and
 * updates every point that has a position (tau gains throughput x dtau).
 * film_rgb is written, not scaled: per pixel tau / (pi r^2 x all photons) +
 * emission sum / passes.  params: max_num_vertices (-1 unbounded or 2..32),
 * seed, uniform_table (rows passes x W x H: pass q's pixel i at row
 * q W H + i, width >= 8), sample_out (16 floats per pixel: film C rgb,
 * emission sum rgb, tau rgb, N, radius, passes with a point, 0 x 4);
 * num_samples, raster_mode and queue_capacity are not used.  The photon
 * parameters, tables and rows are lmx_render_ppm's. */
int lmx_render_sppm(lmx_scene *scene, const lmx_render_params *params, const lmx_ppm_params *ppm, float *film_rgb,
                    void *stream, lmx_render_stats *stats);

/* renderer::ltdirect (renderer_ltdirect.cpp:72-250): num_samples light
 * subpaths (a vertex j >= 1 is kept when the roulette 0.5 after reaching it
 * passes, its throughput doubled), every kept vertex k with
 * k + 1 < max_num_vertices (-1: unbounded, subpaths cut at 32 vertices)
 * connected to the pinhole: throughput x f(towards the camera) x G x We,
 * splatted where it projects when the shadow ray is clear; film scaled by
 * W H / num_samples.  Uniforms: lmx_render_pm's photon layout (table width
 * >= 8 + 8 (M - 1), M = max_num_vertices or 32).  queue_capacity: light
 * subpaths per batch (0: 2^21).  sample_out: M records of 8 floats per
 * sample, record k: raster x, y, C rgb (before the film scale), 1, 0, 0.
 * renderer::lt (renderer_lt.cpp) only gathers light by hitting the sensor's
 * geometry, which a pinhole has none of: its image is black. */
int lmx_render_ltdirect(lmx_scene *scene, const lmx_render_params *params, float *film_rgb, void *stream,
                        lmx_render_stats *stats);

/* renderer::raycast (renderer_raycast.cpp:59-111): one ray through each
 * pixel centre; film_rgb (W x H x 3, row 0 = bottom) gets |sn . d| times the
 * hit surface's reflectance R (1 for a surface without a BSDF), black on a
 * miss.  Synchronous on `stream`. */
int lmx_render_raycast(lmx_scene *scene, float *film_rgb, void *stream, lmx_render_stats *stats);

/* ---------------------------------------------------------------------------
 * Learned directional sampler.  No counterpart in the reference: a third
 * renderer beside renderer::pt and renderer::vcm, for the callers of
 * lmx_render_pt that want path guiding.  At every diffuse or glossy vertex an
 * MLP (7 inputs: position in the scene bounds, shading normal and direction
 * to the previous vertex as theta/pi, phi/2pi; 12-frequency sin/cos
 * encoding; 4 hidden layers of 64 ReLU units; BF16 weights and activations,
 * FP32 accumulation) gives 8 von Mises-Fisher lobes and a BSDF selection
 * probability alpha in [0.5, 0.9]; the next direction is drawn from
 * alpha BSDF + (1 - alpha) lobes (one-sample MIS, so the image stays
 * unbiased).  Parameters: LMX_GUIDE_NPARAM floats, layout
 * W1[169][64] (row 168 = bias), W2[64][64], b2[64], W3, b3, W4, b4,
 * W5[64][41], b5[41], with W[k][n] the weight from input k to output n.
 * Outputs y[41]: lobe means y[3k..3k+2] (normalised), log kappa y[24+k]
 * (clamped to [-4, 9]), lobe weight logits y[32+k], alpha = 0.5 + 0.4
 * sigmoid(y[40]).  Trained online with Adam on the loss
 * -w log(alpha p_bsdf + (1 - alpha) p_vmf), w = radiance x BSDF x cosine /
 * sampling density of each recorded direction, normalised to mean one over
 * the records of a pass that carry radiance (minibatches are drawn from
 * those) and clamped to wclamp. */
#define LMX_GUIDE_NPARAM 25961
typedef struct lmx_guide lmx_guide;
typedef struct lmx_guide_opts {
    float lr, beta1, beta2, eps;   /* Adam */
    uint32_t batch;                /* records per training step */
    uint32_t steps_per_pass;       /* Adam steps after each training pass */
    float wclamp;                  /* loss weights (mean one) clamped to this */
    uint64_t seed;                 /* initial weights and minibatch draws */
} lmx_guide_opts;
typedef struct lmx_guide_stats {
    uint32_t records, records_nonzero, steps;
    float loss;                    /* mean loss of the last step */
    double render_ms, train_ms;
} lmx_guide_stats;
void lmx_guide_default_opts(lmx_guide_opts *opts);   /* lr 1e-3, 0.9, 0.999, 1e-8, 2^15, 16, 50, 7 */
int lmx_guide_create(lmx_scene *scene, const lmx_guide_opts *opts, lmx_guide **out);
void lmx_guide_free(lmx_guide *guide);
/* host copy of the FP32 parameters in (set != 0) or out */
int lmx_guide_params(lmx_guide *guide, float *params, int set);
/* renders num_samples guided paths (as lmx_render_pt; uniform slot 2 of each
 * vertex block selects BSDF or lobe); with train != 0 it also records every
 * sampled direction and afterwards runs steps_per_pass Adam steps on them */
int lmx_render_guided(lmx_scene *scene, lmx_guide *guide, const lmx_render_params *params, int train,
                      float *film_rgb, void *stream, lmx_guide_stats *gstats, lmx_render_stats *stats);
/* network outputs y (n x 41) of inputs x (n x 7), device arrays */
int lmx_guide_eval(lmx_guide *guide, const float *x, uint32_t n, float *y, void *stream);
/* gradient (device, LMX_GUIDE_NPARAM; NULL = internal) and mean loss (host,
 * may be NULL) over n device records of 20 floats: x0..x3 | x4 x5 x6 p_bsdf |
 * dir xyz, unused | unused x3, w | unused x4; adam != 0 also applies one step */
int lmx_guide_grad(lmx_guide *guide, const float *records, uint32_t n, float *grad, float *loss, int adam,
                   void *stream);
/* test hook: from outputs y (n x 41) and uniforms u (n x 3: lobe, u1, u2)
 * draw a mixture direction; out n x 8: dir xyz, p_vmf(dir), alpha, lobe */
int lmx_guide_mix_sample(const float *y, const float *u, uint32_t n, float *out, void *stream);

/* ---------------------------------------------------------------------------
 * Film output (Film_HDR::Save, film_hdr.cpp:260-280 and SaveImage :47-164,
 * without FreeImage).  film_rgb is a device film of W x H x 3 floats (row 0 =
 * bottom, as the renderers write it); the file is path + ".hdr" / ".exr" /
 * ".png" (the reference appends the type's suffix) / ".pfm".  radiancehdr:
 * RGBE with run-length encoded scanlines; openexr: uncompressed 32-bit float
 * B, G, R scanlines; png: 8-bit RGB of (int)(v^(1/2.2) 255) clamped to
 * [0, 255]; radiancehdr and openexr clamp negative values to 0, and all three
 * store the top row first.  pfm (no reference counterpart): little-endian
 * "PF", rows bottom to top, values as rendered.  The pixels are encoded on
 * `stream`; the call returns once the file is written. */
#define LMX_IMAGE_RADIANCEHDR 0
#define LMX_IMAGE_OPENEXR 1
#define LMX_IMAGE_PNG 2
#define LMX_IMAGE_PFM 3
int lmx_film_save(const float *film_rgb, int32_t width, int32_t height, int32_t type, const char *path,
                  void *stream);
/* Reads a .hdr (flat or run-length scanlines, "-Y H +X W"), an .exr as
 * lmx_film_save writes it, or a little-endian .pfm into rgb (host,
 * W x H x 3 floats, row 0 = bottom); rgb NULL: only the size.  No reference
 * counterpart: for exchanging golden images. */
int lmx_image_load(const char *path, float *rgb, int32_t *width, int32_t *height);

/* Test hook, no reference counterpart: the FP32 functions the samplers are
 * built with (lightmetrica-v2_amd/csrc/lmx_fmath.h), evaluated on the GPU
 * over n device values -- out[i] = f(x[i]) or, for LMX_FMATH_POW,
 * x[i]^y[i] -- so tests can check them bit for bit against the host build. */
#define LMX_FMATH_SIN 0
#define LMX_FMATH_COS 1
#define LMX_FMATH_TAN 2
#define LMX_FMATH_EXP 3
#define LMX_FMATH_LOG 4
#define LMX_FMATH_POW 5
int lmx_fmath_eval(int32_t kind, const float *x, const float *y, uint32_t n, float *out, void *stream);

/* Copies the host-side BVH out for inspection: num_nodes x 80-byte nodes and
 * num_tris x 48-byte triangle records (layouts in
 * lightmetrica-v2_amd/csrc/wbvh.h).  Either pointer may be NULL. */
int lmx_accel_export(const lmx_accel *accel, void *nodes, void *tris);

const char *lmx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
