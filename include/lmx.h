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
    uint32_t num_nodes;      /* 128-byte wide nodes */
    uint32_t num_tris;
    uint32_t max_depth;      /* wide levels, root = 1 */
    int32_t isect;
    uint64_t node_bytes;
    uint64_t tri_bytes;
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

/* The same query on the host (the scalar Accel3::Intersect path), walking
 * the same wide BVH in the same order as the kernel.  Host arrays. */
int lmx_accel_intersect_host(const lmx_accel *accel, const lmx_ray *rays, uint32_t n,
                             lmx_hit *hits, int32_t nthreads);

/* Counting walker: total wide-node visits and triangle tests the traversal
 * kernel performs for these rays (host arrays).  Feeds the kernel's
 * algorithmic byte count. */
int lmx_accel_count(const lmx_accel *accel, const lmx_ray *rays, uint32_t n,
                    uint64_t *node_visits, uint64_t *tri_tests, int32_t nthreads);

/* Copies the host-side BVH out for inspection: num_nodes x 128-byte nodes and
 * num_tris x 48-byte triangle records (layouts in
 * lightmetrica-v2_amd/csrc/wbvh.h).  Either pointer may be NULL. */
int lmx_accel_export(const lmx_accel *accel, void *nodes, void *tris);

const char *lmx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
