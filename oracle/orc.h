/*
 * orc.h -- CPU oracle for the Lightmetrica ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against; it is linked or loaded only by tests/,
 * by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  The
 * product library never calls into it.
 *
 * Every function restates the reference algorithm it names (file:line under
 * hi2p-perim/lightmetrica-v2).  Arithmetic is FP32, compiled with
 * -ffp-contract=off so that every product and sum rounds exactly as written.
 *
 * Parity pins (what checks each function against something independent):
 *   orc_triaccel_load / orc_triaccel_intersect
 *       bit-exact against the reference's own triaccel.h compiled into
 *       oracle/_ref (tests/test_oracle_ref.py), and the analytic quads of
 *       src/lightmetrica-test/test_accel3.cpp:272-345 (tests/test_oracle.py).
 *   orc_bound_intersect   bit-exact against the reference's bound.h via _ref.
 *   orc_watertight_intersect
 *       closed-form ray/plane hits, the FP64 arbiter (t and barycentrics of the
 *       exact solution), and the watertightness invariant: a ray through a
 *       shared edge or vertex of a closed mesh is never lost.
 *   orc_closest_hit_naive brute force; qbvh traversal must equal it.
 *   orc_qbvh_*            hit index equal to brute force on meshes of up to
 *                         1e5 triangles (tests/test_oracle.py).
 *   orc_sample_*          bit-exact against the reference's sampler.h via _ref.
 *   orc_dsfmt_*           the reference's dSFMT-19937 stream: equal to the
 *                         golden streams of seeds 1 and 42 written by its
 *                         vendored dSFMT.c (tests/test_oracle.py).
 *   orc_fmath_eval        the shared FP32 math of lmx_fmath.h, which the
 *                         render oracle uses: against double precision
 *                         (ulp and relative bounds, special values) in
 *                         tests/test_fmath.py.
 */
#ifndef LMX_ORC_H
#define LMX_ORC_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Wald projection record; same field order and size (48 B) as
 * include/lightmetrica/triaccel.h:35-48. */
typedef struct orc_triaccel {
    uint32_t k;
    float n_u, n_v, n_d;
    float a_u, a_v, b_nu, b_nv;
    float c_nu, c_nv;
    uint32_t face_index;
    uint32_t prim_index;
} orc_triaccel;

/* Ray record used at the C boundary: origin, tmin, direction, tmax (32 B). */
typedef struct orc_ray {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} orc_ray;

/* Hit record: t, b0 (weight of the 2nd vertex), b1 (weight of the 3rd
 * vertex), global triangle id (0xffffffff = miss).  16 B. */
typedef struct orc_hit {
    float t, b0, b1;
    uint32_t id;
} orc_hit;

/* Tie rule for equal t. */
enum {
    ORC_TIE_LAST_WINS = 0,   /* reference: t > maxT rejects only strictly greater */
    ORC_TIE_MIN_ID = 1       /* canonical rule shared with the HIP kernel */
};

enum {
    ORC_TEST_TRIACCEL = 0,
    ORC_TEST_WATERTIGHT = 1
};

int orc_triaccel_load(const float A[3], const float B[3], const float C[3], orc_triaccel *out);
int orc_triaccel_intersect(const orc_triaccel *tri, const float o[3], const float d[3],
                           float mint, float maxt, float *u, float *v, float *t);
int orc_watertight_intersect(const float v0[3], const float v1[3], const float v2[3],
                             const float o[3], const float d[3], float mint, float maxt,
                             float *b0, float *b1, float *t);
int orc_bound_intersect(const float bmin[3], const float bmax[3], const float o[3],
                        const float d[3], float tmin, float tmax);

/* Brute-force closest hit over all triangles (accel_naive.cpp:92-124). */
void orc_closest_hit_naive(const float *verts, const uint32_t *tris, uint32_t ntri,
                           const orc_ray *rays, uint32_t nray, int test, int tie,
                           orc_hit *hits, int nthreads);

/* 4-wide QBVH (accel_qbvh.cpp) */
typedef struct orc_qbvh orc_qbvh;
orc_qbvh *orc_qbvh_build(const float *verts, const uint32_t *tris, uint32_t ntri);
void orc_qbvh_free(orc_qbvh *q);
uint32_t orc_qbvh_num_nodes(const orc_qbvh *q);
void orc_qbvh_intersect(const orc_qbvh *q, const orc_ray *rays, uint32_t nray,
                        orc_hit *hits, int nthreads);
/* Same traversal with a chosen triangle test and tie rule; with
 * ORC_TIE_MIN_ID the result equals orc_closest_hit_naive's. */
void orc_qbvh_intersect_ex(const orc_qbvh *q, const orc_ray *rays, uint32_t nray,
                           orc_hit *hits, int test, int tie, int nthreads);

/* Binary binned-SAH BVH (accel_bvh_sahbin.cpp) */
typedef struct orc_bvh orc_bvh;
orc_bvh *orc_bvh_build(const float *verts, const uint32_t *tris, uint32_t ntri);
void orc_bvh_free(orc_bvh *b);
uint32_t orc_bvh_num_nodes(const orc_bvh *b);
void orc_bvh_intersect(const orc_bvh *b, const orc_ray *rays, uint32_t nray,
                       orc_hit *hits, int nthreads);

/* Sampling routines of include/lightmetrica/sampler.h */
void orc_sample_concentric_disk(const float u[2], float out[2]);
void orc_sample_cosine_hemisphere(const float u[2], float out[3]);
void orc_sample_uniform_sphere(const float u[2], float out[3]);
void orc_sample_uniform_triangle(const float u[2], float out[2]);

/* Host build of lmx_fmath.h (orc_fmath.c): kind ORC_FM_*, y used by POW */
#define ORC_FM_SIN 0
#define ORC_FM_COS 1
#define ORC_FM_TAN 2
#define ORC_FM_EXP 3
#define ORC_FM_LOG 4
#define ORC_FM_POW 5
void orc_fmath_eval(int kind, const float *x, const float *y, size_t n, float *out);

/* dSFMT-19937, the reference's Random (orc_dsfmt.c) */
typedef struct orc_dsfmt orc_dsfmt;
size_t orc_dsfmt_size(void);
void orc_dsfmt_seed(orc_dsfmt *g, uint32_t seed);
double orc_dsfmt_next(orc_dsfmt *g);
uint32_t orc_dsfmt_uint32(orc_dsfmt *g);
void orc_dsfmt_stream(uint32_t seed, size_t n, double *next, uint32_t *uints);

/* ---------------------------------------------------------------------- */
/* Path samplers (orc_render.c) */

/* Philox4x32-10 (Salmon et al., SC 2011) and the uniform addressed by
 * (seed, sample, slot): block slot/4 of counter (sample lo, sample hi,
 * slot/4, ORC_PHILOX_TAG) under key (seed lo, seed hi), top 24 bits. */
#define ORC_PHILOX_TAG 0x6c6d7870u
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_uniform(uint64_t seed, uint64_t sample, uint32_t slot);

/* Per-sample uniform slots of renderer::pt: slot 0 selects the sensor,
 * 1-2 place the raster position, 3-4 are the sensor's second 2-D sample
 * (unused by a pinhole), 5-7 padding; segment j (j = 0 is the camera ray)
 * owns 8+4j .. 8+4j+3: direction 2-D sample, component sample, Russian
 * roulette after the segment's hit.  Width = 8 + 4 (max_num_vertices - 1). */
#define ORC_PT_RASTER 1u
#define ORC_PT_SEG0 8u

enum { ORC_MAT_NULL = 0, ORC_MAT_DIFFUSE = 1, ORC_MAT_PHONG = 2, ORC_MAT_GGX = 3 };
enum { ORC_RASTER_RANDOM = 0, ORC_RASTER_STRATIFIED = 1 };

typedef struct orc_material {
    int32_t type;        /* ORC_MAT_* */
    float R[3];          /* reflectance */
    float exponent;      /* Phong exponent */
    float Le[3];         /* emitted radiance (area light when > 0) */
    float roughness;     /* GGX (bsdf_cooktorrance.cpp) */
    float eta[3], k[3];
    float reserved;
} orc_material;

typedef struct orc_scene_desc {
    const float *verts; uint32_t nvert;
    const uint32_t *tris; uint32_t ntri;
    const float *normals;            /* per vertex, or NULL */
    const uint32_t *tri_mat;         /* material index per triangle */
    const orc_material *mats; uint32_t nmat;
    const uint32_t *light_prims;     /* nlight x (first triangle, count) */
    uint32_t nlight;
    float cam_pos[3], cam_vx[3], cam_vy[3], cam_vz[3];
    float fov_deg;                   /* vertical, degrees */
    int32_t width, height;
} orc_scene_desc;

typedef struct orc_pt_params {
    int32_t max_num_vertices;        /* -1: unbounded (renderer_pt.cpp:59) */
    int32_t min_num_vertices;
    int32_t rr_depth;                /* Russian roulette after hits of segments >= rr_depth-1 */
    float rr_prob;                   /* continuation probability (0.5 in the reference) */
    uint32_t seed_lo, seed_hi;
    int32_t raster_mode;             /* ORC_RASTER_* */
    int32_t rr_mode;                 /* 0: rr_prob; 1: min(1, max throughput channel) */
} orc_pt_params;

typedef struct orc_scene orc_scene;
orc_scene *orc_scene_new(const orc_scene_desc *d);
void orc_scene_free(orc_scene *s);
/* Samples [first, first+n): out[6 i ..] = raster x, y, C rgb, #splats.
 * table: n rows of table_width uniforms (row 0 = sample `first`), or NULL
 * for Philox. */
void orc_pt_samples(const orc_scene *s, const orc_pt_params *p, const float *table, uint32_t table_width,
                    uint64_t first_sample, uint64_t n, float *out, int nthreads);
void orc_film_splat(const float *samples, uint64_t n, int32_t width, int32_t height, double *film);
/* renderer::ptdirect (mode 1) and renderer::ptmis (mode 2): n x 16 floats,
 * layout in orc_render.c (nee_sample) */
void orc_ptnee_samples(const orc_scene *s, const orc_pt_params *p, int mode, const float *table,
                       uint32_t table_width, uint64_t first_sample, uint64_t n, float *out, int nthreads);
/* renderer::vcm, mode bdpt (orc_render.c).  max_num_vertices >= 2 bounds
 * both subpaths and every connected path.  Output per sample: (maxv+1)^2
 * records of 8 floats indexed t (maxv+1) + s: raster x, y, C rgb (= f w / p,
 * before the W H / N film scale), 1 if the strategy contributed, w, 0. */
typedef struct orc_bdpt_params {
    int32_t max_num_vertices;
    int32_t min_num_vertices;
    uint32_t seed_lo, seed_hi;
    int32_t raster_mode;
} orc_bdpt_params;
void orc_bdpt_samples(const orc_scene *s, const orc_bdpt_params *p, const float *table, uint32_t table_width,
                      uint64_t first_sample, uint64_t n, float *out, int nthreads);

/* renderer::vcm, modes vcm and bdpm, one pass with merge radius `radius`
 * over `num_photons` photon subpaths (orc_render.c).  Eye uniforms as
 * orc_bdpt_samples (table, or Philox with seed); photon subpath i uses the
 * light block of sample i of its own source (ptable, or Philox with pseed).
 * Output per eye sample: 2 (maxv+1)^2 records of 8 floats at
 * 2 (t (maxv+1) + s) + merge: raster x, y, C rgb (f w / p before the
 * W H / N film scale; merges summed over their photon vertices), count, sum
 * of w, 0.  A merge of eye vertex t-1 with photon vertex j is recorded at
 * s = j + 1. */
#define ORC_VCM_VCM 0
#define ORC_VCM_BDPM 2
typedef struct orc_vcm_params {
    int32_t max_num_vertices;
    int32_t min_num_vertices;
    uint32_t seed_lo, seed_hi;
    uint32_t pseed_lo, pseed_hi;
    int32_t raster_mode;
    int32_t mode;
    uint32_t num_photons;
    float radius;
} orc_vcm_params;
void orc_vcm_samples(const orc_scene *s, const orc_vcm_params *p, const float *table, uint32_t table_width,
                     const float *ptable, uint32_t ptable_width, uint64_t first_sample, uint64_t n, float *out,
                     int nthreads);

/* renderer::pm (orc_render.c): num_photons photon subpaths (uniform rows of
 * ptable, or Philox with pseed), then eye samples first_sample .. +n (table
 * or seed; renderer::vcm's eye layout).  max_num_vertices -1 (unbounded) or
 * >= 2.  Output per eye sample 16 floats: raster x, y, emission C rgb,
 * emission count, density-estimate C rgb, photons gathered, vertex count at
 * the gather vertex (0: none), 0 x 5 (before the W H / N film scale). */
typedef struct orc_pm_params {
    int32_t max_num_vertices;
    uint32_t seed_lo, seed_hi;
    uint32_t pseed_lo, pseed_hi;
    int32_t raster_mode;
    int32_t finalgather;
    uint32_t num_photons;
    float radius;
} orc_pm_params;
void orc_pm_samples(const orc_scene *s, const orc_pm_params *p, const float *table, uint32_t table_width,
                    const float *ptable, uint32_t ptable_width, uint64_t first_sample, uint64_t n, float *out,
                    int nthreads);

/* renderer::ppm (orc_render.c): n measurement points (eye uniforms: table
 * or seed, renderer::vcm's eye layout), num_passes passes of num_photons
 * photon subpaths (pass q uses photon rows q num_photons + i of ptable, or
 * Philox with pseed).  max_num_vertices -1 (unbounded) or >= 2.  Output per
 * point 16 floats: raster x, y, C rgb (throughput x tau / (pi r^2 P Np) +
 * emission, before the W H / n film scale), emission rgb, tau rgb, N,
 * radius, 1 if the sample made a point, 0 x 2. */
typedef struct orc_ppm_params {
    int32_t max_num_vertices;
    uint32_t seed_lo, seed_hi;
    uint32_t pseed_lo, pseed_hi;
    int32_t raster_mode;
    uint32_t num_passes;
    uint32_t num_photons;
    float initial_radius;
    float alpha;
} orc_ppm_params;
void orc_ppm_points(const orc_scene *s, const orc_ppm_params *p, const float *table, uint32_t table_width,
                    const float *ptable, uint32_t ptable_width, uint64_t n, float *out, int nthreads);

/* renderer::sppm (orc_render.c): a point per pixel over num_passes passes;
 * pass q's eye subpath of pixel i uses uniform index q W H + i (table rows
 * P W H), its photons rows q num_photons + i as orc_ppm_points.  Output per
 * pixel 16 floats: film C rgb, emission sum rgb, tau rgb, N, radius, passes
 * with a point, 0 x 4. */
void orc_sppm_pixels(const orc_scene *s, const orc_ppm_params *p, const float *table, uint32_t table_width,
                     const float *ptable, uint32_t ptable_width, float *out, int nthreads);

/* renderer::ltdirect (orc_render.c): samples first_sample .. +n (table
 * rows from first_sample, or Philox with seed).  max_num_vertices -1
 * (unbounded) or >= 2.  Output per sample M = max_num_vertices (32 when
 * unbounded) records of 8 floats, record k (light vertex k to the camera):
 * raster x, y, C rgb, contributed flag, 0, 0. */
typedef struct orc_lt_params {
    int32_t max_num_vertices;
    uint32_t seed_lo, seed_hi;
} orc_lt_params;
void orc_ltdirect_samples(const orc_scene *s, const orc_lt_params *p, const float *table, uint32_t table_width,
                          uint64_t first_sample, uint64_t n, float *out, int nthreads);

/* renderer::raycast (orc_render.c): W H x 3 floats, pixel y W + x. */
void orc_raycast(const orc_scene *s, float *out);

/* renderer::bdpt (orc_render.c).  max_num_vertices: -1 (unbounded) or >= 2;
 * min_num_vertices applies only when bounded (renderer_bdpt.cpp:887);
 * mis: 0 power heuristic, 1 simple.  Output per sample: 2 (M+1)^2 records of
 * 8 floats (M = max_num_vertices, 32 when unbounded) indexed
 * 2 (t (M+1) + s) + direct: raster x, y, C rgb (= C* w / SelectionPDF,
 * before the W H / N film scale), 1 if the strategy contributed, w, 0. */
typedef struct orc_bdptref_params {
    int32_t max_num_vertices;
    int32_t min_num_vertices;
    uint32_t seed_lo, seed_hi;
    int32_t raster_mode;
    int32_t mis;
} orc_bdptref_params;
void orc_bdptref_samples(const orc_scene *s, const orc_bdptref_params *p, const float *table, uint32_t table_width,
                         uint64_t first_sample, uint64_t n, float *out, int nthreads);

/* Film output transforms of Film_HDR::Save (orc_image.c), file row order:
 * RGBE bytes (4 per pixel) and PNG tone-mapped bytes (3 per pixel). */
void orc_film_rgbe(const float *film, int w, int h, uint8_t *out);
void orc_film_png(const float *film, int w, int h, uint8_t *out);

/* BSDF test hooks on a flat frame with normal n */
int orc_bsdf_query(const orc_material *m, const float n[3], const float wi[3], const float u[2],
                   float wo[3], float *pdf, float f[3]);
void orc_bsdf_eval(const orc_material *m, const float n[3], const float wi[3], const float wo[3],
                   float *pdf, float f[3]);

/* Learned directional sampler (no counterpart in the reference; orc_guide.c).
 * FP64 restatement of the MLP, the vMF mixture and the training loss:
 *   input x[7] in [0,1]: position (scene bounds), normal (theta/pi,
 *   phi/2pi), direction to the previous vertex (theta/pi, phi/2pi);
 *   features sin/cos(2^f pi x_d), f < 12, index d*24 + 2f (+1 for cos);
 *   168 -> 64 -> 64 -> 64 -> 64 (ReLU) -> 41;
 *   y[3k..3k+2] lobe mean (normalised), y[24+k] log kappa (clamped to
 *   [-4, 9]), y[32+k] lobe weight logits (softmax), y[40] BSDF selection
 *   probability 0.5 + 0.4 sigmoid(y40);
 *   parameters (canonical flat layout, W[k][n] = input k -> output n):
 *   W1[169][64] (row 168 = bias), W2[64][64], b2[64], W3, b3, W4, b4,
 *   W5[64][41], b5[41] -- ORC_GUIDE_NPARAM values.
 * bf16 != 0 rounds weights (except the biases of layers 2-5) and every
 * layer's input activations to bfloat16 (round to nearest even), as the
 * device kernels do; accumulation stays FP64.
 * Pinned by: vMF normalisation and sampling in closed form (Jakob 2012),
 * central finite differences of the loss, and the Adam update formula. */
#define ORC_GUIDE_NIN 7
#define ORC_GUIDE_NFEAT 168
#define ORC_GUIDE_H 64
#define ORC_GUIDE_NOUT 41
#define ORC_GUIDE_NLOBE 8
#define ORC_GUIDE_NPARAM 25961
void orc_guide_features(const double x[7], double feat[168]);
void orc_guide_forward(const double *params, const double x[7], double y[41], int bf16);
void orc_guide_forward_batch(const double *params, const double *x, uint64_t n, double *y, int bf16);
double orc_guide_vmf_pdf(const double mu[3], double kappa, const double w[3]);
void orc_guide_vmf_sample(const double mu[3], double kappa, double u1, double u2, double out[3]);
/* vMF mixture of the outputs y: pdf of direction w (solid angle), and the
 * BSDF selection probability */
double orc_guide_mix_pdf(const double y[41], const double w[3], double *alpha);
/* a direction from the mixture: lobe by u (cumulative weights), then the lobe
 * sampled with (u1, u2); returns the lobe */
int orc_guide_mix_sample(const double y[41], double u, double u1, double u2, double out[3]);
/* mean over n records of -w log(alpha pb + (1 - alpha) p_vmf(dir)); grad
 * (ORC_GUIDE_NPARAM, may be NULL) receives its gradient */
double orc_guide_loss(const double *params, uint64_t n, const double *x, const double *dir, const double *pb,
                      const double *w, double *grad, int bf16);
void orc_guide_adam(double *params, double *m, double *v, const double *grad, uint64_t n, int step, double lr,
                    double beta1, double beta2, double eps);
double orc_bf16_round(double x);

#ifdef __cplusplus
}
#endif
#endif
