/*
 * orc_isect.c -- CPU oracle: triangle tests, brute force, QBVH and binary
 * binned-SAH BVH restated from the Lightmetrica v2 reference.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h).  Compiled with -ffp-contract=off:
 * every expression rounds exactly as written, which is what makes bit-exact
 * comparison with the reference headers (oracle/_ref) and with the HIP
 * kernels meaningful.
 */
#include "orc.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <immintrin.h>

#define ORC_INF FLT_MAX          /* Math::Inf() is numeric_limits<float>::max(), math.h:1663 */
#define ORC_EPS 1e-4f            /* Math::Eps<float>(), math.h:1667 */
#define ORC_EPS_LARGE 1e-3f      /* Math::EpsLarge<float>(), math.h:1665 */

/* ---------------------------------------------------------------------- */
/* Small FP32 vector helpers.  Dot follows _mm_dp_ps(...,0x71): the products
 * of x and y are summed first, then z (math.h:1728-1731). */

static inline void v3sub(float r[3], const float a[3], const float b[3])
{
    r[0] = a[0] - b[0]; r[1] = a[1] - b[1]; r[2] = a[2] - b[2];
}
static inline void v3cross(float r[3], const float a[3], const float b[3])
{
    /* math.h:1784-1787 (the SSE variant at 1791-1799 forms the same products) */
    float x = a[1] * b[2] - b[1] * a[2];
    float y = a[2] * b[0] - b[2] * a[0];
    float z = a[0] * b[1] - b[0] * a[1];
    r[0] = x; r[1] = y; r[2] = z;
}
static inline float v3dot(const float a[3], const float b[3])
{
    float xy = a[0] * b[0] + a[1] * b[1];
    return xy + a[2] * b[2];
}

/* ---------------------------------------------------------------------- */
/* TriAccel (Wald projection test), include/lightmetrica/triaccel.h:50-151 */

int orc_triaccel_load(const float A[3], const float B[3], const float C[3], orc_triaccel *out)
{
    static const int wald_modulo[4] = {1, 2, 0, 1};          /* triaccel.h:52 */
    float b[3], c[3], N[3];
    v3sub(b, C, A);                                            /* :54 */
    v3sub(c, B, A);                                            /* :55 */
    v3cross(N, c, b);                                          /* :56 */

    uint32_t k = 0;                                            /* :58-66 */
    for (int j = 0; j < 3; j++)
        if (fabsf(N[j]) > fabsf(N[k])) k = (uint32_t)j;

    const uint32_t u = (uint32_t)wald_modulo[k];
    const uint32_t v = (uint32_t)wald_modulo[k + 1];
    const float n_k = N[k];
    const float denom = b[u] * c[v] - b[v] * c[u];            /* :71 */
    if (denom == 0.0f) {                                       /* :73-77 */
        out->k = 3;
        return 1;
    }
    out->k = k;
    out->n_u = N[u] / n_k;                                     /* :80-89 */
    out->n_v = N[v] / n_k;
    out->n_d = v3dot(A, N) / n_k;
    out->b_nu = b[u] / denom;
    out->b_nv = -b[v] / denom;
    out->a_u = A[u];
    out->a_v = A[v];
    out->c_nu = c[v] / denom;
    out->c_nv = -c[u] / denom;
    return 0;
}

int orc_triaccel_intersect(const orc_triaccel *tri, const float o[3], const float d[3],
                           float mint, float maxt, float *u, float *v, float *t)
{
    float o_u, o_v, o_k, d_u, d_v, d_k;
    switch (tri->k) {                                          /* :96-125 */
    case 0: o_u = o[1]; o_v = o[2]; o_k = o[0]; d_u = d[1]; d_v = d[2]; d_k = d[0]; break;
    case 1: o_u = o[2]; o_v = o[0]; o_k = o[1]; d_u = d[2]; d_v = d[0]; d_k = d[1]; break;
    case 2: o_u = o[0]; o_v = o[1]; o_k = o[2]; d_u = d[0]; d_v = d[1]; d_k = d[2]; break;
    default: return 0;
    }
    const float demon = d_u * tri->n_u + d_v * tri->n_v + d_k;     /* :127 */
    if (demon == 0.0f) return 0;
    const float tt = (tri->n_d - o_u * tri->n_u - o_v * tri->n_v - o_k) / demon;   /* :134 */
    *t = tt;
    if (tt < mint || tt > maxt) return 0;                      /* :135-138 */
    const float hu = o_u + tt * d_u - tri->a_u;                /* :141-142 */
    const float hv = o_v + tt * d_v - tri->a_v;
    const float uu = hv * tri->b_nu + hu * tri->b_nv;          /* :145-146 */
    const float vv = hu * tri->c_nu + hv * tri->c_nv;
    *u = uu;
    *v = vv;
    return uu >= 0.0f && vv >= 0.0f && uu + vv <= 1.0f;        /* :148 */
}

/* ---------------------------------------------------------------------- */
/* Watertight ray/triangle test.  Restated from Woop, Benthin and Wald,
 * "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013, with one change:
 * the three edge functions are evaluated with Kahan's FMA difference of
 * products instead of an FP64 fallback.  That value is exactly zero iff the
 * exact edge function is zero and otherwise carries its exact sign, so no
 * FP64 is needed for watertightness.  The HIP kernel evaluates the same
 * operation sequence (explicit fmaf everywhere, contraction off), so hit
 * index, t and barycentrics agree bit for bit. */

static inline float dop(float a, float b, float c, float d)
{
    /* a*b - c*d */
    const float w = c * d;
    const float e = fmaf(-c, d, w);
    const float f = fmaf(a, b, -w);
    return f + e;
}

typedef struct wt_ray {
    int kx, ky, kz;
    float Sx, Sy, Sz;
} wt_ray;

static inline void wt_setup(wt_ray *r, const float d[3])
{
    int kz = 0;
    if (fabsf(d[1]) > fabsf(d[kz])) kz = 1;
    if (fabsf(d[2]) > fabsf(d[kz])) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) { int tmp = kx; kx = ky; ky = tmp; }
    r->kx = kx; r->ky = ky; r->kz = kz;
    r->Sx = d[kx] / d[kz];
    r->Sy = d[ky] / d[kz];
    r->Sz = 1.0f / d[kz];
}

/* Returns 1 on a hit with t in [mint, maxt]; writes t and the unnormalised
 * edge functions (U,V,W) and det so a caller can defer the divisions. */
static inline int wt_test(const wt_ray *r, const float o[3], const float v0[3],
                          const float v1[3], const float v2[3], float mint, float maxt,
                          float *t_out, float *V_out, float *W_out, float *det_out)
{
    float A[3], B[3], C[3];
    v3sub(A, v0, o);
    v3sub(B, v1, o);
    v3sub(C, v2, o);
    const float Ax = fmaf(-r->Sx, A[r->kz], A[r->kx]);
    const float Ay = fmaf(-r->Sy, A[r->kz], A[r->ky]);
    const float Bx = fmaf(-r->Sx, B[r->kz], B[r->kx]);
    const float By = fmaf(-r->Sy, B[r->kz], B[r->ky]);
    const float Cx = fmaf(-r->Sx, C[r->kz], C[r->kx]);
    const float Cy = fmaf(-r->Sy, C[r->kz], C[r->ky]);
    const float U = dop(Cx, By, Cy, Bx);
    const float V = dop(Ax, Cy, Ay, Cx);
    const float W = dop(Bx, Ay, By, Ax);
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    const float det = (U + V) + W;
    if (det == 0.0f) return 0;
    const float Az = r->Sz * A[r->kz];
    const float Bz = r->Sz * B[r->kz];
    const float Cz = r->Sz * C[r->kz];
    const float T = fmaf(U, Az, fmaf(V, Bz, W * Cz));
    const float t = T / det;
    if (!(t >= mint && t <= maxt)) return 0;
    *t_out = t; *V_out = V; *W_out = W; *det_out = det;
    return 1;
}

int orc_watertight_intersect(const float v0[3], const float v1[3], const float v2[3],
                             const float o[3], const float d[3], float mint, float maxt,
                             float *b0, float *b1, float *t)
{
    wt_ray r;
    wt_setup(&r, d);
    float tt, V, W, det;
    if (!wt_test(&r, o, v0, v1, v2, mint, maxt, &tt, &V, &W, &det)) return 0;
    *t = tt;
    *b0 = V / det;     /* weight of the 2nd vertex */
    *b1 = W / det;     /* weight of the 3rd vertex */
    return 1;
}

/* ---------------------------------------------------------------------- */
/* Bound::Intersect, include/lightmetrica/bound.h:67-89 */

int orc_bound_intersect(const float bmin[3], const float bmax[3], const float o[3],
                        const float d[3], float tMin, float tMax)
{
    const float *bb[2] = {bmin, bmax};
    float tmin = 0.0f, tmax = 0.0f;
    for (int a = 0; a < 3; a++) {
        const int neg = d[a] < 0.0f;
        const float inv_max = d[a] == 0.0f ? ORC_INF : 1.0f / d[a];
        const float inv_min = d[a] == 0.0f ? 0.0f : 1.0f / d[a];
        const float tamax = (bb[1 - neg][a] - o[a]) * inv_max;
        const float tamin = (bb[neg][a] - o[a]) * inv_min;
        if (a == 0) {
            tmin = tamin;
            tmax = tamax;
        } else {
            if ((tmin > tamax) || (tamin > tmax)) return 0;
            if (tamin > tmin) tmin = tamin;
            if (tamax < tmax) tmax = tamax;
        }
    }
    return (tmin < tMax) && (tmax > tMin);
}

/* ---------------------------------------------------------------------- */
/* Thread helper: split [0,n) into contiguous chunks over nthreads. */

typedef void (*range_fn)(void *ctx, uint32_t begin, uint32_t end);
typedef struct { range_fn fn; void *ctx; uint32_t b, e; } range_job;

static void *range_thread(void *p)
{
    range_job *j = (range_job *)p;
    j->fn(j->ctx, j->b, j->e);
    return NULL;
}

static void parallel_range(range_fn fn, void *ctx, uint32_t n, int nthreads)
{
    if (nthreads <= 1 || n < 64) { fn(ctx, 0, n); return; }
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    range_job jobs[256];
    /* interleaved chunks of 1024 keep the load balanced on mixed rays */
    const uint32_t grain = 1024;
    (void)grain;
    for (int i = 0; i < nthreads; i++) {
        jobs[i].fn = fn; jobs[i].ctx = ctx;
        jobs[i].b = (uint32_t)((uint64_t)n * (uint64_t)i / (uint64_t)nthreads);
        jobs[i].e = (uint32_t)((uint64_t)n * (uint64_t)(i + 1) / (uint64_t)nthreads);
        pthread_create(&th[i], NULL, range_thread, &jobs[i]);
    }
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
}

/* ---------------------------------------------------------------------- */
/* Brute force, accel_naive.cpp:92-124.  The reference keeps the LAST
 * triangle with t <= maxT (t > maxt rejects only strictly greater,
 * triaccel.h:135); ORC_TIE_MIN_ID keeps the smallest id among equal t. */

typedef struct {
    const float *verts; const uint32_t *tris; uint32_t ntri;
    const orc_ray *rays; orc_hit *hits; int test, tie;
    orc_triaccel *ta;
} naive_ctx;

static void naive_range(void *p, uint32_t b, uint32_t e)
{
    naive_ctx *c = (naive_ctx *)p;
    for (uint32_t r = b; r < e; r++) {
        const orc_ray *ray = &c->rays[r];
        float best = ray->tmax;
        uint32_t best_id = 0xffffffffu;
        float bu = 0, bv = 0;
        wt_ray wr = {0, 0, 0, 0.0f, 0.0f, 0.0f};
        if (c->test == ORC_TEST_WATERTIGHT) wt_setup(&wr, ray->d);
        for (uint32_t i = 0; i < c->ntri; i++) {
            float t, u, v;
            int hit;
            if (c->test == ORC_TEST_TRIACCEL) {
                hit = orc_triaccel_intersect(&c->ta[i], ray->o, ray->d, ray->tmin, best, &u, &v, &t);
            } else {
                const uint32_t *f = &c->tris[3 * i];
                float V, W, det;
                hit = wt_test(&wr, ray->o, &c->verts[3 * f[0]], &c->verts[3 * f[1]],
                              &c->verts[3 * f[2]], ray->tmin, best, &t, &V, &W, &det);
                if (hit) { u = V / det; v = W / det; }
            }
            if (!hit) continue;
            if (c->tie == ORC_TIE_MIN_ID && t == best && best_id != 0xffffffffu && i > best_id)
                continue;
            best = t; best_id = i; bu = u; bv = v;
        }
        orc_hit *h = &c->hits[r];
        h->id = best_id;
        if (best_id != 0xffffffffu) { h->t = best; h->b0 = bu; h->b1 = bv; }
        else { h->t = 0; h->b0 = 0; h->b1 = 0; }
    }
}

void orc_closest_hit_naive(const float *verts, const uint32_t *tris, uint32_t ntri,
                           const orc_ray *rays, uint32_t nray, int test, int tie,
                           orc_hit *hits, int nthreads)
{
    naive_ctx c = {verts, tris, ntri, rays, hits, test, tie, NULL};
    if (test == ORC_TEST_TRIACCEL) {
        c.ta = (orc_triaccel *)calloc(ntri ? ntri : 1, sizeof(orc_triaccel));
        for (uint32_t i = 0; i < ntri; i++) {
            const uint32_t *f = &tris[3 * i];
            orc_triaccel_load(&verts[3 * f[0]], &verts[3 * f[1]], &verts[3 * f[2]], &c.ta[i]);
            c.ta[i].face_index = i;
        }
    }
    parallel_range(naive_range, &c, nray, nthreads);
    free(c.ta);
}

/* ---------------------------------------------------------------------- */
/* Shared binned-SAH machinery (accel_qbvh.cpp:202-389 and
 * accel_bvh_sahbin.cpp:120-220).  Bound defaults are (+Inf, -Inf) with
 * Inf = FLT_MAX (bound.h:43-44, math.h:1663). */

typedef struct { float mn[3], mx[3]; } bnd;

static inline void bnd_empty(bnd *b)
{
    for (int a = 0; a < 3; a++) { b->mn[a] = ORC_INF; b->mx[a] = -ORC_INF; }
}
static inline void bnd_union(bnd *r, const bnd *b)
{
    for (int a = 0; a < 3; a++) {
        r->mn[a] = fminf(r->mn[a], b->mn[a]);
        r->mx[a] = fmaxf(r->mx[a], b->mx[a]);
    }
}
static inline void bnd_union_pt(bnd *r, const float p[3])
{
    for (int a = 0; a < 3; a++) {
        r->mn[a] = fminf(r->mn[a], p[a]);
        r->mx[a] = fmaxf(r->mx[a], p[a]);
    }
}
static inline float bnd_area(const bnd *b)
{
    /* bound.h:52-56 */
    const float dx = b->mx[0] - b->mn[0], dy = b->mx[1] - b->mn[1], dz = b->mx[2] - b->mn[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}
static inline float bnd_centroid(const bnd *b, int a)
{
    return (b->mn[a] + b->mx[a]) * 0.5f;
}
static inline int bnd_longest(const bnd *b)
{
    /* bound.h:46-50 */
    const float dx = b->mx[0] - b->mn[0], dy = b->mx[1] - b->mn[1], dz = b->mx[2] - b->mn[2];
    return dx > dy && dx > dz ? 0 : dy > dz ? 1 : 2;
}

typedef struct {
    uint32_t ntri;
    orc_triaccel *ta;     /* per triangle, in enumeration order */
    bnd *tb;              /* padded per-triangle bounds */
    int *indices;
    float *tv;            /* per triangle: 9 floats, the three vertices */
} sah_input;

static void sah_input_init(sah_input *in, const float *verts, const uint32_t *tris, uint32_t ntri)
{
    in->ntri = ntri;
    in->ta = (orc_triaccel *)calloc(ntri ? ntri : 1, sizeof(orc_triaccel));
    in->tb = (bnd *)calloc(ntri ? ntri : 1, sizeof(bnd));
    in->indices = (int *)calloc(ntri ? ntri : 1, sizeof(int));
    in->tv = (float *)calloc(ntri ? 9 * (size_t)ntri : 1, sizeof(float));
    for (uint32_t i = 0; i < ntri; i++) {
        const uint32_t *f = &tris[3 * i];
        const float *p1 = &verts[3 * f[0]], *p2 = &verts[3 * f[1]], *p3 = &verts[3 * f[2]];
        memcpy(&in->tv[9 * (size_t)i], p1, 3 * sizeof(float));
        memcpy(&in->tv[9 * (size_t)i + 3], p2, 3 * sizeof(float));
        memcpy(&in->tv[9 * (size_t)i + 6], p3, 3 * sizeof(float));
        orc_triaccel_load(p1, p2, p3, &in->ta[i]);
        in->ta[i].face_index = i;
        in->ta[i].prim_index = 0;
        bnd b;
        bnd_empty(&b);
        bnd_union_pt(&b, p1); bnd_union_pt(&b, p2); bnd_union_pt(&b, p3);
        for (int a = 0; a < 3; a++) { b.mn[a] -= ORC_EPS; b.mx[a] += ORC_EPS; }   /* accel_qbvh.cpp:183-184 */
        in->tb[i] = b;
        in->indices[i] = (int)i;
    }
}

static void sah_input_free(sah_input *in)
{
    free(in->ta); free(in->tb); free(in->indices); free(in->tv);
}

#define NBINS 100

static inline int bin_of(float c, float mn, float mx)
{
    int idx = (int)((c - mn) / (mx - mn) * (float)NBINS);
    return idx < NBINS - 1 ? idx : NBINS - 1;
}

/* Evaluates the reference's SAH over 100 bins.  The reference recomputes
 * each prefix/suffix union in an O(bins^2) loop; unions and integer counts
 * are exact, so the running prefix/suffix below yields identical costs.
 * Returns the split bin (first minimum, std::min_element) and its cost. */
static int sah_best_split(const sah_input *in, int begin, int end, int axis,
                          const bnd *cb, const bnd *node_bound, float *best_cost)
{
    bnd bins[NBINS];
    int counts[NBINS];
    for (int i = 0; i < NBINS; i++) { bnd_empty(&bins[i]); counts[i] = 0; }
    const float mn = cb->mn[axis], mx = cb->mx[axis];
    for (int i = begin; i < end; i++) {
        const bnd *b = &in->tb[in->indices[i]];
        const int idx = bin_of(bnd_centroid(b, axis), mn, mx);
        bnd_union(&bins[idx], b);
        counts[idx]++;
    }
    bnd suffix[NBINS];
    int nsuffix[NBINS];
    bnd acc; bnd_empty(&acc);
    int nacc = 0;
    for (int i = NBINS - 1; i >= 0; i--) {
        bnd_union(&acc, &bins[i]); nacc += counts[i];
        suffix[i] = acc; nsuffix[i] = nacc;
    }
    const float node_area = bnd_area(node_bound);
    bnd pre; bnd_empty(&pre);
    int npre = 0;
    int best = 0;
    float bestc = 0.0f;
    for (int split = 0; split < NBINS - 1; split++) {
        bnd_union(&pre, &bins[split]); npre += counts[split];
        const int n2 = nsuffix[split + 1];
        const float Cb = 0.125f;
        const float C1 = npre > 0 ? bnd_area(&pre) * (float)npre : 0.0f;
        const float C2 = n2 > 0 ? bnd_area(&suffix[split + 1]) * (float)n2 : 0.0f;
        const float cost = Cb + (C1 + C2) / node_area;
        if (split == 0 || cost < bestc) { bestc = cost; best = split; }
    }
    *best_cost = bestc;
    return best;
}

static int sah_partition(sah_input *in, int begin, int end, int axis, const bnd *cb, int split)
{
    /* std::partition semantics: elements whose bin <= split come first.  The
     * order inside each side is not a parity target (it only affects which
     * of two exactly equal t wins under the reference tie rule). */
    int i = begin, j = end - 1;
    const float mn = cb->mn[axis], mx = cb->mx[axis];
    while (1) {
        while (i <= j && bin_of(bnd_centroid(&in->tb[in->indices[i]], axis), mn, mx) <= split) i++;
        while (i <= j && bin_of(bnd_centroid(&in->tb[in->indices[j]], axis), mn, mx) > split) j--;
        if (i >= j) break;
        int t = in->indices[i]; in->indices[i] = in->indices[j]; in->indices[j] = t;
        i++; j--;
    }
    return i;
}

static void range_bounds(const sah_input *in, int begin, int end, bnd *nb, bnd *cb)
{
    bnd_empty(nb); bnd_empty(cb);
    for (int i = begin; i < end; i++) {
        const bnd *b = &in->tb[in->indices[i]];
        bnd_union(nb, b);
        float c[3] = {bnd_centroid(b, 0), bnd_centroid(b, 1), bnd_centroid(b, 2)};
        bnd_union_pt(cb, c);
    }
}

/* ---------------------------------------------------------------------- */
/* QBVH, accel_qbvh.cpp:54-137 (node), 202-389 (build), 392-481 (traversal) */

typedef struct {
    __m128 bounds[2][3];
    int children[4];
} qnode;                                   /* 112 B, accel_qbvh.cpp:60-72 */

#define QBVH_EMPTY ((int)0xffffffff)

struct orc_qbvh {
    sah_input in;
    qnode *nodes;
    uint32_t nnodes, cap;
};

static void qnode_init(qnode *n)
{
    for (int i = 0; i < 3; i++) {
        n->bounds[0][i] = _mm_set1_ps(INFINITY);
        n->bounds[1][i] = _mm_set1_ps(-INFINITY);
    }
    for (int i = 0; i < 4; i++) n->children[i] = QBVH_EMPTY;
}

static void qnode_set_bound(qnode *n, int child, const bnd *b)
{
    for (int a = 0; a < 3; a++) {
        ((float *)&n->bounds[0][a])[child] = b->mn[a];
        ((float *)&n->bounds[1][a])[child] = b->mx[a];
    }
}

static void qnode_leaf(qnode *n, int child, unsigned size, unsigned offset)
{
    if (size == 0) { n->children[child] = QBVH_EMPTY; return; }
    unsigned v = 0x80000000u;
    v |= ((size - 1) & 0xfu) << 27;
    v |= offset & 0x07ffffffu;
    n->children[child] = (int)v;
}

static uint32_t qbvh_new_node(orc_qbvh *q)
{
    if (q->nnodes == q->cap) {
        q->cap = q->cap ? q->cap * 2 : 1024;
        q->nodes = (qnode *)realloc(q->nodes, sizeof(qnode) * q->cap);
    }
    qnode_init(&q->nodes[q->nnodes]);
    return q->nnodes++;
}

static void qbvh_build_rec(orc_qbvh *q, int begin, int end, int parent, int child, int depth)
{
    sah_input *in = &q->in;
    bnd bound, cb;
    range_bounds(in, begin, end, &bound, &cb);
    if (end - begin < 10) {                                   /* LeafNumNodes, :232-239 */
        qnode_set_bound(&q->nodes[parent], child, &bound);
        qnode_leaf(&q->nodes[parent], child, (unsigned)(end - begin), (unsigned)begin);
        return;
    }
    const int axis = bnd_longest(&cb);                         /* centroid bound axis, :259 */
    int mid;
    if (!(cb.mx[axis] > cb.mn[axis])) {
        /* Coincident centroids: the reference divides by zero here; split at
         * the middle instead (documented oracle guard). */
        mid = begin + (end - begin) / 2;
    } else {
        float cost;
        const int split = sah_best_split(in, begin, end, axis, &cb, &bound, &cost);
        mid = sah_partition(in, begin, end, axis, &cb, split);
        if (mid == begin || mid == end) mid = begin + (end - begin) / 2;
    }
    int current, c1, c2;
    if (depth % 2 == 1) {                                      /* :340-349 */
        current = parent; c1 = child; c2 = child + 1;
    } else {                                                   /* :350-367 */
        current = (int)qbvh_new_node(q);
        q->nodes[parent].children[child] = current;
        qnode_set_bound(&q->nodes[parent], child, &bound);
        c1 = 0; c2 = 2;
    }
    qbvh_build_rec(q, begin, mid, current, c1, depth + 1);
    qbvh_build_rec(q, mid, end, current, c2, depth + 1);
}

orc_qbvh *orc_qbvh_build(const float *verts, const uint32_t *tris, uint32_t ntri)
{
    orc_qbvh *q = (orc_qbvh *)calloc(1, sizeof(orc_qbvh));
    sah_input_init(&q->in, verts, tris, ntri);
    q->nodes = NULL; q->nnodes = 0; q->cap = 0;
    qbvh_new_node(q);
    if (ntri > 0) qbvh_build_rec(q, 0, (int)ntri, 0, 0, 0);
    return q;
}

void orc_qbvh_free(orc_qbvh *q)
{
    if (!q) return;
    sah_input_free(&q->in);
    free(q->nodes);
    free(q);
}

uint32_t orc_qbvh_num_nodes(const orc_qbvh *q) { return q->nnodes; }

static inline int qnode_hit(const qnode *n, const __m128 o[3], const __m128 inv_min[3],
                            const __m128 inv_max[3], const int sign[3], float mint_s, float maxt_s)
{
    /* accel_qbvh.cpp:123-136; operand order of max/min kept for NaN behaviour */
    __m128 minT = _mm_set1_ps(mint_s);
    __m128 maxT = _mm_set1_ps(maxt_s);
    for (int a = 0; a < 3; a++) {
        minT = _mm_max_ps(minT, _mm_mul_ps(_mm_sub_ps(n->bounds[sign[a]][a], o[a]), inv_min[a]));
        maxT = _mm_min_ps(maxT, _mm_mul_ps(_mm_sub_ps(n->bounds[1 - sign[a]][a], o[a]), inv_max[a]));
    }
    return _mm_movemask_ps(_mm_cmpge_ps(maxT, minT));
}

typedef struct { const orc_qbvh *q; const orc_ray *rays; orc_hit *hits; int test, tie; } qbvh_ctx;

static void qbvh_range(void *p, uint32_t b, uint32_t e)
{
    qbvh_ctx *c = (qbvh_ctx *)p;
    const orc_qbvh *q = c->q;
    for (uint32_t r = b; r < e; r++) {
        const orc_ray *ray = &c->rays[r];
        float minT = ray->tmin, maxT = ray->tmax;
        int hit = 0, min_index = -1;
        float mb0 = 0, mb1 = 0;
        __m128 o[3], inv_min[3], inv_max[3];
        int sign[3];
        for (int a = 0; a < 3; a++) {                          /* :403-424 */
            o[a] = _mm_set1_ps(ray->o[a]);
            inv_min[a] = _mm_set1_ps(ray->d[a] == 0.0f ? ORC_EPS_LARGE : 1.0f / ray->d[a]);
            inv_max[a] = _mm_set1_ps(ray->d[a] == 0.0f ? ORC_INF : 1.0f / ray->d[a]);
            sign[a] = ray->d[a] < 0.0f;
        }
        wt_ray wr = {0, 0, 0, 0.0f, 0.0f, 0.0f};
        if (c->test == ORC_TEST_WATERTIGHT) wt_setup(&wr, ray->d);
        int stack[64];                                         /* :429-434 */
        int sp = 0;
        stack[0] = 0;
        while (sp >= 0) {
            const int data = stack[sp--];
            if (data < 0) {
                if (data == QBVH_EMPTY) continue;
                const unsigned size = (unsigned)(((data >> 27) & 0xf) + 1);
                const unsigned offset = (unsigned)data & 0x07ffffffu;
                for (unsigned i = offset; i < offset + size; i++) {
                    const int ti = q->in.indices[i];
                    float t, u, v;
                    int h;
                    if (c->test == ORC_TEST_TRIACCEL) {
                        h = orc_triaccel_intersect(&q->in.ta[ti], ray->o, ray->d, minT, maxT, &u, &v, &t);
                    } else {
                        const float *tv = &q->in.tv[9 * (size_t)ti];
                        float V, W, det;
                        h = wt_test(&wr, ray->o, tv, tv + 3, tv + 6, minT, maxT, &t, &V, &W, &det);
                        if (h) { u = V / det; v = W / det; }
                    }
                    if (!h) continue;
                    if (c->tie == ORC_TIE_MIN_ID && hit && t == maxT && ti > min_index) continue;
                    hit = 1; maxT = t; min_index = ti; mb0 = u; mb1 = v;
                }
            } else {
                const qnode *n = &q->nodes[data];
                const int mask = qnode_hit(n, o, inv_min, inv_max, sign, minT, maxT);
                if (mask & 1) stack[++sp] = n->children[0];
                if (mask & 2) stack[++sp] = n->children[1];
                if (mask & 4) stack[++sp] = n->children[2];
                if (mask & 8) stack[++sp] = n->children[3];
            }
        }
        orc_hit *h = &c->hits[r];
        if (hit) { h->t = maxT; h->b0 = mb0; h->b1 = mb1; h->id = (uint32_t)min_index; }
        else { h->t = 0; h->b0 = 0; h->b1 = 0; h->id = 0xffffffffu; }
    }
}

void orc_qbvh_intersect(const orc_qbvh *q, const orc_ray *rays, uint32_t nray,
                        orc_hit *hits, int nthreads)
{
    qbvh_ctx c = {q, rays, hits, ORC_TEST_TRIACCEL, ORC_TIE_LAST_WINS};
    parallel_range(qbvh_range, &c, nray, nthreads);
}

void orc_qbvh_intersect_ex(const orc_qbvh *q, const orc_ray *rays, uint32_t nray,
                           orc_hit *hits, int test, int tie, int nthreads)
{
    qbvh_ctx c = {q, rays, hits, test, tie};
    parallel_range(qbvh_range, &c, nray, nthreads);
}

/* ---------------------------------------------------------------------- */
/* Binary BVH, accel_bvh_sahbin.cpp:36-55 (node), 120-220 (build),
 * 234-288 (traversal: child1 then child2, no ordering) */

typedef struct {
    bnd b;
    int isleaf;
    int a, c;            /* leaf: begin,end; internal: child1, child2 */
} bnode;

struct orc_bvh {
    sah_input in;
    bnode *nodes;
    uint32_t nnodes, cap;
};

static int bvh_build_rec(orc_bvh *t, int begin, int end)
{
    if (t->nnodes == t->cap) {
        t->cap = t->cap ? t->cap * 2 : 1024;
        t->nodes = (bnode *)realloc(t->nodes, sizeof(bnode) * t->cap);
    }
    const int idx = (int)t->nnodes++;
    sah_input *in = &t->in;
    bnd nb, cb;
    range_bounds(in, begin, end, &nb, &cb);
    t->nodes[idx].b = nb;
    if (end - begin < 10) {                                    /* :138-145 */
        t->nodes[idx].isleaf = 1; t->nodes[idx].a = begin; t->nodes[idx].c = end;
        return idx;
    }
    const int axis = bnd_longest(&nb);                         /* node bound axis, :148 */
    int mid;
    if (!(cb.mx[axis] > cb.mn[axis])) {
        mid = begin + (end - begin) / 2;                       /* oracle guard, see qbvh */
    } else {
        float cost;
        const int split = sah_best_split(in, begin, end, axis, &cb, &nb, &cost);
        if (cost > (float)(end - begin)) {                     /* :194-200 */
            t->nodes[idx].isleaf = 1; t->nodes[idx].a = begin; t->nodes[idx].c = end;
            return idx;
        }
        mid = sah_partition(in, begin, end, axis, &cb, split);
        if (mid == begin || mid == end) mid = begin + (end - begin) / 2;
    }
    t->nodes[idx].isleaf = 0;
    const int c1 = bvh_build_rec(t, begin, mid);
    const int c2 = bvh_build_rec(t, mid, end);
    t->nodes[idx].a = c1;
    t->nodes[idx].c = c2;
    return idx;
}

orc_bvh *orc_bvh_build(const float *verts, const uint32_t *tris, uint32_t ntri)
{
    orc_bvh *t = (orc_bvh *)calloc(1, sizeof(orc_bvh));
    sah_input_init(&t->in, verts, tris, ntri);
    if (ntri > 0) bvh_build_rec(t, 0, (int)ntri);
    return t;
}

void orc_bvh_free(orc_bvh *b)
{
    if (!b) return;
    sah_input_free(&b->in);
    free(b->nodes);
    free(b);
}

uint32_t orc_bvh_num_nodes(const orc_bvh *b) { return b->nnodes; }

typedef struct { const orc_bvh *t; const orc_ray *rays; orc_hit *hits; } bvh_ctx;

static void bvh_range(void *p, uint32_t b, uint32_t e)
{
    bvh_ctx *c = (bvh_ctx *)p;
    const orc_bvh *t = c->t;
    for (uint32_t r = b; r < e; r++) {
        const orc_ray *ray = &c->rays[r];
        float minT = ray->tmin, maxT = ray->tmax;
        int hit = 0, min_index = -1;
        float mb0 = 0, mb1 = 0;
        /* explicit stack reproducing the recursion order child1, child2 */
        int stack[128];
        int sp = 0;
        if (t->nnodes > 0) stack[sp++] = 0;
        while (sp > 0) {
            const bnode *n = &t->nodes[stack[--sp]];
            if (!orc_bound_intersect(n->b.mn, n->b.mx, ray->o, ray->d, minT, maxT)) continue;
            if (n->isleaf) {
                for (int i = n->a; i < n->c; i++) {
                    const int ti = t->in.indices[i];
                    float tt, u, v;
                    if (orc_triaccel_intersect(&t->in.ta[ti], ray->o, ray->d, minT, maxT, &u, &v, &tt)) {
                        hit = 1; maxT = tt; min_index = ti; mb0 = u; mb1 = v;
                    }
                }
            } else {
                stack[sp++] = n->c;     /* child2 visited after child1 */
                stack[sp++] = n->a;
            }
        }
        orc_hit *h = &c->hits[r];
        if (hit) { h->t = maxT; h->b0 = mb0; h->b1 = mb1; h->id = (uint32_t)min_index; }
        else { h->t = 0; h->b0 = 0; h->b1 = 0; h->id = 0xffffffffu; }
    }
}

void orc_bvh_intersect(const orc_bvh *b, const orc_ray *rays, uint32_t nray,
                       orc_hit *hits, int nthreads)
{
    bvh_ctx c = {b, rays, hits};
    parallel_range(bvh_range, &c, nray, nthreads);
}

/* ---------------------------------------------------------------------- */
/* include/lightmetrica/sampler.h:44-101 */

void orc_sample_concentric_disk(const float u[2], float out[2])
{
    const float vx = 2.0f * u[0] - 1.0f, vy = 2.0f * u[1] - 1.0f;
    if (vx == 0.0f && vy == 0.0f) { out[0] = 0.0f; out[1] = 0.0f; return; }
    const float pi4 = (float)3.141592653589793238462643383279502884e+00 * 0.25f;
    float r, theta;
    if (vx > -vy) {
        if (vx > vy) { r = vx; theta = pi4 * vy / vx; }   /* (pi/4 * y) / x */
        else { r = vy; theta = pi4 * (2.0f - vx / vy); }
    } else {
        if (vx < vy) { r = -vx; theta = pi4 * (4.0f + vy / vx); }
        else { r = -vy; theta = pi4 * (6.0f - vx / vy); }
    }
    out[0] = r * cosf(theta);
    out[1] = r * sinf(theta);
}

void orc_sample_cosine_hemisphere(const float u[2], float out[3])
{
    float s[2];
    orc_sample_concentric_disk(u, s);
    out[0] = s[0]; out[1] = s[1];
    out[2] = sqrtf(fmaxf(0.0f, 1.0f - s[0] * s[0] - s[1] * s[1]));
}

void orc_sample_uniform_sphere(const float u[2], float out[3])
{
    const float z = 1.0f - 2.0f * u[0];
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float phi = 2.0f * (float)3.141592653589793238462643383279502884e+00 * u[1];
    out[0] = r * cosf(phi); out[1] = r * sinf(phi); out[2] = z;
}

void orc_sample_uniform_triangle(const float u[2], float out[2])
{
    const float s = sqrtf(fmaxf(0.0f, u[0]));
    out[0] = 1.0f - s;
    out[1] = u[1] * s;
}
