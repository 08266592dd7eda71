// ref_harness.cpp -- exposes pieces of the reference's own header-only
// hot-path code (include/lightmetrica/{triaccel,bound,sampler}.h) and its
// vendored dSFMT behind extern "C" entry points, so tests can check the CPU
// oracle against the reference compiled from its sources where they lie.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) into
// oracle/_ref/libref.so, in the development container only; nothing here is
// part of the product.  The reference headers need <limits> before math.h
// (math.h:1663 uses std::numeric_limits without including it).
#include <limits>
#include <cstdint>
#include <cstring>
#include <lightmetrica/triaccel.h>
#include <lightmetrica/bound.h>
#include <lightmetrica/sampler.h>
#include <lightmetrica/dist.h>
extern "C" {
#include <dSFMT.h>
}

using namespace lightmetrica_v2;

extern "C" {

struct ref_triaccel_rec {
    uint32_t k;
    float f[9];
    uint32_t face, prim;
};

int ref_triaccel_load(const float* A, const float* B, const float* C, ref_triaccel_rec* out)
{
    TriAccelTriangle t;
    std::memset(&t, 0, sizeof(t));
    const int r = t.Load(Vec3(A[0], A[1], A[2]), Vec3(B[0], B[1], B[2]), Vec3(C[0], C[1], C[2]));
    out->k = t.k;
    out->f[0] = t.n_u; out->f[1] = t.n_v; out->f[2] = t.n_d;
    out->f[3] = t.a_u; out->f[4] = t.a_v; out->f[5] = t.b_nu; out->f[6] = t.b_nv;
    out->f[7] = t.c_nu; out->f[8] = t.c_nv;
    out->face = 0; out->prim = 0;
    return r;
}

int ref_triaccel_intersect(const ref_triaccel_rec* rec, const float* o, const float* d,
                           float mint, float maxt, float* u, float* v, float* t)
{
    TriAccelTriangle tri;
    tri.k = rec->k;
    tri.n_u = rec->f[0]; tri.n_v = rec->f[1]; tri.n_d = rec->f[2];
    tri.a_u = rec->f[3]; tri.a_v = rec->f[4]; tri.b_nu = rec->f[5]; tri.b_nv = rec->f[6];
    tri.c_nu = rec->f[7]; tri.c_nv = rec->f[8];
    Ray ray;
    ray.o = Vec3(o[0], o[1], o[2]);
    ray.d = Vec3(d[0], d[1], d[2]);
    Float uu = 0, vv = 0, tt = 0;
    const bool hit = tri.Intersect(ray, mint, maxt, uu, vv, tt);
    *u = uu; *v = vv; *t = tt;
    return hit ? 1 : 0;
}

int ref_bound_intersect(const float* bmin, const float* bmax, const float* o, const float* d,
                        float tmin, float tmax)
{
    Bound b;
    b.min = Vec3(bmin[0], bmin[1], bmin[2]);
    b.max = Vec3(bmax[0], bmax[1], bmax[2]);
    Ray ray;
    ray.o = Vec3(o[0], o[1], o[2]);
    ray.d = Vec3(d[0], d[1], d[2]);
    return b.Intersect(ray, tmin, tmax) ? 1 : 0;
}

void ref_sample_concentric_disk(const float* u, float* out)
{
    const auto r = Sampler::UniformConcentricDiskSample(Vec2(u[0], u[1]));
    out[0] = r.x; out[1] = r.y;
}

void ref_sample_cosine_hemisphere(const float* u, float* out)
{
    const auto r = Sampler::CosineSampleHemisphere(Vec2(u[0], u[1]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void ref_sample_uniform_sphere(const float* u, float* out)
{
    const auto r = Sampler::UniformSampleSphere(Vec2(u[0], u[1]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void ref_sample_uniform_triangle(const float* u, float* out)
{
    const auto r = Sampler::UniformSampleTriangle(Vec2(u[0], u[1]));
    out[0] = r.x; out[1] = r.y;
}

// Distribution1D::SampleReuse over a weight list (dist.h:63-82)
int ref_dist_sample_reuse(const float* w, int n, float u, float* u2)
{
    Distribution1D dist;
    for (int i = 0; i < n; i++) dist.Add(w[i]);
    dist.Normalize();
    Float uu = 0;
    const int i = dist.SampleReuse(u, uu);
    *u2 = uu;
    return i;
}

// dSFMT stream as Random::SetSeed / Next / NextUInt use it (random.cpp:65-90)
void ref_dsfmt_stream(uint32_t seed, int n, double* out_next, uint32_t* out_uint)
{
    static dsfmt_t st;
    dsfmt_init_gen_rand(&st, seed);
    for (int i = 0; i < n; i++) out_next[i] = dsfmt_genrand_close_open(&st);
    dsfmt_init_gen_rand(&st, seed);
    for (int i = 0; i < n; i++) out_uint[i] = dsfmt_genrand_uint32(&st);
}

}  // extern "C"
