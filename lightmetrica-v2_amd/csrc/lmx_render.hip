// lmx_render.hip -- renderer::pt as a wavefront pipeline on gfx950.
//
// Replaces the per-sample CPU loop of the reference
// (src/liblightmetrica/renderer/renderer_pt.cpp:64-231 driven by
//  Scheduler::Process, src/liblightmetrica/scheduler.cpp:78-295).  Paths live
// in two queues (structure of arrays: ray, throughput, metadata); one
// iteration is
//   trace   -- closest hit of every queued ray (the persistent traversal
//              kernel of lmx_accel.hip, ray count read on the device),
//   shade   -- emission, Russian roulette and BSDF sampling per hit; paths
//              that continue are appended to the other queue with one
//              atomic per wave (ballot + mbcnt), so the queue stays dense
//              whatever the path lengths,
//   gen     -- new camera paths fill the free capacity, so paths of all
//              lengths share every launch,
//   commit  -- one thread swaps the queue counts.
// Two iterations (A->B, B->A) are captured once in a hipGraph and replayed;
// every count lives on the device, so nothing waits for the host between
// iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "lmx.h"
#include "lmx_internal.h"
#include "rmath.h"
#include "wbvh.h"

using namespace lmx;

namespace {

struct RMat {
    int32_t type;
    float R[3];
    float exponent;
    float Le[3];
};
static_assert(sizeof(RMat) == 32, "RMat layout");

struct RScene {
    const uint32_t* tris;
    const float* verts;
    const float* normals;      // may be null
    const uint32_t* tri_mat;
    const RMat* mats;
    f3 cam_pos, cam_vx, cam_vy, cam_vz;
    float tan_fov, aspect;
    int32_t width, height;
};

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
};

// Device-side control block.
struct Ctl {
    uint32_t n_trace;          // rays in the queue traced next
    uint32_t n_next;           // continuing paths appended by shade
    uint32_t issued;           // camera samples started
    uint32_t iterations;
    unsigned long long rays;   // extension rays traced so far
};

struct Queue {
    float4* ray;               // 2 per path: (o, tmin), (d, tmax)
    float4* thr;               // throughput rgb
    uint4* meta;               // sample, segment, pixel, flags
};

#define META_DEAD 1u

// Uniform of (sample, slot) from the table or from Philox.
struct Uniforms {
    const RParams& p;
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
        const float ca = powf(u0, 1.0f / (m.exponent + 1.0f));
        const float sa = sqrtf(fmaxf(0.0f, 1.0f - ca * ca));
        const float phi = 2.0f * LMX_PI_F * u1;
        const f3 l = mk3(sa * cosf(phi), sa * sinf(phi), ca);
        wo = add(add(mul(ru, l.x), mul(rv, l.y)), mul(r, l.z));
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
        const float pdf_sa = (m.exponent + 1.0f) * (0.5f * LMX_INV_PI_F) * powf(ca, m.exponent);
        return pdf_sa / lo.z;
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
        const float k = (m.exponent + 2.0f) * (0.5f * LMX_INV_PI_F) * powf(ca, m.exponent);
        return mk3(m.R[0] * k * corr, m.R[1] * k * corr, m.R[2] * k * corr);
    }
    return mk3(0.0f, 0.0f, 0.0f);
}

// Appends `take` lanes of this wave to a queue with one atomic per wave;
// returns each taking lane's slot.
RD uint32_t wave_append(uint32_t* counter, bool take)
{
    const uint64_t m = __ballot(take);
    if (m == 0) return 0;                      // wave-uniform
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    uint32_t base = 0;
    const uint32_t first = (uint32_t)(__ffsll((unsigned long long)m) - 1);
    if (lane == first) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, (int)first);
    return base + rank;
}

// ---------------------------------------------------------------------------
// kernels

__global__ __launch_bounds__(256) void k_shade(RScene s, RParams p, Ctl* __restrict__ ctl, Queue in,
                                               const float4* __restrict__ hits, Queue out)
{
    const uint32_t n = ctl->n_trace;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool cont = false;
        float4 o_ray0, o_ray1, o_thr;
        uint4 o_meta;
        if (i < n) {
            const uint4 meta = in.meta[i];
            const float4 h = hits[i];
            const uint32_t id = __float_as_uint(h.w);
            if (!(meta.w & META_DEAD) && id != WBVH_MISS) {
                const float4 r0 = in.ray[2 * i], r1 = in.ray[2 * i + 1];
                const float4 t4 = in.thr[i];
                f3 thr = mk3(t4.x, t4.y, t4.z);
                const f3 o = mk3(r0.x, r0.y, r0.z), d = mk3(r1.x, r1.y, r1.z);
                const uint32_t sample = meta.x, seg = meta.y, pixel = meta.z;
                Frame fr;
                surface_at(s, o, d, h.x, h.y, h.z, id, fr);
                const RMat m = s.mats[s.tri_mat[id]];
                int nv = (int)seg + 1;                  // numVertices when this segment was traced
                // hit with a light (renderer_pt.cpp:183-194)
                if (m.Le[0] > 0.0f || m.Le[1] > 0.0f || m.Le[2] > 0.0f) {
                    if (nv + 1 >= p.min_nv && to_local(fr, neg(d)).z > 0.0f) {
                        const f3 C = mk3(thr.x * m.Le[0], thr.y * m.Le[1], thr.z * m.Le[2]);
                        float* fp = p.film + 3 * (size_t)pixel;
                        atomicAdd(fp, C.x); atomicAdd(fp + 1, C.y); atomicAdd(fp + 2, C.z);
                        if (p.sample_out) {
                            float* so = p.sample_out + 8 * (size_t)sample;
                            so[2] += C.x; so[3] += C.y; so[4] += C.z; so[5] += 1.0f;
                        }
                    }
                }
                const Uniforms U{p, sample};
                bool alive = true;
                // Russian roulette (renderer_pt.cpp:207-215), from rr_depth on
                if ((int)seg + 1 >= p.rr_depth) {
                    if (U.one(LMX_PT_SLOT_SEG0 + 4u * seg + 3u) > p.rr_prob) alive = false;
                    else thr = mk3(thr.x / p.rr_prob, thr.y / p.rr_prob, thr.z / p.rr_prob);
                }
                nv++;
                if (alive && p.max_nv != -1 && nv >= p.max_nv) alive = false;
                if (alive && (m.type == LMX_MAT_DIFFUSE || m.type == LMX_MAT_PHONG)) {
                    const uint32_t seg2 = seg + 1u;
                    const float4 u = U.block(LMX_PT_SLOT_SEG0 + 4u * seg2);
                    const f3 wi = neg(d);
                    f3 wo;
                    if (bsdf_sample(m, fr, wi, u.x, u.y, wo)) {
                        const float pdf = bsdf_pdf(m, fr, wi, wo);
                        const f3 fs = bsdf_eval(m, fr, wi, wo);
                        if (!(fs.x == 0.0f && fs.y == 0.0f && fs.z == 0.0f)) {
                            thr = mk3(thr.x * (fs.x / pdf), thr.y * (fs.y / pdf), thr.z * (fs.z / pdf));
                            o_ray0 = make_float4(fr.p.x, fr.p.y, fr.p.z, LMX_EPS_ISECT);
                            o_ray1 = make_float4(wo.x, wo.y, wo.z, LMX_FLT_MAX);
                            o_thr = make_float4(thr.x, thr.y, thr.z, 0.0f);
                            o_meta = make_uint4(sample, seg2, pixel, 0u);
                            cont = true;
                        }
                    }
                }
            }
        }
        const uint32_t slot = wave_append(&ctl->n_next, cont);
        if (cont) {
            out.ray[2 * slot] = o_ray0;
            out.ray[2 * slot + 1] = o_ray1;
            out.thr[slot] = o_thr;
            out.meta[slot] = o_meta;
        }
    }
}

RD uint32_t gen_count(const Ctl* ctl, uint32_t cap, uint32_t total)
{
    const uint32_t free_slots = cap - ctl->n_next;
    const uint32_t left = total - ctl->issued;
    return free_slots < left ? free_slots : left;
}

__global__ __launch_bounds__(256) void k_gen(RScene s, RParams p, const Ctl* __restrict__ ctl, uint32_t cap, Queue out)
{
    const uint32_t want = gen_count(ctl, cap, p.total);
    const uint32_t base_slot = ctl->n_next, first = ctl->issued;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < want; t += gridDim.x * blockDim.x) {
        const uint32_t sample = first + t, slot = base_slot + t;
        const Uniforms U{p, sample};
        const float ux = U.one(LMX_PT_SLOT_RASTER), uy = U.one(LMX_PT_SLOT_RASTER + 1);
        float rx, ry;
        f3 wo;
        uint32_t pixel, flags = 0;
        float out_rx, out_ry;
        if (p.raster_mode == LMX_RASTER_STRATIFIED) {
            const uint32_t npix = (uint32_t)s.width * (uint32_t)s.height;
            pixel = sample % npix;
            const uint32_t px = pixel % (uint32_t)s.width, py = pixel / (uint32_t)s.width;
            rx = ((float)px + ux) / (float)s.width;
            ry = ((float)py + uy) / (float)s.height;
            wo = camera_dir(s, rx, ry);
            out_rx = ((float)px + 0.5f) / (float)s.width;
            out_ry = ((float)py + 0.5f) / (float)s.height;
        } else {
            wo = camera_dir(s, ux, uy);
            if (!camera_raster(s, wo, rx, ry)) { flags = META_DEAD; rx = ry = 0.0f; }
            pixel = raster_pixel(s, rx, ry);
            out_rx = rx; out_ry = ry;
        }
        if (flags & META_DEAD) { out_rx = 0.0f; out_ry = 0.0f; }
        out.ray[2 * slot] = make_float4(s.cam_pos.x, s.cam_pos.y, s.cam_pos.z, LMX_EPS_ISECT);
        out.ray[2 * slot + 1] = make_float4(wo.x, wo.y, wo.z, (flags & META_DEAD) ? -1.0f : LMX_FLT_MAX);
        out.thr[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        out.meta[slot] = make_uint4(sample, 0u, pixel, flags);
        if (p.sample_out) {
            float4* so = reinterpret_cast<float4*>(p.sample_out + 8 * (size_t)sample);
            so[0] = make_float4(out_rx, out_ry, 0.0f, 0.0f);
            so[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
}

__global__ void k_commit(Ctl* ctl, uint32_t cap, uint32_t total)
{
    if (threadIdx.x != 0) return;
    const uint32_t want = gen_count(ctl, cap, total);
    ctl->rays += ctl->n_trace;
    ctl->n_trace = ctl->n_next + want;
    ctl->n_next = 0;
    ctl->issued += want;
    ctl->iterations += 1;
}

__global__ void k_init_ctl(Ctl* ctl)
{
    if (threadIdx.x == 0) {
        ctl->n_trace = 0; ctl->n_next = 0; ctl->issued = 0; ctl->iterations = 0; ctl->rays = 0;
    }
}

__global__ void k_zero(float* a, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0.0f;
}

__global__ void k_scale(float* a, size_t n, float w)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] *= w;
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t count)
    {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = count;
        return hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T));
    }
};

}  // namespace

struct lmx_scene {
    lmx_accel* accel = nullptr;
    DevBuf<uint32_t> tris, tri_mat;
    DevBuf<float> verts, normals;
    DevBuf<RMat> mats;
    RScene rs{};
    int width = 0, height = 0;
    int cus = 0;
    // wavefront buffers, reused across renders of the same capacity
    uint32_t cap = 0;
    DevBuf<float4> ray[2], thr[2];
    DevBuf<uint4> meta[2];
    DevBuf<float4> hits;
    DevBuf<Ctl> ctl;
    Ctl* h_ctl = nullptr;      // pinned
    uint32_t tuned_cap = 0;
    ~lmx_scene()
    {
        if (h_ctl) (void)hipHostFree(h_ctl);
        lmx_accel_free(accel);
    }
};

namespace {

int ensure_queues(lmx_scene* sc, uint32_t cap)
{
    if (sc->cap == cap) return LMX_OK;
    hipError_t e;
    for (int q = 0; q < 2; q++) {
        if ((e = sc->ray[q].alloc(2 * (size_t)cap)) != hipSuccess) return hip_error(e, "queue alloc");
        if ((e = sc->thr[q].alloc(cap)) != hipSuccess) return hip_error(e, "queue alloc");
        if ((e = sc->meta[q].alloc(cap)) != hipSuccess) return hip_error(e, "queue alloc");
    }
    if ((e = sc->hits.alloc(cap)) != hipSuccess) return hip_error(e, "hit buffer alloc");
    sc->cap = cap;
    return LMX_OK;
}

// Records one wavefront iteration tracing queue `a` and filling queue `b`.
int enqueue_iteration(lmx_scene* sc, const RParams& rp, int a, int b, hipStream_t st)
{
    Queue qa{sc->ray[a].p, sc->thr[a].p, sc->meta[a].p};
    Queue qb{sc->ray[b].p, sc->thr[b].p, sc->meta[b].p};
    int rc = launch_trace(sc->accel, reinterpret_cast<const WRay*>(qa.ray), 0, &sc->ctl.p->n_trace, sc->cap,
                          reinterpret_cast<WHit*>(sc->hits.p), st);
    if (rc != LMX_OK) return rc;
    const int grid = std::max(1, std::min((int)((sc->cap + 255) / 256), sc->cus * 8));
    k_shade<<<grid, 256, 0, st>>>(sc->rs, rp, sc->ctl.p, qa, sc->hits.p, qb);
    k_gen<<<grid, 256, 0, st>>>(sc->rs, rp, sc->ctl.p, sc->cap, qb);
    k_commit<<<1, 64, 0, st>>>(sc->ctl.p, sc->cap, rp.total);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? LMX_OK : hip_error(e, "wavefront launch");
}

int render_with_capacity(lmx_scene* sc, const RParams& rp, uint32_t cap, hipStream_t st, uint64_t* rays,
                         uint32_t* iters)
{
    int rc = ensure_queues(sc, cap);
    if (rc != LMX_OK) return rc;
    k_init_ctl<<<1, 64, 0, st>>>(sc->ctl.p);
    // the first generation: queue 0 gets camera paths, then the counts commit
    k_gen<<<std::max(1, std::min((int)((cap + 255) / 256), sc->cus * 8)), 256, 0, st>>>(
        sc->rs, rp, sc->ctl.p, cap, Queue{sc->ray[0].p, sc->thr[0].p, sc->meta[0].p});
    k_commit<<<1, 64, 0, st>>>(sc->ctl.p, cap, rp.total);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "wavefront start");
    // capture two iterations (0 -> 1, 1 -> 0) in a graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hip_error(e, "graph capture begin");
    rc = enqueue_iteration(sc, rp, 0, 1, st);
    if (rc == LMX_OK) rc = enqueue_iteration(sc, rp, 1, 0, st);
    e = hipStreamEndCapture(st, &graph);
    if (rc != LMX_OK) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return hip_error(e, "graph capture end");
    if ((e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return hip_error(e, "graph instantiate");
    }
    // replay until every sample is issued and every queue is empty; the
    // control block is read back every `batch` replays
    const int batch = 4;
    rc = LMX_OK;
    for (int guard = 0; guard < 1000000; guard++) {
        for (int k = 0; k < batch; k++) {
            if ((e = hipGraphLaunch(exec, st)) != hipSuccess) { rc = hip_error(e, "graph launch"); break; }
        }
        if (rc != LMX_OK) break;
        if ((e = hipMemcpyAsync(sc->h_ctl, sc->ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess) {
            rc = hip_error(e, "control readback");
            break;
        }
        if (sc->h_ctl->n_trace == 0 && sc->h_ctl->issued >= rp.total) break;
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (rays) *rays = sc->h_ctl->rays;
    if (iters) *iters = sc->h_ctl->iterations;
    return rc;
}

// Queue-capacity autotuner: renders the same prefix of samples with each
// candidate capacity (paths in flight) and keeps the fastest in samples per
// second.  More paths in flight fill the GPU better but cost memory and make
// the tail (the last iterations with few long paths) longer; the best point
// depends on the scene's path-length distribution.  The result is kept per
// scene.
int autotune_capacity(lmx_scene* sc, const RParams& rp, hipStream_t st)
{
    static const uint32_t cands[] = {1u << 20, 1u << 21, 1u << 22, 1u << 23};
    RParams probe = rp;
    probe.sample_out = nullptr;
    probe.total = (uint32_t)std::min<uint64_t>(rp.total, 1ull << 24);
    uint32_t best = cands[1];
    double best_rate = -1.0;
    hipEvent_t e0, e1;
    hipError_t e;
    if ((e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess)
        return hip_error(e, "autotune events");
    int rc = LMX_OK;
    for (uint32_t cap : cands) {
        if (cap > 4 * (uint64_t)probe.total && cap != cands[0]) break;
        if ((rc = render_with_capacity(sc, probe, cap, st, nullptr, nullptr)) != LMX_OK) break;   // warm
        (void)hipEventRecord(e0, st);
        if ((rc = render_with_capacity(sc, probe, cap, st, nullptr, nullptr)) != LMX_OK) break;
        (void)hipEventRecord(e1, st);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double rate = probe.total / (double)std::max(ms, 1e-3f);
        if (rate > best_rate) { best_rate = rate; best = cap; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == LMX_OK) sc->tuned_cap = best;
    return rc;
}

}  // namespace

extern "C" {

int lmx_scene_create(const lmx_scene_desc* d, lmx_scene** out)
{
    if (!out) return set_error(LMX_EINVAL, "lmx_scene_create: out is NULL");
    *out = nullptr;
    if (!d || !d->verts || !d->tris || !d->tri_mat || !d->mats)
        return set_error(LMX_EINVAL, "lmx_scene_create: missing scene arrays");
    if (d->camera.width <= 0 || d->camera.height <= 0)
        return set_error(LMX_EINVAL, "lmx_scene_create: bad film size");
    for (uint32_t i = 0; i < d->ntri; i++)
        if (d->tri_mat[i] >= d->nmat) return set_error(LMX_EINVAL, "lmx_scene_create: material index out of range");
    for (uint32_t l = 0; l < d->nlight; l++)
        if ((uint64_t)d->light_prims[2 * l] + d->light_prims[2 * l + 1] > d->ntri)
            return set_error(LMX_EINVAL, "lmx_scene_create: light primitive out of range");
    std::unique_ptr<lmx_scene> sc(new (std::nothrow) lmx_scene());
    if (!sc) return set_error(LMX_ENOMEM, "lmx_scene_create: out of host memory");
    lmx_build_opts bo = {LMX_ISECT_WATERTIGHT, 0};
    int rc = lmx_accel_build(d->verts, d->nvert, d->tris, d->ntri, &bo, &sc->accel);
    if (rc != LMX_OK) return rc;
    if ((rc = lmx_accel_upload(sc->accel)) != LMX_OK) return rc;
    hipError_t e;
    auto up = [&](auto& buf, const auto* src, size_t count) -> hipError_t {
        hipError_t r = buf.alloc(count);
        if (r != hipSuccess) return r;
        return hipMemcpy(buf.p, src, count * sizeof(*src), hipMemcpyHostToDevice);
    };
    std::vector<RMat> mats(d->nmat);
    for (uint32_t i = 0; i < d->nmat; i++) std::memcpy(&mats[i], &d->mats[i], sizeof(RMat));
    if ((e = up(sc->tris, d->tris, 3 * (size_t)d->ntri)) != hipSuccess ||
        (e = up(sc->verts, d->verts, 3 * (size_t)d->nvert)) != hipSuccess ||
        (e = up(sc->tri_mat, d->tri_mat, (size_t)d->ntri)) != hipSuccess ||
        (e = up(sc->mats, mats.data(), mats.size())) != hipSuccess)
        return hip_error(e, "lmx_scene_create: upload");
    if (d->normals && (e = up(sc->normals, d->normals, 3 * (size_t)d->nvert)) != hipSuccess)
        return hip_error(e, "lmx_scene_create: upload normals");
    if ((e = sc->ctl.alloc(1)) != hipSuccess) return hip_error(e, "lmx_scene_create: control block");
    if ((e = hipHostMalloc((void**)&sc->h_ctl, sizeof(Ctl), hipHostMallocDefault)) != hipSuccess)
        return hip_error(e, "lmx_scene_create: pinned control block");
    int dev = 0;
    (void)hipGetDevice(&dev);
    if ((e = hipDeviceGetAttribute(&sc->cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
        return hip_error(e, "lmx_scene_create: device attribute");
    RScene& rs = sc->rs;
    rs.tris = sc->tris.p; rs.verts = sc->verts.p; rs.normals = d->normals ? sc->normals.p : nullptr;
    rs.tri_mat = sc->tri_mat.p; rs.mats = sc->mats.p;
    const lmx_camera& c = d->camera;
    rs.cam_pos = mk3(c.pos[0], c.pos[1], c.pos[2]);
    rs.cam_vx = mk3(c.vx[0], c.vx[1], c.vx[2]);
    rs.cam_vy = mk3(c.vy[0], c.vy[1], c.vy[2]);
    rs.cam_vz = mk3(c.vz[0], c.vz[1], c.vz[2]);
    const float fov = c.fov_deg * LMX_PI_F / 180.0f;          // Math::Radians (math.h:1674)
    rs.tan_fov = tanf(fov * 0.5f);
    rs.aspect = (float)c.width / (float)c.height;
    rs.width = c.width; rs.height = c.height;
    sc->width = c.width; sc->height = c.height;
    *out = sc.release();
    return LMX_OK;
}

void lmx_scene_free(lmx_scene* s) { delete s; }

int lmx_render_pt(lmx_scene* sc, const lmx_render_params* p, float* film_rgb, void* stream, lmx_render_stats* stats)
{
    if (!sc || !p || !film_rgb) return set_error(LMX_EINVAL, "lmx_render_pt: NULL argument");
    if (p->num_samples == 0 || p->num_samples > 0xffffffffull)
        return set_error(LMX_EINVAL, "lmx_render_pt: num_samples must be in [1, 2^32)");
    if (p->max_num_vertices != -1 && p->max_num_vertices < 2)
        return set_error(LMX_EINVAL, "lmx_render_pt: max_num_vertices must be -1 or >= 2");
    if (p->max_num_vertices == -1 && p->uniform_table)
        return set_error(LMX_EINVAL, "lmx_render_pt: a uniform table needs a bounded max_num_vertices");
    if (p->uniform_table && p->table_width < (uint32_t)(LMX_PT_SLOT_SEG0 + 4 * (p->max_num_vertices - 1)))
        return set_error(LMX_EINVAL, "lmx_render_pt: uniform table narrower than 8 + 4 (max_num_vertices - 1)");
    if (!(p->rr_prob > 0.0f && p->rr_prob <= 1.0f)) return set_error(LMX_EINVAL, "lmx_render_pt: rr_prob must be in (0, 1]");
    if (p->raster_mode != LMX_RASTER_RANDOM && p->raster_mode != LMX_RASTER_STRATIFIED)
        return set_error(LMX_EINVAL, "lmx_render_pt: unknown raster_mode");
    hipStream_t st = (hipStream_t)stream;
    hipStream_t own = nullptr;
    hipError_t e;
    if (!st) {   // graph capture needs a non-legacy stream
        if ((e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking)) != hipSuccess) return hip_error(e, "stream create");
        st = own;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const size_t npix = (size_t)sc->width * (size_t)sc->height;
    RParams rp;
    rp.max_nv = p->max_num_vertices; rp.min_nv = p->min_num_vertices; rp.rr_depth = p->rr_depth;
    rp.rr_prob = p->rr_prob;
    rp.key = make_uint2((uint32_t)p->seed, (uint32_t)(p->seed >> 32));
    rp.total = (uint32_t)p->num_samples;
    rp.raster_mode = p->raster_mode;
    rp.table = p->uniform_table; rp.table_width = p->table_width;
    rp.sample_out = p->sample_out;
    rp.film = film_rgb;
    const int zgrid = std::max(1, std::min((int)((3 * npix + 255) / 256), sc->cus * 8));
    int rc = LMX_OK;
    if (p->queue_capacity == 0 && sc->tuned_cap == 0) rc = autotune_capacity(sc, rp, st);
    uint32_t cap = p->queue_capacity ? p->queue_capacity : sc->tuned_cap;
    cap = (uint32_t)std::min<uint64_t>(cap, std::max<uint64_t>(p->num_samples, 64));
    uint64_t rays = 0;
    uint32_t iters = 0;
    if (rc == LMX_OK) {
        k_zero<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix);
        rc = render_with_capacity(sc, rp, cap, st, &rays, &iters);
    }
    if (rc == LMX_OK) {
        k_scale<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix, (float)((double)npix / (double)p->num_samples));
        if ((e = hipStreamSynchronize(st)) != hipSuccess) rc = hip_error(e, "lmx_render_pt: sync");
    }
    if (own) (void)hipStreamDestroy(own);
    if (stats) {
        stats->rays = rays;
        stats->iterations = iters;
        stats->queue_capacity = cap;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return rc;
}

}  // extern "C"
