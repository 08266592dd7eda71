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
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "lmx.h"
#include "lmx_internal.h"
#include "trace_cursors.h"
#include "rmath.h"
#include "rshade.h"
#include "wbvh.h"

using namespace lmx;

namespace {



// ---------------------------------------------------------------------------
// kernels

// Next-event estimation from one vertex (renderer_ptmis.cpp:130-195,
// renderer_ptdirect.cpp:109-150): a light by u[0] (uniform over the area
// lights, Scene3::SampleEmitter), a point on it by u[1..2]; contribution
// thr fs G Le / (p_light p_area), weighted in MIS mode by the balance
// heuristic against BSDF sampling of the same direction.  Camera vertex:
// fs = the sensor importance, raster position of the connection returned.
struct NeeOut {
    bool emit;
    float4 r0, r1;
    f3 C;
    float rx, ry;
};

RD NeeOut nee_connect(const RScene& s, const RParams& p, bool camera, const Frame& fr, const RMat& m, f3 wi, f3 thr,
                      float u0, float u1, float u2)
{
    NeeOut o;
    o.emit = false;
    o.rx = o.ry = 0.0f;
    if (s.nlight == 0) return o;
    int l = (int)(u0 * (float)s.nlight);
    l = l < 0 ? 0 : (l > (int)s.nlight - 1 ? (int)s.nlight - 1 : l);
    f3 lp, lgn;
    uint32_t tri;
    sample_light_point(s, l, u1, u2, lp, lgn, tri);
    const f3 pv = camera ? s.cam_pos : fr.p;
    const f3 ppL = dir_to(pv, lp);
    const RMat lm = s.mats[s.tri_mat[tri]];
    if (!(dot(lgn, neg(ppL)) > 0.0f)) return o;                      // one-sided emitter
    f3 fsE;
    float pB;
    if (camera) {
        const float W = importance(s, ppL);
        if (!(W > 0.0f) || !camera_raster(s, ppL, o.rx, o.ry)) return o;
        fsE = mk3(W, W, W);
        pB = W;
    } else {
        fsE = bsdf_eval(m, fr, wi, ppL);
        pB = bsdf_pdf(m, fr, wi, ppL);
    }
    const float G = gterm(pv, fr.sn, camera, lp, lgn, false);
    const float inv_pl = (float)s.nlight / s.lights[l].inv_area;       // 1 / (p_light p_area)
    f3 C = mk3(thr.x * fsE.x * G * lm.Le[0] * inv_pl, thr.y * fsE.y * G * lm.Le[1] * inv_pl,
               thr.z * fsE.z * G * lm.Le[2] * inv_pl);
    // a black contribution is dropped before the weight (renderer_ptmis.cpp:171),
    // which also keeps G = 0 (an infinite light density) out of it
    if (C.x == 0.0f && C.y == 0.0f && C.z == 0.0f) return o;
    if (p.nee == NEE_MIS) {
        const float pL = s.lights[l].inv_area / G * (1.0f / (float)s.nlight);
        const float w = pL / (pL + pB);
        C = mk3(C.x * w, C.y * w, C.z * w);
    }
    // Scene3::Visible (scene3.h:107-116)
    f3 d = sub(lp, pv);
    const float len = sqrtf(dot(d, d));
    d = mk3(d.x / len, d.y / len, d.z / len);
    o.r0 = make_float4(pv.x, pv.y, pv.z, LMX_EPS_ISECT);
    o.r1 = make_float4(d.x, d.y, d.z, len * (1.0f - LMX_EPS_ISECT));
    o.C = C;
    o.emit = true;
    return o;
}

RD void nee_append(const RParams& p, Ctl* ctl, uint32_t* lds, const NeeOut& o, uint32_t sample, uint32_t pixel,
                   bool camera)
{
    const uint32_t slot = block_reserve<4>(&ctl->n_shadow, o.emit ? 1u : 0u, lds);
    if (o.emit) {
        p.nq.ray[2 * slot] = o.r0;
        p.nq.ray[2 * slot + 1] = o.r1;
        p.nq.c[slot] = make_float4(o.C.x, o.C.y, o.C.z, 0.0f);
        p.nq.meta[slot] = make_uint4(sample, pixel, camera ? 1u : 0u, __float_as_uint(o.rx));
        p.nq.ry[slot] = o.ry;
    }
}

// Material-binned shading: the traced paths grouped by the class of the
// surface they hit, so that a wave of the shading kernel runs one BSDF's
// code.  One thread per queued path; per wave a ballot per class ranks the
// lanes, per workgroup one atomic per class reserves the class's range of
// perm (class c owns perm[c cap, (c + 1) cap)).
RD uint32_t shade_class(const RScene& s, uint32_t id, bool dead)
{
    if (dead || id == WBVH_MISS) return 0;
    const int32_t t = s.mats[s.tri_mat[id]].type;
    return t == LMX_MAT_DIFFUSE ? 1u : (t == LMX_MAT_PHONG ? 2u : (t == LMX_MAT_GGX ? 3u : 0u));
}

__global__ __launch_bounds__(256) void k_bin(RScene s, Ctl* __restrict__ ctl, const uint4* __restrict__ meta,
                                             const float4* __restrict__ hits, uint32_t* __restrict__ perm,
                                             uint32_t cap)
{
    __shared__ uint32_t wcount[SHADE_CLASSES][4];
    __shared__ uint32_t wbase[SHADE_CLASSES][4];
    const uint32_t n = ctl->n_trace;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        uint32_t c = SHADE_CLASSES;                  // no path
        if (i < n) c = shade_class(s, __float_as_uint(hits[i].w), (meta[i].w & META_DEAD) != 0);
        uint32_t rank = 0;
#pragma unroll
        for (uint32_t k = 0; k < SHADE_CLASSES; k++) {
            const uint64_t m = __ballot(c == k);
            if (lane == 0) wcount[k][wave] = (uint32_t)__popcll(m);
            if (c == k) rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        }
        __syncthreads();
        if (threadIdx.x < SHADE_CLASSES) {
            const uint32_t k = threadIdx.x;
            const uint32_t tot = wcount[k][0] + wcount[k][1] + wcount[k][2] + wcount[k][3];
            uint32_t b = tot ? atomicAdd(&ctl->bin[k], tot) : 0u;
            for (int w = 0; w < 4; w++) { wbase[k][w] = b; b += wcount[k][w]; }
        }
        __syncthreads();
        if (c < SHADE_CLASSES) perm[(size_t)c * cap + wbase[c][wave] + rank] = i;
        __syncthreads();
    }
}

__global__ void k_bin_clear(Ctl* ctl)
{
    if (threadIdx.x < SHADE_CLASSES) ctl->bin[threadIdx.x] = 0;
}

__global__ __launch_bounds__(256) void k_shade(RScene s, RParams p, Ctl* __restrict__ ctl, Queue in,
                                               const float4* __restrict__ hits, Queue out)
{
    __shared__ uint32_t lds[5];
    const uint32_t n = ctl->n_trace;
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t bend[SHADE_CLASSES];                    // class ranges in the binned order
    if (p.perm) {
        uint32_t acc = 0;
#pragma unroll
        for (int k = 0; k < SHADE_CLASSES; k++) { acc += ctl->bin[k]; bend[k] = acc; }
    }
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        uint32_t i = base + threadIdx.x;
        if (p.perm && i < n) {
            uint32_t k = 0, start = 0;
            while (k + 1 < SHADE_CLASSES && i >= bend[k]) { start = bend[k]; k++; }
            i = p.perm[(size_t)k * p.perm_cap + (i - start)];
        }
        bool cont = false;
        float4 o_ray0, o_ray1, o_thr;
        uint4 o_meta;
        NeeOut no;
        no.emit = false;
        uint32_t no_sample = 0, no_pixel = 0;
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
                // hit with a light (renderer_pt.cpp:183-194; renderer_ptmis.cpp:229-257
                // weighs it against light sampling; renderer::ptdirect ignores it)
                if (p.nee != NEE_DIRECT && (m.Le[0] > 0.0f || m.Le[1] > 0.0f || m.Le[2] > 0.0f)) {
                    const float cl = to_local(fr, neg(d)).z;
                    if (nv + 1 >= p.min_nv && cl > 0.0f) {
                        float w = 1.0f;
                        if (p.nee == NEE_MIS) {
                            // t4.w: the BSDF density of d at the previous vertex times its cosine there
                            const int l = s.tri_light[id];
                            const float pL = s.lights[l].inv_area * h.x * h.x / cl * (1.0f / (float)s.nlight);
                            w = t4.w / (t4.w + pL);
                        }
                        const f3 C = mk3(thr.x * m.Le[0] * w, thr.y * m.Le[1] * w, thr.z * m.Le[2] * w);
                        float* fp = p.film + 3 * (size_t)pixel;
                        atomicAdd(fp, C.x); atomicAdd(fp + 1, C.y); atomicAdd(fp + 2, C.z);
                        if (p.sample_out) {
                            float* so = p.sample_out + p.so_stride * (size_t)sample;
                            so[2] += C.x; so[3] += C.y; so[4] += C.z; so[5] += 1.0f;
                        }
                    }
                }
                const Uniforms<RParams> U{p, sample};
                bool alive = true;
                // Russian roulette (renderer_pt.cpp:207-215), from rr_depth on
                if ((int)seg + 1 >= p.rr_depth) {
                    const float q = rr_continue(p, thr);
                    if (!(q > 0.0f) || U.one(LMX_PT_SLOT_SEG0 + p.seg_stride * seg + 3u) > q) alive = false;
                    else thr = mk3(thr.x / q, thr.y / q, thr.z / q);
                }
                nv++;
                if (alive && p.max_nv != -1 && nv >= p.max_nv) alive = false;
                if (alive && is_scatterer(m)) {
                    const uint32_t seg2 = seg + 1u;
                    const f3 wi = neg(d);
                    if (p.nee != NEE_NONE && (p.nee == NEE_DIRECT || nv + 1 >= p.min_nv)) {
                        const float4 un = U.block(LMX_PT_SLOT_SEG0 + p.seg_stride * seg2 + 4u);
                        no = nee_connect(s, p, false, fr, m, wi, thr, un.x, un.y, un.z);
                        no_sample = sample;
                        no_pixel = pixel;
                    }
                    const float4 u = U.block(LMX_PT_SLOT_SEG0 + p.seg_stride * seg2);
                    f3 wo;
                    if (bsdf_sample(m, fr, wi, u.x, u.y, wo)) {
                        const float pdf = bsdf_pdf(m, fr, wi, wo);
                        const f3 fs = bsdf_eval(m, fr, wi, wo);
                        if (!(fs.x == 0.0f && fs.y == 0.0f && fs.z == 0.0f)) {
                            thr = mk3(thr.x * (fs.x / pdf), thr.y * (fs.y / pdf), thr.z * (fs.z / pdf));
                            o_ray0 = make_float4(fr.p.x, fr.p.y, fr.p.z, LMX_EPS_ISECT);
                            o_ray1 = make_float4(wo.x, wo.y, wo.z, LMX_FLT_MAX);
                            o_thr = make_float4(thr.x, thr.y, thr.z, pdf * fabsf(to_local(fr, wo).z));
                            o_meta = make_uint4(sample, seg2, pixel, 0u);
                            cont = true;
                        }
                    }
                }
            }
        }
        const uint32_t slot = block_reserve<4>(&ctl->n_next, cont ? 1u : 0u, lds);
        if (cont) {
            out.ray[2 * slot] = o_ray0;
            out.ray[2 * slot + 1] = o_ray1;
            out.thr[slot] = o_thr;
            out.meta[slot] = o_meta;
        }
        if (p.nee != NEE_NONE) nee_append(p, ctl, lds, no, no_sample, no_pixel, false);
    }
}

RD uint32_t gen_count(const Ctl* ctl, uint32_t cap, uint32_t total)
{
    if (ctl->stop) return 0;
    const uint32_t free_slots = cap - ctl->n_next;
    const uint32_t left = total - ctl->issued;
    return free_slots < left ? free_slots : left;
}

__global__ __launch_bounds__(256) void k_gen(RScene s, RParams p, Ctl* __restrict__ ctl, uint32_t cap, Queue out)
{
    __shared__ uint32_t lds[5];
    const uint32_t want = gen_count(ctl, cap, p.total);
    const uint32_t base_slot = ctl->n_next, first = ctl->issued;
    // workgroup-uniform loop: the camera-vertex connections append per workgroup
    for (uint32_t base = blockIdx.x * blockDim.x; base < want; base += gridDim.x * blockDim.x) {
        const uint32_t t = base + threadIdx.x;
        NeeOut no;
        no.emit = false;
        uint32_t sample = 0, pixel = 0;
        if (t < want) {
            sample = first + t;
            const uint32_t slot = base_slot + t;
            const Uniforms<RParams> U{p, sample};
            const float ux = U.one(LMX_PT_SLOT_RASTER), uy = U.one(LMX_PT_SLOT_RASTER + 1);
            float rx, ry;
            f3 wo;
            uint32_t flags = 0;
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
            // thr.w: the direction's density at the camera (importance), for MIS at a light hit
            out.ray[2 * slot] = make_float4(s.cam_pos.x, s.cam_pos.y, s.cam_pos.z, LMX_EPS_ISECT);
            out.ray[2 * slot + 1] = make_float4(wo.x, wo.y, wo.z, (flags & META_DEAD) ? -1.0f : LMX_FLT_MAX);
            out.thr[slot] = make_float4(1.0f, 1.0f, 1.0f, p.nee == NEE_MIS ? importance(s, wo) : 0.0f);
            out.meta[slot] = make_uint4(sample, 0u, pixel, flags);
            if (p.sample_out) {
                float4* so = reinterpret_cast<float4*>(p.sample_out + p.so_stride * (size_t)sample);
                so[0] = make_float4(out_rx, out_ry, 0.0f, 0.0f);
                so[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (p.so_stride == 16) { so[2] = so[1]; so[3] = so[1]; }
            }
            // next-event estimation at the camera vertex
            // (none for a sample whose raster position failed, renderer_ptmis.cpp:95-103)
            if (p.nee != NEE_NONE && !(flags & META_DEAD) && (p.nee == NEE_DIRECT || 2 >= p.min_nv)) {
                const float4 un = U.block(LMX_PT_SLOT_SEG0 + 4u);
                Frame fr;
                RMat m;
                no = nee_connect(s, p, true, fr, m, mk3(0, 0, 0), mk3(1.0f, 1.0f, 1.0f), un.x, un.y, un.z);
                if (no.emit) pixel = raster_pixel(s, no.rx, no.ry);
            }
        }
        if (p.nee != NEE_NONE) nee_append(p, ctl, lds, no, sample, pixel, true);
    }
}

// Splats the unoccluded next-event connections of the iteration.
__global__ __launch_bounds__(256) void k_nee_splat(RParams p, Ctl* __restrict__ ctl)
{
    const uint32_t n = ctl->n_shadow;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        if (__float_as_uint(p.nq.hits[j].w) != WBVH_MISS) continue;          // occluded
        const float4 c = p.nq.c[j];
        const uint4 m = p.nq.meta[j];
        float* fp = p.film + 3 * (size_t)m.y;
        atomicAdd(fp, c.x); atomicAdd(fp + 1, c.y); atomicAdd(fp + 2, c.z);
        if (p.sample_out) {
            float* so = p.sample_out + p.so_stride * (size_t)m.x;
            if (m.z) {
                so[8] = __uint_as_float(m.w); so[9] = p.nq.ry[j];
                so[10] += c.x; so[11] += c.y; so[12] += c.z; so[13] += 1.0f;
            } else {
                so[2] += c.x; so[3] += c.y; so[4] += c.z; so[5] += 1.0f;
            }
        }
    }
}

// Commits the generation: the queue just written becomes the one to trace.
// Also sets the next trace launch's cursors (so that launch needs no cursor
// kernel of its own), for `grid` workgroups.
__global__ void k_commit(Ctl* ctl, uint32_t cap, uint32_t total, uint32_t* cursors, uint32_t grid)
{
    __shared__ uint32_t n_next;
    if (threadIdx.x == 0) {
        const uint32_t want = gen_count(ctl, cap, total);
        ctl->shadow_rays += ctl->n_shadow;
        ctl->n_shadow = 0;
        ctl->rays += ctl->n_trace;
        ctl->n_trace = ctl->n_next + want;
        n_next = ctl->n_trace;
        ctl->n_next = 0;
        ctl->issued += want;
        ctl->iterations += 1;
    }
    __syncthreads();
    trace_cursors_set(cursors, n_next, grid, threadIdx.x);
}

__global__ void k_init_ctl(Ctl* ctl)
{
    if (threadIdx.x == 0) {
        ctl->n_trace = 0; ctl->n_next = 0; ctl->issued = 0; ctl->iterations = 0; ctl->rays = 0;
        ctl->n_shadow = 0; ctl->shadow_rays = 0; ctl->stop = 0; ctl->ray_out_n = 0; ctl->film_scaled = 0;
    }
}

// Measurement hook: copies the queue about to be traced (the first
// ray_out_cap rays of the render, in launch order) out, then advances.
__global__ void k_capture(const Ctl* __restrict__ ctl, const float4* __restrict__ ray, float4* __restrict__ out,
                          uint32_t cap)
{
    const uint32_t n = ctl->n_trace, at = ctl->ray_out_n;
    const uint32_t m = at >= cap ? 0u : min(n, cap - at);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        out[2 * (size_t)(at + i)] = ray[2 * (size_t)i];
        out[2 * (size_t)(at + i) + 1] = ray[2 * (size_t)i + 1];
    }
}

__global__ void k_capture_advance(Ctl* ctl, uint32_t cap)
{
    if (threadIdx.x == 0) ctl->ray_out_n = min(cap, ctl->ray_out_n + ctl->n_trace);
}

// Time mode: camera samples stop starting from the next iteration on.
__global__ void k_stop(Ctl* ctl)
{
    if (threadIdx.x == 0) ctl->stop = 1;
}

__global__ void k_zero(float* a, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0.0f;
}

__global__ void k_scale(float* a, size_t n, float w)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] *= w;
}

// Film::Rescale on the device, queued before each control readback of a
// render whose sample count is fixed: it scales only once the render is
// complete (every sample issued, no path left), which that readback then
// finds, so the host launches no scaling after it.
__global__ void k_scale_done(Ctl* __restrict__ ctl, float* a, size_t n, uint32_t total, float w)
{
    const bool done = ctl->n_trace == 0 && ctl->issued >= total;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->film_scaled = done ? 1u : 0u;
    if (!done) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] *= w;
}

}  // namespace


namespace {

int ensure_queues(lmx_scene* sc, uint32_t cap)
{
    if (sc->cap == cap) return LMX_OK;
    sc->cap = 0;                       // set again only once every buffer is in place
    hipError_t e;
    // a failure frees what was allocated so far
    auto fail = [&](hipError_t err, const char* what) {
        for (int q = 0; q < 2; q++) { sc->ray[q].release(); sc->thr[q].release(); sc->meta[q].release(); }
        sc->hits.release();
        return hip_error(err, what);
    };
    for (int q = 0; q < 2; q++) {
        if ((e = sc->ray[q].alloc(2 * (size_t)cap)) != hipSuccess) return fail(e, "queue alloc");
        if ((e = sc->thr[q].alloc(cap)) != hipSuccess) return fail(e, "queue alloc");
        if ((e = sc->meta[q].alloc(cap)) != hipSuccess) return fail(e, "queue alloc");
    }
    if ((e = sc->hits.alloc(cap)) != hipSuccess) return fail(e, "hit buffer alloc");
    sc->cap = cap;
    return LMX_OK;
}

int ensure_nee(lmx_scene* sc, uint32_t cap)
{
    if (sc->nee_cap >= cap) return LMX_OK;
    sc->nee_cap = 0;
    hipError_t e;
    const size_t n = 2 * (size_t)cap;
    if ((e = sc->nee_ray.alloc(2 * n)) != hipSuccess || (e = sc->nee_c.alloc(n)) != hipSuccess ||
        (e = sc->nee_hits.alloc(n)) != hipSuccess || (e = sc->nee_meta.alloc(n)) != hipSuccess ||
        (e = sc->nee_ry.alloc(n)) != hipSuccess) {
        sc->nee_ray.release(); sc->nee_c.release(); sc->nee_hits.release(); sc->nee_meta.release(); sc->nee_ry.release();
        return hip_error(e, "shadow queue alloc");
    }
    sc->nee_cap = cap;
    return LMX_OK;
}

// The next-event connections of one iteration: any-hit trace, then splat.
int enqueue_nee(lmx_scene* sc, const RParams& rp, hipStream_t st)
{
    int rc = launch_trace(sc->accel, reinterpret_cast<const WRay*>(rp.nq.ray), 0, &sc->ctl.p->n_shadow,
                          2 * sc->nee_cap, reinterpret_cast<WHit*>(rp.nq.hits), st, true);
    if (rc != LMX_OK) return rc;
    k_nee_splat<<<std::max(1, std::min((int)((2 * sc->nee_cap + 255) / 256), sc->cus * 8)), 256, 0, st>>>(rp, sc->ctl.p);
    return LMX_OK;
}

// Records one wavefront iteration tracing queue `a` and filling queue `b`.
int enqueue_iteration(lmx_scene* sc, const RParams& rp, int a, int b, hipStream_t st)
{
    Queue qa{sc->ray[a].p, sc->thr[a].p, sc->meta[a].p};
    Queue qb{sc->ray[b].p, sc->thr[b].p, sc->meta[b].p};
    if (rp.ray_out) {
        k_capture<<<std::max(1, std::min((int)((sc->cap + 255) / 256), sc->cus * 4)), 256, 0, st>>>(
            sc->ctl.p, qa.ray, rp.ray_out, rp.ray_out_cap);
        k_capture_advance<<<1, 64, 0, st>>>(sc->ctl.p, rp.ray_out_cap);
    }
    int rc = launch_trace(sc->accel, reinterpret_cast<const WRay*>(qa.ray), 0, &sc->ctl.p->n_trace, sc->cap,
                          reinterpret_cast<WHit*>(sc->hits.p), st, false, true);   // cursors set by k_commit
    if (rc != LMX_OK) return rc;
    const int grid = std::max(1, std::min((int)((sc->cap + 255) / 256), sc->cus * 8));
    if (rp.guide) {
        if ((rc = launch_guided_shade(sc, rp, qa, qb, st)) != LMX_OK) return rc;
    } else {
        if (rp.perm) {
            k_bin_clear<<<1, 64, 0, st>>>(sc->ctl.p);
            k_bin<<<grid, 256, 0, st>>>(sc->rs, sc->ctl.p, qa.meta, sc->hits.p, sc->perm.p, sc->cap);
        }
        k_shade<<<grid, 256, 0, st>>>(sc->rs, rp, sc->ctl.p, qa, sc->hits.p, qb);
    }
    k_gen<<<grid, 256, 0, st>>>(sc->rs, rp, sc->ctl.p, sc->cap, qb);
    if (rp.nee != NEE_NONE && (rc = enqueue_nee(sc, rp, st)) != LMX_OK) return rc;
    uint32_t* cursors = nullptr;
    uint32_t tgrid = 0;
    if ((rc = trace_shape(sc->accel, sc->cap, &cursors, &tgrid)) != LMX_OK) return rc;
    k_commit<<<1, 64, 0, st>>>(sc->ctl.p, sc->cap, rp.total, cursors, tgrid);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? LMX_OK : hip_error(e, "wavefront launch");
}

}  // namespace

int wavefront_render(lmx_scene* sc, const RParams& rp_in, uint32_t cap, hipStream_t st, uint64_t* rays,
                     uint32_t* iters, double render_time, uint64_t* samples, size_t film_n, float film_w)
{
    int rc = ensure_queues(sc, cap);
    if (rc != LMX_OK) return rc;
    RParams rp = rp_in;
    if (rp.perm_cap) {                       // material-binned shading requested
        if (sc->perm.n < (size_t)SHADE_CLASSES * cap) {
            const hipError_t ea = sc->perm.alloc((size_t)SHADE_CLASSES * cap);
            if (ea != hipSuccess) return hip_error(ea, "shading bins");
        }
        rp.perm = sc->perm.p;
        rp.perm_cap = cap;
    }
    if (rp.nee != NEE_NONE) {
        if ((rc = ensure_nee(sc, cap)) != LMX_OK) return rc;
        rp.nq = NeeQ{sc->nee_ray.p, sc->nee_c.p, sc->nee_meta.p, sc->nee_ry.p, sc->nee_hits.p};
    }
    k_init_ctl<<<1, 64, 0, st>>>(sc->ctl.p);
    // the first generation: queue 0 gets camera paths (and their camera-vertex
    // connections), then the counts commit
    k_gen<<<std::max(1, std::min((int)((cap + 255) / 256), sc->cus * 8)), 256, 0, st>>>(
        sc->rs, rp, sc->ctl.p, cap, Queue{sc->ray[0].p, sc->thr[0].p, sc->meta[0].p});
    if (rp.nee != NEE_NONE && (rc = enqueue_nee(sc, rp, st)) != LMX_OK) return rc;
    uint32_t* cursors = nullptr;
    uint32_t tgrid = 0;
    if ((rc = trace_shape(sc->accel, cap, &cursors, &tgrid)) != LMX_OK) return rc;
    k_commit<<<1, 64, 0, st>>>(sc->ctl.p, cap, rp.total, cursors, tgrid);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "wavefront start");
    // The first two iterations (0 -> 1, 1 -> 0) run at once on the render
    // stream, and the host records the same two iterations in a graph on a
    // stream of its own while the GPU works through them, so the recording
    // costs the GPU no idle time (about 36 us per render of a 2^20-sample
    // frame).  The scene keeps the executable graph of its last render and
    // updates it in place with the new recording (same kernels, new
    // arguments), which costs a fraction of an instantiation.
    if ((rc = enqueue_iteration(sc, rp, 0, 1, st)) != LMX_OK) return rc;
    if ((rc = enqueue_iteration(sc, rp, 1, 0, st)) != LMX_OK) return rc;
    if (!sc->cap_stream && (e = hipStreamCreateWithFlags(&sc->cap_stream, hipStreamNonBlocking)) != hipSuccess) {
        sc->cap_stream = nullptr;
        return hip_error(e, "capture stream");
    }
    hipStream_t cs = sc->cap_stream;
    hipGraph_t graph = nullptr;
    if ((e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hip_error(e, "graph capture begin");
    rc = enqueue_iteration(sc, rp, 0, 1, cs);
    if (rc == LMX_OK) rc = enqueue_iteration(sc, rp, 1, 0, cs);
    e = hipStreamEndCapture(cs, &graph);
    if (rc != LMX_OK) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return hip_error(e, "graph capture end");
    if (sc->wf_exec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(sc->wf_exec, graph, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(sc->wf_exec);
            sc->wf_exec = nullptr;
        }
    }
    if (!sc->wf_exec && (e = hipGraphInstantiate(&sc->wf_exec, graph, nullptr, nullptr, 0)) != hipSuccess) {
        sc->wf_exec = nullptr;
        (void)hipGraphDestroy(graph);
        return hip_error(e, "graph instantiate");
    }
    hipGraphExec_t exec = sc->wf_exec;
    // replay until every sample is issued and every queue is empty; the
    // control block is read back every `batch` replays.  Time mode: once the
    // deadline has passed no new camera sample starts, and the replays go on
    // until the samples already started are finished.
    int batch = 3;                           // 2 eager + 3 x 2 replayed: 8 iterations before the first readback
    const auto t_start = std::chrono::steady_clock::now();
    bool stopped = false;
    rc = LMX_OK;
    for (int guard = 0; guard < 1000000; guard++) {
        for (int k = 0; k < batch; k++) {
            if ((e = hipGraphLaunch(exec, st)) != hipSuccess) { rc = hip_error(e, "graph launch"); break; }
        }
        if (rc != LMX_OK) break;
        batch = 4;
        if (film_n && render_time <= 0.0)
            k_scale_done<<<std::max(1, std::min((int)((film_n + 255) / 256), sc->cus * 8)), 256, 0, st>>>(
                sc->ctl.p, rp.film, film_n, rp.total, film_w);
        if ((e = hipMemcpyAsync(sc->h_ctl, sc->ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess) {
            rc = hip_error(e, "control readback");
            break;
        }
        if (sc->h_ctl->n_trace == 0 && (sc->h_ctl->issued >= rp.total || stopped)) break;
        if (render_time > 0.0 && !stopped &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() >= render_time) {
            k_stop<<<1, 64, 0, st>>>(sc->ctl.p);
            stopped = true;
        }
    }
    (void)hipGraphDestroy(graph);
    if (rays) *rays = sc->h_ctl->rays + sc->h_ctl->shadow_rays;
    if (iters) *iters = sc->h_ctl->iterations;
    if (samples) *samples = sc->h_ctl->issued;
    return rc;
}

namespace {

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
    probe.ray_out = nullptr;
    probe.ray_out_cap = 0;
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
        if ((rc = wavefront_render(sc, probe, cap, st, nullptr, nullptr)) != LMX_OK) break;   // warm
        (void)hipEventRecord(e0, st);
        if ((rc = wavefront_render(sc, probe, cap, st, nullptr, nullptr)) != LMX_OK) break;
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

static int render_unidir(lmx_scene* sc, const lmx_render_params* p, int nee, const char* name, float* film_rgb, void* stream,
                         lmx_render_stats* stats)
{
    if (!sc || !p || !film_rgb) return set_error(LMX_EINVAL, std::string(name) + ": NULL argument");
    if (p->num_samples == 0 || p->num_samples > 0xffffffffull)
        return set_error(LMX_EINVAL, std::string(name) + ": num_samples must be in [1, 2^32)");
    if (p->max_num_vertices != -1 && p->max_num_vertices < 2)
        return set_error(LMX_EINVAL, std::string(name) + ": max_num_vertices must be -1 or >= 2");
    if (p->max_num_vertices == -1 && p->uniform_table)
        return set_error(LMX_EINVAL, std::string(name) + ": a uniform table needs a bounded max_num_vertices");
    const uint32_t stride = nee == NEE_NONE ? 4u : 8u;
    if (p->uniform_table && p->table_width < (uint32_t)(LMX_PT_SLOT_SEG0 + stride * (p->max_num_vertices - 1)))
        return set_error(LMX_EINVAL, std::string(name) + ": uniform table narrower than 8 + " + std::to_string(stride) +
                                     " (max_num_vertices - 1)");
    if (!(p->rr_prob > 0.0f && p->rr_prob <= 1.0f)) return set_error(LMX_EINVAL, std::string(name) + ": rr_prob must be in (0, 1]");
    if (p->rr_mode != LMX_RR_FIXED && p->rr_mode != LMX_RR_THROUGHPUT)
        return set_error(LMX_EINVAL, std::string(name) + ": unknown rr_mode");
    if (p->raster_mode != LMX_RASTER_RANDOM && p->raster_mode != LMX_RASTER_STRATIFIED)
        return set_error(LMX_EINVAL, std::string(name) + ": unknown raster_mode");
    const bool timed = p->render_time > 0.0;
    if (timed && (p->uniform_table || p->sample_out))
        return set_error(LMX_EINVAL, std::string(name) + ": render_time takes neither a uniform table nor sample_out");
    if (p->shade_order != LMX_SHADE_QUEUE && p->shade_order != LMX_SHADE_BY_MATERIAL)
        return set_error(LMX_EINVAL, std::string(name) + ": unknown shade_order");
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
    rp.total = timed ? 0xffffffffu : (uint32_t)p->num_samples;
    rp.raster_mode = p->raster_mode;
    rp.table = p->uniform_table; rp.table_width = p->table_width;
    rp.sample_out = p->sample_out;
    rp.film = film_rgb;
    rp.guide = nullptr;
    rp.rr_mode = p->rr_mode;
    rp.nee = nee;
    rp.seg_stride = stride;
    rp.so_stride = nee == NEE_NONE ? 8u : 16u;
    rp.perm = nullptr;
    rp.perm_cap = p->shade_order == LMX_SHADE_BY_MATERIAL ? 1u : 0u;   // set to the capacity when the queues exist
    rp.ray_out = p->ray_out && p->ray_out_cap ? reinterpret_cast<float4*>(p->ray_out) : nullptr;
    rp.ray_out_cap = rp.ray_out ? p->ray_out_cap : 0u;
    const int zgrid = std::max(1, std::min((int)((3 * npix + 255) / 256), sc->cus * 8));
    int rc = LMX_OK;
    if (p->queue_capacity == 0 && sc->tuned_cap == 0) rc = autotune_capacity(sc, rp, st);
    uint32_t cap = p->queue_capacity ? p->queue_capacity : sc->tuned_cap;
    if (!timed) cap = (uint32_t)std::min<uint64_t>(cap, std::max<uint64_t>(p->num_samples, 64));
    uint64_t rays = 0, taken = 0;
    uint32_t iters = 0;
    if (rc == LMX_OK) {
        k_zero<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix);
        // a fixed sample count is rescaled on the device as the render ends
        rc = wavefront_render(sc, rp, cap, st, &rays, &iters, timed ? p->render_time : 0.0, &taken,
                              timed ? 0 : 3 * npix, (float)((double)npix / (double)p->num_samples));
    }
    if (rc == LMX_OK) {
        // Film::Rescale by W H / the samples taken (scheduler.cpp:281-288)
        if (!sc->h_ctl->film_scaled)
            k_scale<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix, (float)((double)npix / (double)std::max<uint64_t>(taken, 1)));
        if ((e = hipStreamSynchronize(st)) != hipSuccess) rc = hip_error(e, "render sync");
    }
    if (own) (void)hipStreamDestroy(own);
    if (stats) {
        stats->rays = rays;
        stats->iterations = iters;
        stats->queue_capacity = cap;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->samples = rc == LMX_OK ? taken : 0;
    }
    return rc;
}


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
    // light primitives: per-triangle owner and area CDFs
    // (TriangleUtils::CreateTriangleAreaDist, triangleutils.h:47-70)
    std::vector<int32_t> tri_light(d->ntri, -1);
    std::vector<RLight> lights(d->nlight);
    std::vector<float> cdf;
    for (uint32_t l = 0; l < d->nlight; l++) {
        const uint32_t first = d->light_prims[2 * l], n = d->light_prims[2 * l + 1];
        if (n == 0) return set_error(LMX_EINVAL, "lmx_scene_create: empty light primitive");
        lights[l].first = first; lights[l].count = n; lights[l].cdf_off = (uint32_t)cdf.size();
        const size_t base = cdf.size();
        cdf.push_back(0.0f);
        float sum = 0.0f;
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t* t = d->tris + 3 * (size_t)(first + i);
            tri_light[first + i] = (int32_t)l;
            const float* a = d->verts + 3 * (size_t)t[0];
            const float* b = d->verts + 3 * (size_t)t[1];
            const float* c = d->verts + 3 * (size_t)t[2];
            const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            const float cr[3] = {e1[1] * e2[2] - e2[1] * e1[2], e1[2] * e2[0] - e2[2] * e1[0], e1[0] * e2[1] - e2[0] * e1[1]};
            const float dd = (cr[0] * cr[0] + cr[1] * cr[1]) + cr[2] * cr[2];
            const float area = std::sqrt(dd) * 0.5f;
            cdf.push_back(cdf.back() + area);
            sum += area;
        }
        const float inv = 1.0f / cdf.back();
        for (size_t i = base; i < cdf.size(); i++) cdf[i] *= inv;
        lights[l].inv_area = 1.0f / sum;
    }
    if ((e = up(sc->tri_light, tri_light.data(), tri_light.size())) != hipSuccess ||
        (e = up(sc->lights, lights.data(), lights.size())) != hipSuccess ||
        (e = up(sc->light_cdf, cdf.data(), cdf.size())) != hipSuccess)
        return hip_error(e, "lmx_scene_create: upload lights");
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
    rs.tri_light = sc->tri_light.p; rs.lights = sc->lights.p; rs.light_cdf = sc->light_cdf.p;
    rs.nlight = d->nlight;
    const lmx_camera& c = d->camera;
    rs.cam_pos = mk3(c.pos[0], c.pos[1], c.pos[2]);
    rs.cam_vx = mk3(c.vx[0], c.vx[1], c.vx[2]);
    rs.cam_vy = mk3(c.vy[0], c.vy[1], c.vy[2]);
    rs.cam_vz = mk3(c.vz[0], c.vz[1], c.vz[2]);
    const float fov = c.fov_deg * LMX_PI_F / 180.0f;          // Math::Radians (math.h:1674)
    rs.tan_fov = lmx_tanf(fov * 0.5f);
    rs.aspect = (float)c.width / (float)c.height;
    rs.width = c.width; rs.height = c.height;
    sc->width = c.width; sc->height = c.height;
    for (int a = 0; a < 3; a++) { sc->bounds[a] = 3.4e38f; sc->bounds[3 + a] = -3.4e38f; }
    for (uint32_t v = 0; v < d->nvert; v++)
        for (int a = 0; a < 3; a++) {
            sc->bounds[a] = std::min(sc->bounds[a], d->verts[3 * v + a]);
            sc->bounds[3 + a] = std::max(sc->bounds[3 + a], d->verts[3 * v + a]);
        }
    *out = sc.release();
    return LMX_OK;
}

void lmx_scene_free(lmx_scene* s) { delete s; }


int lmx_render_pt(lmx_scene* sc, const lmx_render_params* p, float* film_rgb, void* stream, lmx_render_stats* stats)
{
    return render_unidir(sc, p, NEE_NONE, "lmx_render_pt", film_rgb, stream, stats);
}

int lmx_render_ptdirect(lmx_scene* sc, const lmx_render_params* p, float* film_rgb, void* stream,
                        lmx_render_stats* stats)
{
    return render_unidir(sc, p, NEE_DIRECT, "lmx_render_ptdirect", film_rgb, stream, stats);
}

int lmx_render_ptmis(lmx_scene* sc, const lmx_render_params* p, float* film_rgb, void* stream, lmx_render_stats* stats)
{
    return render_unidir(sc, p, NEE_MIS, "lmx_render_ptmis", film_rgb, stream, stats);
}

}  // extern "C"

namespace {

__global__ void k_fmath(int kind, const float* __restrict__ x, const float* __restrict__ y, uint32_t n,
                        float* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = x[i];
    float r;
    switch (kind) {
    case LMX_FMATH_SIN: r = lmx_sinf(a); break;
    case LMX_FMATH_COS: r = lmx_cosf(a); break;
    case LMX_FMATH_TAN: r = lmx_tanf(a); break;
    case LMX_FMATH_EXP: r = lmx_expf(a); break;
    case LMX_FMATH_LOG: r = a > 0.0f ? lmx_logf_pos(a) : lmx_u2f(0x7fc00000u); break;
    default: r = lmx_powf(a, y[i]); break;
    }
    out[i] = r;
}

}  // namespace

extern "C" int lmx_fmath_eval(int32_t kind, const float* x, const float* y, uint32_t n, float* out, void* stream)
{
    if (kind < LMX_FMATH_SIN || kind > LMX_FMATH_POW) return set_error(LMX_EINVAL, "lmx_fmath_eval: unknown kind");
    if (n == 0) return LMX_OK;
    if (!x || !out || (kind == LMX_FMATH_POW && !y)) return set_error(LMX_EINVAL, "lmx_fmath_eval: NULL array");
    k_fmath<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(kind, x, y, n, out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? LMX_OK : hip_error(e, "lmx_fmath_eval");
}
