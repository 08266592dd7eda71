// lmx_bdpt.hip -- the bidirectional path samplers as one batched wavefront
// pipeline on gfx950, in two variants:
//   vcm  -- renderer::vcm in bdpt mode (src/liblightmetrica/renderer/
//           renderer_vcm.cpp:51-341,662-713, with SubpathSampler::TraceSubpath_,
//           subpathsampler.cpp:42-203): subpaths without roulette;
//   bdpt -- renderer::bdpt (renderer_bdpt.cpp:50-1002): roulette 0.5 after
//           every subpath vertex but the endpoint (compensated by
//           1 / SelectionPDF, here a factor 2 per survived roulette in the
//           subpath throughput), plus the "direct" strategy of every eye
//           vertex: a light point sampled from it (next-event estimation,
//           s = 0 with a direct vertex), with a power-heuristic or
//           1 / #strategies weight over all strategies including those.
//           The light subpath's direct vertices (a point on the sensor) are
//           not sampled: with a pinhole, whose position is a delta, their
//           strategies carry nothing (EvaluateCst evaluates the sensor's
//           position with evalDelta and gets 0, :389-393);
//   vcm merging -- renderer::vcm in modes vcm and bdpm (renderer_vcm.cpp:
//           584-790): per pass, N photon subpaths (light subpaths of their
//           own uniform source) traced by the same gen/trace/shade kernels,
//           a hash grid over their vertices past the light, then the eye
//           batches with (mode vcm) every connection and, at every eye
//           vertex past the pinhole, a merge with every photon vertex within
//           the pass's radius.  A merge of photon vertex j with eye vertex
//           t-1 makes the path of the connection (s = j, t) between photon
//           vertex j-1 and eye vertex t-1, so it is evaluated as that
//           connection without a shadow ray: its pdf is the connection's
//           times pf_L(eye vertex) pi r^2, and every strategy's term of the
//           power-heuristic denominator (connection s' and merge at x_s')
//           folds into mu_s' = c_conn + N (pi r^2 pf_L(x_s'))^2 times the
//           connection ratio, c_conn 1 (vcm) or 0 (bdpm).  The O(1) ratio
//           chains below carry mu instead of 1; with N = 0 they are the
//           bdpt-mode chains.
//
// Per batch of B samples:
//   gen      -- the eye and light endpoints of every sample and their first
//               rays, into one queue;
//   (trace, shade, commit) x (max_vertices - 1)
//            -- subpaths of both directions advance together in one queue, so
//               paths of every length and direction share each launch; every
//               vertex is stored with its throughput and with the area pdfs
//               of being generated from either side (forward pf, reverse pr),
//               evaluated from vertex positions as the reference does;
//   connect  -- one thread per (sample, t) evaluates every s: the connection
//               term, the power-heuristic weight from the stored pdfs in O(n)
//               (ratio chain instead of the reference's O(n^2) pdf
//               re-evaluation per strategy), and appends a shadow ray;
//   trace    -- any-hit traversal of the shadow rays;
//   splat    -- unoccluded contributions to the film.
// The whole batch is captured once as a hipGraph and replayed per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lmx.h"
#include "lmx_internal.h"
#include "rmath.h"
#include "rshade.h"
#include "wbvh.h"

using namespace lmx;

namespace {

// Vertex record: 4 float4
//   [0] p.xyz,  flags (bits 0-1 kind, bit 2 degenerate, bits 3-15 light+1,
//                      bits 16-31 material id)
//   [1] sn.xyz, pf  (area pdf of generating this vertex from its own side)
//   [2] gn.xyz, pr  (area pdf of generating it from the opposite side, given
//                    the next two vertices of its own subpath)
//   [3] alpha rgb (subpath throughput up to the vertex), U
// U of vertex k is S_{k-2}, where S_j = sum over i <= j of the products
// prod_{m=i..j} (pr_m / pf_m)^2 -- the part of the power-heuristic
// denominator that lies beyond the connection's two nearest vertices on this
// side (the camera vertex is excluded: a pinhole is never hit).  It obeys
// S_j = (pr_j / pf_j)^2 (1 + S_{j-1}), so every connection's MIS weight
// takes O(1) work; U < 0 marks a subpath with a zero pf in that range.
// With vertex merging the 1 becomes mu of the vertex where that strategy's
// merge would happen: on the light side S_j = rho_j^2 (mu(x_j) + S_{j-1}),
// on the eye side S_j = rho_j^2 (mu(x_{j-1}) + S_{j-1}), mu of the camera and
// of the light endpoint being c_conn (neither can be merged at).
// With the simple weight the square becomes the indicator (ratio != 0), and
// the same recursion counts the strategies with a nonzero pdf.
// The bdpt variant also keeps, per light vertex k >= 2, V_k = prod over
// m = 1 .. k-2 of (pr_m / pf_m)^2: with it the ratio of the direct
// strategy's pdf (equal to that of s = 1) to any s is O(1) as well.
#define VK_CAMERA 1u
#define VK_LIGHT 2u
#define VK_SURFACE 3u
#define VF_DEGENERATE 4u

#define BD_VCM 0
#define BD_BDPT 1
#define BD_PM 2                // renderer::pm photon subpaths: roulette after every vertex, the endpoint included
#define BD_PPM 3               // renderer::ppm photon subpaths: roulette after every vertex but the endpoint

struct BParams {
    int32_t variant;           // BD_VCM, BD_BDPT, BD_PM or BD_PPM
    int32_t merging;           // vcm/bdpm: merge records, photon grid, merge kernel
    int32_t connect;           // connections and s = 0 strategies evaluated (all but bdpm)
    int32_t phase;             // merging: 0 eye samples, 1 photon subpaths
    float c_conn;              // mu's connection term: 1, or 0 (bdpm)
    int32_t mis;               // LMX_MIS_POWER or LMX_MIS_SIMPLE (bdpt)
    int32_t bounded;           // bdpt: max_num_vertices given (else capped at maxv)
    int32_t maxv, minv;
    uint2 key;
    uint32_t total;
    int32_t raster_mode;
    const float* table;
    uint32_t table_width;
    uint32_t wsub;             // uniform block width of one subpath
    uint32_t batch;            // samples per batch
    float* sample_out;         // (maxv+1)^2 x 8 floats per sample, or null
    float* film;
    float4* ray_out = nullptr; // measurement hook: extension rays at [0, cap/2), shadow rays at [cap/2, cap)
    uint32_t ray_out_cap = 0;
};

// Measurement hook: copies the n rays about to be traced to out[at, ...),
// within [lo, hi) (the slot of BCtl::cap_n[kind] counts them).
__global__ void k_bd_capture(BCtl* __restrict__ ctl, int kind, const float4* __restrict__ ray, float4* __restrict__ out,
                             uint32_t lo, uint32_t hi)
{
    const uint32_t n = kind == 0 ? ctl->n_trace : ctl->n_shadow;
    const uint32_t at = lo + ctl->cap_n[kind];
    const uint32_t m = at >= hi ? 0u : min(n, hi - at);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        out[2 * (size_t)(at + i)] = ray[2 * (size_t)i];
        out[2 * (size_t)(at + i) + 1] = ray[2 * (size_t)i + 1];
    }
}

__global__ void k_bd_capture_advance(BCtl* ctl, int kind, uint32_t room)
{
    if (threadIdx.x == 0) ctl->cap_n[kind] = min(room, ctl->cap_n[kind] + (kind == 0 ? ctl->n_trace : ctl->n_shadow));
}

struct VB {
    float4* v[2];              // eye, light
    uint32_t* nsub[2];
    uint32_t* epix;
    float2* eraster;
    float* V;                  // bdpt: V_k of the light vertices
    uint32_t* mq;              // merging: queue of eye vertices to merge at (sample * 32 + vertex)
    uint32_t* off;             // connection slots per sample, scanned into slot offsets (batch + 1 words)
    uint32_t* dmask;           // bdpt: eye vertices k with a direct strategy (bit k), per sample
};

struct Vtx {
    f3 p, sn, gn;
    uint32_t kind;
    bool degenerate;
    int light;
    float pf, pr;
    f3 alpha;
    float U;
    uint32_t mat;
    Frame fr;
};

#define VB_MAX_MATERIALS 65536u
#define VB_MAX_LIGHTS 8190u

RD uint32_t vflags(uint32_t kind, bool degenerate, int light, uint32_t mat)
{
    return kind | (degenerate ? VF_DEGENERATE : 0u) | ((uint32_t)(light + 1) << 3) | (mat << 16);
}

RD Vtx load_v(const VB& vb, const BParams& p, int kind, uint32_t i, uint32_t k)
{
    const float4* r = vb.v[kind] + ((size_t)i * (size_t)p.maxv + k) * 4;
    const float4 a = r[0], b = r[1], c = r[2], d = r[3];
    Vtx v;
    v.p = mk3(a.x, a.y, a.z);
    const uint32_t fl = __float_as_uint(a.w);
    v.kind = fl & 3u;
    v.degenerate = (fl & VF_DEGENERATE) != 0;
    v.light = (int)((fl >> 3) & 0x1fffu) - 1;
    v.mat = fl >> 16;
    v.sn = mk3(b.x, b.y, b.z); v.pf = b.w;
    v.gn = mk3(c.x, c.y, c.z); v.pr = c.w;
    v.alpha = mk3(d.x, d.y, d.z);
    v.U = d.w;
    v.fr.p = v.p; v.fr.sn = v.sn; v.fr.gn = v.gn;
    if (!v.degenerate) onb(v.sn, v.fr.u, v.fr.v);
    return v;
}

RD void store_v(const VB& vb, const BParams& p, int kind, uint32_t i, uint32_t k, f3 pos, uint32_t flags, f3 sn,
                float pf, f3 gn, float pr, f3 alpha, float U)
{
    float4* r = vb.v[kind] + ((size_t)i * (size_t)p.maxv + k) * 4;
    r[0] = make_float4(pos.x, pos.y, pos.z, __uint_as_float(flags));
    r[1] = make_float4(sn.x, sn.y, sn.z, pf);
    r[2] = make_float4(gn.x, gn.y, gn.z, pr);
    r[3] = make_float4(alpha.x, alpha.y, alpha.z, U);
}

RD void store_pr(const VB& vb, const BParams& p, int kind, uint32_t i, uint32_t k, float pr)
{
    float* r = reinterpret_cast<float*>(vb.v[kind] + ((size_t)i * (size_t)p.maxv + k) * 4 + 2);
    r[3] = pr;
}

RD float load_pr(const VB& vb, const BParams& p, int kind, uint32_t i, uint32_t k)
{
    return vb.v[kind][((size_t)i * (size_t)p.maxv + k) * 4 + 2].w;
}

RD float gterm(const Vtx& a, const Vtx& b) { return gterm(a.p, a.sn, a.degenerate, b.p, b.sn, b.degenerate); }

// the weight heuristic's transform of one pdf ratio: its square (power
// heuristic) or whether it is nonzero (simple weight)
RD float sqm(float x, int mis) { return mis == LMX_MIS_SIMPLE ? (x != 0.0f ? 1.0f : 0.0f) : x * x; }

// mu of a path vertex: c_conn, plus N (pi r^2 pf_L)^2 where a merge can
// happen there
RD float mu_of(float c_conn, float c_merge, bool mergeable, float pfL)
{
    return (mergeable && c_merge != 0.0f) ? c_conn + c_merge * (pfL * pfL) : c_conn;
}

// a^2 (m1 + b^2 (m2 + U)): the sum over one side of the squared pdf ratios
// of the strategies that move the junction 1, 2, 3, ... vertices into it,
// each times mu of its merge vertex (m1, m2 for the first two, the rest
// inside U); a zero factor ends the chain (the later ratios contain it).
RD float chain_mu(float a, float b, float m1, float m2, float U, int mis)
{
    if (a == 0.0f) return 0.0f;
    const float a2 = sqm(a, mis);
    // a strategy that does not exist (mu 0: no merge there and no connection
    // term) adds 0 even when its ratio overflowed to infinity
    const float rest = (b == 0.0f || m2 + U == 0.0f) ? 0.0f : sqm(b, mis) * (m2 + U);
    const float inner = m1 + rest;
    return inner == 0.0f ? 0.0f : a2 * inner;
}

// uniform slots of the bdpt variant (within a subpath block): step j >= 1
// owns 8 + 8 (j - 1) .. +7
RD uint32_t rb_step(uint32_t j) { return LMX_PT_SLOT_SEG0 + 8u * (j - 1u); }

// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_bd_gen(RScene s, BParams p, BCtl* __restrict__ ctl, VB vb, Queue q)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t sample = ctl->batch_first + i;
    const bool valid = i < p.batch && sample - ctl->pass_first < p.total;
    const bool eye = p.phase == 0;             // the photon phase traces light subpaths only
    const bool lsub = p.connect != 0 || p.phase == 1;   // bdpm's eye samples use no light subpath
    bool eye_ray = false, light_ray = false;
    f3 ewo = mk3(0, 0, 0), lwo = mk3(0, 0, 0), lp = mk3(0, 0, 0), la = mk3(0, 0, 0);
    float epdf = 0.0f, lpdf = 0.0f;
    if (valid) {
        const Uniforms<BParams> U{p, sample};
        // eye endpoint (pinhole; We = 1, pdf 1)
        if (eye) {
            const float ux = U.one(4), uy = U.one(5);
            uint32_t pix;
            float rx, ry;
            bool alive = true;
            if (p.raster_mode == LMX_RASTER_STRATIFIED) {
                const uint32_t npix = (uint32_t)s.width * (uint32_t)s.height;
                pix = sample % npix;
                const uint32_t px = pix % (uint32_t)s.width, py = pix / (uint32_t)s.width;
                ewo = camera_dir(s, ((float)px + ux) / (float)s.width, ((float)py + uy) / (float)s.height);
                rx = ((float)px + 0.5f) / (float)s.width;
                ry = ((float)py + 0.5f) / (float)s.height;
            } else {
                ewo = camera_dir(s, ux, uy);
                alive = camera_raster(s, ewo, rx, ry);
                if (!alive) { rx = ry = 0.0f; }
                pix = raster_pixel(s, rx, ry);
            }
            vb.epix[i] = pix;
            vb.eraster[i] = make_float2(rx, ry);
            store_v(vb, p, 0, i, 0, s.cam_pos, vflags(VK_CAMERA, true, -1, 0u), mk3(0, 0, 0), 1.0f, mk3(0, 0, 0), 0.0f,
                    mk3(1, 1, 1), 0.0f);
            if (alive) {
                epdf = importance(s, ewo);
                eye_ray = epdf > 0.0f;
            }
            vb.nsub[0][i] = 1;
        }
        // light endpoint: primitive from uEP_C, point from uEP, direction from uED
        if (s.nlight > 0 && lsub) {
            const uint32_t b = p.wsub;
            const float e0 = U.one(b + 1), e1 = U.one(b + 2), ec = U.one(b + 3);
            const float d0 = U.one(b + 4), d1 = U.one(b + 5);
            int l = (int)(ec * (float)s.nlight);
            l = l < 0 ? 0 : (l > (int)s.nlight - 1 ? (int)s.nlight - 1 : l);
            f3 gn;
            uint32_t tri;
            sample_light_point(s, l, e0, e1, lp, gn, tri);
            const float ppdf = pos_pdf_light(s, l);
            const float a0 = 1.0f / ppdf;
            const uint32_t mat = s.tri_mat[tri];
            store_v(vb, p, 1, i, 0, lp, vflags(VK_LIGHT, false, l, mat), gn, ppdf, gn, 0.0f, mk3(a0, a0, a0), 0.0f);
            vb.nsub[1][i] = 1;
            Frame fr;
            fr.p = lp; fr.gn = gn; fr.sn = gn;
            onb(gn, fr.u, fr.v);
            lwo = to_world(fr, cosine_hemisphere(d0, d1));
            if (emits_towards(fr, lwo)) {
                const RMat m = s.mats[mat];
                lpdf = LMX_INV_PI_F;
                la = mk3(a0 * (m.Le[0] / lpdf), a0 * (m.Le[1] / lpdf), a0 * (m.Le[2] / lpdf));
                light_ray = !(m.Le[0] == 0.0f && m.Le[1] == 0.0f && m.Le[2] == 0.0f);
            }
            // renderer::pm: roulette after the endpoint (renderer_pm.cpp:111-120),
            // the survivor's throughput doubled
            if (p.variant == BD_PM && light_ray) {
                if (U.one(b + 6) > 0.5f) light_ray = false;
                else la = mk3(la.x * 2.0f, la.y * 2.0f, la.z * 2.0f);
            }
        } else {
            vb.nsub[1][i] = 0;
        }
    } else if (i < p.batch) {
        if (eye) vb.nsub[0][i] = 0;
        vb.nsub[1][i] = 0;
    }
    __shared__ uint32_t lds[5];
    const uint32_t se = block_reserve<4>(&ctl->n_trace, (uint32_t)eye_ray + (uint32_t)light_ray, lds);
    const uint32_t sl = se + (uint32_t)eye_ray;
    if (eye_ray) {
        q.ray[2 * se] = make_float4(s.cam_pos.x, s.cam_pos.y, s.cam_pos.z, LMX_EPS_ISECT);
        q.ray[2 * se + 1] = make_float4(ewo.x, ewo.y, ewo.z, LMX_FLT_MAX);
        q.thr[se] = make_float4(1.0f, 1.0f, 1.0f, epdf);
        q.meta[se] = make_uint4(i, 0u, 1u, 0u);
    }
    if (light_ray) {
        q.ray[2 * sl] = make_float4(lp.x, lp.y, lp.z, LMX_EPS_ISECT);
        q.ray[2 * sl + 1] = make_float4(lwo.x, lwo.y, lwo.z, LMX_FLT_MAX);
        q.thr[sl] = make_float4(la.x, la.y, la.z, lpdf);
        q.meta[sl] = make_uint4(i, 1u, 1u, 0u);
    }
}

__global__ __launch_bounds__(256) void k_bd_shade(RScene s, BParams p, BCtl* __restrict__ ctl, Queue in,
                                                  const float4* __restrict__ hits, Queue out, VB vb)
{
    // one thread per queue entry (the grid covers the queue's capacity;
    // workgroups past the current count leave at once)
    __shared__ uint32_t lds[5];
    const uint32_t n = ctl->n_trace;
    {
        const uint32_t base = blockIdx.x * blockDim.x;
        if (base >= n) return;
        const uint32_t j = base + threadIdx.x;
        bool cont = false;
        float4 o0, o1, ot;
        uint4 om;
        if (j < n) {
            const uint4 meta = in.meta[j];
            const float4 h = hits[j];
            const uint32_t id = __float_as_uint(h.w);
            if (id != WBVH_MISS) {
                const uint32_t i = meta.x, k = meta.z;
                const int kind = (int)meta.y;
                const float4 r0 = in.ray[2 * j], r1 = in.ray[2 * j + 1], t4 = in.thr[j];
                Frame fr;
                surface_at(s, mk3(r0.x, r0.y, r0.z), mk3(r1.x, r1.y, r1.z), h.x, h.y, h.z, id, fr);
                const uint32_t mat = s.tri_mat[id];
                const int light = s.tri_light[id];
                const Vtx prev = load_v(vb, p, kind, i, k - 1);
                const float G = gterm(prev.p, prev.sn, prev.degenerate, fr.p, fr.sn, false);
                const f3 alpha = mk3(t4.x, t4.y, t4.z);
                // area pdfs from the vertex positions (as EvaluatePathPDF
                // evaluates them, renderer_vcm.cpp:227-253): this vertex from
                // its own side, and the vertex two back from the opposite side
                const f3 dpn = dir_to(prev.p, fr.p);
                float pf, U = 0.0f;
                if (prev.kind == VK_CAMERA) {
                    pf = importance(s, dpn) * G;
                } else if (prev.kind == VK_LIGHT) {
                    pf = emits_towards(prev.fr, dpn) ? LMX_INV_PI_F * G : 0.0f;
                } else {
                    const Vtx prev2 = load_v(vb, p, kind, i, k - 2);
                    const RMat pm = s.mats[prev.mat];
                    const f3 dpp = dir_to(prev.p, prev2.p);
                    pf = bsdf_pdf(pm, prev.fr, dpp, dpn) * G;
                    const float pr2 = bsdf_pdf(pm, prev.fr, dpn, dpp) * gterm(prev, prev2);
                    store_pr(vb, p, kind, i, k - 2, pr2);
                    // U_k = S_{k-2}; on the eye side S_0 (the camera) is excluded.
                    // mu of the strategy's merge vertex: light side x_{k-2}
                    // (the endpoint for k = 2), eye side x_{k-3} (the camera
                    // for k = 3)
                    const bool bad = prev2.pf == 0.0f || prev.U < 0.0f;
                    const float rho = bad ? 0.0f : pr2 / prev2.pf;
                    const float cm = p.merging ? ctl->c_merge : 0.0f;
                    float mu = p.c_conn;
                    if (cm != 0.0f) {
                        if (kind == 1) mu = mu_of(p.c_conn, cm, k >= 3, prev2.pf);
                        else if (k >= 4) mu = mu_of(p.c_conn, cm, true, load_pr(vb, p, 0, i, k - 3));
                    }
                    if (bad && (k > 2 || kind == 1)) U = -1.0f;
                    else if (k == 2) U = kind == 1 ? sqm(rho, p.mis) * p.c_conn : 0.0f;
                    else U = rho == 0.0f ? 0.0f : sqm(rho, p.mis) * (mu + prev.U);
                    // V_k = V_{k-1} (pr_{k-2} / pf_{k-2})^2, V_2 = 1
                    if (p.variant == BD_BDPT && kind == 1) {
                        float* Vi = vb.V + (size_t)i * (size_t)p.maxv;
                        Vi[k] = k == 2 ? 1.0f : (bad ? 0.0f : Vi[k - 1] * sqm(rho, p.mis));
                    }
                }
                store_v(vb, p, kind, i, k, fr.p, vflags(VK_SURFACE, false, light, mat), fr.sn, pf, fr.gn,
                        0.0f, alpha, U);
                vb.nsub[kind][i] = k + 1;
                const RMat m = s.mats[mat];
                const Uniforms<BParams> Un{p, ctl->batch_first + i};
                const uint32_t blk = kind ? p.wsub : 0u;
                // bdpt: roulette after the vertex (renderer_bdpt.cpp:205-211),
                // the survivors' throughput doubled (1 / SelectionPDF)
                const bool rr8 = p.variant == BD_BDPT || p.variant == BD_PM || p.variant == BD_PPM;
                const bool rr_ok = !rr8 || Un.one(blk + rb_step(k) + 3u) <= 0.5f;
                if ((int)k + 1 < p.maxv && is_scatterer(m) && rr_ok) {
                    const f3 wi = dir_to(fr.p, prev.p);
                    const float4 u = Un.block(blk + (rr8 ? rb_step(k + 1u) : LMX_PT_SLOT_SEG0 + 4u * (k - 1u)));
                    f3 wo;
                    if (bsdf_sample(m, fr, wi, u.x, u.y, wo)) {
                        const f3 fs = bsdf_eval_t(m, fr, wi, wo, kind == 1);
                        if (!(fs.x == 0.0f && fs.y == 0.0f && fs.z == 0.0f)) {
                            const float pdf = bsdf_pdf(m, fr, wi, wo);
                            const float sel = rr8 ? 2.0f : 1.0f;
                            const f3 a2 = mk3(alpha.x * (fs.x / pdf) * sel, alpha.y * (fs.y / pdf) * sel,
                                              alpha.z * (fs.z / pdf) * sel);
                            o0 = make_float4(fr.p.x, fr.p.y, fr.p.z, LMX_EPS_ISECT);
                            o1 = make_float4(wo.x, wo.y, wo.z, LMX_FLT_MAX);
                            ot = make_float4(a2.x, a2.y, a2.z, pdf);
                            om = make_uint4(i, (uint32_t)kind, k + 1u, 0u);
                            cont = true;
                        }
                    }
                }
            }
        }
        const uint32_t slot = block_reserve<4>(&ctl->n_next, cont ? 1u : 0u, lds);
        if (cont) {
            out.ray[2 * slot] = o0;
            out.ray[2 * slot + 1] = o1;
            out.thr[slot] = ot;
            out.meta[slot] = om;
        }
    }
}

__global__ void k_bd_commit(BCtl* ctl)
{
    if (threadIdx.x == 0) {
        ctl->rays += ctl->n_trace;
        ctl->n_trace = ctl->n_next;
        ctl->n_next = 0;
    }
}

struct ShadowQ {
    float4* ray;               // 2 per connection
    float4* c;                 // weighted contribution rgb, MIS weight
    float4* aux;               // raster x, y
    uint4* meta;               // sample, s, t, pixel
};

// sample_out record of strategy (s, t, direct or merge): vcm (M+1)^2
// records at t (M+1) + s; bdpt and vcm merging 2 (M+1)^2 at
// 2 (t (M+1) + s) + direct / merge
RD float4* record_at(const BParams& p, uint32_t sample, int s, int t, int flag)
{
    const size_t M1 = (size_t)(p.maxv + 1);
    const bool two = p.variant == BD_BDPT || p.merging;
    const size_t idx = two ? 2 * ((size_t)t * M1 + (size_t)s) + (size_t)flag : (size_t)t * M1 + (size_t)s;
    const size_t per = two ? 2 * M1 * M1 : M1 * M1;
    return reinterpret_cast<float4*>(p.sample_out + 8 * ((size_t)sample * per + idx));
}

RD void record(const BParams& p, uint32_t sample, int s, int t, int direct, float rx, float ry, f3 C, float w)
{
    if (!p.sample_out) return;
    float4* o = record_at(p, sample, s, t, direct);
    o[0] = make_float4(rx, ry, C.x, C.y);
    o[1] = make_float4(C.z, 1.0f, w, 0.0f);
}

// merges of one (sample, t, s) add up: C, count and weight sums
RD void record_merge(const BParams& p, uint32_t sample, int s, int t, float rx, float ry, f3 C, float w)
{
    if (!p.sample_out) return;
    float4* o = record_at(p, sample, s, t, 1);
    const float4 a = o[0], b = o[1];
    o[0] = make_float4(rx, ry, a.z + C.x, a.w + C.y);
    o[1] = make_float4(b.x + C.z, b.y + 1.0f, b.z + w, 0.0f);
}

RD void splat(const BParams& p, uint32_t pixel, f3 C)
{
    float* fp = p.film + 3 * (size_t)pixel;
    atomicAdd(fp, C.x); atomicAdd(fp + 1, C.y); atomicAdd(fp + 2, C.z);
}

// the smallest path length a strategy may have (renderer_bdpt.cpp:887 applies
// min_num_vertices only when max_num_vertices is given)
RD int min_len(const BParams& p) { return (p.variant == BD_BDPT && !p.bounded) ? 0 : p.minv; }

// bdpt: does eye vertex k sample a direct (light) vertex?  Step k + 1 runs
// (k + 1 < maxv, and the roulette after vertex k >= 1 passed), and the
// vertex scatters (the camera's importance for k = 0).
RD bool eye_direct(const RScene& s, const BParams& p, const VB& vb, uint32_t i, uint32_t sample, int k)
{
    if (s.nlight == 0 || k + 1 >= p.maxv || k + 2 < min_len(p)) return false;
    if (k == 0) return true;
    const Uniforms<BParams> Un{p, sample};
    if (!(Un.one(rb_step((uint32_t)k) + 3u) <= 0.5f)) return false;
    const float4 a = vb.v[0][((size_t)i * (size_t)p.maxv + (size_t)k) * 4];
    return is_scatterer(s.mats[__float_as_uint(a.w) >> 16]);
}

// One junction of light vertex vL (index sv - 1 of its subpath, vLp the one
// before) with eye vertex vE (index t - 1, vEp the one before): the
// unweighted contribution alpha_L f_L G f_E alpha_E, the area pdf pLs of vE
// from the light side, and the power-heuristic denominator over every
// strategy of the path relative to this connection, the two nearest vertices
// of each side from the junction pdfs and the rest from the running sums U.
// With merging (cm = N (pi r^2)^2 != 0) every term carries mu of its merge
// vertex; prE3 is the stored pr of eye vertex t - 3 (its pdf from the light
// side), used when t >= 4.
struct Junction {
    f3 Cs;
    float pLs;
    float aL, bL;      // the light side's ratios at vertices sv - 1, sv - 2
    float aE, bE;      // the eye side's ratios at vertices t - 2, t - 3 (t >= 2)
    float pLsp1;       // pdf of eye vertex t - 2 from the junction (t >= 3)
    float pEsm1, pEsm2;  // pdfs of light vertices sv - 1, sv - 2 from the eye side
    float sum;
    bool ok;           // every pdf the sum divides by is nonzero
};

RD Junction junction(const RScene& s, const BParams& p, float cm, const Vtx& vL, const Vtx& vLp, int sv,
                     const Vtx& vE, const Vtx& vEp, int t, float prE3)
{
    Junction J;
    const f3 dLE = dir_to(vL.p, vE.p);
    const f3 dEL = dir_to(vE.p, vL.p);
    const float G = gterm(vL, vE);
    f3 fsL = mk3(0, 0, 0), fsE = mk3(0, 0, 0);
    float pLs = 0.0f, pEsm1 = 0.0f, pEsm2 = 0.0f, pLsp1 = 0.0f;
    // light side of the connection
    if (sv == 1) {
        if (emits_towards(vL.fr, dLE)) {
            const RMat m = s.mats[vL.mat];
            fsL = mk3(m.Le[0], m.Le[1], m.Le[2]);
            pLs = LMX_INV_PI_F * G;
        }
    } else {
        const RMat m = s.mats[vL.mat];
        if (is_scatterer(m)) {
            const f3 wiL = dir_to(vL.p, vLp.p);
            fsL = bsdf_eval_t(m, vL.fr, wiL, dLE, true);
            pLs = bsdf_pdf(m, vL.fr, wiL, dLE) * G;
            pEsm2 = bsdf_pdf(m, vL.fr, dLE, wiL) * gterm(vL, vLp);
        }
    }
    // eye side
    if (t == 1) {
        const float imp = importance(s, dEL);
        fsE = mk3(imp, imp, imp);
        pEsm1 = imp * G;
    } else {
        const RMat m = s.mats[vE.mat];
        if (is_scatterer(m)) {
            const f3 wiE = dir_to(vE.p, vEp.p);
            fsE = bsdf_eval_t(m, vE.fr, wiE, dEL, false);
            pEsm1 = bsdf_pdf(m, vE.fr, wiE, dEL) * G;
            pLsp1 = bsdf_pdf(m, vE.fr, dEL, wiE) * gterm(vE, vEp);
        }
    }
    J.Cs = mk3(vL.alpha.x * fsL.x * G * fsE.x * vE.alpha.x, vL.alpha.y * fsL.y * G * fsE.y * vE.alpha.y,
               vL.alpha.z * fsL.z * G * fsE.z * vE.alpha.z);
    J.pLs = pLs;
    J.ok = true;
    if (vL.pf == 0.0f || (sv >= 2 && (vLp.pf == 0.0f || vL.U < 0.0f))) J.ok = false;
    if (t >= 2 && (vE.pf == 0.0f || (t >= 3 && (vEp.pf == 0.0f || vE.U < 0.0f)))) J.ok = false;
    J.aL = J.bL = J.aE = J.bE = 0.0f;
    J.pLsp1 = pLsp1;
    J.pEsm1 = pEsm1;
    J.pEsm2 = pEsm2;
    J.sum = p.c_conn;
    if (J.ok) {
        J.aL = pEsm1 / vL.pf;
        J.bL = sv >= 2 ? pEsm2 / vLp.pf : 0.0f;
        if (t >= 2) {
            J.aE = pLs / vE.pf;
            J.bE = t >= 3 ? pLsp1 / vEp.pf : 0.0f;
        }
        // this strategy's own term: mu of vE (no merge at the pinhole)
        J.sum = mu_of(p.c_conn, cm, t >= 2, pLs);
        // light side: mu of light vertices sv - 1, sv - 2 (the endpoint is no merge)
        J.sum += chain_mu(J.aL, J.bL, mu_of(p.c_conn, cm, sv >= 2, vL.pf),
                          mu_of(p.c_conn, cm, sv >= 3, sv >= 3 ? vLp.pf : 0.0f), vL.U, p.mis);
        // eye side: mu of eye vertices t - 2, t - 3 (the camera is no merge)
        if (t >= 2)
            J.sum += chain_mu(J.aE, J.bE, mu_of(p.c_conn, cm, t >= 3, pLsp1),
                              mu_of(p.c_conn, cm, t >= 4, prE3), vE.U, p.mis);
    }
    return J;
}

// A float with a separate exponent, x = m 2^e (m from frexpf, in [0.5, 1) or
// 0, inf, nan): products and quotients of pdfs whose factors would overflow
// or underflow FP32 on their own, while the result is in range.
struct XF {
    float m;
    int e;
};
RD XF xf(float v) { XF r; r.m = frexpf(v, &r.e); return r; }
RD XF xmul(XF a, XF b) { XF r = xf(a.m * b.m); r.e += a.e + b.e; return r; }
RD XF xdiv(XF a, XF b) { XF r = xf(a.m / b.m); r.e += a.e - b.e; return r; }
// x^2 as an FP32 value: +inf above FLT_MAX, 0 below the denormals
RD float xsq(XF a)
{
    const int e = 2 * a.e;
    return e > 256 ? INFINITY : (e < -320 ? 0.0f : ldexpf(a.m * a.m, e));
}

// The weight denominator of a merge with the connection terms off (mode
// bdpm, c_conn = 0), divided by the merge's own term N (pi r^2 pLs)^2: every
// mu = N (pi r^2 pf)^2 becomes (pf / pLs)^2 and the stored chains U are
// divided by N (pi r^2 pLs)^2 as well.  The power-heuristic terms are written
// out as products of the pdfs they consist of, so that factors that cancel
// (the light side's pdfs against pLs, which carry the same glossy lobe; pLs
// against the eye side's ratio a_E = pLs / pf) never pass through FP32 on
// their own:
//   light side  (pEsm1 / pLs)^2 [sv >= 2],
//               (pEsm1 pEsm2 / (pf_{sv-1} pLs))^2 [sv >= 3],
//               (pEsm1 pEsm2 / (pf_{sv-1} pf_{sv-2} pLs))^2 U_L / N(pi r^2)^2;
//   eye side    (pLsp1 / pf_{t-1})^2 [t >= 3],
//               (pLsp1 prE3 / (pf_{t-1} pf_{t-2}))^2 [t >= 4],
//               (pLsp1 / (pf_{t-1} pf_{t-2}))^2 U_E / N(pi r^2)^2
// (a_L a_L m1 and so on of the chains, the zero-factor ends included).  The
// sum is >= 1 and overflows only where the merge's weight truly vanishes.
RD float merge_den_bdpm(const Junction& J, float cm, const Vtx& vL, const Vtx& vLp, int sv, const Vtx& vE,
                        const Vtx& vEp, int t, float prE3)
{
    const XF pLs = xf(J.pLs);
    float d = 1.0f;                                          // the merge's own term
    if (J.pEsm1 != 0.0f) {
        const XF e1 = xf(J.pEsm1);
        if (sv >= 2) d += xsq(xdiv(e1, pLs));
        if (J.pEsm2 != 0.0f) {
            const XF e12 = xdiv(xmul(e1, xf(J.pEsm2)), xmul(xf(vL.pf), pLs));
            if (sv >= 3) d += xsq(e12);
            if (vL.U > 0.0f) d += xsq(xdiv(xmul(e12, xf(sqrtf(vL.U / cm))), xf(vLp.pf)));
        }
    }
    if (t >= 3 && J.pLsp1 != 0.0f) {
        const XF e0 = xdiv(xf(J.pLsp1), xf(vE.pf));
        d += xsq(e0);
        const XF e01 = xdiv(e0, xf(vEp.pf));
        if (t >= 4 && prE3 != 0.0f) d += xsq(xmul(e01, xf(prE3)));
        if (vE.U > 0.0f) d += xsq(xmul(e01, xf(sqrtf(vE.U / cm))));
    }
    return d;
}

// s range [s1, s2] of the connections with eye vertex t - 1:
// [max(2 - t, minv - t, 1), min(maxv - t, nL)]; the unbounded bdpt variant
// takes every s <= nL
RD int conn_s1(int t, int minv) { return max(max(1, 2 - t), minv - t); }
RD int conn_s2(const BParams& p, int t, int nL)
{
    return (p.variant == BD_BDPT && !p.bounded) ? nL : min(p.maxv - t, nL);
}

// Connection plan: one thread per sample.  Strategies s = 0 (the eye
// subpath hit a light) are evaluated and splatted here; every structurally
// possible s >= 1 needs a shadow-queue slot, and so does (bdpt) every
// direct strategy: their number goes to off[i], which a scan turns into the
// sample's first slot.  The slots themselves are written round by round
// (k_bd_round, k_bd_slots), so the shadow queue holds what one round
// connects rather than the worst case of every sample.
__global__ __launch_bounds__(256) void k_bd_plan(RScene s, BParams p, BCtl* __restrict__ ctl, VB vb)
{
    __shared__ uint32_t lds[5];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t sample = ctl->batch_first + i;
    const bool valid = i < p.batch && sample - ctl->pass_first < p.total;
    const int nE = valid ? (int)vb.nsub[0][i] : 0;
    const int nL = valid ? (int)vb.nsub[1][i] : 0;
    const int minv = min_len(p);
    const bool bdpt = p.variant == BD_BDPT;
    uint32_t cnt = 0;
    for (int t = 1; t <= nE; t++) {
        const int s1 = conn_s1(t, minv), s2 = conn_s2(p, t, nL);
        if (s2 >= s1) cnt += (uint32_t)(s2 - s1 + 1);
    }
    uint32_t dm = 0;
    if (bdpt)
        for (int k = 0; k < nE; k++) dm |= eye_direct(s, p, vb, i, sample, k) ? 1u << k : 0u;
    cnt += (uint32_t)__popc(dm);
    if (i < p.batch) {
        vb.off[i] = cnt;
        if (bdpt) vb.dmask[i] = dm;
    }
    if (i == 0) vb.off[p.batch] = 0u;          // becomes the batch's total
    if (p.merging) {
        // eye vertices k >= 1 that scatter and leave room for a merged path
        // of at most maxv vertices (photon vertex j >= 1, length j + k + 1)
        uint32_t mmask = 0;
        for (int k = 1; k < nE && k + 2 <= p.maxv; k++) {
            const float4 a = vb.v[0][((size_t)i * (size_t)p.maxv + (size_t)k) * 4];
            if (is_scatterer(s.mats[__float_as_uint(a.w) >> 16])) mmask |= 1u << k;
        }
        uint32_t mslot = block_reserve<4>(&ctl->n_mq, (uint32_t)__popc(mmask), lds);
        for (int k = 1; k < 32; k++)
            if (mmask & (1u << k)) vb.mq[mslot++] = i * 32u + (uint32_t)k;
    }
    if (!valid || !p.connect) return;
    const uint32_t pixel = vb.epix[i];
    const float2 eras = vb.eraster[i];
    const float cm = p.merging ? ctl->c_merge : 0.0f;
    for (int t = 2; t <= nE; t++) {
        if (t < minv) continue;                        // s = 0 excluded by min_num_vertices
        const Vtx vE = load_v(vb, p, 0, i, (uint32_t)t - 1);
        if (!(vE.light >= 0 && vE.kind == VK_SURFACE)) continue;
        const Vtx vEp = load_v(vb, p, 0, i, (uint32_t)t - 2);
        // the eye subpath hit a light (t >= 2)
        const f3 d = dir_to(vE.p, vEp.p);
        if (!emits_towards(vE.fr, d)) continue;
        const RMat m = s.mats[vE.mat];
        const f3 Cs = mk3(vE.alpha.x * m.Le[0], vE.alpha.y * m.Le[1], vE.alpha.z * m.Le[2]);
        // MIS: strategies s' = 1 .. n-1 relative to s = 0, and (bdpt) the
        // direct strategy, whose pdf is that of s' = 1
        const bool ok = vE.pf > 0.0f && (t < 3 || (vEp.pf > 0.0f && vE.U >= 0.0f));
        if (!ok) continue;
        const float a = pos_pdf_light(s, vE.light) / vE.pf;
        // with merging: mu of eye vertices t-2 (from the light's emission
        // pdf) and t-3 (its stored pr); the light point itself is no merge
        const float pL1 = t >= 3 ? LMX_INV_PI_F * gterm(vE, vEp) : 0.0f;
        const float m1 = mu_of(p.c_conn, cm, t >= 3, pL1);
        const float m2 = mu_of(p.c_conn, cm, t >= 4, t >= 4 && cm != 0.0f ? load_pr(vb, p, 0, i, t - 3) : 0.0f);
        float sum = p.c_conn + chain_mu(a, t >= 3 ? pL1 / vEp.pf : 0.0f, m1, m2, vE.U, p.mis);
        if (bdpt) sum += sqm(a, p.mis);
        const float w = 1.0f / sum;
        const f3 C = mk3(Cs.x * w, Cs.y * w, Cs.z * w);
        splat(p, pixel, C);
        record(p, sample, 0, t, 0, eras.x, eras.y, C, w);
    }
}

// Start of the next connection round: the samples [lo, hi) whose slots fit
// the shadow queue's `cap` slots from lo's first slot on, hi as large as
// possible.  The offsets are non-decreasing and no sample needs more than
// cap slots, so every round takes at least one sample while any remain.
// One wave: a 64-way search, each level probing 64 indices at once.
__global__ void k_bd_round(BCtl* ctl, const uint32_t* __restrict__ off, uint32_t batch, uint32_t cap)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t lo = ctl->conn_next;
    if (lo >= batch) {
        if (lane == 0) { ctl->n_shadow = 0; ctl->conn_lo = batch; ctl->conn_hi = batch; }
        return;
    }
    const uint32_t base = off[lo];
    // the answer lies in [a, b]; off[a] - base <= cap holds throughout
    uint32_t a = lo, b = batch;
    while (a < b) {
        const uint32_t step = (b - a + 63u) / 64u;
        const uint32_t x = min(b, a + (lane + 1u) * step);
        const uint32_t k = (uint32_t)__popcll(__ballot(off[x] - base <= cap));   // lanes 0 .. k-1 fit
        const uint32_t na = a + k * step;
        b = k == 64u ? b : min(b, a + (k + 1u) * step) - 1u;
        a = min(na, b);
    }
    if (lane == 0) {
        ctl->conn_lo = lo;
        ctl->conn_hi = a;
        ctl->conn_base = base;
        ctl->n_shadow = off[a] - base;
        ctl->conn_next = a;
    }
}

// The shadow-queue slots of the round's samples: (sample, s, t, direct) of
// every connection counted by k_bd_plan, from the sample's own offset on.
__global__ __launch_bounds__(256) void k_bd_slots(RScene s, BParams p, const BCtl* __restrict__ ctl, VB vb, ShadowQ sq)
{
    const uint32_t lo = ctl->conn_lo, hi = ctl->conn_hi, base = ctl->conn_base;
    const int minv = min_len(p);
    for (uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += gridDim.x * blockDim.x) {
        const uint32_t sample = ctl->batch_first + i;
        if (!(sample - ctl->pass_first < p.total)) continue;
        const int nE = (int)vb.nsub[0][i], nL = (int)vb.nsub[1][i];
        uint32_t slot = vb.off[i] - base;
        for (int t = 1; t <= nE; t++) {
            const int s1 = conn_s1(t, minv), s2 = conn_s2(p, t, nL);
            for (int sv = s1; sv <= s2; sv++) sq.meta[slot++] = make_uint4(i, (uint32_t)sv, (uint32_t)t, 0u);
        }
        // a direct strategy is evaluated as the s = 1 connection of a fresh
        // light point with eye vertex k (its pdf equals that of s = 1); flag in .w
        if (p.variant == BD_BDPT)
            for (uint32_t dm = vb.dmask[i]; dm; dm &= dm - 1u)
                sq.meta[slot++] = make_uint4(i, 1u, (uint32_t)__ffs(dm), 1u);
    }
}

// Connection evaluation: one thread per shadow-queue slot.  The
// waves-per-EU hints here and on the photon gathers below cap the register
// budget (4 waves: 128 VGPRs, 5: 96) at the cost of a few spilled floats:
// more waves hide more of their random gathers (measured 1-5 % faster).  Writes the
// shadow ray with the weighted contribution, or a ray that cannot hit
// (tmax < tmin) with s = 0 when the connection carries nothing.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_bd_connect(RScene s, BParams p, BCtl* __restrict__ ctl, VB vb, ShadowQ sq)
{
    const uint32_t nslot = ctl->n_shadow;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nslot; j += gridDim.x * blockDim.x) {
        const uint4 pm = sq.meta[j];
        const uint32_t i = pm.x;
        const int sv = (int)pm.y, t = (int)pm.z;
        const bool direct = pm.w == 1u;                  // bdpt: light point sampled from eye vertex t - 1
        const uint32_t sample = ctl->batch_first + i;
        const Vtx vE = load_v(vb, p, 0, i, (uint32_t)t - 1);
        Vtx vEp;
        if (t >= 2) vEp = load_v(vb, p, 0, i, (uint32_t)t - 2);
        uint32_t pixel = vb.epix[i];
        const float2 eras = vb.eraster[i];
        bool emit = false;
        f3 C = mk3(0, 0, 0);
        float w = 0.0f, rx = 0.0f, ry = 0.0f;
        float4 r0, r1;
        Vtx vL;
        if (direct) {
            // the direct vertex of step t (renderer_bdpt.cpp:166-190): light by
            // uniform selection, point by area (light_area.cpp:67-70)
            const Uniforms<BParams> Un{p, sample};
            const uint32_t b = rb_step((uint32_t)t);
            int l = (int)(Un.one(b + 4u) * (float)s.nlight);
            l = l < 0 ? 0 : (l > (int)s.nlight - 1 ? (int)s.nlight - 1 : l);
            f3 gn;
            uint32_t tri;
            sample_light_point(s, l, Un.one(b + 5u), Un.one(b + 6u), vL.p, gn, tri);
            vL.sn = gn; vL.gn = gn;
            vL.kind = VK_LIGHT; vL.degenerate = false; vL.light = l;
            vL.mat = s.tri_mat[tri];
            vL.pf = pos_pdf_light(s, l); vL.pr = 0.0f;
            const float a0 = 1.0f / vL.pf;
            vL.alpha = mk3(a0, a0, a0);
            vL.U = 0.0f;
            vL.fr.p = vL.p; vL.fr.sn = gn; vL.fr.gn = gn;
            onb(gn, vL.fr.u, vL.fr.v);
        } else {
            vL = load_v(vb, p, 1, i, (uint32_t)sv - 1);
        }
        Vtx vLp;
        if (sv >= 2) vLp = load_v(vb, p, 1, i, (uint32_t)sv - 2);
        const float cm = p.merging ? ctl->c_merge : 0.0f;
        const float prE3 = (t >= 4 && cm != 0.0f) ? load_pr(vb, p, 0, i, (uint32_t)t - 3) : 0.0f;
        const Junction J = junction(s, p, cm, vL, vLp, sv, vE, vEp, t, prE3);
        const f3 Cs = J.Cs;
        bool ok = !(Cs.x == 0.0f && Cs.y == 0.0f && Cs.z == 0.0f);
        if (ok && t == 1) {
            ok = camera_raster(s, dir_to(vE.p, vL.p), rx, ry);
            pixel = raster_pixel(s, rx, ry);
        } else {
            rx = eras.x; ry = eras.y;
        }
        ok = ok && J.ok;
        if (ok) {
            float sum = J.sum;
            // bdpt: the direct strategy, whose pdf equals that of s = 1
            // (p_1 / p_s = the light side's ratios down to vertex 1)
            if (p.variant == BD_BDPT) {
                if (sv == 1) sum += 1.0f;
                else if (sv == 2) sum += sqm(J.aL, p.mis);
                else sum += sqm(J.aL, p.mis) * sqm(J.bL, p.mis) * vb.V[(size_t)i * (size_t)p.maxv + (size_t)(sv - 1)];
            }
            w = 1.0f / sum;
            // a direct strategy from eye vertex t - 1 >= 1 ran one roulette
            // more than that vertex's throughput holds
            const float sel = (direct && t >= 2) ? 2.0f : 1.0f;
            C = mk3(Cs.x * w * sel, Cs.y * w * sel, Cs.z * w * sel);
            // Scene3::Visible (scene3.h:107-116)
            f3 d = sub(vE.p, vL.p);
            const float l = sqrtf(dot(d, d));
            d = mk3(d.x / l, d.y / l, d.z / l);
            r0 = make_float4(vL.p.x, vL.p.y, vL.p.z, LMX_EPS_ISECT);
            r1 = make_float4(d.x, d.y, d.z, l * (1.0f - LMX_EPS_ISECT));
            emit = true;
        }
        if (emit) {
            sq.ray[2 * j] = r0;
            sq.ray[2 * j + 1] = r1;
            sq.c[j] = make_float4(C.x, C.y, C.z, w);
            sq.aux[j] = make_float4(rx, ry, 0.0f, 0.0f);
            sq.meta[j] = make_uint4(sample, (uint32_t)sv, (uint32_t)t | (direct ? 0x10000u : 0u), pixel);
        } else {
            sq.ray[2 * j] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
            sq.ray[2 * j + 1] = make_float4(0.0f, 0.0f, 1.0f, -1.0f);
            sq.meta[j] = make_uint4(sample, 0u, (uint32_t)t, 0u);
        }
    }
}

__global__ __launch_bounds__(256) void k_bd_splat(BParams p, BCtl* __restrict__ ctl, ShadowQ sq,
                                                  const float4* __restrict__ hits)
{
    const uint32_t n = ctl->n_shadow;
    uint32_t live = 0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const uint4 m = sq.meta[j];
        if (m.y == 0) continue;                                      // no contribution
        live++;
        if (__float_as_uint(hits[j].w) != WBVH_MISS) continue;      // occluded
        const float4 c = sq.c[j];
        const float4 a = sq.aux[j];
        splat(p, m.w, mk3(c.x, c.y, c.z));
        // a direct strategy's path has the light point as vertex 0 and the
        // eye vertices 0 .. t-1: recorded as (s = 0, t + 1, direct)
        const int t = (int)(m.z & 0xffffu), direct = (int)(m.z >> 16);
        record(p, m.x, direct ? 0 : (int)m.y, direct ? t + 1 : t, direct, a.x, a.y, mk3(c.x, c.y, c.z), c.w);
    }
    // live shadow rays, one atomic per wave
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) live += __shfl_xor(live, off);
    if ((threadIdx.x & 63u) == 0 && live) atomicAdd(&ctl->shadow_rays, (unsigned long long)live);
}

__global__ void k_bd_next_batch(BCtl* ctl, uint32_t batch)
{
    if (threadIdx.x == 0) {
        ctl->batch_first += batch;
        ctl->n_trace = 0; ctl->n_next = 0; ctl->n_shadow = 0; ctl->n_mq = 0;
    }
}

__global__ void k_bd_conn_start(BCtl* ctl)
{
    if (threadIdx.x == 0) ctl->conn_next = 0;
}

__global__ void k_bd_init(BCtl* ctl)
{
    if (threadIdx.x == 0) {
        ctl->conn_next = 0; ctl->conn_lo = 0; ctl->conn_hi = 0; ctl->conn_base = 0;
        ctl->n_trace = 0; ctl->n_next = 0; ctl->n_shadow = 0; ctl->batch_first = 0;
        ctl->rays = 0; ctl->shadow_rays = 0;
        ctl->pass_first = 0; ctl->radius = 0.0f; ctl->cell = 0.0f; ctl->c_merge = 0.0f; ctl->merges = 0; ctl->dropped = 0; ctl->cap_n[0] = 0; ctl->cap_n[1] = 0;
        ctl->n_mq = 0;
    }
}

// start of a vertex-merging phase: its first sample index, the pass's merge
// radius and N (pi r^2)^2
__global__ void k_bd_phase(BCtl* ctl, uint32_t first, float radius, float cell, float c_merge)
{
    if (threadIdx.x == 0) {
        ctl->n_trace = 0; ctl->n_next = 0; ctl->n_shadow = 0; ctl->n_mq = 0;
        ctl->batch_first = first; ctl->pass_first = first;
        ctl->radius = radius; ctl->cell = cell; ctl->c_merge = c_merge;
    }
}

// ---------------------------------------------------------------------------
// Vertex merging.  A hash grid of cell edge 2r over the photon vertices past
// the light: count per bucket, exclusive scan, scatter of (position, photon
// * maxv + vertex) records.  A query at x visits the buckets of the (at most
// 2 x 2 x 2) cells that [x - r, x + r] touches and keeps the records of those
// very cells (two cells may share a bucket) within r of x: the set is that of
// the reference's kd-tree range query (renderer_vcm.cpp:356-505).

struct Grid {
    float4* items;
    uint32_t* start;           // H + 1 bucket starts after the scan
    uint32_t* cursor;          // H fill counters
    uint32_t* part;            // scan partial sums
    uint32_t H;                // power of two
};

RD int cell_of(float x, float cell)
{
    const float c = floorf(x / cell);
    return (int)fminf(fmaxf(c, -1.0e9f), 1.0e9f);
}

RD uint32_t bucket_of(int cx, int cy, int cz, uint32_t H)
{
    return (((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u)) & (H - 1u);
}

__global__ void k_vc_grid_clear(Grid g)
{
    for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h <= g.H; h += gridDim.x * blockDim.x) {
        g.start[h] = 0u;
        if (h < g.H) g.cursor[h] = 0u;
    }
}

// one thread per (photon, vertex >= 1) slot; phase 0 counts, phase 1 scatters
template <int PHASE>
__global__ __launch_bounds__(256) void k_vc_grid_fill(RScene s, BParams p, const BCtl* __restrict__ ctl, VB ph,
                                                      Grid g)
{
    const uint32_t np = p.batch;
    const size_t n = (size_t)np * (size_t)(p.maxv - 1);
    const float cell = ctl->cell;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
        // neighbouring lanes take neighbouring vertices of one photon subpath
        // (64 B apart, not maxv x 64 B): 8 % off both phases, which the bucket
        // atomics bound
        const uint32_t i = (uint32_t)(idx / (uint32_t)(p.maxv - 1)), j = 1u + (uint32_t)(idx % (uint32_t)(p.maxv - 1));
        if (j >= ph.nsub[1][i]) continue;
        const float4 a = ph.v[1][((size_t)i * (size_t)p.maxv + j) * 4];
        // renderer::pm / ppm store photons at diffuse and glossy vertices only
        // (renderer_pm.cpp:98-110, renderer_ppm.cpp:214-226)
        if ((p.variant == BD_PM || p.variant == BD_PPM) && !is_scatterer(s.mats[__float_as_uint(a.w) >> 16])) continue;
        const uint32_t h = bucket_of(cell_of(a.x, cell), cell_of(a.y, cell), cell_of(a.z, cell), g.H);
        if (PHASE == 0) {
            atomicAdd(&g.start[h], 1u);
        } else {
            const uint32_t slot = g.start[h] + atomicAdd(&g.cursor[h], 1u);
            g.items[slot] = make_float4(a.x, a.y, a.z, __uint_as_float(i * (uint32_t)p.maxv + j));
        }
    }
}

// exclusive scan of a 256-thread block's values; `tot` gets the block total
RD uint32_t block_scan(uint32_t v, uint32_t* lds, uint32_t& tot)
{
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63u) lds[w] = x;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t q = 0; q < w; q++) base += lds[q];
    tot = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return base + x - v;
}

#define SCAN_PER 16u
#define SCAN_TILE (256u * SCAN_PER)

// tile-local exclusive scan of a[0, n), tile totals into part.  The tile is
// staged through LDS so that global reads and writes are coalesced (lane l
// of a pass touches word pass * 256 + l); each thread then scans its own 16
// consecutive words there.  Word i sits at i + i / 32, so when a wave's 64
// lanes read words 16 l + k no bank serves more than two of them.
// (Lanes reading their 16 words straight from global memory took 4.7x the
// time of k_scan_add over the same 2^24 counts.)
__global__ __launch_bounds__(256) void k_scan_tiles(uint32_t* a, uint32_t n, uint32_t* part)
{
    __shared__ uint32_t tile[SCAN_TILE + SCAN_TILE / 32u];
    __shared__ uint32_t lds[4];
    const uint32_t t0 = blockIdx.x * SCAN_TILE;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER; k++) {
        const uint32_t i = k * 256u + threadIdx.x;
        tile[i + (i >> 5)] = t0 + i < n ? a[t0 + i] : 0u;
    }
    __syncthreads();
    const uint32_t b = threadIdx.x * SCAN_PER;
    uint32_t v[SCAN_PER], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER; k++) {
        v[k] = tile[b + k + ((b + k) >> 5)];
        sum += v[k];
    }
    uint32_t tot;
    uint32_t run = block_scan(sum, lds, tot);
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER; k++) {
        tile[b + k + ((b + k) >> 5)] = run;
        run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < SCAN_PER; k++) {
        const uint32_t i = k * 256u + threadIdx.x;
        if (t0 + i < n) a[t0 + i] = tile[i + (i >> 5)];
    }
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// exclusive scan of the tile totals (one workgroup; nt <= 256 * 64)
__global__ __launch_bounds__(256) void k_scan_part(uint32_t* part, uint32_t nt)
{
    __shared__ uint32_t lds[4];
    const uint32_t per = (nt + 255u) / 256u;
    const uint32_t base = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < per; k++) sum += base + k < nt ? part[base + k] : 0u;
    uint32_t tot;
    uint32_t run = block_scan(sum, lds, tot);
    for (uint32_t k = 0; k < per; k++)
        if (base + k < nt) {
            const uint32_t v = part[base + k];
            part[base + k] = run;
            run += v;
        }
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* a, uint32_t n, const uint32_t* part)
{
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
        a[j] += part[j / SCAN_TILE];
}

// Merges: one thread per queued eye vertex k = t - 1 >= 1.  Every photon
// vertex j within r makes the path of the connection (s = j, t) between
// photon vertex j - 1 and eye vertex t - 1, kept when its length is within
// [min, max] vertices; its contribution f w / p is the connection's
// unweighted term times pLs pi r^2 / D and its weight (pi r^2 pLs)^2 / D, D
// the mu-weighted denominator (EvaluateF / EvaluatePathPDF with merging and
// EvaluateMISWeight_VCM / _BDPM, renderer_vcm.cpp:116-340).  The thread
// alternates a gather of up to MERGE_LIST photon vertices within r (cheap
// distance tests) into its LDS list with the evaluation of the list, so the
// lanes of a wave evaluate together instead of idling behind each other's
// candidate tests.  32 entries (32 KB of LDS per workgroup: 4 workgroups
// per CU) beat 16 by 4-8 % and 48 (3 workgroups per CU) by 1-3 %.
#define MERGE_LIST 32

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_vc_merge(RScene s, BParams p, BCtl* __restrict__ ctl, VB vb, VB ph,
                                                  Grid g)
{
    __shared__ uint32_t hl[MERGE_LIST * 256];          // [hit][thread]
    const uint32_t nq = ctl->n_mq;
    const float r = ctl->radius, cell = ctl->cell, cm = ctl->c_merge;
    const float r2 = r * r, area = LMX_PI_F * r * r;
    uint32_t nmerge = 0, ndrop = 0;
    for (uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x; qi < nq; qi += gridDim.x * blockDim.x) {
        const uint32_t e = vb.mq[qi];
        const uint32_t i = e >> 5, k = e & 31u;
        const uint32_t sample = ctl->batch_first + i;
        const int t = (int)k + 1;
        const Vtx vE = load_v(vb, p, 0, i, k);
        const Vtx vEp = load_v(vb, p, 0, i, k - 1u);
        const float prE3 = t >= 4 ? load_pr(vb, p, 0, i, (uint32_t)t - 3) : 0.0f;
        const float2 eras = vb.eraster[i];
        int lo[3], nc[3];
        const float q[3] = {vE.p.x, vE.p.y, vE.p.z};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            lo[a] = cell_of(q[a] - r, cell);
            nc[a] = cell_of(q[a] + r, cell) - lo[a] + 1;
        }
        const int ncell = nc[0] * nc[1] * nc[2];
        // photon vertex j may merge only if the merged path fits: j in [minv - t, maxv - t]
        const int jmax = p.maxv - t, jmin = p.minv - t;
        int c = 0;
        uint32_t it = 0, end = 0;
        bool fresh = true;                               // `it` not yet set for cell c
        f3 Csum = mk3(0, 0, 0);
        while (c < ncell) {
            // gather
            uint32_t nh = 0;
            while (c < ncell && nh < MERGE_LIST) {
                const int cx = lo[0] + c % nc[0], cy = lo[1] + (c / nc[0]) % nc[1], cz = lo[2] + c / (nc[0] * nc[1]);
                if (fresh) {
                    const uint32_t h = bucket_of(cx, cy, cz, g.H);
                    it = g.start[h];
                    end = g.start[h + 1];
                    fresh = false;
                }
                for (; it < end && nh < MERGE_LIST; it++) {
                    const float4 cd = g.items[it];
                    const int j = (int)(__float_as_uint(cd.w) % (uint32_t)p.maxv);
                    if (j > jmax || j < jmin) continue;
                    if (cell_of(cd.x, cell) != cx || cell_of(cd.y, cell) != cy || cell_of(cd.z, cell) != cz) continue;
                    const f3 d = sub(mk3(cd.x, cd.y, cd.z), vE.p);
                    if (!(dot(d, d) < r2)) continue;
                    hl[nh * 256u + threadIdx.x] = __float_as_uint(cd.w);
                    nh++;
                }
                if (it >= end) { c++; fresh = true; }
            }
            // evaluate
            for (uint32_t h = 0; h < nh; h++) {
                const uint32_t id = hl[h * 256u + threadIdx.x];
                const uint32_t pi = id / (uint32_t)p.maxv;
                const int j = (int)(id % (uint32_t)p.maxv);
                const Vtx vL = load_v(ph, p, 1, pi, (uint32_t)j - 1u);
                Vtx vLp;
                if (j >= 2) vLp = load_v(ph, p, 1, pi, (uint32_t)j - 2u);
                const Junction J = junction(s, p, cm, vL, vLp, j, vE, vEp, t, prE3);
                if (!J.ok || J.pLs == 0.0f) continue;
                if (J.Cs.x == 0.0f && J.Cs.y == 0.0f && J.Cs.z == 0.0f) continue;
                // C = Cs pLs pi r^2 / sum and w = (pi r^2 pLs)^2 / sum.  With
                // the connection terms on, sum >= c_conn keeps both in range;
                // in mode bdpm every term of sum scales with squared pdfs, so
                // it is taken relative to the merge's own term instead.
                f3 C;
                float w;
                if (p.c_conn != 0.0f) {
                    const float f = J.pLs * area / J.sum;
                    C = mk3(J.Cs.x * f, J.Cs.y * f, J.Cs.z * f);
                    const float pm = area * J.pLs;
                    w = pm * pm / J.sum;
                } else {
                    // C = Cs / (N pi r^2 pLs d), with separate exponents:
                    // Cs and pLs share the geometry term and the light
                    // side's glossy lobe, so their quotient is in range
                    // where each of them underflows
                    const float d = merge_den_bdpm(J, cm, vL, vLp, j, vE, vEp, t, prE3);
                    const float ac = area / cm;
                    const XF pd = xmul(xf(J.pLs), xf(d));
                    auto c1 = [&](float c) {
                        const XF q = xdiv(xmul(xf(c), xf(ac)), pd);
                        return ldexpf(q.m, q.e);
                    };
                    C = mk3(c1(J.Cs.x), c1(J.Cs.y), c1(J.Cs.z));
                    w = ac * area / d;
                }
                if (!isfinite(C.x) || !isfinite(C.y) || !isfinite(C.z)) {
                    ndrop++;                                     // counted and reported, never silent
                    continue;
                }
                Csum = add(Csum, C);
                nmerge++;
                record_merge(p, sample, j + 1, t, eras.x, eras.y, C, w);
            }
        }
        if (!(Csum.x == 0.0f && Csum.y == 0.0f && Csum.z == 0.0f)) splat(p, vb.epix[i], Csum);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        nmerge += __shfl_xor(nmerge, off);
        ndrop += __shfl_xor(ndrop, off);
    }
    if ((threadIdx.x & 63u) == 0 && nmerge) atomicAdd(&ctl->merges, (unsigned long long)nmerge);
    if ((threadIdx.x & 63u) == 0 && ndrop) atomicAdd(&ctl->dropped, (unsigned long long)ndrop);
}

// renderer::pm's eye side (renderer_pm.cpp:145-205): one thread per eye
// sample walks its subpath (traced without roulette up to the gather
// vertex); an emitter vertex adds throughput x Le towards the previous vertex,
// and the first (finalgather 0) or second (finalgather 1) diffuse or glossy
// vertex gathers every photon within r: kernel 3/pi (1 - d^2/r^2)^2 over
// r^2 N, times the BSDF towards the photon's incoming direction and the
// photon's throughput, for photons whose merged path has at most
// max_num_vertices vertices.  The gather alternates a list of up to
// MERGE_LIST photons within r (distance tests only) with their evaluation.
struct PMArgs {
    int32_t pmaxv;             // photon record stride (vertices per photon subpath)
    int32_t maxv;              // max_num_vertices, -1 unbounded
    int32_t finalgather;
    float nphoton;
};

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_pm_gather(RScene s, BParams p, BCtl* __restrict__ ctl, VB vb, VB ph,
                                                   Grid g, PMArgs a)
{
    __shared__ uint32_t hl[MERGE_LIST * 256];
    const float r = ctl->radius, cell = ctl->cell;
    const float r2 = r * r;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.batch; i += gridDim.x * blockDim.x) {
        const uint32_t sample = ctl->batch_first + i;
        if (sample - ctl->pass_first >= p.total) continue;
        const int nE = (int)vb.nsub[0][i];
        const float2 eras = vb.eraster[i];
        f3 Cem = mk3(0, 0, 0), Cde = mk3(0, 0, 0);
        float nem = 0.0f, nph = 0.0f, gv = 0.0f;
        bool gather_next = a.finalgather == 0;
        for (int k = 1; k < nE; k++) {
            const Vtx v = load_v(vb, p, 0, i, (uint32_t)k);
            const float4 pq = vb.v[0][((size_t)i * (size_t)p.maxv + (size_t)(k - 1)) * 4];
            const f3 toprev = dir_to(v.p, mk3(pq.x, pq.y, pq.z));
            const RMat m = s.mats[v.mat];
            if (v.light >= 0) {
                if (emits_towards(v.fr, toprev))
                    Cem = add(Cem, mk3(v.alpha.x * m.Le[0], v.alpha.y * m.Le[1], v.alpha.z * m.Le[2]));
                nem += 1.0f;
            }
            if (!is_scatterer(m)) continue;
            if (!gather_next) { gather_next = true; continue; }
            gv = (float)(k + 1);
            // range query: cells of edge 2r around v
            int lo[3], nc[3];
            const float q[3] = {v.p.x, v.p.y, v.p.z};
#pragma unroll
            for (int ax = 0; ax < 3; ax++) {
                lo[ax] = cell_of(q[ax] - r, cell);
                nc[ax] = cell_of(q[ax] + r, cell) - lo[ax] + 1;
            }
            const int ncell = nc[0] * nc[1] * nc[2];
            // photon vertex j may join: (k + 1) + (j + 1) - 1 <= maxv
            const int jmax = a.maxv == -1 ? 0x7fffffff : a.maxv - k - 1;
            int c = 0;
            uint32_t it = 0, end = 0;
            bool fresh = true;
            while (c < ncell) {
                uint32_t nh = 0;
                while (c < ncell && nh < MERGE_LIST) {
                    const int cx = lo[0] + c % nc[0], cy = lo[1] + (c / nc[0]) % nc[1], cz = lo[2] + c / (nc[0] * nc[1]);
                    if (fresh) {
                        const uint32_t h = bucket_of(cx, cy, cz, g.H);
                        it = g.start[h];
                        end = g.start[h + 1];
                        fresh = false;
                    }
                    for (; it < end && nh < MERGE_LIST; it++) {
                        const float4 cd = g.items[it];
                        const int j = (int)(__float_as_uint(cd.w) % (uint32_t)a.pmaxv);
                        if (j > jmax) continue;
                        if (cell_of(cd.x, cell) != cx || cell_of(cd.y, cell) != cy || cell_of(cd.z, cell) != cz) continue;
                        const f3 d = sub(mk3(cd.x, cd.y, cd.z), v.p);
                        if (!(dot(d, d) < r2)) continue;
                        hl[nh * 256u + threadIdx.x] = __float_as_uint(cd.w);
                        nh++;
                    }
                    if (it >= end) { c++; fresh = true; }
                }
                for (uint32_t h = 0; h < nh; h++) {
                    const uint32_t id = hl[h * 256u + threadIdx.x];
                    const float4* rec = ph.v[1] + (size_t)id * 4;          // photon vertex j of subpath id / pmaxv
                    const float4 r0 = rec[0], r3 = rec[3], rp = rec[-4];   // its position, throughput; vertex j - 1
                    const f3 pp = mk3(r0.x, r0.y, r0.z);
                    const f3 d = sub(pp, v.p);
                    const float sk = 1.0f - dot(d, d) / r / r;
                    const float kern = 3.0f * LMX_INV_PI_F * sk * sk;
                    const float pden = kern / (r * r * a.nphoton);
                    const f3 f = bsdf_eval_t(m, v.fr, toprev, dir_to(pp, mk3(rp.x, rp.y, rp.z)), false);
                    Cde = add(Cde, mk3(v.alpha.x * pden * f.x * r3.x, v.alpha.y * pden * f.y * r3.y,
                                       v.alpha.z * pden * f.z * r3.z));
                    nph += 1.0f;
                }
            }
            break;
        }
        const f3 C = add(Cem, Cde);
        if (!(C.x == 0.0f && C.y == 0.0f && C.z == 0.0f)) splat(p, vb.epix[i], C);
        if (p.sample_out) {
            float4* o = reinterpret_cast<float4*>(p.sample_out + 16 * (size_t)sample);
            o[0] = make_float4(eras.x, eras.y, Cem.x, Cem.y);
            o[1] = make_float4(Cem.z, nem, Cde.x, Cde.y);
            o[2] = make_float4(Cde.z, nph, gv, 0.0f);
            o[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
}

// renderer::ppm (renderer_ppm.cpp:84-320).  Measurement point i is eye
// vertex 1 of sample i when it is diffuse or glossy (an eye subpath reaches
// no further point: other surfaces end it); its state is st[2 i] = (tau,
// N) and st[2 i + 1] = (emission, radius), radius < 0 for no point.
__global__ __launch_bounds__(256) void k_ppm_init(RScene s, BParams p, VB vb, float4* __restrict__ st, float r0)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.batch; i += gridDim.x * blockDim.x) {
        float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f), b = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        if (vb.nsub[0][i] >= 2) {
            const Vtx v = load_v(vb, p, 0, i, 1u);
            const RMat m = s.mats[v.mat];
            if (is_scatterer(m)) {
                b.w = r0;
                const float4 c = vb.v[0][(size_t)i * (size_t)p.maxv * 4];
                // emission at the point (:154-162)
                if (v.light >= 0 && emits_towards(v.fr, dir_to(v.p, mk3(c.x, c.y, c.z)))) {
                    b.x = v.alpha.x * m.Le[0]; b.y = v.alpha.y * m.Le[1]; b.z = v.alpha.z * m.Le[2];
                }
            }
        }
        st[2 * (size_t)i] = a;
        st[2 * (size_t)i + 1] = b;
    }
}

// The photons within r of a point v (direction wi to its previous vertex,
// material m) whose vertex index is at most jmax: M and dtau = sum of f x
// photon throughput (renderer_ppm.cpp:274-290), through the gather list of
// k_vc_merge (hl: the workgroup's MERGE_LIST x 256 LDS list).
RD void ppm_gather(const Vtx& v, f3 wi, const RMat& m, float r, float cell, const Grid& g, const VB& ph, int pmaxv,
                   int jmax, uint32_t* hl, f3& dtau, float& M)
{
    const float r2 = r * r;
    int lo[3], nc[3];
    const float q[3] = {v.p.x, v.p.y, v.p.z};
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        lo[ax] = cell_of(q[ax] - r, cell);
        nc[ax] = cell_of(q[ax] + r, cell) - lo[ax] + 1;
    }
    const int ncell = nc[0] * nc[1] * nc[2];
    dtau = mk3(0, 0, 0);
    M = 0.0f;
    int c = 0;
    uint32_t it = 0, end = 0;
    bool fresh = true;
    while (c < ncell) {
        uint32_t nh = 0;
        while (c < ncell && nh < MERGE_LIST) {
            const int cx = lo[0] + c % nc[0], cy = lo[1] + (c / nc[0]) % nc[1], cz = lo[2] + c / (nc[0] * nc[1]);
            if (fresh) {
                const uint32_t h = bucket_of(cx, cy, cz, g.H);
                it = g.start[h];
                end = g.start[h + 1];
                fresh = false;
            }
            for (; it < end && nh < MERGE_LIST; it++) {
                const float4 cd = g.items[it];
                const int j = (int)(__float_as_uint(cd.w) % (uint32_t)pmaxv);
                if (j > jmax) continue;
                if (cell_of(cd.x, cell) != cx || cell_of(cd.y, cell) != cy || cell_of(cd.z, cell) != cz) continue;
                const f3 d = sub(mk3(cd.x, cd.y, cd.z), v.p);
                if (!(dot(d, d) < r2)) continue;
                hl[nh * 256u + threadIdx.x] = __float_as_uint(cd.w);
                nh++;
            }
            if (it >= end) { c++; fresh = true; }
        }
        for (uint32_t h = 0; h < nh; h++) {
            const uint32_t id = hl[h * 256u + threadIdx.x];
            const float4* rec = ph.v[1] + (size_t)id * 4;
            const float4 r0 = rec[0], r3 = rec[3], rp = rec[-4];
            const f3 pp = mk3(r0.x, r0.y, r0.z);
            const f3 f = bsdf_eval_t(m, v.fr, wi, dir_to(pp, mk3(rp.x, rp.y, rp.z)), false);
            dtau = add(dtau, mk3(f.x * r3.x, f.y * r3.y, f.z * r3.z));
            M += 1.0f;
        }
    }
}

// One progressive density-estimation step per measurement point
// (:270-297): the pass's photons within the point's radius whose merged
// path has at most max_num_vertices vertices give M and dtau; then ratio =
// (N + alpha M) / (N + M), tau = (tau + dtau) ratio, r = r sqrt(ratio),
// N = N + alpha M.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_ppm_update(RScene s, BParams p, const BCtl* __restrict__ ctl, VB vb, VB ph,
                                                    Grid g, PMArgs a, float4* __restrict__ st, float alpha)
{
    __shared__ uint32_t hl[MERGE_LIST * 256];
    const float cell = ctl->cell;
    // photon vertex j joins when 2 + (j + 1) - 1 <= maxv
    const int jmax = a.maxv == -1 ? 0x7fffffff : a.maxv - 2;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.batch; i += gridDim.x * blockDim.x) {
        float4 s0 = st[2 * (size_t)i];
        const float4 s1 = st[2 * (size_t)i + 1];
        const float r = s1.w;
        if (!(r > 0.0f)) continue;
        const Vtx v = load_v(vb, p, 0, i, 1u);
        const float4 c0 = vb.v[0][(size_t)i * (size_t)p.maxv * 4];
        f3 dtau;
        float M;
        ppm_gather(v, dir_to(v.p, mk3(c0.x, c0.y, c0.z)), s.mats[v.mat], r, cell, g, ph, a.pmaxv, jmax, hl, dtau, M);
        if (s0.w + M == 0.0f) continue;
        const float ratio = (s0.w + alpha * M) / (s0.w + M);
        s0 = make_float4((s0.x + dtau.x) * ratio, (s0.y + dtau.y) * ratio, (s0.z + dtau.z) * ratio,
                         s0.w + alpha * M);
        st[2 * (size_t)i] = s0;
        st[2 * (size_t)i + 1] = make_float4(s1.x, s1.y, s1.z, r * sqrtf(ratio));
    }
}

// renderer::sppm's pass (renderer_sppm.cpp:159-316) per pixel i: this
// pass's eye vertex 1 is the point when it is diffuse or glossy (else the
// pixel skips the pass); its emission adds to the pixel's sum; then the
// update of k_ppm_update with tau gaining throughput x dtau.  State per
// pixel: st[3 i] = (tau, N), st[3 i + 1] = (emission sum, radius),
// st[3 i + 2] = (passes with a point, 0, 0, 0).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_sppm_update(RScene s, BParams p, const BCtl* __restrict__ ctl, VB vb,
                                                     VB ph, Grid g, PMArgs a, float4* __restrict__ st, float alpha)
{
    __shared__ uint32_t hl[MERGE_LIST * 256];
    const float cell = ctl->cell;
    const int jmax = a.maxv == -1 ? 0x7fffffff : a.maxv - 2;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.batch; i += gridDim.x * blockDim.x) {
        if (vb.nsub[0][i] < 2) continue;
        const Vtx v = load_v(vb, p, 0, i, 1u);
        const RMat m = s.mats[v.mat];
        if (!is_scatterer(m)) continue;
        float4 s0 = st[3 * (size_t)i], s1 = st[3 * (size_t)i + 1], s2 = st[3 * (size_t)i + 2];
        const float4 c0 = vb.v[0][(size_t)i * (size_t)p.maxv * 4];
        const f3 wi = dir_to(v.p, mk3(c0.x, c0.y, c0.z));
        // emission at the point (:183-191)
        if (v.light >= 0 && emits_towards(v.fr, wi)) {
            s1.x += v.alpha.x * m.Le[0]; s1.y += v.alpha.y * m.Le[1]; s1.z += v.alpha.z * m.Le[2];
        }
        s2.x += 1.0f;
        f3 dtau;
        float M;
        ppm_gather(v, wi, m, s1.w, cell, g, ph, a.pmaxv, jmax, hl, dtau, M);
        if (s0.w + M != 0.0f) {
            const float ratio = (s0.w + alpha * M) / (s0.w + M);
            s0 = make_float4((s0.x + v.alpha.x * dtau.x) * ratio, (s0.y + v.alpha.y * dtau.y) * ratio,
                             (s0.z + v.alpha.z * dtau.z) * ratio, s0.w + alpha * M);
            s1.w = s1.w * sqrtf(ratio);
        }
        st[3 * (size_t)i] = s0;
        st[3 * (size_t)i + 1] = s1;
        st[3 * (size_t)i + 2] = s2;
    }
}

// renderer::sppm's film (:321-327): pixel i = tau / (pi r^2 N_total) +
// emission / passes, written (not splatted).
__global__ __launch_bounds__(256) void k_sppm_film(BParams p, const float4* __restrict__ st, float total,
                                                   float passes)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.batch; i += gridDim.x * blockDim.x) {
        const float4 s0 = st[3 * (size_t)i], s1 = st[3 * (size_t)i + 1], s2 = st[3 * (size_t)i + 2];
        const float d = s1.w * s1.w * LMX_PI_F * total;
        const f3 C = mk3(s0.x / d + s1.x / passes, s0.y / d + s1.y / passes, s0.z / d + s1.z / passes);
        float* fp = p.film + 3 * (size_t)i;
        fp[0] = C.x; fp[1] = C.y; fp[2] = C.z;
        if (p.sample_out) {
            float4* o = reinterpret_cast<float4*>(p.sample_out + 16 * (size_t)i);
            o[0] = make_float4(C.x, C.y, C.z, s1.x);
            o[1] = make_float4(s1.y, s1.z, s0.x, s0.y);
            o[2] = make_float4(s0.z, s0.w, s1.w, s2.x);
            o[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
}

// renderer::ltdirect (renderer_ltdirect.cpp:72-250): every kept vertex k of
// light subpath i (k + 1 < max_num_vertices) connects to the pinhole:
// C = throughput x f_L(towards the camera) x G x We, with a shadow ray to
// the camera, appended to the shadow queue (one reservation per workgroup).
__global__ __launch_bounds__(256) void k_lt_connect(RScene s, BParams p, BCtl* __restrict__ ctl, VB ph, ShadowQ sq,
                                                    int32_t maxv_user)
{
    __shared__ uint32_t lds[5];
    const size_t n = (size_t)p.batch * (size_t)p.maxv;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    // whole workgroups iterate together (the reservation synchronises them)
    for (size_t base = (size_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const size_t idx = base + threadIdx.x;
        bool emit = false;
        float4 r0, r1, cc;
        float rx = 0.0f, ry = 0.0f;
        uint32_t i = 0, k = 0, pixel = 0;
        if (idx < n) {
            i = (uint32_t)(idx % p.batch);
            k = (uint32_t)(idx / p.batch);
            const uint32_t sample = ctl->batch_first + i;
            if (sample < p.total && k < ph.nsub[1][i] && (maxv_user == -1 || (int)k + 1 < maxv_user)) {
                const Vtx v = load_v(ph, p, 1, i, k);
                const f3 cam = s.cam_pos;
                const f3 ppE = dir_to(v.p, cam);
                f3 fsL = mk3(0, 0, 0);
                if (k == 0) {
                    if (emits_towards(v.fr, ppE)) {
                        const RMat m = s.mats[v.mat];
                        fsL = mk3(m.Le[0], m.Le[1], m.Le[2]);
                    }
                } else {
                    const RMat m = s.mats[v.mat];
                    if (is_scatterer(m)) {
                        const float4 q = ph.v[1][((size_t)i * (size_t)p.maxv + (size_t)(k - 1)) * 4];
                        fsL = bsdf_eval_t(m, v.fr, dir_to(v.p, mk3(q.x, q.y, q.z)), ppE, true);
                    }
                }
                const f3 toward = neg(ppE);
                const float fsE = importance(s, toward);
                const float G = gterm(v.p, v.sn, v.degenerate, cam, mk3(0, 0, 0), true);
                if (fsE != 0.0f && G != 0.0f && !(fsL.x == 0.0f && fsL.y == 0.0f && fsL.z == 0.0f) &&
                    camera_raster(s, toward, rx, ry)) {
                    pixel = raster_pixel(s, rx, ry);
                    cc = make_float4(v.alpha.x * fsL.x * G * fsE, v.alpha.y * fsL.y * G * fsE,
                                     v.alpha.z * fsL.z * G * fsE, 0.0f);
                    // Scene3::Visible (scene3.h:107-116)
                    f3 d = sub(cam, v.p);
                    const float l = sqrtf(dot(d, d));
                    d = mk3(d.x / l, d.y / l, d.z / l);
                    r0 = make_float4(v.p.x, v.p.y, v.p.z, LMX_EPS_ISECT);
                    r1 = make_float4(d.x, d.y, d.z, l * (1.0f - LMX_EPS_ISECT));
                    emit = true;
                }
            }
        }
        const uint32_t slot = block_reserve<4>(&ctl->n_shadow, emit ? 1u : 0u, lds);
        if (emit) {
            sq.ray[2 * slot] = r0;
            sq.ray[2 * slot + 1] = r1;
            sq.c[slot] = cc;
            sq.aux[slot] = make_float4(rx, ry, 0.0f, 0.0f);
            sq.meta[slot] = make_uint4(ctl->batch_first + i, k, 1u, pixel);
        }
    }
}

// Unoccluded camera connections to the film (and, when asked, record k of
// the sample: raster x, y, C rgb, 1, 0, 0).
__global__ __launch_bounds__(256) void k_lt_splat(BParams p, BCtl* __restrict__ ctl, ShadowQ sq,
                                                  const float4* __restrict__ hits, int32_t M)
{
    const uint32_t n = ctl->n_shadow;
    uint32_t live = 0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        live++;
        if (__float_as_uint(hits[j].w) != WBVH_MISS) continue;
        const uint4 m = sq.meta[j];
        const float4 c = sq.c[j], a = sq.aux[j];
        splat(p, m.w, mk3(c.x, c.y, c.z));
        if (p.sample_out) {
            float4* o = reinterpret_cast<float4*>(p.sample_out + 8 * ((size_t)m.x * (size_t)M + m.y));
            o[0] = make_float4(a.x, a.y, c.x, c.y);
            o[1] = make_float4(c.z, 1.0f, 0.0f, 0.0f);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) live += __shfl_xor(live, off);
    if ((threadIdx.x & 63u) == 0 && live) atomicAdd(&ctl->shadow_rays, (unsigned long long)live);
}

__global__ void k_lt_batch(BCtl* ctl, uint32_t first)
{
    if (threadIdx.x == 0) {
        ctl->n_trace = 0; ctl->n_next = 0; ctl->n_shadow = 0;
        ctl->batch_first = first;
    }
}

// renderer::raycast (renderer_raycast.cpp:59-111): a ray through every pixel
// centre; a hit shows |sn . d| x R (the reflectance of diffuse, glossy and
// cook-torrance surfaces, 1 for a surface without a BSDF), a miss black.
__global__ __launch_bounds__(256) void k_rc_gen(RScene s, float4* __restrict__ ray, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = i % (uint32_t)s.width, y = i / (uint32_t)s.width;
    const f3 d = camera_dir(s, ((float)x + 0.5f) / (float)s.width, ((float)y + 0.5f) / (float)s.height);
    ray[2 * i] = make_float4(s.cam_pos.x, s.cam_pos.y, s.cam_pos.z, LMX_EPS_ISECT);
    ray[2 * i + 1] = make_float4(d.x, d.y, d.z, LMX_FLT_MAX);
}

__global__ __launch_bounds__(256) void k_rc_shade(RScene s, const float4* __restrict__ ray,
                                                  const float4* __restrict__ hits, float* __restrict__ film, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 h = hits[i];
    const uint32_t id = __float_as_uint(h.w);
    f3 c = mk3(0, 0, 0);
    if (id != WBVH_MISS) {
        const float4 r0 = ray[2 * i], r1 = ray[2 * i + 1];
        const f3 d = mk3(r1.x, r1.y, r1.z);
        Frame fr;
        surface_at(s, mk3(r0.x, r0.y, r0.z), d, h.x, h.y, h.z, id, fr);
        const RMat m = s.mats[s.tri_mat[id]];
        const float k = fabsf(dot(fr.sn, neg(d)));
        c = is_scatterer(m) ? mk3(k * m.R[0], k * m.R[1], k * m.R[2]) : mk3(k, k, k);
    }
    film[3 * (size_t)i] = c.x;
    film[3 * (size_t)i + 1] = c.y;
    film[3 * (size_t)i + 2] = c.z;
}

__global__ void k_sppm_init(float4* __restrict__ st, uint32_t n, float r0)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        st[3 * (size_t)i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        st[3 * (size_t)i + 1] = make_float4(0.0f, 0.0f, 0.0f, r0);
        st[3 * (size_t)i + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// The film after the last pass (:300-310): throughput x tau / (pi r^2 N_total)
// + emission per point.
__global__ __launch_bounds__(256) void k_ppm_splat(BParams p, VB vb, const float4* __restrict__ st, float total)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < p.batch; i += gridDim.x * blockDim.x) {
        const float4 s0 = st[2 * (size_t)i], s1 = st[2 * (size_t)i + 1];
        if (!(s1.w > 0.0f)) continue;
        const float4 d = vb.v[0][((size_t)i * (size_t)p.maxv + 1) * 4 + 3];          // throughput at the point
        const float pd = 1.0f / (s1.w * s1.w * LMX_PI_F * total);
        const f3 C = mk3(d.x * pd * s0.x + s1.x, d.y * pd * s0.y + s1.y, d.z * pd * s0.z + s1.z);
        splat(p, vb.epix[i], C);
        if (p.sample_out) {
            const float2 e = vb.eraster[i];
            float4* o = reinterpret_cast<float4*>(p.sample_out + 16 * (size_t)i);
            o[0] = make_float4(e.x, e.y, C.x, C.y);
            o[1] = make_float4(C.z, s1.x, s1.y, s1.z);
            o[2] = make_float4(s0.x, s0.y, s0.z, s0.w);
            o[3] = make_float4(s1.w, 1.0f, 0.0f, 0.0f);
        }
    }
}


__global__ void k_bd_zero(float* a, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0.0f;
}

__global__ void k_bd_scale(float* a, size_t n, float w)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] *= w;
}

// Shadow-queue slots one sample may need: the (s >= 1, t >= 1) connections
// with s + t <= maxv (vcm and bounded bdpt) or s, t <= maxv (unbounded bdpt),
// plus one s = 0 marker (vcm) or up to maxv - 1 direct strategies (bdpt).
uint32_t bd_conn_per_sample(uint32_t maxv, int variant, bool bounded, bool connect = true)
{
    if (!connect) return 1;
    if (variant == BD_VCM) return maxv * (maxv - 1) / 2 + 1;
    return (bounded ? maxv * (maxv - 1) / 2 : maxv * maxv) + maxv;
}

// Shadow-queue slots per sample of one connection round: a sample whose
// subpaths hold more connections than this spills into further rounds
// instead of every sample reserving its worst case (1,056 slots, 101 KB,
// for unbounded renderer::bdpt).
#define BD_ROUND_SLOTS 24u

uint32_t bd_round_slots(uint32_t worst) { return std::min(worst, BD_ROUND_SLOTS); }

// Device bytes of one batch slot: two subpaths of vertex records, their
// queues and hits, the connection count, and the shadow-ray records of one
// round's connections.
size_t bd_bytes_per_sample(uint32_t maxv, uint32_t worst)
{
    return 2 * ((size_t)maxv * 64 + 4 + 4 * 16 + 2 * 16 + 2 * 16) + (size_t)maxv * 4 + 4 + 8 + 2 * 16 + 8 +
           (size_t)bd_round_slots(worst) * (2 * 16 + 4 * 16);
}

// Buffers for batches of B samples of up to maxv vertices and conn
// connections each; kept when the current ones are large enough, else
// replaced by buffers of exactly this shape (the old ones are freed first).
void bd_release(lmx_scene* sc)
{
    sc->bd_batch = sc->bd_maxv = sc->bd_conn = 0;
    for (int k = 0; k < 2; k++) {
        sc->bd_vert[k].release(); sc->bd_nsub[k].release(); sc->bd_ray[k].release();
        sc->bd_thr[k].release(); sc->bd_meta[k].release();
    }
    sc->bd_V.release(); sc->bd_epix.release(); sc->bd_eraster.release(); sc->bd_hits.release();
    sc->bd_sray.release(); sc->bd_sc.release(); sc->bd_saux.release(); sc->bd_shits.release();
    sc->bd_smeta.release(); sc->bd_off.release(); sc->bd_part.release(); sc->bd_dmask.release();
    sc->bd_nconn = 0;
}

// Device bytes the scene's batch buffers hold.
size_t bd_held_bytes(const lmx_scene* sc)
{
    const size_t B = sc->bd_batch, nv = B * sc->bd_maxv * 4;
    return 2 * (nv * 16 + B * 4 + 4 * B * 16 + 2 * B * 16 + 2 * B * 16) + B * sc->bd_maxv * 4 + B * 4 + B * 8 +
           2 * B * 16 + (B + 1) * 4 + B * 4 + (size_t)sc->bd_nconn * (2 * 16 + 4 * 16);
}

// Buffers for batches of B samples of up to maxv vertices and `conn`
// shadow-queue slots per sample.  With `rounds` (the samplers that connect
// in rounds), conn is a sample's worst case and the queue holds one round:
// B x bd_round_slots(conn) slots, never fewer than one sample's worst case.
int bd_buffers(lmx_scene* sc, uint32_t B, uint32_t maxv, uint32_t conn, bool rounds = false)
{
    const size_t nconn = rounds ? std::max((size_t)B * bd_round_slots(conn), (size_t)conn) : (size_t)B * conn;
    if (nconn > 0xffffffffull) return set_error(LMX_EINVAL, "bdpt buffers: more than 2^32 - 1 connection slots");
    if (B <= sc->bd_batch && maxv <= sc->bd_maxv && nconn <= sc->bd_nconn) return LMX_OK;
    bd_release(sc);
    hipError_t e;
    const size_t nv = (size_t)B * maxv * 4;
    // a failure frees whatever was allocated so far: the scene then holds no
    // batch buffers (and reallocates them on its next render)
    auto fail = [&](hipError_t err, const char* what) { bd_release(sc); return hip_error(err, what); };
    for (int k = 0; k < 2; k++) {
        if ((e = sc->bd_vert[k].alloc(nv)) != hipSuccess) return fail(e, "bdpt vertex buffers");
        if ((e = sc->bd_nsub[k].alloc(B)) != hipSuccess) return fail(e, "bdpt buffers");
        if ((e = sc->bd_ray[k].alloc(4 * (size_t)B)) != hipSuccess) return fail(e, "bdpt queues");
        if ((e = sc->bd_thr[k].alloc(2 * (size_t)B)) != hipSuccess) return fail(e, "bdpt queues");
        if ((e = sc->bd_meta[k].alloc(2 * (size_t)B)) != hipSuccess) return fail(e, "bdpt queues");
    }
    if ((e = sc->bd_V.alloc((size_t)B * maxv)) != hipSuccess || (e = sc->bd_epix.alloc(B)) != hipSuccess ||
        (e = sc->bd_eraster.alloc(B)) != hipSuccess ||
        (e = sc->bd_hits.alloc(2 * (size_t)B)) != hipSuccess || (e = sc->bd_sray.alloc(2 * nconn)) != hipSuccess ||
        (e = sc->bd_sc.alloc(nconn)) != hipSuccess || (e = sc->bd_saux.alloc(nconn)) != hipSuccess ||
        (e = sc->bd_shits.alloc(nconn)) != hipSuccess || (e = sc->bd_smeta.alloc(nconn)) != hipSuccess ||
        (e = sc->bd_off.alloc((size_t)B + 1)) != hipSuccess || (e = sc->bd_dmask.alloc(B)) != hipSuccess ||
        (e = sc->bd_part.alloc(((size_t)B + SCAN_TILE) / SCAN_TILE)) != hipSuccess)
        return fail(e, "bdpt buffers");
    if (!sc->bd_ctl.p && (e = sc->bd_ctl.alloc(1)) != hipSuccess) return hip_error(e, "bdpt control block");
    if (!sc->bd_hctl && (e = hipHostMalloc((void**)&sc->bd_hctl, sizeof(BCtl), hipHostMallocDefault)) != hipSuccess)
        return hip_error(e, "bdpt pinned control block");
    sc->bd_batch = B;
    sc->bd_maxv = maxv;
    sc->bd_conn = conn;
    sc->bd_nconn = (uint32_t)nconn;
    return LMX_OK;
}

// Records one connection round: the next range of samples whose slots fit
// the shadow queue, their slots, the connection terms, the any-hit trace of
// the shadow rays, and the splat of the unoccluded ones.  A round with no
// sample left launches kernels that find nothing to do.
int bd_enqueue_round(lmx_scene* sc, const BParams& bp, const VB& vb, hipStream_t st)
{
    BCtl* ctl = sc->bd_ctl.p;
    const int cap_grid = sc->cus * 8;
    const uint32_t nconn = sc->bd_nconn;
    ShadowQ sq{sc->bd_sray.p, sc->bd_sc.p, sc->bd_saux.p, sc->bd_smeta.p};
    k_bd_round<<<1, 64, 0, st>>>(ctl, vb.off, bp.batch, nconn);
    k_bd_slots<<<cap_grid, 256, 0, st>>>(sc->rs, bp, ctl, vb, sq);
    k_bd_connect<<<cap_grid, 256, 0, st>>>(sc->rs, bp, ctl, vb, sq);
    if (bp.ray_out) {
        const uint32_t half = bp.ray_out_cap / 2;
        k_bd_capture<<<cap_grid, 256, 0, st>>>(ctl, 1, sq.ray, bp.ray_out, half, bp.ray_out_cap);
        k_bd_capture_advance<<<1, 64, 0, st>>>(ctl, 1, bp.ray_out_cap - half);
    }
    int rc = launch_trace(sc->accel, reinterpret_cast<const WRay*>(sq.ray), 0, &ctl->n_shadow, nconn,
                          reinterpret_cast<WHit*>(sc->bd_shits.p), st, true);
    if (rc != LMX_OK) return rc;
    k_bd_splat<<<std::max(1, std::min((int)((nconn + 255) / 256), cap_grid)), 256, 0, st>>>(bp, ctl, sq, sc->bd_shits.p);
    return LMX_OK;
}

// Connection rounds a batch graph holds: one when a round can take every
// sample at its worst case, else two, after which the host checks for
// samples left over (rarely any: the round slots exceed the typical count).
int bd_graph_rounds(const lmx_scene* sc, const BParams& bp)
{
    const uint32_t worst = bd_conn_per_sample((uint32_t)bp.maxv, bp.variant, bp.bounded != 0, bp.connect != 0);
    return (uint64_t)sc->bd_nconn >= (uint64_t)bp.batch * worst ? 1 : 2;
}

// Records one batch: generation, subpath iterations, connections, shadow
// rays, splats, advance to the next batch.
int bd_enqueue_batch(lmx_scene* sc, const BParams& bp, hipStream_t st)
{
    const uint32_t B = bp.batch;
    VB vb;
    vb.v[0] = sc->bd_vert[0].p; vb.v[1] = sc->bd_vert[1].p;
    vb.nsub[0] = sc->bd_nsub[0].p; vb.nsub[1] = sc->bd_nsub[1].p;
    vb.epix = sc->bd_epix.p; vb.eraster = sc->bd_eraster.p;
    vb.V = sc->bd_V.p;
    vb.mq = sc->vc_mq.p;
    vb.off = sc->bd_off.p;
    vb.dmask = sc->bd_dmask.p;
    Queue q[2];
    for (int k = 0; k < 2; k++) { q[k].ray = sc->bd_ray[k].p; q[k].thr = sc->bd_thr[k].p; q[k].meta = sc->bd_meta[k].p; }
    BCtl* ctl = sc->bd_ctl.p;
    const int cap_grid = sc->cus * 8;
    k_bd_gen<<<(B + 255) / 256, 256, 0, st>>>(sc->rs, bp, ctl, vb, q[0]);
    int a = 0;
    const uint32_t half = bp.ray_out_cap / 2;
    for (int it = 0; it < bp.maxv - 1; it++) {
        if (bp.ray_out) {
            k_bd_capture<<<cap_grid, 256, 0, st>>>(ctl, 0, q[a].ray, bp.ray_out, 0u, half);
            k_bd_capture_advance<<<1, 64, 0, st>>>(ctl, 0, half);
        }
        int rc = launch_trace(sc->accel, reinterpret_cast<const WRay*>(q[a].ray), 0, &ctl->n_trace, 2 * B,
                              reinterpret_cast<WHit*>(sc->bd_hits.p), st, false);
        if (rc != LMX_OK) return rc;
        k_bd_shade<<<(2 * B + 255) / 256, 256, 0, st>>>(sc->rs, bp, ctl, q[a], sc->bd_hits.p, q[1 - a], vb);
        k_bd_commit<<<1, 64, 0, st>>>(ctl);
        a = 1 - a;
    }
    k_bd_plan<<<(B + 255) / 256, 256, 0, st>>>(sc->rs, bp, ctl, vb);
    // connection counts -> slot offsets; off[B] becomes the batch's total
    const uint32_t nt = (B + 1 + SCAN_TILE - 1) / SCAN_TILE;
    k_scan_tiles<<<nt, 256, 0, st>>>(vb.off, B + 1, sc->bd_part.p);
    if (nt > 1) {
        k_scan_part<<<1, 256, 0, st>>>(sc->bd_part.p, nt);
        k_scan_add<<<std::max(1, std::min((int)((B + 1 + 255) / 256), cap_grid)), 256, 0, st>>>(vb.off, B + 1, sc->bd_part.p);
    }
    k_bd_conn_start<<<1, 64, 0, st>>>(ctl);
    const int rounds = bd_graph_rounds(sc, bp);
    for (int k = 0; k < rounds; k++) {
        const int rc = bd_enqueue_round(sc, bp, vb, st);
        if (rc != LMX_OK) return rc;
    }
    if (bp.merging) {
        VB ph = vb;
        ph.v[1] = sc->vc_vert.p; ph.nsub[1] = sc->vc_nsub.p;
        const Grid g{sc->vc_items.p, sc->vc_start.p, sc->vc_cursor.p, sc->vc_part.p, sc->vc_H};
        const size_t nm = (size_t)B * (size_t)(bp.maxv - 1);
        k_vc_merge<<<std::max(1, (int)std::min<size_t>((nm + 255) / 256, (size_t)cap_grid * 2)), 256, 0, st>>>(
            sc->rs, bp, ctl, vb, ph, g);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? LMX_OK : hip_error(e, "bdpt launch");
}

// Captures one connection round as a graph.
int bd_round_graph(lmx_scene* sc, const BParams& bp, hipStream_t st, hipGraphExec_t* exec)
{
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_error(e, "capture");
    VB vb;
    vb.v[0] = sc->bd_vert[0].p; vb.v[1] = sc->bd_vert[1].p;
    vb.nsub[0] = sc->bd_nsub[0].p; vb.nsub[1] = sc->bd_nsub[1].p;
    vb.epix = sc->bd_epix.p; vb.eraster = sc->bd_eraster.p;
    vb.V = sc->bd_V.p;
    vb.mq = sc->vc_mq.p;
    vb.off = sc->bd_off.p;
    vb.dmask = sc->bd_dmask.p;
    int rc = bd_enqueue_round(sc, bp, vb, st);
    if (rc == LMX_OK && (e = hipGetLastError()) != hipSuccess) rc = hip_error(e, "bdpt launch");
    e = hipStreamEndCapture(st, &graph);
    if (rc == LMX_OK && e != hipSuccess) rc = hip_error(e, "capture end");
    if (rc == LMX_OK && (e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0)) != hipSuccess)
        rc = hip_error(e, "graph instantiate");
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
}

// The graphs of one batch shape: the batch itself (with its in-graph
// connection rounds) and, when those rounds may not be enough, one more
// round to replay until every sample is connected.
struct BGraphs {
    hipGraphExec_t batch = nullptr, round = nullptr;
    uint32_t B = 0;
    ~BGraphs()
    {
        if (batch) (void)hipGraphExecDestroy(batch);
        if (round) (void)hipGraphExecDestroy(round);
    }
};

// Captures one batch as a graph (and the extra round when needed).
int bd_graph(lmx_scene* sc, const BParams& bp, hipStream_t st, BGraphs& g)
{
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_error(e, "capture");
    int rc = bd_enqueue_batch(sc, bp, st);
    e = hipStreamEndCapture(st, &graph);
    if (rc == LMX_OK && e != hipSuccess) rc = hip_error(e, "capture end");
    if (rc == LMX_OK && (e = hipGraphInstantiate(&g.batch, graph, nullptr, nullptr, 0)) != hipSuccess)
        rc = hip_error(e, "graph instantiate");
    if (graph) (void)hipGraphDestroy(graph);
    g.B = bp.batch;
    if (rc == LMX_OK && bd_graph_rounds(sc, bp) > 1) rc = bd_round_graph(sc, bp, st, &g.round);
    return rc;
}

// Runs one batch: its graph; then, when its rounds may have left samples
// unconnected, a read of the round cursor and further rounds until none is
// left; then the advance to the next batch.
int bd_run_batch(lmx_scene* sc, const BGraphs& g, hipStream_t st)
{
    hipError_t e = hipGraphLaunch(g.batch, st);
    if (e != hipSuccess) return hip_error(e, "graph launch");
    while (g.round) {
        if ((e = hipMemcpyAsync(&sc->bd_hctl->conn_next, &sc->bd_ctl.p->conn_next, sizeof(uint32_t),
                                hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess)
            return hip_error(e, "connection rounds");
        if (sc->bd_hctl->conn_next >= g.B) break;
        if ((e = hipGraphLaunch(g.round, st)) != hipSuccess) return hip_error(e, "graph launch");
    }
    k_bd_next_batch<<<1, 64, 0, st>>>(sc->bd_ctl.p, g.B);
    e = hipGetLastError();
    return e == hipSuccess ? LMX_OK : hip_error(e, "bdpt launch");
}

// autotuner key: variant, max_num_vertices, bounded
uint32_t bd_key(const BParams& bp)
{
    return 1u + (uint32_t)bp.maxv + 64u * (uint32_t)bp.variant + 128u * (uint32_t)bp.bounded;
}

// Batch-size autotuner: times one batch of each candidate size on this scene
// (after one untimed batch that warms the graph and the caches) and keeps the
// size with the highest samples/s for this max_num_vertices.  Candidates are
// bounded by half of the free device memory (what this scene's batch buffers
// hold counts as free) and by the render's own sample count rounded up to a
// power of two.  A half rather than a quarter lets unbounded renderer::bdpt
// (105 KB per sample) reach 2^20 samples per batch: +49 %.  A candidate whose
// buffers or graph cannot be made ends the probing and the best size so far
// is kept.  Afterwards the scene holds buffers for the chosen size only.
// The probe batches splat into `film`, which the caller clears afterwards.
int bd_autotune(lmx_scene* sc, BParams bp, uint64_t num_samples, float* film, hipStream_t st)
{
    // 2^23 and 2^24 (110 GB at max_num_vertices 10) lift renderer::bdpt on a
    // 1024^2 x 16 spp frame from 342 to 384 and 403 Msamples/s
    static const uint32_t cands[] = {1u << 19, 1u << 20, 1u << 21, 1u << 22, 1u << 23, 1u << 24};
    size_t free_b = 0, total_b = 0;
    hipError_t e = hipMemGetInfo(&free_b, &total_b);
    if (e != hipSuccess) return hip_error(e, "bdpt autotune: memory info");
    const size_t budget = (free_b + bd_held_bytes(sc)) / 2;
    const uint32_t conn = bd_conn_per_sample((uint32_t)bp.maxv, bp.variant, bp.bounded != 0);
    const size_t per = bd_bytes_per_sample((uint32_t)bp.maxv, conn);
    bp.sample_out = nullptr;
    bp.ray_out = nullptr;
    bp.ray_out_cap = 0;
    bp.table = nullptr;                        // probes draw from the generator
    bp.film = film;
    hipEvent_t e0, e1;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return hip_error(e, "autotune events");
    if ((e = hipEventCreate(&e1)) != hipSuccess) { (void)hipEventDestroy(e0); return hip_error(e, "autotune events"); }
    uint32_t best = cands[0];
    double best_rate = -1.0;
    int rc = LMX_OK;
    uint64_t want = 1;
    while (want < num_samples) want <<= 1;
    for (uint32_t B : cands) {
        if (B != cands[0] && ((size_t)B * per > budget || B > want)) break;
        bp.batch = B;
        bp.total = 2u * B;
        BGraphs g;
        if ((rc = bd_buffers(sc, B, (uint32_t)bp.maxv, conn, true)) != LMX_OK ||
            (rc = bd_graph(sc, bp, st, g)) != LMX_OK) {
            if (best_rate < 0.0) break;        // not even the smallest size fits: report it
            (void)hipGetLastError();           // a larger candidate failed: keep the best so far
            rc = LMX_OK;
            break;
        }
        k_bd_init<<<1, 64, 0, st>>>(sc->bd_ctl.p);
        if ((rc = bd_run_batch(sc, g, st)) != LMX_OK) break;
        (void)hipEventRecord(e0, st);
        rc = bd_run_batch(sc, g, st);
        (void)hipEventRecord(e1, st);
        const hipError_t es = hipEventSynchronize(e1);
        if (rc != LMX_OK) break;
        if (es != hipSuccess) { rc = hip_error(es, "bdpt autotune"); break; }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double rate = B / (double)std::max(ms, 1e-3f);
        if (rate > best_rate) { best_rate = rate; best = B; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == LMX_OK) {
        sc->bd_tuned_batch = best;
        sc->bd_tuned_key = bd_key(bp);
        sc->bd_tuned_limit = want;
        if (sc->bd_batch > best) bd_release(sc);   // the render allocates the chosen size
    }
    return rc;
}

// One render of either variant: batches of B samples, each replayed from one
// captured graph; the film scaled by W H / N at the end.
int render_bidir(lmx_scene* sc, const lmx_render_params* p, int variant, const char* name, float* film_rgb,
                 void* stream, lmx_render_stats* stats)
{
    char msg[160];
    if (!sc || !p || !film_rgb) return set_error(LMX_EINVAL, (std::string(name) + ": NULL argument").c_str());
    if (p->num_samples == 0 || p->num_samples > 0xffffffffull)
        return set_error(LMX_EINVAL, (std::string(name) + ": num_samples must be in [1, 2^32)").c_str());
    const bool bounded = p->max_num_vertices != -1;
    if (bounded && (p->max_num_vertices < 2 || p->max_num_vertices > 32)) {
        snprintf(msg, sizeof msg, "%s: max_num_vertices must be in [2, 32]%s", name,
                 variant == BD_BDPT ? " or -1" : "");
        return set_error(LMX_EINVAL, msg);
    }
    if (!bounded && variant == BD_VCM)
        return set_error(LMX_EINVAL, (std::string(name) + ": max_num_vertices must be in [2, 32]").c_str());
    if (variant == BD_BDPT && p->mis != LMX_MIS_POWER && p->mis != LMX_MIS_SIMPLE)
        return set_error(LMX_EINVAL, (std::string(name) + ": unknown mis").c_str());
    if (sc->mats.n > VB_MAX_MATERIALS || sc->lights.n > VB_MAX_LIGHTS)
        return set_error(LMX_EINVAL, (std::string(name) + ": at most 65536 materials and 8190 lights").c_str());
    const uint32_t maxv = bounded ? (uint32_t)p->max_num_vertices : 32u;
    const uint32_t wsub = LMX_PT_SLOT_SEG0 + (variant == BD_BDPT ? 8u : 4u) * (maxv - 1u);
    if (p->uniform_table && p->table_width < 2 * wsub) {
        snprintf(msg, sizeof msg, "%s: uniform table narrower than %u", name, 2 * wsub);
        return set_error(LMX_EINVAL, msg);
    }
    if (p->raster_mode != LMX_RASTER_RANDOM && p->raster_mode != LMX_RASTER_STRATIFIED)
        return set_error(LMX_EINVAL, (std::string(name) + ": unknown raster_mode").c_str());
    const bool timed = p->render_time > 0.0;
    if (timed && (p->uniform_table || p->sample_out))
        return set_error(LMX_EINVAL, (std::string(name) + ": render_time takes neither a uniform table nor sample_out").c_str());
    hipStream_t st = (hipStream_t)stream;
    hipStream_t own = nullptr;
    hipError_t e;
    if (!st) {
        if ((e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking)) != hipSuccess) return hip_error(e, "stream create");
        st = own;
    }
    const auto t0 = std::chrono::steady_clock::now();
    // time mode: batches of the autotuned size for the largest candidate,
    // started until the deadline (at most 2^32 - 1 samples)
    const uint64_t num_samples = timed ? 0xffffffffull : p->num_samples;
    BParams bp;
    bp.variant = variant;
    bp.merging = 0; bp.connect = 1; bp.phase = 0; bp.c_conn = 1.0f;
    bp.mis = variant == BD_BDPT ? p->mis : LMX_MIS_POWER;
    bp.bounded = bounded ? 1 : 0;
    bp.maxv = (int)maxv; bp.minv = p->min_num_vertices;
    bp.key = make_uint2((uint32_t)p->seed, (uint32_t)(p->seed >> 32));
    bp.total = (uint32_t)num_samples;
    bp.raster_mode = p->raster_mode;
    bp.table = p->uniform_table; bp.table_width = p->table_width;
    bp.wsub = wsub;
    bp.sample_out = p->sample_out;
    bp.film = film_rgb;
    bp.ray_out = p->ray_out && p->ray_out_cap >= 2 ? reinterpret_cast<float4*>(p->ray_out) : nullptr;
    bp.ray_out_cap = bp.ray_out ? p->ray_out_cap : 0u;
    int rc = LMX_OK;
    // batch size: the caller's, else the autotuned one (tuned once per scene,
    // variant and max_num_vertices, for renders of more than the smallest
    // candidate)
    uint32_t B = p->queue_capacity;
    if (!B) {
        if (num_samples <= (1u << 19)) B = (uint32_t)num_samples;
        else {
            // tuned again when an earlier, smaller render capped the candidates
            if (sc->bd_tuned_key != bd_key(bp) ||
                (sc->bd_tuned_limit < (1ull << 24) && sc->bd_tuned_limit < num_samples))
                rc = bd_autotune(sc, bp, num_samples, film_rgb, st);
            B = sc->bd_tuned_batch;
        }
    }
    B = (uint32_t)std::min<uint64_t>(B, num_samples);
    bp.batch = B;
    if (rc == LMX_OK) rc = bd_buffers(sc, B, maxv, bd_conn_per_sample(maxv, variant, bounded), true);
    const size_t npix = (size_t)sc->width * (size_t)sc->height;
    const int zgrid = std::max(1, std::min((int)((3 * npix + 255) / 256), sc->cus * 8));
    BGraphs g;
    if (rc == LMX_OK) {
        k_bd_zero<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix);
        k_bd_init<<<1, 64, 0, st>>>(sc->bd_ctl.p);
        rc = bd_graph(sc, bp, st, g);
    }
    uint64_t nbatch = (num_samples + B - 1) / B;
    uint64_t taken = p->num_samples;
    if (!timed) {
        for (uint64_t b = 0; rc == LMX_OK && b < nbatch; b++) rc = bd_run_batch(sc, g, st);
    } else {
        // the scheduler's time mode: whole batches until the deadline, each
        // finished before the clock is read (scheduler.cpp:108,191-199)
        const auto ts = std::chrono::steady_clock::now();
        uint64_t b = 0;
        while (rc == LMX_OK && b < nbatch) {
            if ((rc = bd_run_batch(sc, g, st)) != LMX_OK) break;
            if ((e = hipStreamSynchronize(st)) != hipSuccess) {
                rc = hip_error(e, "graph launch");
                break;
            }
            b++;
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count() >= p->render_time) break;
        }
        nbatch = b;
        taken = b * (uint64_t)B;
    }
    if (rc == LMX_OK) {
        k_bd_scale<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix, (float)((double)npix / (double)std::max<uint64_t>(taken, 1)));
        if ((e = hipMemcpyAsync(sc->bd_hctl, sc->bd_ctl.p, sizeof(BCtl), hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess)
            rc = hip_error(e, (std::string(name) + ": sync").c_str());
    }
    if (own) (void)hipStreamDestroy(own);
    if (stats) {
        stats->rays = rc == LMX_OK ? sc->bd_hctl->rays + sc->bd_hctl->shadow_rays : 0;
        stats->iterations = (uint32_t)nbatch;
        stats->queue_capacity = B;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->samples = rc == LMX_OK ? taken : 0;
    }
    return rc;
}

// Photon-subpath, photon-queue and grid buffers for np photons of up to
// maxv vertices (per photon: maxv vertex records of 64 B, maxv - 1 grid
// records of 16 B, 144 B of ray queues); kept when large enough, else
// replaced (the old ones freed first).
int vc_buffers(lmx_scene* sc, uint32_t np, uint32_t maxv, uint32_t B)
{
    if ((size_t)B * (maxv - 1) > sc->vc_mq.n) {
        hipError_t e = sc->vc_mq.alloc((size_t)B * (maxv - 1));
        if (e != hipSuccess) return hip_error(e, "vcm merge queue");
    }
    if (np <= sc->vc_np && maxv <= sc->vc_maxv) return LMX_OK;
    sc->vc_np = sc->vc_maxv = sc->vc_H = 0;
    sc->vc_vert.release(); sc->vc_nsub.release(); sc->vc_items.release();
    sc->vc_start.release(); sc->vc_cursor.release(); sc->vc_part.release();
    for (int k = 0; k < 2; k++) { sc->vc_ray[k].release(); sc->vc_thr[k].release(); sc->vc_meta[k].release(); }
    sc->vc_hits.release();
    const size_t items = (size_t)np * (maxv - 1);
    uint32_t H = 1024;
    while (H < items && H < (1u << 25)) H <<= 1;
    hipError_t e;
    auto fail = [&](hipError_t err, const char* what) {
        sc->vc_vert.release(); sc->vc_nsub.release(); sc->vc_items.release();
        sc->vc_start.release(); sc->vc_cursor.release(); sc->vc_part.release();
        for (int k = 0; k < 2; k++) { sc->vc_ray[k].release(); sc->vc_thr[k].release(); sc->vc_meta[k].release(); }
        sc->vc_hits.release();
        return hip_error(err, what);
    };
    if ((e = sc->vc_vert.alloc((size_t)np * maxv * 4)) != hipSuccess || (e = sc->vc_nsub.alloc(np)) != hipSuccess ||
        (e = sc->vc_items.alloc(items)) != hipSuccess || (e = sc->vc_start.alloc((size_t)H + 1)) != hipSuccess ||
        (e = sc->vc_cursor.alloc(H)) != hipSuccess ||
        (e = sc->vc_part.alloc((H + SCAN_TILE) / SCAN_TILE)) != hipSuccess ||
        (e = sc->vc_hits.alloc(np)) != hipSuccess)
        return fail(e, "photon buffers");
    for (int k = 0; k < 2; k++)
        if ((e = sc->vc_ray[k].alloc(2 * (size_t)np)) != hipSuccess || (e = sc->vc_thr[k].alloc(np)) != hipSuccess ||
            (e = sc->vc_meta[k].alloc(np)) != hipSuccess)
            return fail(e, "photon queues");
    sc->vc_np = np;
    sc->vc_maxv = maxv;
    sc->vc_H = H;
    return LMX_OK;
}

// Records one photon phase: the photon subpaths (light subpaths traced by the
// gen / trace / shade kernels into the photon vertex buffer) and the hash
// grid over their vertices.
int vc_enqueue_photons(lmx_scene* sc, const BParams& pp, hipStream_t st, bool grid = true)
{
    const uint32_t np = pp.batch;
    VB ph;
    ph.v[0] = sc->bd_vert[0].p; ph.v[1] = sc->vc_vert.p;
    ph.nsub[0] = sc->bd_nsub[0].p; ph.nsub[1] = sc->vc_nsub.p;
    ph.epix = sc->bd_epix.p; ph.eraster = sc->bd_eraster.p;
    ph.V = sc->bd_V.p;
    ph.mq = nullptr;
    // the photon phase's own queues: one ray per photon subpath
    Queue q[2];
    for (int k = 0; k < 2; k++) { q[k].ray = sc->vc_ray[k].p; q[k].thr = sc->vc_thr[k].p; q[k].meta = sc->vc_meta[k].p; }
    BCtl* ctl = sc->bd_ctl.p;
    const int cap_grid = sc->cus * 8;
    k_bd_gen<<<(np + 255) / 256, 256, 0, st>>>(sc->rs, pp, ctl, ph, q[0]);
    int a = 0;
    for (int it = 0; it < pp.maxv - 1; it++) {
        int rc = launch_trace(sc->accel, reinterpret_cast<const WRay*>(q[a].ray), 0, &ctl->n_trace, np,
                              reinterpret_cast<WHit*>(sc->vc_hits.p), st, false);
        if (rc != LMX_OK) return rc;
        k_bd_shade<<<(np + 255) / 256, 256, 0, st>>>(sc->rs, pp, ctl, q[a], sc->vc_hits.p, q[1 - a], ph);
        k_bd_commit<<<1, 64, 0, st>>>(ctl);
        a = 1 - a;
    }
    if (!grid) {
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? LMX_OK : hip_error(e, "photon launch");
    }
    const Grid g{sc->vc_items.p, sc->vc_start.p, sc->vc_cursor.p, sc->vc_part.p, sc->vc_H};
    const uint32_t n1 = g.H + 1u, ntile = (n1 + SCAN_TILE - 1) / SCAN_TILE;
    const size_t nslot = (size_t)np * (size_t)(pp.maxv - 1);
    const int fgrid = std::max(1, (int)std::min<size_t>((nslot + 255) / 256, (size_t)cap_grid));
    k_vc_grid_clear<<<std::max(1, std::min((int)((n1 + 255) / 256), cap_grid)), 256, 0, st>>>(g);
    k_vc_grid_fill<0><<<fgrid, 256, 0, st>>>(sc->rs, pp, ctl, ph, g);
    k_scan_tiles<<<ntile, 256, 0, st>>>(g.start, n1, g.part);
    k_scan_part<<<1, 256, 0, st>>>(g.part, ntile);
    k_scan_add<<<std::max(1, std::min((int)((n1 + 255) / 256), cap_grid)), 256, 0, st>>>(g.start, n1, g.part);
    k_vc_grid_fill<1><<<fgrid, 256, 0, st>>>(sc->rs, pp, ctl, ph, g);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? LMX_OK : hip_error(e, "vcm photon launch");
}

// Captures a phase (photons, or one eye batch) as a graph.
template <typename F>
int capture(hipStream_t st, hipGraphExec_t* exec, F&& record)
{
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_error(e, "capture");
    int rc = record();
    e = hipStreamEndCapture(st, &graph);
    if (rc == LMX_OK && e != hipSuccess) rc = hip_error(e, "capture end");
    if (rc == LMX_OK && (e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0)) != hipSuccess)
        rc = hip_error(e, "graph instantiate");
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
}

// renderer::vcm, modes vcm and bdpm: per pass the photon graph, then the eye
// batches (each one graph: subpaths, connections, shadow rays, merges); the
// film holds the sum of every pass's splats, scaled at the end by
// W H / (N_eye passes) -- the mean of the per-pass images that the
// reference's running rescale (renderer_vcm.cpp:752-757) keeps.
int render_vcm(lmx_scene* sc, const lmx_render_params* p, const lmx_vcm_params* v, float* film_rgb, void* stream,
               lmx_render_stats* stats)
{
    const char* name = "lmx_render_vcm";
    char msg[160];
    if (!sc || !p || !v || !film_rgb) return set_error(LMX_EINVAL, "lmx_render_vcm: NULL argument");
    if (p->num_samples == 0 || p->num_samples > 0xffffffffull)
        return set_error(LMX_EINVAL, "lmx_render_vcm: num_samples must be in [1, 2^32)");
    if (p->max_num_vertices < 2 || p->max_num_vertices > 32)
        return set_error(LMX_EINVAL, "lmx_render_vcm: max_num_vertices must be in [2, 32]");
    if (v->mode != LMX_VCM_MODE_VCM && v->mode != LMX_VCM_MODE_BDPM)
        return set_error(LMX_EINVAL, "lmx_render_vcm: unknown mode");
    if (v->num_iteration_pass == 0) return set_error(LMX_EINVAL, "lmx_render_vcm: num_iteration_pass must be >= 1");
    if (v->num_photon_trace_samples == 0 || v->num_photon_trace_samples > (1u << 26))
        return set_error(LMX_EINVAL, "lmx_render_vcm: num_photon_trace_samples must be in [1, 2^26]");
    if (!(v->initial_radius > 0.0f) || !(v->initial_radius < 3.0e38f) || !(v->alpha > 0.0f) ||
        !(v->alpha < 3.0e38f))
        return set_error(LMX_EINVAL, "lmx_render_vcm: initial_radius and alpha must be positive and finite");
    if ((uint64_t)v->num_iteration_pass * p->num_samples > 0xffffffffull ||
        (uint64_t)v->num_iteration_pass * v->num_photon_trace_samples > 0xffffffffull)
        return set_error(LMX_EINVAL, "lmx_render_vcm: passes x samples must stay below 2^32");
    if ((p->uniform_table || v->photon_table || p->sample_out) && v->num_iteration_pass != 1)
        return set_error(LMX_EINVAL, "lmx_render_vcm: uniform tables and sample_out need num_iteration_pass 1");
    if (sc->mats.n > VB_MAX_MATERIALS || sc->lights.n > VB_MAX_LIGHTS)
        return set_error(LMX_EINVAL, "lmx_render_vcm: at most 65536 materials and 8190 lights");
    if (p->raster_mode != LMX_RASTER_RANDOM && p->raster_mode != LMX_RASTER_STRATIFIED)
        return set_error(LMX_EINVAL, "lmx_render_vcm: unknown raster_mode");
    const uint32_t maxv = (uint32_t)p->max_num_vertices;
    const uint32_t wsub = LMX_PT_SLOT_SEG0 + 4u * (maxv - 1u);
    if ((p->uniform_table && p->table_width < 2 * wsub) || (v->photon_table && v->photon_table_width < 2 * wsub)) {
        snprintf(msg, sizeof msg, "%s: uniform table narrower than %u", name, 2 * wsub);
        return set_error(LMX_EINVAL, msg);
    }
    hipStream_t st = (hipStream_t)stream;
    hipStream_t own = nullptr;
    hipError_t e;
    if (!st) {
        if ((e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking)) != hipSuccess) return hip_error(e, "stream create");
        st = own;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const bool connect = v->mode == LMX_VCM_MODE_VCM;
    const uint32_t Ne = (uint32_t)p->num_samples, np = v->num_photon_trace_samples;
    BParams bp;
    bp.variant = BD_VCM;
    bp.merging = 1; bp.connect = connect ? 1 : 0; bp.phase = 0; bp.c_conn = connect ? 1.0f : 0.0f;
    bp.mis = LMX_MIS_POWER;
    bp.bounded = 1;
    bp.maxv = (int)maxv; bp.minv = p->min_num_vertices;
    bp.key = make_uint2((uint32_t)p->seed, (uint32_t)(p->seed >> 32));
    bp.total = Ne;
    bp.raster_mode = p->raster_mode;
    bp.table = p->uniform_table; bp.table_width = p->table_width;
    bp.wsub = wsub;
    bp.sample_out = p->sample_out;
    bp.film = film_rgb;
    uint32_t B = p->queue_capacity ? p->queue_capacity : (1u << 21);   // a 1024^2 x 2 spp pass in one batch
    B = std::min(B, Ne);
    bp.batch = B;
    BParams pp = bp;                     // the photon phase
    pp.phase = 1;
    pp.batch = np; pp.total = np;
    pp.key = make_uint2((uint32_t)v->photon_seed, (uint32_t)(v->photon_seed >> 32));
    pp.table = v->photon_table; pp.table_width = v->photon_table_width;
    pp.sample_out = nullptr;
    int rc = bd_buffers(sc, B, maxv, bd_conn_per_sample(maxv, BD_VCM, true, connect), true);
    if (rc == LMX_OK) rc = vc_buffers(sc, np, maxv, B);
    const size_t npix = (size_t)sc->width * (size_t)sc->height;
    const int zgrid = std::max(1, std::min((int)((3 * npix + 255) / 256), sc->cus * 8));
    hipGraphExec_t gph = nullptr;
    BGraphs geye;
    if (rc == LMX_OK) {
        k_bd_zero<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix);
        if (p->sample_out) {
            const size_t ns = (size_t)Ne * 16 * (size_t)(maxv + 1) * (size_t)(maxv + 1);
            k_bd_zero<<<std::max(1, (int)std::min<size_t>((ns + 255) / 256, (size_t)sc->cus * 8)), 256, 0, st>>>(
                p->sample_out, ns);
        }
        k_bd_init<<<1, 64, 0, st>>>(sc->bd_ctl.p);
        rc = capture(st, &gph, [&] { return vc_enqueue_photons(sc, pp, st); });
    }
    if (rc == LMX_OK) rc = bd_graph(sc, bp, st, geye);
    const uint64_t nbatch = ((uint64_t)Ne + B - 1) / B;
    double r = v->initial_radius;
    for (uint32_t q = 0; rc == LMX_OK && q < v->num_iteration_pass; q++) {
        if (q > 0) r = std::sqrt((v->alpha + (double)q) / (1.0 + (double)q)) * r;
        const double a = 3.14159265358979323846 * r * r;
        const float rf = (float)r, cell = (float)(2.0 * r), cm = (float)((double)np * a * a);
        k_bd_phase<<<1, 64, 0, st>>>(sc->bd_ctl.p, q * np, rf, cell, cm);
        if ((e = hipGraphLaunch(gph, st)) != hipSuccess) { rc = hip_error(e, "graph launch"); break; }
        k_bd_phase<<<1, 64, 0, st>>>(sc->bd_ctl.p, q * Ne, rf, cell, cm);
        for (uint64_t b = 0; rc == LMX_OK && b < nbatch; b++) rc = bd_run_batch(sc, geye, st);
    }
    if (rc == LMX_OK) {
        k_bd_scale<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix,
                                          (float)((double)npix / ((double)Ne * (double)v->num_iteration_pass)));
        if ((e = hipMemcpyAsync(sc->bd_hctl, sc->bd_ctl.p, sizeof(BCtl), hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess)
            rc = hip_error(e, "lmx_render_vcm: sync");
    }
    if (gph) (void)hipGraphExecDestroy(gph);
    if (own) (void)hipStreamDestroy(own);
    if (stats) {
        stats->rays = rc == LMX_OK ? sc->bd_hctl->rays + sc->bd_hctl->shadow_rays : 0;
        stats->iterations = v->num_iteration_pass;
        stats->queue_capacity = B;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->samples = rc == LMX_OK ? p->num_samples * v->num_iteration_pass : 0;
        stats->dropped = rc == LMX_OK ? sc->bd_hctl->dropped : 0;
    }
    return rc;
}

// renderer::pm's eye batch: eye endpoints, subpaths up to the gather
// vertex, the gather kernel, advance.
int pm_enqueue_eye(lmx_scene* sc, const BParams& bp, const PMArgs& pa, hipStream_t st)
{
    const uint32_t B = bp.batch;
    VB vb;
    vb.v[0] = sc->bd_vert[0].p; vb.v[1] = sc->bd_vert[1].p;
    vb.nsub[0] = sc->bd_nsub[0].p; vb.nsub[1] = sc->bd_nsub[1].p;
    vb.epix = sc->bd_epix.p; vb.eraster = sc->bd_eraster.p;
    vb.V = sc->bd_V.p;
    vb.mq = nullptr;
    VB ph = vb;
    ph.v[1] = sc->vc_vert.p; ph.nsub[1] = sc->vc_nsub.p;
    Queue q[2];
    for (int k = 0; k < 2; k++) { q[k].ray = sc->bd_ray[k].p; q[k].thr = sc->bd_thr[k].p; q[k].meta = sc->bd_meta[k].p; }
    BCtl* ctl = sc->bd_ctl.p;
    const int cap_grid = sc->cus * 8;
    k_bd_gen<<<(B + 255) / 256, 256, 0, st>>>(sc->rs, bp, ctl, vb, q[0]);
    int a = 0;
    for (int it = 0; it < bp.maxv - 1; it++) {
        int rc = launch_trace(sc->accel, reinterpret_cast<const WRay*>(q[a].ray), 0, &ctl->n_trace, B,
                              reinterpret_cast<WHit*>(sc->bd_hits.p), st, false);
        if (rc != LMX_OK) return rc;
        k_bd_shade<<<(B + 255) / 256, 256, 0, st>>>(sc->rs, bp, ctl, q[a], sc->bd_hits.p, q[1 - a], vb);
        k_bd_commit<<<1, 64, 0, st>>>(ctl);
        a = 1 - a;
    }
    const Grid g{sc->vc_items.p, sc->vc_start.p, sc->vc_cursor.p, sc->vc_part.p, sc->vc_H};
    k_pm_gather<<<std::max(1, std::min((int)((B + 255) / 256), cap_grid * 2)), 256, 0, st>>>(sc->rs, bp, ctl, vb,
                                                                                             ph, g, pa);
    k_bd_next_batch<<<1, 64, 0, st>>>(ctl, B);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? LMX_OK : hip_error(e, "pm launch");
}

// renderer::pm: the photon graph (roulette light subpaths, the grid over
// their diffuse and glossy vertices), then the eye batches; film scaled by
// W H / N as the scheduler does (scheduler.cpp:281-288).
int render_pm(lmx_scene* sc, const lmx_render_params* p, const lmx_pm_params* v, float* film_rgb, void* stream,
              lmx_render_stats* stats)
{
    const char* name = "lmx_render_pm";
    char msg[160];
    if (!sc || !p || !v || !film_rgb) return set_error(LMX_EINVAL, "lmx_render_pm: NULL argument");
    if (p->num_samples == 0 || p->num_samples > 0xffffffffull)
        return set_error(LMX_EINVAL, "lmx_render_pm: num_samples must be in [1, 2^32)");
    const bool bounded = p->max_num_vertices != -1;
    if (bounded && (p->max_num_vertices < 2 || p->max_num_vertices > 32))
        return set_error(LMX_EINVAL, "lmx_render_pm: max_num_vertices must be in [2, 32] or -1");
    if (v->finalgather != 0 && v->finalgather != 1) return set_error(LMX_EINVAL, "lmx_render_pm: finalgather must be 0 or 1");
    if (v->num_photon_trace_samples == 0 || v->num_photon_trace_samples > (1u << 26))
        return set_error(LMX_EINVAL, "lmx_render_pm: num_photon_trace_samples must be in [1, 2^26]");
    if (!(v->radius > 0.0f) || !(v->radius < 3.0e38f))
        return set_error(LMX_EINVAL, "lmx_render_pm: radius must be positive and finite");
    if (sc->mats.n > VB_MAX_MATERIALS || sc->lights.n > VB_MAX_LIGHTS)
        return set_error(LMX_EINVAL, "lmx_render_pm: at most 65536 materials and 8190 lights");
    if (p->raster_mode != LMX_RASTER_RANDOM && p->raster_mode != LMX_RASTER_STRATIFIED)
        return set_error(LMX_EINVAL, "lmx_render_pm: unknown raster_mode");
    const uint32_t pcap = bounded ? (uint32_t)p->max_num_vertices : 32u;
    const uint32_t ecap = std::min<uint32_t>(pcap, 2u + (uint32_t)v->finalgather);
    const uint32_t ewidth = LMX_PT_SLOT_SEG0 + 4u * (ecap - 1u);
    const uint32_t pwidth = LMX_PT_SLOT_SEG0 + 8u * (pcap - 1u);
    if ((p->uniform_table && p->table_width < ewidth) || (v->photon_table && v->photon_table_width < pwidth)) {
        snprintf(msg, sizeof msg, "%s: uniform table narrower than %u (eye) / %u (photon)", name, ewidth, pwidth);
        return set_error(LMX_EINVAL, msg);
    }
    hipStream_t st = (hipStream_t)stream;
    hipStream_t own = nullptr;
    hipError_t e;
    if (!st) {
        if ((e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking)) != hipSuccess) return hip_error(e, "stream create");
        st = own;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t N = (uint32_t)p->num_samples, np = v->num_photon_trace_samples;
    BParams bp;                          // eye samples
    bp.variant = BD_VCM;
    bp.merging = 0; bp.connect = 0; bp.phase = 0; bp.c_conn = 1.0f;
    bp.mis = LMX_MIS_POWER;
    bp.bounded = 1;
    bp.maxv = (int)ecap; bp.minv = 0;
    bp.key = make_uint2((uint32_t)p->seed, (uint32_t)(p->seed >> 32));
    bp.total = N;
    bp.raster_mode = p->raster_mode;
    bp.table = p->uniform_table; bp.table_width = p->table_width;
    bp.wsub = ewidth;
    bp.sample_out = p->sample_out;
    bp.film = film_rgb;
    uint32_t B = p->queue_capacity ? p->queue_capacity : (1u << 23);   // ltdirect 1.5x, pm 1.09x the rate at 2^21
    B = std::min(B, N);
    bp.batch = B;
    BParams pp = bp;                     // photon subpaths
    pp.variant = BD_PM;
    pp.phase = 1;
    pp.bounded = bounded ? 1 : 0;
    pp.maxv = (int)pcap;
    pp.batch = np; pp.total = np;
    pp.key = make_uint2((uint32_t)v->photon_seed, (uint32_t)(v->photon_seed >> 32));
    pp.table = v->photon_table; pp.table_width = v->photon_table_width;
    pp.wsub = 0;
    pp.sample_out = nullptr;
    PMArgs pa{(int32_t)pcap, bounded ? p->max_num_vertices : -1, v->finalgather, (float)np};
    int rc = bd_buffers(sc, B, ecap, 1);
    if (rc == LMX_OK) rc = vc_buffers(sc, np, pcap, 1);
    const size_t npix = (size_t)sc->width * (size_t)sc->height;
    const int zgrid = std::max(1, std::min((int)((3 * npix + 255) / 256), sc->cus * 8));
    hipGraphExec_t gph = nullptr, geye = nullptr;
    if (rc == LMX_OK) {
        k_bd_zero<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix);
        k_bd_init<<<1, 64, 0, st>>>(sc->bd_ctl.p);
        rc = capture(st, &gph, [&] { return vc_enqueue_photons(sc, pp, st); });
    }
    if (rc == LMX_OK) rc = capture(st, &geye, [&] { return pm_enqueue_eye(sc, bp, pa, st); });
    const float r = v->radius;
    if (rc == LMX_OK) {
        k_bd_phase<<<1, 64, 0, st>>>(sc->bd_ctl.p, 0u, r, (float)(2.0 * (double)r), 0.0f);
        if ((e = hipGraphLaunch(gph, st)) != hipSuccess) rc = hip_error(e, "graph launch");
    }
    const uint64_t nbatch = ((uint64_t)N + B - 1) / B;
    if (rc == LMX_OK) k_bd_phase<<<1, 64, 0, st>>>(sc->bd_ctl.p, 0u, r, (float)(2.0 * (double)r), 0.0f);
    for (uint64_t b = 0; rc == LMX_OK && b < nbatch; b++)
        if ((e = hipGraphLaunch(geye, st)) != hipSuccess) rc = hip_error(e, "graph launch");
    if (rc == LMX_OK) {
        k_bd_scale<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix, (float)((double)npix / (double)N));
        if ((e = hipMemcpyAsync(sc->bd_hctl, sc->bd_ctl.p, sizeof(BCtl), hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess)
            rc = hip_error(e, "lmx_render_pm: sync");
    }
    if (gph) (void)hipGraphExecDestroy(gph);
    if (geye) (void)hipGraphExecDestroy(geye);
    if (own) (void)hipStreamDestroy(own);
    if (stats) {
        stats->rays = rc == LMX_OK ? sc->bd_hctl->rays : 0;
        stats->iterations = (uint32_t)nbatch;
        stats->queue_capacity = B;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->samples = rc == LMX_OK ? p->num_samples : 0;
    }
    return rc;
}

// renderer::ppm: the measurement points (one eye batch of every sample,
// subpaths of 2 vertices), then per pass the photon graph and the
// progressive update, then the film.
int render_ppm(lmx_scene* sc, const lmx_render_params* p, const lmx_ppm_params* v, float* film_rgb, void* stream,
               lmx_render_stats* stats)
{
    const char* name = "lmx_render_ppm";
    char msg[160];
    if (!sc || !p || !v || !film_rgb) return set_error(LMX_EINVAL, "lmx_render_ppm: NULL argument");
    if (p->num_samples == 0 || p->num_samples > (1ull << 24))
        return set_error(LMX_EINVAL, "lmx_render_ppm: num_samples (measurement points) must be in [1, 2^24]");
    const bool bounded = p->max_num_vertices != -1;
    if (bounded && (p->max_num_vertices < 2 || p->max_num_vertices > 32))
        return set_error(LMX_EINVAL, "lmx_render_ppm: max_num_vertices must be in [2, 32] or -1");
    if (v->num_iteration_pass == 0) return set_error(LMX_EINVAL, "lmx_render_ppm: num_iteration_pass must be >= 1");
    if (v->num_photon_trace_samples == 0 || v->num_photon_trace_samples > (1u << 26))
        return set_error(LMX_EINVAL, "lmx_render_ppm: num_photon_trace_samples must be in [1, 2^26]");
    if ((uint64_t)v->num_iteration_pass * v->num_photon_trace_samples > 0xffffffffull)
        return set_error(LMX_EINVAL, "lmx_render_ppm: passes x photons must stay below 2^32");
    if (!(v->initial_radius > 0.0f) || !(v->initial_radius < 3.0e38f) || !(v->alpha > 0.0f) || !(v->alpha <= 1.0f))
        return set_error(LMX_EINVAL, "lmx_render_ppm: initial_radius must be positive and finite, alpha in (0, 1]");
    if (sc->mats.n > VB_MAX_MATERIALS || sc->lights.n > VB_MAX_LIGHTS)
        return set_error(LMX_EINVAL, "lmx_render_ppm: at most 65536 materials and 8190 lights");
    if (p->raster_mode != LMX_RASTER_RANDOM && p->raster_mode != LMX_RASTER_STRATIFIED)
        return set_error(LMX_EINVAL, "lmx_render_ppm: unknown raster_mode");
    const uint32_t pcap = bounded ? (uint32_t)p->max_num_vertices : 32u;
    const uint32_t ewidth = LMX_PT_SLOT_SEG0;                 // the eye endpoint's slots
    const uint32_t pwidth = LMX_PT_SLOT_SEG0 + 8u * (pcap - 1u);
    if ((p->uniform_table && p->table_width < ewidth) || (v->photon_table && v->photon_table_width < pwidth)) {
        snprintf(msg, sizeof msg, "%s: uniform table narrower than %u (eye) / %u (photon)", name, ewidth, pwidth);
        return set_error(LMX_EINVAL, msg);
    }
    hipStream_t st = (hipStream_t)stream;
    hipStream_t own = nullptr;
    hipError_t e;
    if (!st) {
        if ((e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking)) != hipSuccess) return hip_error(e, "stream create");
        st = own;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t N = (uint32_t)p->num_samples, np = v->num_photon_trace_samples;
    BParams bp;                          // the measurement points: one batch of every sample
    bp.variant = BD_VCM;
    bp.merging = 0; bp.connect = 0; bp.phase = 0; bp.c_conn = 1.0f;
    bp.mis = LMX_MIS_POWER;
    bp.bounded = 1;
    bp.maxv = 2; bp.minv = 0;
    bp.key = make_uint2((uint32_t)p->seed, (uint32_t)(p->seed >> 32));
    bp.total = N; bp.batch = N;
    bp.raster_mode = p->raster_mode;
    bp.table = p->uniform_table; bp.table_width = p->table_width;
    bp.wsub = ewidth;
    bp.sample_out = p->sample_out;
    bp.film = film_rgb;
    BParams pp = bp;                     // photon subpaths
    pp.variant = BD_PPM;
    pp.phase = 1;
    pp.bounded = bounded ? 1 : 0;
    pp.maxv = (int)pcap;
    pp.batch = np; pp.total = np;
    pp.key = make_uint2((uint32_t)v->photon_seed, (uint32_t)(v->photon_seed >> 32));
    pp.table = v->photon_table; pp.table_width = v->photon_table_width;
    pp.wsub = 0;
    pp.sample_out = nullptr;
    PMArgs pa{(int32_t)pcap, bounded ? p->max_num_vertices : -1, 1, (float)np};
    int rc = bd_buffers(sc, N, 2, 1);
    if (rc == LMX_OK) rc = vc_buffers(sc, np, pcap, 1);
    if (rc == LMX_OK && sc->ppm_st.n < 2 * (size_t)N && (e = sc->ppm_st.alloc(2 * (size_t)N)) != hipSuccess)
        rc = hip_error(e, "ppm measurement points");
    const size_t npix = (size_t)sc->width * (size_t)sc->height;
    const int zgrid = std::max(1, std::min((int)((3 * npix + 255) / 256), sc->cus * 8));
    const int mgrid = std::max(1, std::min((int)((N + 255) / 256), sc->cus * 16));
    const float r0 = v->initial_radius, cell = (float)(2.0 * (double)r0);
    hipGraphExec_t gph = nullptr;
    VB vb;
    vb.v[0] = sc->bd_vert[0].p; vb.v[1] = sc->bd_vert[1].p;
    vb.nsub[0] = sc->bd_nsub[0].p; vb.nsub[1] = sc->bd_nsub[1].p;
    vb.epix = sc->bd_epix.p; vb.eraster = sc->bd_eraster.p;
    vb.V = sc->bd_V.p;
    vb.mq = nullptr;
    VB ph = vb;
    ph.v[1] = sc->vc_vert.p; ph.nsub[1] = sc->vc_nsub.p;
    const Grid g{sc->vc_items.p, sc->vc_start.p, sc->vc_cursor.p, sc->vc_part.p, sc->vc_H};
    if (rc == LMX_OK) {
        k_bd_zero<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix);
        if (p->sample_out)
            k_bd_zero<<<std::max(1, (int)std::min<size_t>((16 * (size_t)N + 255) / 256, (size_t)sc->cus * 8)), 256,
                        0, st>>>(p->sample_out, 16 * (size_t)N);
        k_bd_init<<<1, 64, 0, st>>>(sc->bd_ctl.p);
        // the measurement points: endpoints and the first hit of every sample
        Queue q[2];
        for (int k = 0; k < 2; k++) { q[k].ray = sc->bd_ray[k].p; q[k].thr = sc->bd_thr[k].p; q[k].meta = sc->bd_meta[k].p; }
        BCtl* ctl = sc->bd_ctl.p;
        k_bd_gen<<<(N + 255) / 256, 256, 0, st>>>(sc->rs, bp, ctl, vb, q[0]);
        rc = launch_trace(sc->accel, reinterpret_cast<const WRay*>(q[0].ray), 0, &ctl->n_trace, N,
                          reinterpret_cast<WHit*>(sc->bd_hits.p), st, false);
        if (rc == LMX_OK) {
            k_bd_shade<<<(N + 255) / 256, 256, 0, st>>>(sc->rs, bp, ctl, q[0], sc->bd_hits.p, q[1], vb);
            k_bd_commit<<<1, 64, 0, st>>>(ctl);
            k_ppm_init<<<mgrid, 256, 0, st>>>(sc->rs, bp, vb, sc->ppm_st.p, r0);
            if ((e = hipGetLastError()) != hipSuccess) rc = hip_error(e, "ppm launch");
        }
    }
    if (rc == LMX_OK) rc = capture(st, &gph, [&] { return vc_enqueue_photons(sc, pp, st); });
    for (uint32_t q = 0; rc == LMX_OK && q < v->num_iteration_pass; q++) {
        k_bd_phase<<<1, 64, 0, st>>>(sc->bd_ctl.p, q * np, r0, cell, 0.0f);
        if ((e = hipGraphLaunch(gph, st)) != hipSuccess) { rc = hip_error(e, "graph launch"); break; }
        k_ppm_update<<<mgrid, 256, 0, st>>>(sc->rs, bp, sc->bd_ctl.p, vb, ph, g, pa, sc->ppm_st.p, v->alpha);
    }
    if (rc == LMX_OK) {
        const float total = (float)((double)v->num_iteration_pass * (double)np);
        k_ppm_splat<<<mgrid, 256, 0, st>>>(bp, vb, sc->ppm_st.p, total);
        k_bd_scale<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix, (float)((double)npix / (double)N));
        if ((e = hipGetLastError()) != hipSuccess) rc = hip_error(e, "ppm launch");
    }
    if (rc == LMX_OK &&
        ((e = hipMemcpyAsync(sc->bd_hctl, sc->bd_ctl.p, sizeof(BCtl), hipMemcpyDeviceToHost, st)) != hipSuccess ||
         (e = hipStreamSynchronize(st)) != hipSuccess))
        rc = hip_error(e, "lmx_render_ppm: sync");
    if (gph) (void)hipGraphExecDestroy(gph);
    if (own) (void)hipStreamDestroy(own);
    if (stats) {
        stats->rays = rc == LMX_OK ? sc->bd_hctl->rays : 0;
        stats->iterations = v->num_iteration_pass;
        stats->queue_capacity = N;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->samples = rc == LMX_OK ? p->num_samples : 0;
    }
    return rc;
}

// renderer::sppm: per pass the pixels' points (one eye batch of W H
// samples, 2-vertex subpaths), the photon graph, the per-pixel update; the
// film written after the last pass.
int render_sppm(lmx_scene* sc, const lmx_render_params* p, const lmx_ppm_params* v, float* film_rgb, void* stream,
                lmx_render_stats* stats)
{
    const char* name = "lmx_render_sppm";
    char msg[160];
    if (!sc || !p || !v || !film_rgb) return set_error(LMX_EINVAL, "lmx_render_sppm: NULL argument");
    const bool bounded = p->max_num_vertices != -1;
    if (bounded && (p->max_num_vertices < 2 || p->max_num_vertices > 32))
        return set_error(LMX_EINVAL, "lmx_render_sppm: max_num_vertices must be in [2, 32] or -1");
    if (v->num_iteration_pass == 0) return set_error(LMX_EINVAL, "lmx_render_sppm: num_iteration_pass must be >= 1");
    if (v->num_photon_trace_samples == 0 || v->num_photon_trace_samples > (1u << 26))
        return set_error(LMX_EINVAL, "lmx_render_sppm: num_photon_trace_samples must be in [1, 2^26]");
    const uint32_t npix = (uint32_t)sc->width * (uint32_t)sc->height;
    if ((uint64_t)v->num_iteration_pass * v->num_photon_trace_samples > 0xffffffffull ||
        (uint64_t)v->num_iteration_pass * npix > 0xffffffffull)
        return set_error(LMX_EINVAL, "lmx_render_sppm: passes x photons and passes x pixels must stay below 2^32");
    if (!(v->initial_radius > 0.0f) || !(v->initial_radius < 3.0e38f) || !(v->alpha > 0.0f) || !(v->alpha <= 1.0f))
        return set_error(LMX_EINVAL, "lmx_render_sppm: initial_radius must be positive and finite, alpha in (0, 1]");
    if (sc->mats.n > VB_MAX_MATERIALS || sc->lights.n > VB_MAX_LIGHTS)
        return set_error(LMX_EINVAL, "lmx_render_sppm: at most 65536 materials and 8190 lights");
    const uint32_t pcap = bounded ? (uint32_t)p->max_num_vertices : 32u;
    const uint32_t ewidth = LMX_PT_SLOT_SEG0;
    const uint32_t pwidth = LMX_PT_SLOT_SEG0 + 8u * (pcap - 1u);
    if ((p->uniform_table && p->table_width < ewidth) || (v->photon_table && v->photon_table_width < pwidth)) {
        snprintf(msg, sizeof msg, "%s: uniform table narrower than %u (eye) / %u (photon)", name, ewidth, pwidth);
        return set_error(LMX_EINVAL, msg);
    }
    hipStream_t st = (hipStream_t)stream;
    hipStream_t own = nullptr;
    hipError_t e;
    if (!st) {
        if ((e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking)) != hipSuccess) return hip_error(e, "stream create");
        st = own;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t np = v->num_photon_trace_samples;
    BParams bp;                          // the pixels' points of one pass
    bp.variant = BD_VCM;
    bp.merging = 0; bp.connect = 0; bp.phase = 0; bp.c_conn = 1.0f;
    bp.mis = LMX_MIS_POWER;
    bp.bounded = 1;
    bp.maxv = 2; bp.minv = 0;
    bp.key = make_uint2((uint32_t)p->seed, (uint32_t)(p->seed >> 32));
    bp.total = npix; bp.batch = npix;
    bp.raster_mode = LMX_RASTER_STRATIFIED;          // pixel i, jittered inside it (renderer_sppm.cpp:162)
    bp.table = p->uniform_table; bp.table_width = p->table_width;
    bp.wsub = ewidth;
    bp.sample_out = p->sample_out;
    bp.film = film_rgb;
    BParams pp = bp;                     // photon subpaths
    pp.variant = BD_PPM;
    pp.phase = 1;
    pp.bounded = bounded ? 1 : 0;
    pp.maxv = (int)pcap;
    pp.batch = np; pp.total = np;
    pp.key = make_uint2((uint32_t)v->photon_seed, (uint32_t)(v->photon_seed >> 32));
    pp.table = v->photon_table; pp.table_width = v->photon_table_width;
    pp.wsub = 0;
    pp.sample_out = nullptr;
    PMArgs pa{(int32_t)pcap, bounded ? p->max_num_vertices : -1, 1, (float)np};
    int rc = bd_buffers(sc, npix, 2, 1);
    if (rc == LMX_OK) rc = vc_buffers(sc, np, pcap, 1);
    if (rc == LMX_OK && sc->ppm_st.n < 3 * (size_t)npix && (e = sc->ppm_st.alloc(3 * (size_t)npix)) != hipSuccess)
        rc = hip_error(e, "sppm pixel points");
    const int mgrid = std::max(1, std::min((int)((npix + 255) / 256), sc->cus * 16));
    const float r0 = v->initial_radius, cell = (float)(2.0 * (double)r0);
    VB vb;
    vb.v[0] = sc->bd_vert[0].p; vb.v[1] = sc->bd_vert[1].p;
    vb.nsub[0] = sc->bd_nsub[0].p; vb.nsub[1] = sc->bd_nsub[1].p;
    vb.epix = sc->bd_epix.p; vb.eraster = sc->bd_eraster.p;
    vb.V = sc->bd_V.p;
    vb.mq = nullptr;
    VB ph = vb;
    ph.v[1] = sc->vc_vert.p; ph.nsub[1] = sc->vc_nsub.p;
    const Grid g{sc->vc_items.p, sc->vc_start.p, sc->vc_cursor.p, sc->vc_part.p, sc->vc_H};
    hipGraphExec_t gph = nullptr, geye = nullptr;
    if (rc == LMX_OK) {
        k_bd_init<<<1, 64, 0, st>>>(sc->bd_ctl.p);
        k_sppm_init<<<mgrid, 256, 0, st>>>(sc->ppm_st.p, npix, r0);
        rc = capture(st, &gph, [&] { return vc_enqueue_photons(sc, pp, st); });
    }
    if (rc == LMX_OK)
        rc = capture(st, &geye, [&] {
            Queue q[2];
            for (int k = 0; k < 2; k++) { q[k].ray = sc->bd_ray[k].p; q[k].thr = sc->bd_thr[k].p; q[k].meta = sc->bd_meta[k].p; }
            BCtl* ctl = sc->bd_ctl.p;
            k_bd_gen<<<(npix + 255) / 256, 256, 0, st>>>(sc->rs, bp, ctl, vb, q[0]);
            int rc2 = launch_trace(sc->accel, reinterpret_cast<const WRay*>(q[0].ray), 0, &ctl->n_trace, npix,
                                   reinterpret_cast<WHit*>(sc->bd_hits.p), st, false);
            if (rc2 != LMX_OK) return rc2;
            k_bd_shade<<<(npix + 255) / 256, 256, 0, st>>>(sc->rs, bp, ctl, q[0], sc->bd_hits.p, q[1], vb);
            k_bd_commit<<<1, 64, 0, st>>>(ctl);
            const hipError_t e2 = hipGetLastError();
            return e2 == hipSuccess ? LMX_OK : hip_error(e2, "sppm launch");
        });
    for (uint32_t q = 0; rc == LMX_OK && q < v->num_iteration_pass; q++) {
        k_bd_phase<<<1, 64, 0, st>>>(sc->bd_ctl.p, q * npix, r0, cell, 0.0f);
        if ((e = hipGraphLaunch(geye, st)) != hipSuccess) { rc = hip_error(e, "graph launch"); break; }
        k_bd_phase<<<1, 64, 0, st>>>(sc->bd_ctl.p, q * np, r0, cell, 0.0f);
        if ((e = hipGraphLaunch(gph, st)) != hipSuccess) { rc = hip_error(e, "graph launch"); break; }
        k_sppm_update<<<mgrid, 256, 0, st>>>(sc->rs, bp, sc->bd_ctl.p, vb, ph, g, pa, sc->ppm_st.p, v->alpha);
    }
    if (rc == LMX_OK) {
        const float total = (float)((double)v->num_iteration_pass * (double)np);
        k_sppm_film<<<mgrid, 256, 0, st>>>(bp, sc->ppm_st.p, total, (float)v->num_iteration_pass);
        if ((e = hipGetLastError()) != hipSuccess) rc = hip_error(e, "sppm launch");
    }
    if (rc == LMX_OK &&
        ((e = hipMemcpyAsync(sc->bd_hctl, sc->bd_ctl.p, sizeof(BCtl), hipMemcpyDeviceToHost, st)) != hipSuccess ||
         (e = hipStreamSynchronize(st)) != hipSuccess))
        rc = hip_error(e, "lmx_render_sppm: sync");
    if (gph) (void)hipGraphExecDestroy(gph);
    if (geye) (void)hipGraphExecDestroy(geye);
    if (own) (void)hipStreamDestroy(own);
    if (stats) {
        stats->rays = rc == LMX_OK ? sc->bd_hctl->rays : 0;
        stats->iterations = v->num_iteration_pass;
        stats->queue_capacity = npix;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->samples = rc == LMX_OK ? (uint64_t)npix * v->num_iteration_pass : 0;
    }
    return rc;
}

// renderer::ltdirect: batches of B light subpaths through the photon graph
// (no grid), their camera connections, the any-hit trace, the splat; film
// scaled by W H / N.
int render_ltdirect(lmx_scene* sc, const lmx_render_params* p, float* film_rgb, void* stream, lmx_render_stats* stats)
{
    const char* name = "lmx_render_ltdirect";
    char msg[160];
    if (!sc || !p || !film_rgb) return set_error(LMX_EINVAL, "lmx_render_ltdirect: NULL argument");
    if (p->num_samples == 0 || p->num_samples > 0xffffffffull)
        return set_error(LMX_EINVAL, "lmx_render_ltdirect: num_samples must be in [1, 2^32)");
    const bool bounded = p->max_num_vertices != -1;
    if (bounded && (p->max_num_vertices < 2 || p->max_num_vertices > 32))
        return set_error(LMX_EINVAL, "lmx_render_ltdirect: max_num_vertices must be in [2, 32] or -1");
    if (sc->mats.n > VB_MAX_MATERIALS || sc->lights.n > VB_MAX_LIGHTS)
        return set_error(LMX_EINVAL, "lmx_render_ltdirect: at most 65536 materials and 8190 lights");
    const uint32_t pcap = bounded ? (uint32_t)p->max_num_vertices : 32u;
    const uint32_t width = LMX_PT_SLOT_SEG0 + 8u * (pcap - 1u);
    if (p->uniform_table && p->table_width < width) {
        snprintf(msg, sizeof msg, "%s: uniform table narrower than %u", name, width);
        return set_error(LMX_EINVAL, msg);
    }
    hipStream_t st = (hipStream_t)stream;
    hipStream_t own = nullptr;
    hipError_t e;
    if (!st) {
        if ((e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking)) != hipSuccess) return hip_error(e, "stream create");
        st = own;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t N = (uint32_t)p->num_samples;
    uint32_t B = p->queue_capacity ? p->queue_capacity : (1u << 23);   // ltdirect 1.5x, pm 1.09x the rate at 2^21
    B = std::min(B, N);
    BParams pp;
    pp.variant = BD_PM;
    pp.merging = 0; pp.connect = 0; pp.phase = 1; pp.c_conn = 1.0f;
    pp.mis = LMX_MIS_POWER;
    pp.bounded = bounded ? 1 : 0;
    pp.maxv = (int)pcap; pp.minv = 0;
    pp.key = make_uint2((uint32_t)p->seed, (uint32_t)(p->seed >> 32));
    pp.total = N; pp.batch = B;
    pp.raster_mode = LMX_RASTER_RANDOM;
    pp.table = p->uniform_table; pp.table_width = p->table_width;
    pp.wsub = 0;
    pp.sample_out = p->sample_out;
    pp.film = film_rgb;
    int rc = bd_buffers(sc, B, 2, pcap);
    if (rc == LMX_OK) rc = vc_buffers(sc, B, pcap, 1);
    const size_t npix = (size_t)sc->width * (size_t)sc->height;
    const int zgrid = std::max(1, std::min((int)((3 * npix + 255) / 256), sc->cus * 8));
    const int cap_grid = sc->cus * 8;
    hipGraphExec_t gb = nullptr;
    if (rc == LMX_OK) {
        k_bd_zero<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix);
        if (p->sample_out) {
            const size_t ns = (size_t)N * pcap * 8;
            k_bd_zero<<<std::max(1, (int)std::min<size_t>((ns + 255) / 256, (size_t)cap_grid)), 256, 0, st>>>(
                p->sample_out, ns);
        }
        k_bd_init<<<1, 64, 0, st>>>(sc->bd_ctl.p);
        rc = capture(st, &gb, [&] {
            int rc2 = vc_enqueue_photons(sc, pp, st, false);
            if (rc2 != LMX_OK) return rc2;
            VB ph;
            ph.v[0] = sc->bd_vert[0].p; ph.v[1] = sc->vc_vert.p;
            ph.nsub[0] = sc->bd_nsub[0].p; ph.nsub[1] = sc->vc_nsub.p;
            ph.epix = sc->bd_epix.p; ph.eraster = sc->bd_eraster.p;
            ph.V = sc->bd_V.p;
            ph.mq = nullptr;
            ShadowQ sq{sc->bd_sray.p, sc->bd_sc.p, sc->bd_saux.p, sc->bd_smeta.p};
            BCtl* ctl = sc->bd_ctl.p;
            const size_t nslot = (size_t)B * pcap;
            const int g = std::max(1, (int)std::min<size_t>((nslot + 255) / 256, (size_t)cap_grid * 2));
            k_lt_connect<<<g, 256, 0, st>>>(sc->rs, pp, ctl, ph, sq, bounded ? p->max_num_vertices : -1);
            rc2 = launch_trace(sc->accel, reinterpret_cast<const WRay*>(sq.ray), 0, &ctl->n_shadow, (uint32_t)nslot,
                               reinterpret_cast<WHit*>(sc->bd_shits.p), st, true);
            if (rc2 != LMX_OK) return rc2;
            k_lt_splat<<<g, 256, 0, st>>>(pp, ctl, sq, sc->bd_shits.p, (int32_t)pcap);
            const hipError_t e2 = hipGetLastError();
            return e2 == hipSuccess ? LMX_OK : hip_error(e2, "ltdirect launch");
        });
    }
    const uint64_t nbatch = ((uint64_t)N + B - 1) / B;
    for (uint64_t b = 0; rc == LMX_OK && b < nbatch; b++) {
        k_lt_batch<<<1, 64, 0, st>>>(sc->bd_ctl.p, (uint32_t)(b * B));
        if ((e = hipGraphLaunch(gb, st)) != hipSuccess) rc = hip_error(e, "graph launch");
    }
    if (rc == LMX_OK) {
        k_bd_scale<<<zgrid, 256, 0, st>>>(film_rgb, 3 * npix, (float)((double)npix / (double)N));
        if ((e = hipMemcpyAsync(sc->bd_hctl, sc->bd_ctl.p, sizeof(BCtl), hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipStreamSynchronize(st)) != hipSuccess)
            rc = hip_error(e, "lmx_render_ltdirect: sync");
    }
    if (gb) (void)hipGraphExecDestroy(gb);
    if (own) (void)hipStreamDestroy(own);
    if (stats) {
        stats->rays = rc == LMX_OK ? sc->bd_hctl->rays + sc->bd_hctl->shadow_rays : 0;
        stats->iterations = (uint32_t)nbatch;
        stats->queue_capacity = B;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->samples = rc == LMX_OK ? p->num_samples : 0;
    }
    return rc;
}

int render_raycast(lmx_scene* sc, float* film_rgb, void* stream, lmx_render_stats* stats)
{
    if (!sc || !film_rgb) return set_error(LMX_EINVAL, "lmx_render_raycast: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    hipStream_t own = nullptr;
    hipError_t e;
    if (!st) {
        if ((e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking)) != hipSuccess) return hip_error(e, "stream create");
        st = own;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t n = (uint32_t)sc->width * (uint32_t)sc->height;
    int rc = bd_buffers(sc, (n + 1) / 2, 2, 1);           // queue 0 holds 2 (n + 1) / 2 >= n rays
    if (rc == LMX_OK) {
        k_rc_gen<<<(n + 255) / 256, 256, 0, st>>>(sc->rs, sc->bd_ray[0].p, n);
        rc = launch_trace(sc->accel, reinterpret_cast<const WRay*>(sc->bd_ray[0].p), n, nullptr, n,
                          reinterpret_cast<WHit*>(sc->bd_hits.p), st, false);
    }
    if (rc == LMX_OK) {
        k_rc_shade<<<(n + 255) / 256, 256, 0, st>>>(sc->rs, sc->bd_ray[0].p, sc->bd_hits.p, film_rgb, n);
        if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(st)) != hipSuccess)
            rc = hip_error(e, "lmx_render_raycast");
    }
    if (own) (void)hipStreamDestroy(own);
    if (stats) {
        stats->rays = rc == LMX_OK ? n : 0;
        stats->iterations = 1;
        stats->queue_capacity = n;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->samples = rc == LMX_OK ? n : 0;
    }
    return rc;
}

}  // namespace

extern "C" {

int lmx_render_vcm_bdpt(lmx_scene* sc, const lmx_render_params* p, float* film_rgb, void* stream,
                        lmx_render_stats* stats)
{
    return render_bidir(sc, p, BD_VCM, "lmx_render_vcm_bdpt", film_rgb, stream, stats);
}

int lmx_render_bdpt(lmx_scene* sc, const lmx_render_params* p, float* film_rgb, void* stream, lmx_render_stats* stats)
{
    return render_bidir(sc, p, BD_BDPT, "lmx_render_bdpt", film_rgb, stream, stats);
}

int lmx_render_vcm(lmx_scene* sc, const lmx_render_params* p, const lmx_vcm_params* v, float* film_rgb, void* stream,
                   lmx_render_stats* stats)
{
    return render_vcm(sc, p, v, film_rgb, stream, stats);
}

int lmx_render_pm(lmx_scene* sc, const lmx_render_params* p, const lmx_pm_params* v, float* film_rgb, void* stream,
                  lmx_render_stats* stats)
{
    return render_pm(sc, p, v, film_rgb, stream, stats);
}

int lmx_render_ppm(lmx_scene* sc, const lmx_render_params* p, const lmx_ppm_params* v, float* film_rgb,
                   void* stream, lmx_render_stats* stats)
{
    return render_ppm(sc, p, v, film_rgb, stream, stats);
}

int lmx_render_sppm(lmx_scene* sc, const lmx_render_params* p, const lmx_ppm_params* v, float* film_rgb,
                    void* stream, lmx_render_stats* stats)
{
    return render_sppm(sc, p, v, film_rgb, stream, stats);
}

int lmx_render_ltdirect(lmx_scene* sc, const lmx_render_params* p, float* film_rgb, void* stream,
                        lmx_render_stats* stats)
{
    return render_ltdirect(sc, p, film_rgb, stream, stats);
}

int lmx_render_raycast(lmx_scene* sc, float* film_rgb, void* stream, lmx_render_stats* stats)
{
    return render_raycast(sc, film_rgb, stream, stats);
}

}  // extern "C"
