// lmx_accel.hip -- closest-hit traversal of the compressed 8-wide BVH on
// gfx950, the host walker over the same structure, and the lmx_accel_* C ABI.
//
// Replaces the per-ray CPU traversal loops of the reference
// (src/liblightmetrica/accel/accel_qbvh.cpp:392-481,
//  src/liblightmetrica/accel/accel_bvh_sahbin.cpp:234-288) with a persistent
// wave-per-workgroup kernel:
//   * one ray per lane, 64 lanes per wave, one wave per workgroup;
//   * a per-lane stack of node groups in LDS (a group = the not yet visited
//     hit children of one node: child base, hit mask, internal mask), so the
//     stack never holds more entries than the tree has levels;
//   * the next node is chosen right after a node's box test and its load is
//     issued before that node's triangles are tested, so node latency
//     overlaps triangle work; each lane loads exactly one node per loop
//     iteration at one site, which keeps the 80-byte node in registers
//     across iterations without copies;
//   * lanes whose ray finished are refilled as a batch once at least
//     LMX_REFILL_MIN lanes are idle: one ballot counts them, mbcnt ranks give
//     every idle lane its ray, and the rays come from a per-wave pool that
//     one atomic refills up to LMX_CLAIM_CHUNK rays at a time; the ray
//     set-up that depends on the direction alone (reciprocals, octant, the
//     watertight shear with its divisions) is computed when the pool is
//     filled, a ray per lane at the full width of the wave, and handed to
//     the refilled lanes by shuffles;
//   * XCD-aware work split: workgroup b runs on XCD (b mod 8); the rays are
//     cut into 8 contiguous regions and a wave first drains the region of
//     its own XCD (neighbouring rays share BVH nodes in that XCD's L2), then
//     helps the others; its first chunk is fixed by its index, and only
//     what the fixed chunks leave is claimed through the regions' atomic
//     cursors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lmx.h"
#include "lmx_internal.h"
#include "wbvh_build.h"
#include "wbvh_math.h"
#include "trace_cursors.h"

namespace lmx {

// ---------------------------------------------------------------------------
// errors

static thread_local std::string g_err;

int set_error(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

// A failed HIP call is reported through the library's own error string;
// the runtime's last-error slot is cleared so that the caller's next launch
// check does not report it a second time.
int hip_error(hipError_t e, const char* what)
{
    (void)hipGetLastError();
    return set_error(LMX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace lmx

using namespace lmx;

struct lmx_accel {
    WbvhHost host;
    int device = -1;
    WNode* d_nodes = nullptr;
    WTri* d_tris = nullptr;
    uint32_t* d_counters = nullptr;   // 8 region cursors, the ray count, the stack-overflow flag; CURSOR_STRIDE words apart
    int grid = 0;
    int stack_depth = 0;              // LDS stack entries per lane
    size_t lds_bytes = 0;
};

// ---------------------------------------------------------------------------
// host walker (same visiting order as the kernel)

namespace lmx {

// Returns false when a node group did not fit the stack (a BVH deeper than
// WBVH_STACK, which the builder refuses); the hit is then a miss.  The walk
// is the kernel's per-lane state machine, step for step (node test while the
// two-group triangle queue has room, one triangle per step), so the node
// visits and triangle tests it counts, and an any-hit ray's reported
// triangle, are the kernel's.
template <int ISECT, bool ANYHIT>
static bool walk_impl(const WbvhHost& b, const WRay& ray, WHit& hit, uint64_t* nv, uint64_t* nt)
{
    WRayPre r;
    wb_ray_setup(r, ray.o[0], ray.o[1], ray.o[2], ray.tmin, ray.d[0], ray.d[1], ray.d[2], ray.tmax);
    float best_t = ray.tmax, h0 = 0.0f, h1 = 0.0f, hdet = 1.0f;
    uint32_t best_id = WBVH_MISS;
    uint32_t stack_base[WBVH_STACK], stack_bits[WBVH_STACK];
    int sp = 0;
    uint32_t g_base = 0, g_hits = 0, g_imask = 0;
    uint32_t pend = 0, tbase = 0, tvalid = 0, pend2 = 0, tbase2 = 0, tvalid2 = 0;
    uint32_t cur = 0;                   // the node to test next (the root first)
    bool more = true;
    uint64_t visits = 0, tests = 0;
    while (true) {
        uint32_t child = 0;
        if (pend2 == 0 && more && !(ANYHIT && best_id != WBVH_MISS)) {
            const WNode& nd = b.nodes[cur];
            visits++;
            uint32_t tmask = 0;
            const uint32_t inner = wb_node_hits(nd, r, best_t, tmask);
            if (pend == 0) {
                pend = tmask; tbase = nd.tri_base; tvalid = nd.tvalid;
            } else {
                pend2 = tmask; tbase2 = nd.tri_base; tvalid2 = nd.tvalid;
            }
            g_base = nd.child_base; g_hits = inner; g_imask = nd.imask;
            if (g_hits == 0) {
                if (sp == 0) {
                    more = false;
                } else {
                    sp--;
                    g_base = stack_base[sp];
                    g_hits = stack_bits[sp] & 0xffu;
                    g_imask = stack_bits[sp] >> 8;
                }
            }
            if (more) {
                const uint32_t p = (uint32_t)wb_ctz(g_hits);
                g_hits &= g_hits - 1u;
                const uint32_t s = p ^ r.oct;
                child = g_base + (uint32_t)wb_popc(g_imask & ((1u << s) - 1u));
                if (g_hits) {
                    if (sp >= WBVH_STACK) {
                        hit.t = 0.0f; hit.b0 = 0.0f; hit.b1 = 0.0f; hit.id = WBVH_MISS;
                        return false;
                    }
                    stack_base[sp] = g_base;
                    stack_bits[sp] = g_hits | (g_imask << 8);
                    sp++;
                }
            }
        } else if (pend2 != 0 && more) {
            child = cur;
        }
        cur = child;
        if (pend) {
            const uint32_t k = (uint32_t)wb_ctz(pend);
            pend &= pend - 1u;
            const WTri& tri = b.tris[wb_tri_index(tbase, tvalid, k)];
            const uint32_t id = wb_tri_id(tri, ISECT);
            tests++;
            float t, a0, a1, dt = 1.0f;
            bool h;
            if (ISECT == WBVH_ISECT_WATERTIGHT) h = wb_tri_watertight(tri, r, best_t, t, a0, a1, dt);
            else h = wb_tri_triaccel(tri, r, best_t, t, a0, a1);
            if (h && (t < best_t || id < best_id)) {
                best_t = t; best_id = id; h0 = a0; h1 = a1; hdet = dt;
            }
            if (ANYHIT && best_id != WBVH_MISS) { pend = 0; pend2 = 0; more = false; }
        }
        if (pend == 0) {
            pend = pend2; tbase = tbase2; tvalid = tvalid2;
            pend2 = 0;
        }
        if (pend == 0 && !more && cur == 0) break;
    }
    if (best_id == WBVH_MISS) {
        hit.t = 0.0f; hit.b0 = 0.0f; hit.b1 = 0.0f; hit.id = WBVH_MISS;
    } else {
        hit.t = best_t; hit.id = best_id;
        if (ISECT == WBVH_ISECT_WATERTIGHT) { hit.b0 = h0 / hdet; hit.b1 = h1 / hdet; }
        else { hit.b0 = h0; hit.b1 = h1; }
    }
    if (nv) *nv += visits;
    if (nt) *nt += tests;
    return true;
}

bool wbvh_walk(const WbvhHost& b, const WRay& ray, WHit& hit, uint64_t* nv, uint64_t* nt)
{
    if (b.isect == WBVH_ISECT_WATERTIGHT) return walk_impl<WBVH_ISECT_WATERTIGHT, false>(b, ray, hit, nv, nt);
    return walk_impl<WBVH_ISECT_TRIACCEL, false>(b, ray, hit, nv, nt);
}

bool wbvh_walk_anyhit(const WbvhHost& b, const WRay& ray, WHit& hit, uint64_t* nv, uint64_t* nt)
{
    if (b.isect == WBVH_ISECT_WATERTIGHT) return walk_impl<WBVH_ISECT_WATERTIGHT, true>(b, ray, hit, nv, nt);
    return walk_impl<WBVH_ISECT_TRIACCEL, true>(b, ray, hit, nv, nt);
}

}  // namespace lmx

// ---------------------------------------------------------------------------
// device kernels

#define TRACE_WAVE 64
// lmx_accel_intersect_host walks batches below this size on the calling thread
#define LMX_HOST_INLINE_RAYS 1024u
#ifndef LMX_REFILL_MIN
// refill when at least this many lanes of a wave are idle (8, 16, 24, 32 and
// 48 measured on 2048^2 coherent + 2048^2 incoherent rays over 1.3M
// triangles: 24 is 1.6 % faster than 8; 48 is 9 % slower)
#define LMX_REFILL_MIN 24
#endif
#ifndef LMX_TRACE_WAVES_PER_EU
#define LMX_TRACE_WAVES_PER_EU 1   // occupancy hint for the register allocator (5 measured 2 % slower)
#endif
#define REFILL_MIN LMX_REFILL_MIN

// Sets the 8 region cursors for a launch of `grid` waves.  The ray count
// comes from n_dev when given (a queue length written by an earlier
// kernel), else from n; the count the trace kernel uses is stored after the
// cursors.
__global__ void k_init_cursors(uint32_t* cursors, uint32_t n, const uint32_t* n_dev, uint32_t grid)
{
    trace_cursors_set(cursors, n_dev ? *n_dev : n, grid, threadIdx.x);
}

// The node test's bit shuffles as tables (the internal-child hit mask in
// each octant's order; leaf slot bits spread to triangle bits), built at
// compile time and copied into each workgroup's LDS at launch.
struct alignas(16) TraceTables {
    uint8_t order[8 * 256];
    uint32_t spread[256];
};

constexpr TraceTables make_trace_tables()
{
    TraceTables t{};
    for (uint32_t i = 0; i < 8u * 256u; i++) t.order[i] = (uint8_t)wb_octant_order(i & 255u, i >> 8);
    for (uint32_t i = 0; i < 256u; i++) t.spread[i] = wb_leaf_spread(i);
    return t;
}

__constant__ TraceTables g_trace_tables = make_trace_tables();

// Node idx through a raw buffer resource over the node array: one 32-bit
// offset per lane instead of a 64-bit address (3 VGPRs and an address
// computation less per node, measured +1 %), and loads past the array's
// bytes read zero.
__device__ __forceinline__ void load_node(WNode& nd, __amdgpu_buffer_rsrc_t rs, uint32_t idx)
{
    uint4* d = reinterpret_cast<uint4*>(&nd);
    const int off = (int)(idx * (uint32_t)sizeof(WNode));
#pragma unroll
    for (int i = 0; i < (int)(sizeof(WNode) / 16); i++) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off + 16 * i, 0, 0);
        d[i] = make_uint4(v[0], v[1], v[2], v[3]);
    }
}

__device__ __forceinline__ void load_tri(WTri& t, const WTri* src)
{
    const float4* s = reinterpret_cast<const float4*>(src);
    float4* d = reinterpret_cast<float4*>(&t);
    d[0] = s[0]; d[1] = s[1]; d[2] = s[2];
}

// Closest-hit update with the canonical tie rule (smallest id among equal t).
template <int ISECT>
__device__ __forceinline__ void test_tri(const WTri& tri, const WRayPre& r, float& best_t, uint32_t& best_id,
                                         float& h0, float& h1, float& hdet)
{
    const uint32_t id = wb_tri_id(tri, ISECT);
    float t, a0, a1, dt = 1.0f;
    bool h;
    if (ISECT == WBVH_ISECT_WATERTIGHT) h = wb_tri_watertight(tri, r, best_t, t, a0, a1, dt);
    else h = wb_tri_triaccel(tri, r, best_t, t, a0, a1);
    if (h && (t < best_t || id < best_id)) {
        best_t = t; best_id = id; h0 = a0; h1 = a1; hdet = dt;
    }
}

// Hit record of a finished ray (b0 = V/det, b1 = W/det for the watertight
// test; a miss is all zeros but the id).
template <int ISECT>
__device__ __forceinline__ void write_hit(WHit* __restrict__ hits, uint32_t ray_id, float best_t, uint32_t best_id,
                                          float h0, float h1, float hdet)
{
    float4 o;
    if (best_id == WBVH_MISS) {
        o = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(WBVH_MISS));
    } else if (ISECT == WBVH_ISECT_WATERTIGHT) {
        o = make_float4(best_t, h0 / hdet, h1 / hdet, __uint_as_float(best_id));
    } else {
        o = make_float4(best_t, h0, h1, __uint_as_float(best_id));
    }
    reinterpret_cast<float4*>(hits)[ray_id] = o;
}

// The LDS stack is sized per launch to the BVH's depth (stack_depth entries
// of 8 bytes per lane); with the 3 KB of node-test tables it bounds the
// workgroups per CU (22 at depth 8, where the registers allow 20).
// ANYHIT: stop at the first accepted triangle (shadow rays); the hit record
// then names that triangle, not necessarily the closest.
template <int ISECT, bool ANYHIT>
__global__ __launch_bounds__(TRACE_WAVE) __attribute__((amdgpu_waves_per_eu(LMX_TRACE_WAVES_PER_EU)))
void k_trace(const WNode* __restrict__ nodes,
                                                      const WTri* __restrict__ tris,
                                                      const WRay* __restrict__ rays,
                                                      WHit* __restrict__ hits,
                                                      uint32_t* __restrict__ cursors,
                                                      uint32_t* __restrict__ overflow,
                                                      int stack_depth, uint32_t refill_min, uint32_t node_bytes)
{
    extern __shared__ uint2 stack[];     // [stack_depth][TRACE_WAVE]
    const uint32_t lane = threadIdx.x;
    const uint32_t nray = cursors[8 * CURSOR_STRIDE];
    if (nray == 0) return;                  // uniform over the grid
    const uint32_t claim = trace_claim(nray, gridDim.x);
    __shared__ TraceTables s_tab;
    {
        const uint4* src = reinterpret_cast<const uint4*>(&g_trace_tables);
        uint4* dst = reinterpret_cast<uint4*>(&s_tab);
        for (uint32_t i = lane; i < (uint32_t)(sizeof(TraceTables) / 16); i += TRACE_WAVE) dst[i] = src[i];
    }
    __syncthreads();

    uint32_t region = blockIdx.x & 7u;
    uint32_t regions_left = 8;          // regions not yet found exhausted
    uint32_t pool_base = 0, pool_cnt = 0;   // claimed, not yet started rays
    // direction set-up of the pool's rays: lane pool_lane + k holds that of
    // ray pool_base + k (octant | kx << 3 | ky << 5 | kz << 7 in p_meta)
    uint32_t pool_lane = 0;
    float p_ix = 0.0f, p_iy = 0.0f, p_iz = 0.0f, p_sx = 0.0f, p_sy = 0.0f, p_sz = 0.0f;
    uint32_t p_meta = 0;
    bool first = true;                  // the fixed first chunk is still to take
    bool valid = false;
    bool done = false;                  // finished, hit not yet written
    uint32_t ray_id = 0;
    WRayPre r;
    float best_t = 0.0f, h0 = 0.0f, h1 = 0.0f, hdet = 1.0f;
    uint32_t best_id = WBVH_MISS;
    uint32_t g_base = 0, g_hits = 0, g_imask = 0;
    int sp = 0;
    // Leaf triangles still to test, one per iteration, in a queue of two
    // node groups (bits of a node's triangle mask, with the node's triangle
    // base and valid mask): a lane goes on testing nodes while one group is
    // pending and stops only when a second one is queued, and the wave
    // tests nodes and triangles in the same iteration instead of looping as
    // many times as its lane with the most triangles.  `more`: a next node
    // was chosen (it is `cur`, loaded into nd).
    uint32_t pend = 0, tbase = 0, tvalid = 0, cur = 0;
    uint32_t pend2 = 0, tbase2 = 0, tvalid2 = 0;
    bool more = false;
    // raw buffer (stride 0) over the node_bytes of the nodes; dword 3 of the
    // resource is the gfx9-family raw-buffer format
    const __amdgpu_buffer_rsrc_t nrs = __builtin_amdgcn_make_buffer_rsrc((void*)nodes, (short)0, (int)node_bytes,
                                                                         0x00020000);
    WNode nd;                           // node to test in the next iteration
    load_node(nd, nrs, 0u);             // idle lanes always hold the root

    while (true) {
        // ---- batch refill of idle lanes from this wave's pool of claimed
        // rays; an empty pool is refilled first with a fresh chunk, whose
        // direction set-up (reciprocals, octant, shear: the divisions) the
        // whole wave computes at once, a ray per lane, so that a refill
        // costs a few lane shuffles instead of the set-up at the width of
        // the lanes refilled
        const uint64_t idle = __ballot(!valid);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if ((regions_left != 0 || pool_cnt != 0) && (nidle >= refill_min || nidle == TRACE_WAVE)) {
            if (done) {
                write_hit<ISECT>(hits, ray_id, best_t, best_id, h0, h1, hdet);
                done = false;
            }
            if (pool_cnt == 0) {
                uint32_t cb = 0, cg = 0;
                if (first) {                       // this wave's fixed chunk
                    first = false;
                    const uint32_t r0 = blockIdx.x & 7u;
                    const uint64_t s0 = (uint64_t)region_start(r0, nray) + (uint64_t)(blockIdx.x >> 3) * claim;
                    const uint32_t e0 = region_end(r0, nray);
                    if (s0 < e0) {
                        cb = (uint32_t)s0;
                        cg = min(claim, e0 - cb);
                    }
                }
                if (cg == 0) {
                    if (lane == 0) {
                        while (regions_left != 0) {
                            const uint32_t end = region_end(region, nray);
                            // a region whose fixed chunks cover it has
                            // nothing for the cursor
                            if (region_dyn_start(region, nray, gridDim.x, claim) < end) {
                                const uint32_t old = atomicAdd(&cursors[region * CURSOR_STRIDE], claim);
                                if (old < end) {
                                    cb = old;
                                    cg = min(claim, end - old);
                                    break;
                                }
                            }
                            region = (region + 1u) & 7u;
                            regions_left--;
                        }
                    }
                    cb = __shfl(cb, 0);
                    cg = __shfl(cg, 0);
                    regions_left = __shfl(regions_left, 0);
                    region = __shfl(region, 0);
                }
                pool_base = cb;
                pool_cnt = cg;
                pool_lane = 0;
                if (lane < cg) {
                    const float4 b = reinterpret_cast<const float4*>(rays)[2 * (cb + lane) + 1];
                    WRayPre q;
                    wb_ray_setup(q, 0.0f, 0.0f, 0.0f, 0.0f, b.x, b.y, b.z, 0.0f);
                    p_ix = q.ix; p_iy = q.iy; p_iz = q.iz;
                    p_sx = q.Sx; p_sy = q.Sy; p_sz = q.Sz;
                    p_meta = q.oct | ((uint32_t)q.kx << 3) | ((uint32_t)q.ky << 5) | ((uint32_t)q.kz << 7);
                }
            }
            const uint32_t take = min(nidle, pool_cnt);
            if (take != 0) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                       __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                const int src = (int)(pool_lane + min(rank, take - 1u));
                const float ix = __shfl(p_ix, src), iy = __shfl(p_iy, src), iz = __shfl(p_iz, src);
                const float sx = __shfl(p_sx, src), sy = __shfl(p_sy, src), sz = __shfl(p_sz, src);
                const uint32_t meta = (uint32_t)__shfl((int)p_meta, src);
                if (!valid && rank < take) {
                    ray_id = pool_base + rank;
                    const float4 a = reinterpret_cast<const float4*>(rays)[2 * ray_id];
                    const float4 b = reinterpret_cast<const float4*>(rays)[2 * ray_id + 1];
                    r.ox = a.x; r.oy = a.y; r.oz = a.z; r.tmin = a.w;
                    r.dx = b.x; r.dy = b.y; r.dz = b.z; r.tmax = b.w;
                    r.ix = ix; r.iy = iy; r.iz = iz;
                    r.Sx = sx; r.Sy = sy; r.Sz = sz;
                    r.oct = meta & 7u;
                    r.kx = (int)((meta >> 3) & 3u); r.ky = (int)((meta >> 5) & 3u); r.kz = (int)((meta >> 7) & 3u);
                    best_t = b.w; best_id = WBVH_MISS; h0 = h1 = 0.0f; hdet = 1.0f;
                    sp = 0;                        // nd already holds the root
                    pend = 0; pend2 = 0; more = true;
                    valid = true;
                }
                pool_base += take;
                pool_cnt -= take;
                pool_lane += take;
            }
        }
        // single loop back-edge: with no ray in any lane and work left, the
        // body below does nothing but reload the root
        if (__ballot(valid) == 0 && regions_left == 0 && pool_cnt == 0) break;

        // 1. lanes with room in their triangle queue: box test of the node
        //    loaded in the previous iteration, its leaf triangles queued, and
        //    the choice of the next node; lanes with a full queue keep their
        //    node (it is loaded again).  Every lane loads exactly one node
        //    per iteration at one site (lanes without a next node load the
        //    root, which is what a refilled lane needs first).
        uint32_t child = 0;
        if (valid && pend2 == 0 && more && !(ANYHIT && best_id != WBVH_MISS)) {
            const uint32_t hb = wb_slot_hits(nd, r, best_t);
            const uint32_t inner = s_tab.order[(r.oct << 8) | (hb & (uint32_t)nd.imask)];
            const uint32_t tmask = s_tab.spread[hb & ~(uint32_t)nd.imask & 0xffu] & nd.tvalid;
            if (pend == 0) {
                pend = tmask; tbase = nd.tri_base; tvalid = nd.tvalid;
            } else {
                pend2 = tmask; tbase2 = nd.tri_base; tvalid2 = nd.tvalid;
            }
            g_base = nd.child_base; g_hits = inner; g_imask = nd.imask;
            if (g_hits == 0) {
                if (sp == 0) {
                    more = false;
                } else {
                    sp--;
                    const uint2 e = stack[sp * TRACE_WAVE + lane];
                    g_base = e.x; g_hits = e.y & 0xffu; g_imask = e.y >> 8;
                }
            }
            if (more) {
                const uint32_t p = (uint32_t)(__ffs(g_hits) - 1);
                g_hits &= g_hits - 1u;
                const uint32_t s = p ^ r.oct;
                child = g_base + (uint32_t)__popc(g_imask & ((1u << s) - 1u));
                if (g_hits) {
                    if (sp < stack_depth) {
                        stack[sp * TRACE_WAVE + lane] = make_uint2(g_base, g_hits | (g_imask << 8));
                        sp++;
                    } else {
                        // unreachable for a tree the builder accepts (the
                        // stack holds max_depth groups); flagged, never silent
                        atomicOr(overflow, 1u);
                    }
                }
            }
        } else if (valid && pend2 != 0 && more) {
            child = cur;                       // the node chosen before the queued triangles
        }
        // 2. load the next node; its latency overlaps the triangle test below
        cur = child;
        load_node(nd, nrs, child);
        if (valid) {
            // 3. one triangle of the tested node
            if (pend) {
                const uint32_t k1 = (uint32_t)(__ffs(pend) - 1);
                pend &= pend - 1u;
                WTri t1;
                load_tri(t1, tris + wb_tri_index(tbase, tvalid, k1));
                test_tri<ISECT>(t1, r, best_t, best_id, h0, h1, hdet);
                if (ANYHIT && best_id != WBVH_MISS) { pend = 0; pend2 = 0; more = false; }
            }
            if (pend == 0) {                   // the queued group moves up
                pend = pend2; tbase = tbase2; tvalid = tvalid2;
                pend2 = 0;
            }
            // 4. ray finished: free the lane; its hit is written with those
            //    of the other lanes finished by then at the next refill (or
            //    at the end), so the wave runs the divisions of the hit
            //    record once per refill rather than once per iteration.  An
            //    any-hit ray that found a triangle with a next node already
            //    loaded stops one iteration later, after the root has been
            //    loaded.
            if (pend == 0 && !more && cur == 0) {
                valid = false;
                done = true;
            }
        }
    }
    if (done) write_hit<ISECT>(hits, ray_id, best_t, best_id, h0, h1, hdet);
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char* lmx_last_error(void) { return g_err.c_str(); }

int lmx_accel_build(const float* verts, uint32_t nvert, const uint32_t* tris, uint32_t ntri,
                    const lmx_build_opts* opts, lmx_accel** out)
{
    if (!out) return set_error(LMX_EINVAL, "lmx_accel_build: out is NULL");
    *out = nullptr;
    if (!verts || !tris) return set_error(LMX_EINVAL, "lmx_accel_build: NULL mesh arrays");
    int isect = LMX_ISECT_WATERTIGHT;
    int nthreads = 0;
    if (opts) { isect = opts->isect; nthreads = opts->nthreads; }
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    lmx_accel* a = new (std::nothrow) lmx_accel();
    if (!a) return set_error(LMX_ENOMEM, "lmx_accel_build: out of host memory");
    const char* err = wbvh_build(verts, nvert, tris, ntri, isect, nthreads, a->host);
    if (err) {
        delete a;
        return set_error(LMX_EBUILD, std::string("lmx_accel_build: ") + err);
    }
    *out = a;
    return LMX_OK;
}

void lmx_accel_free(lmx_accel* a)
{
    if (!a) return;
    if (a->device >= 0) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(a->device);
        (void)hipFree(a->d_nodes);
        (void)hipFree(a->d_tris);
        (void)hipFree(a->d_counters);
        (void)hipSetDevice(cur);
    }
    delete a;
}

int lmx_accel_info_get(const lmx_accel* a, lmx_accel_info* info)
{
    if (!a || !info) return set_error(LMX_EINVAL, "lmx_accel_info_get: NULL argument");
    info->num_nodes = (uint32_t)a->host.nodes.size();
    info->num_tris = (uint32_t)a->host.tris.size();
    info->max_depth = a->host.max_depth;
    info->isect = a->host.isect;
    info->node_bytes = (uint64_t)a->host.nodes.size() * sizeof(WNode);
    info->tri_bytes = (uint64_t)a->host.tris.size() * sizeof(WTri);
    info->stack_overflow = 0;
    info->reserved = 0;
    if (a->device >= 0 && a->d_counters) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(a->device);
        const hipError_t e = hipMemcpy(&info->stack_overflow, a->d_counters + 9 * CURSOR_STRIDE, sizeof(uint32_t),
                                       hipMemcpyDeviceToHost);
        (void)hipSetDevice(cur);
        if (e != hipSuccess) return hip_error(e, "lmx_accel_info_get: overflow flag");
    }
    return LMX_OK;
}

int lmx_accel_export(const lmx_accel* a, void* nodes, void* tris)
{
    if (!a) return set_error(LMX_EINVAL, "lmx_accel_export: NULL accel");
    if (nodes) std::memcpy(nodes, a->host.nodes.data(), a->host.nodes.size() * sizeof(WNode));
    if (tris) std::memcpy(tris, a->host.tris.data(), a->host.tris.size() * sizeof(WTri));
    return LMX_OK;
}

int lmx_accel_upload(lmx_accel* a)
{
    if (!a) return set_error(LMX_EINVAL, "lmx_accel_upload: NULL accel");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_error(e, "lmx_accel_upload: hipGetDevice");
    if (a->device == dev) return LMX_OK;
    if (a->device >= 0) return set_error(LMX_EINVAL, "lmx_accel_upload: accel already lives on another device");
    const size_t nb = a->host.nodes.size() * sizeof(WNode), tb = a->host.tris.size() * sizeof(WTri);
    if ((e = hipMalloc(&a->d_nodes, nb)) != hipSuccess) return hip_error(e, "lmx_accel_upload: hipMalloc nodes");
    if ((e = hipMalloc(&a->d_tris, tb)) != hipSuccess) {
        (void)hipFree(a->d_nodes); a->d_nodes = nullptr;
        return hip_error(e, "lmx_accel_upload: hipMalloc triangles");
    }
    if ((e = hipMalloc(&a->d_counters, LMX_TRACE_COUNTER_WORDS * sizeof(uint32_t))) != hipSuccess) {
        (void)hipFree(a->d_nodes); (void)hipFree(a->d_tris); a->d_nodes = nullptr; a->d_tris = nullptr;
        return hip_error(e, "lmx_accel_upload: hipMalloc counters");
    }
    if ((e = hipMemcpy(a->d_nodes, a->host.nodes.data(), nb, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(a->d_tris, a->host.tris.data(), tb, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(a->d_counters, 0, LMX_TRACE_COUNTER_WORDS * sizeof(uint32_t))) != hipSuccess) {
        (void)hipFree(a->d_nodes); (void)hipFree(a->d_tris); (void)hipFree(a->d_counters);
        a->d_nodes = nullptr; a->d_tris = nullptr; a->d_counters = nullptr;
        return hip_error(e, "lmx_accel_upload: hipMemcpy");
    }
    int cus = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
        return hip_error(e, "lmx_accel_upload: device attribute");
    // A node group is pushed only when descending one level, so the stack
    // never holds more entries than the tree has levels below the root.
    a->stack_depth = (int)std::max(1u, a->host.max_depth);
    a->lds_bytes = (size_t)a->stack_depth * TRACE_WAVE * sizeof(uint2);
    int per_cu = 0;
    if (a->host.isect == WBVH_ISECT_WATERTIGHT)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<WBVH_ISECT_WATERTIGHT, false>, TRACE_WAVE, a->lds_bytes);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<WBVH_ISECT_TRIACCEL, false>, TRACE_WAVE, a->lds_bytes);
    if (e != hipSuccess) return hip_error(e, "lmx_accel_upload: occupancy query");
    a->grid = std::max(8, cus * std::max(1, per_cu));
    a->device = dev;
    return LMX_OK;
}

}  // extern "C"

namespace lmx {

// Enqueues one traversal launch of an accel already resident on the current
// device, with its region cursors and ray count at `counters` (9 words,
// CURSOR_STRIDE apart, of device memory); the stack-overflow flag is always
// the accel's own.
static int trace_grid(const lmx_accel* a, uint32_t n_max)
{
    return std::max(1, std::min(a->grid, (int)((n_max + TRACE_WAVE - 1) / TRACE_WAVE)));
}

static int enqueue_trace(const lmx_accel* a, uint32_t* counters, const WRay* rays, uint32_t n,
                         const uint32_t* n_dev, uint32_t n_max, WHit* hits, hipStream_t s, bool anyhit,
                         bool cursors_set = false)
{
    uint32_t* overflow = a->d_counters + 9 * CURSOR_STRIDE;
    const int grid = trace_grid(a, n_max);
    const uint32_t node_bytes = (uint32_t)(a->host.nodes.size() * sizeof(WNode));   // < 4 GB, the builder's bound
    if (!cursors_set) k_init_cursors<<<1, 64, 0, s>>>(counters, n, n_dev, (uint32_t)grid);
    const bool wt = a->host.isect == WBVH_ISECT_WATERTIGHT;
    if (wt && !anyhit)
        k_trace<WBVH_ISECT_WATERTIGHT, false><<<grid, TRACE_WAVE, a->lds_bytes, s>>>(
            a->d_nodes, a->d_tris, rays, hits, counters, overflow, a->stack_depth, REFILL_MIN, node_bytes);
    else if (wt)
        k_trace<WBVH_ISECT_WATERTIGHT, true><<<grid, TRACE_WAVE, a->lds_bytes, s>>>(
            a->d_nodes, a->d_tris, rays, hits, counters, overflow, a->stack_depth, REFILL_MIN, node_bytes);
    else if (!anyhit)
        k_trace<WBVH_ISECT_TRIACCEL, false><<<grid, TRACE_WAVE, a->lds_bytes, s>>>(
            a->d_nodes, a->d_tris, rays, hits, counters, overflow, a->stack_depth, REFILL_MIN, node_bytes);
    else
        k_trace<WBVH_ISECT_TRIACCEL, true><<<grid, TRACE_WAVE, a->lds_bytes, s>>>(
            a->d_nodes, a->d_tris, rays, hits, counters, overflow, a->stack_depth, REFILL_MIN, node_bytes);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "trace launch");
    return LMX_OK;
}

int launch_trace(lmx_accel* a, const WRay* rays, uint32_t n, const uint32_t* n_dev, uint32_t n_max,
                 WHit* hits, hipStream_t s, bool anyhit, bool cursors_set)
{
    int rc = lmx_accel_upload(a);
    if (rc != LMX_OK) return rc;
    return enqueue_trace(a, a->d_counters, rays, n, n_dev, n_max, hits, s, anyhit, cursors_set);
}

int trace_shape(lmx_accel* a, uint32_t n_max, uint32_t** counters, uint32_t* grid)
{
    int rc = lmx_accel_upload(a);
    if (rc != LMX_OK) return rc;
    *counters = a->d_counters;
    *grid = (uint32_t)trace_grid(a, n_max);
    return LMX_OK;
}

}  // namespace lmx

// ---------------------------------------------------------------------------
// Allocation-free closest hit (lmx_accel_intersect_noalloc)
//
// The same launch as lmx_accel_intersect, with the launch's scratch -- the 8
// region cursors and the ray count, each on a 256-byte line of its own,
// 2,304 bytes, whatever the number of rays -- in this file-scope array
// instead of the accel's counter block.  Every GPU the library runs on holds
// its own copy until the process ends, shared by every call on that GPU;
// k_init_cursors rewrites all 9 words at the start of each launch, so no
// value carries over from the previous call (two calls that overlap on one
// GPU, on different streams, race on it).  On the host, `scratch` only
// stands in for the device symbol: its device address comes from
// hipGetSymbolAddress, taken once per device with that device current, and
// only that address is handed to kernels.
#define LMX_NOALLOC_SCRATCH_WORDS (9 * CURSOR_STRIDE)
static __device__ __align__(256) uint32_t scratch[LMX_NOALLOC_SCRATCH_WORDS];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::atomic<uint32_t*> g_scratch_dev[SCRATCH_MAX_DEVICES];
std::mutex g_scratch_mu;

// Device address of `scratch` on device `dev`, which is current.
int scratch_on(int dev, uint32_t** out)
{
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
        return set_error(LMX_EINVAL, "lmx_accel_intersect_noalloc: device index beyond 63");
    uint32_t* p = g_scratch_dev[dev].load(std::memory_order_acquire);
    if (!p) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        p = g_scratch_dev[dev].load(std::memory_order_relaxed);
        if (!p) {
            void* q = nullptr;
            const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
            if (e != hipSuccess) return hip_error(e, "lmx_accel_intersect_noalloc: hipGetSymbolAddress(scratch)");
            if (!q) return set_error(LMX_EHIP, "lmx_accel_intersect_noalloc: no device address for scratch");
            p = static_cast<uint32_t*>(q);
            g_scratch_dev[dev].store(p, std::memory_order_release);
        }
    }
    *out = p;
    return LMX_OK;
}

}  // namespace

extern "C" {

int lmx_accel_intersect_noalloc(lmx_accel* a, const lmx_ray* rays, uint32_t n, lmx_hit* hits, void* stream)
{
    static_assert(LMX_NOALLOC_SCRATCH_WORDS * sizeof(uint32_t) == 2304, "scratch size stated in lmx.h");
    if (!a) return set_error(LMX_EINVAL, "lmx_accel_intersect_noalloc: NULL accel");
    if (n == 0) return LMX_OK;
    if (!rays || !hits) return set_error(LMX_EINVAL, "lmx_accel_intersect_noalloc: NULL ray or hit array");
    if (n > 0x7fffffffu)
        return set_error(LMX_EINVAL, "lmx_accel_intersect_noalloc: too many rays (at most 2^31 - 1 per call)");
    int dev = 0;
    const hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_error(e, "lmx_accel_intersect_noalloc: hipGetDevice");
    if (a->device != dev || !a->d_nodes)
        return set_error(LMX_EINVAL, "lmx_accel_intersect_noalloc: the accel is not resident on the current "
                                     "device; call lmx_accel_upload first (this entry point allocates nothing)");
    uint32_t* cur = nullptr;
    const int rc = scratch_on(dev, &cur);
    if (rc != LMX_OK) return rc;
    return enqueue_trace(a, cur, reinterpret_cast<const WRay*>(rays), n, nullptr, n, reinterpret_cast<WHit*>(hits),
                         (hipStream_t)stream, false);
}

int lmx_accel_intersect(lmx_accel* a, const lmx_ray* rays, uint32_t n, lmx_hit* hits, void* stream)
{
    if (!a) return set_error(LMX_EINVAL, "lmx_accel_intersect: NULL accel");
    if (n == 0) return LMX_OK;
    if (!rays || !hits) return set_error(LMX_EINVAL, "lmx_accel_intersect: NULL ray or hit array");
    if (n > 0x7fffffffu) return set_error(LMX_EINVAL, "lmx_accel_intersect: too many rays");
    return launch_trace(a, reinterpret_cast<const WRay*>(rays), n, nullptr, n, reinterpret_cast<WHit*>(hits),
                        (hipStream_t)stream, false);
}

int lmx_accel_occluded(lmx_accel* a, const lmx_ray* rays, uint32_t n, lmx_hit* hits, void* stream)
{
    if (!a) return set_error(LMX_EINVAL, "lmx_accel_occluded: NULL accel");
    if (n == 0) return LMX_OK;
    if (!rays || !hits) return set_error(LMX_EINVAL, "lmx_accel_occluded: NULL ray or hit array");
    if (n > 0x7fffffffu) return set_error(LMX_EINVAL, "lmx_accel_occluded: too many rays");
    return launch_trace(a, reinterpret_cast<const WRay*>(rays), n, nullptr, n, reinterpret_cast<WHit*>(hits),
                        (hipStream_t)stream, true);
}

int lmx_accel_intersect_host(const lmx_accel* a, const lmx_ray* rays, uint32_t n, lmx_hit* hits,
                             int32_t nthreads)
{
    if (!a) return set_error(LMX_EINVAL, "lmx_accel_intersect_host: NULL accel");
    if (n == 0) return LMX_OK;
    if (!rays || !hits) return set_error(LMX_EINVAL, "lmx_accel_intersect_host: NULL array");
    const WRay* r = reinterpret_cast<const WRay*>(rays);
    WHit* h = reinterpret_cast<WHit*>(hits);
    // The per-ray path (Accel3::Intersect from every scheduler thread at
    // once): walked on the calling thread, no allocation, no lock; reads of
    // the accel only, so any number of threads may call it concurrently.
    if (nthreads == 1 || n < LMX_HOST_INLINE_RAYS) {
        bool ok = true;
        for (uint32_t i = 0; i < n; i++) ok &= wbvh_walk(a->host, r[i], h[i], nullptr, nullptr);
        return ok ? LMX_OK : set_error(LMX_EINVAL, "lmx_accel_intersect_host: traversal stack overflow");
    }
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = (int)std::min<uint32_t>((uint32_t)nthreads, (n + LMX_HOST_INLINE_RAYS - 1) / LMX_HOST_INLINE_RAYS);
    std::vector<uint8_t> ok(nthreads, 1);
    auto work = [&](int t, uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; i++)
            if (!wbvh_walk(a->host, r[i], h[i], nullptr, nullptr)) ok[t] = 0;
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; t++) {
        const uint32_t b = (uint32_t)((uint64_t)n * t / nthreads), e = (uint32_t)((uint64_t)n * (t + 1) / nthreads);
        th.emplace_back(work, t, b, e);
    }
    work(0, 0, (uint32_t)((uint64_t)n / nthreads));
    for (auto& x : th) x.join();
    for (int t = 0; t < nthreads; t++)
        if (!ok[t]) return set_error(LMX_EINVAL, "lmx_accel_intersect_host: traversal stack overflow");
    return LMX_OK;
}

static int count_impl(const lmx_accel* a, const lmx_ray* rays, uint32_t n, uint64_t* node_visits,
                      uint64_t* tri_tests, int32_t nthreads, bool anyhit);

int lmx_accel_count(const lmx_accel* a, const lmx_ray* rays, uint32_t n, uint64_t* node_visits,
                    uint64_t* tri_tests, int32_t nthreads)
{
    return count_impl(a, rays, n, node_visits, tri_tests, nthreads, false);
}

int lmx_accel_count_occluded(const lmx_accel* a, const lmx_ray* rays, uint32_t n, uint64_t* node_visits,
                             uint64_t* tri_tests, int32_t nthreads)
{
    return count_impl(a, rays, n, node_visits, tri_tests, nthreads, true);
}

}  // extern "C"

static int count_impl(const lmx_accel* a, const lmx_ray* rays, uint32_t n, uint64_t* node_visits,
                      uint64_t* tri_tests, int32_t nthreads, bool anyhit)
{
    if (!a || !node_visits || !tri_tests) return set_error(LMX_EINVAL, "lmx_accel_count: NULL argument");
    *node_visits = 0; *tri_tests = 0;
    if (n == 0) return LMX_OK;
    if (!rays) return set_error(LMX_EINVAL, "lmx_accel_count: NULL ray array");
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<uint64_t> nv(nthreads, 0), nt(nthreads, 0);
    auto work = [&](int t, uint32_t b, uint32_t e) {
        WHit h;
        for (uint32_t i = b; i < e; i++) {
            if (anyhit) wbvh_walk_anyhit(a->host, reinterpret_cast<const WRay*>(rays)[i], h, &nv[t], &nt[t]);
            else wbvh_walk(a->host, reinterpret_cast<const WRay*>(rays)[i], h, &nv[t], &nt[t]);
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        const uint32_t b = (uint32_t)((uint64_t)n * t / nthreads), e = (uint32_t)((uint64_t)n * (t + 1) / nthreads);
        th.emplace_back(work, t, b, e);
    }
    for (auto& x : th) x.join();
    for (int t = 0; t < nthreads; t++) { *node_visits += nv[t]; *tri_tests += nt[t]; }
    return LMX_OK;
}


