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
//   * lanes whose ray finished are refilled from the work queue as a batch:
//     one ballot counts them, one atomic per wave claims that many rays and
//     mbcnt ranks give every idle lane its ray;
//   * XCD-aware work split: workgroup b runs on XCD (b mod 8); the rays are
//     cut into 8 contiguous regions and a wave first drains the region of
//     its own XCD (neighbouring rays share BVH nodes in that XCD's L2), then
//     helps the others.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "lmx.h"
#include "lmx_internal.h"
#include "wbvh_build.h"
#include "wbvh_math.h"

namespace lmx {

// ---------------------------------------------------------------------------
// errors

static thread_local std::string g_err;

int set_error(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int hip_error(hipError_t e, const char* what)
{
    return set_error(LMX_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace lmx

using namespace lmx;

struct lmx_accel {
    WbvhHost host;
    int device = -1;
    WNode* d_nodes = nullptr;
    WTri* d_tris = nullptr;
    uint32_t* d_counters = nullptr;   // 8 region cursors
    int grid = 0;
};

// ---------------------------------------------------------------------------
// host walker (same visiting order as the kernel)

namespace lmx {

template <int ISECT>
static void walk_impl(const WbvhHost& b, const WRay& ray, WHit& hit, uint64_t* nv, uint64_t* nt)
{
    WRayPre r;
    wb_ray_setup(r, ray.o[0], ray.o[1], ray.o[2], ray.tmin, ray.d[0], ray.d[1], ray.d[2], ray.tmax);
    float best_t = ray.tmax, h0 = 0.0f, h1 = 0.0f, hdet = 1.0f;
    uint32_t best_id = WBVH_MISS;
    uint32_t stack_base[WBVH_STACK], stack_bits[WBVH_STACK];
    int sp = 0;
    uint32_t g_base = 0, g_hits = 1u << (0 ^ r.oct), g_imask = 1u;
    uint64_t visits = 0, tests = 0;
    while (true) {
        if (g_hits == 0) {
            if (sp == 0) break;
            sp--;
            g_base = stack_base[sp];
            g_hits = stack_bits[sp] & 0xffu;
            g_imask = stack_bits[sp] >> 8;
        }
        const uint32_t p = (uint32_t)wb_ctz(g_hits);
        g_hits &= g_hits - 1u;
        const uint32_t s = p ^ r.oct;
        const uint32_t child = g_base + (uint32_t)wb_popc(g_imask & ((1u << s) - 1u));
        if (g_hits) {
            stack_base[sp] = g_base;
            stack_bits[sp] = g_hits | (g_imask << 8);
            sp++;
        }
        const WNode& nd = b.nodes[child];
        visits++;
        uint32_t tmask = 0;
        const uint32_t inner = wb_node_hits(nd, r, best_t, tmask);
        g_base = nd.child_base; g_hits = inner; g_imask = nd.imask;
        while (tmask) {
            const uint32_t k = (uint32_t)wb_ctz(tmask);
            tmask &= tmask - 1u;
            const WTri& tri = b.tris[nd.tri_base + k];
            const uint32_t id = wb_tri_id(tri, ISECT);
            tests++;
            float t, a0, a1, dt = 1.0f;
            bool h;
            if (ISECT == WBVH_ISECT_WATERTIGHT) h = wb_tri_watertight(tri, r, best_t, t, a0, a1, dt);
            else h = wb_tri_triaccel(tri, r, best_t, t, a0, a1);
            if (h && (t < best_t || id < best_id)) {
                best_t = t; best_id = id; h0 = a0; h1 = a1; hdet = dt;
            }
        }
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
}

void wbvh_walk(const WbvhHost& b, const WRay& ray, WHit& hit, uint64_t* nv, uint64_t* nt)
{
    if (b.isect == WBVH_ISECT_WATERTIGHT) walk_impl<WBVH_ISECT_WATERTIGHT>(b, ray, hit, nv, nt);
    else walk_impl<WBVH_ISECT_TRIACCEL>(b, ray, hit, nv, nt);
}

}  // namespace lmx

// ---------------------------------------------------------------------------
// device kernels

#define TRACE_WAVE 64
#define REFILL_MIN 16   // refill when at least this many lanes are idle

__global__ void k_init_cursors(uint32_t* cursors, uint32_t n)
{
    const uint32_t j = threadIdx.x;
    if (j < 8) cursors[j] = (uint32_t)(((uint64_t)n * j) / 8u);
}

__device__ __forceinline__ uint32_t region_end(uint32_t j, uint32_t n)
{
    return (uint32_t)(((uint64_t)n * (j + 1u)) / 8u);
}

template <int ISECT>
__global__ __launch_bounds__(TRACE_WAVE) void k_trace(const WNode* __restrict__ nodes,
                                                      const WTri* __restrict__ tris,
                                                      const WRay* __restrict__ rays,
                                                      WHit* __restrict__ hits, uint32_t nray,
                                                      uint32_t* __restrict__ cursors)
{
    __shared__ uint2 stack[WBVH_STACK][TRACE_WAVE];
    const uint32_t lane = threadIdx.x;

    uint32_t region = blockIdx.x & 7u;
    uint32_t regions_left = 8;          // regions not yet found exhausted
    bool valid = false;
    uint32_t ray_id = 0;
    WRayPre r;
    float best_t = 0.0f, h0 = 0.0f, h1 = 0.0f, hdet = 1.0f;
    uint32_t best_id = WBVH_MISS;
    uint32_t g_base = 0, g_hits = 0, g_imask = 0;
    int sp = 0;

    while (true) {
        // ---- batch refill of idle lanes
        const uint64_t idle = __ballot(!valid);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (regions_left != 0 && nidle >= REFILL_MIN) {
            uint32_t base = 0, got = 0;
            if (lane == 0) {
                while (regions_left != 0) {
                    const uint32_t end = region_end(region, nray);
                    const uint32_t old = atomicAdd(&cursors[region], nidle);
                    if (old < end) {
                        base = old;
                        got = min(nidle, end - old);
                        break;
                    }
                    region = (region + 1u) & 7u;
                    regions_left--;
                }
            }
            base = __shfl(base, 0);
            got = __shfl(got, 0);
            regions_left = __shfl(regions_left, 0);
            region = __shfl(region, 0);
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            if (!valid && rank < got) {
                ray_id = base + rank;
                const float4 a = reinterpret_cast<const float4*>(rays)[2 * ray_id];
                const float4 b = reinterpret_cast<const float4*>(rays)[2 * ray_id + 1];
                wb_ray_setup(r, a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w);
                best_t = b.w; best_id = WBVH_MISS; h0 = h1 = 0.0f; hdet = 1.0f;
                g_base = 0; g_hits = 1u << (0u ^ r.oct); g_imask = 1u;
                sp = 0;
                valid = true;
            }
        }
        if (__ballot(valid) == 0) {
            if (regions_left == 0) break;
            continue;
        }

        if (valid) {
            if (g_hits == 0) {
                if (sp == 0) {
                    float4 o;
                    if (best_id == WBVH_MISS) {
                        o = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(WBVH_MISS));
                    } else if (ISECT == WBVH_ISECT_WATERTIGHT) {
                        o = make_float4(best_t, h0 / hdet, h1 / hdet, __uint_as_float(best_id));
                    } else {
                        o = make_float4(best_t, h0, h1, __uint_as_float(best_id));
                    }
                    reinterpret_cast<float4*>(hits)[ray_id] = o;
                    valid = false;
                } else {
                    sp--;
                    const uint2 e = stack[sp][lane];
                    g_base = e.x; g_hits = e.y & 0xffu; g_imask = e.y >> 8;
                }
            }
        }
        if (valid) {
            const uint32_t p = (uint32_t)(__ffs(g_hits) - 1);
            g_hits &= g_hits - 1u;
            const uint32_t s = p ^ r.oct;
            const uint32_t child = g_base + (uint32_t)__popc(g_imask & ((1u << s) - 1u));
            if (g_hits && sp < WBVH_STACK) {
                stack[sp][lane] = make_uint2(g_base, g_hits | (g_imask << 8));
                sp++;
            }
            WNode nd;
            {
                const uint4* src = reinterpret_cast<const uint4*>(nodes + child);
                uint4* dst = reinterpret_cast<uint4*>(&nd);
#pragma unroll
                for (int i = 0; i < 8; i++) dst[i] = src[i];
            }
            uint32_t tmask = 0;
            const uint32_t inner = wb_node_hits(nd, r, best_t, tmask);
            g_base = nd.child_base; g_hits = inner; g_imask = nd.imask;
            const uint32_t tbase = nd.tri_base;
            while (tmask) {
                const uint32_t k = (uint32_t)(__ffs(tmask) - 1);
                tmask &= tmask - 1u;
                WTri tri;
                {
                    const float4* src = reinterpret_cast<const float4*>(tris + tbase + k);
                    float4* dst = reinterpret_cast<float4*>(&tri);
                    dst[0] = src[0]; dst[1] = src[1]; dst[2] = src[2];
                }
                const uint32_t id = wb_tri_id(tri, ISECT);
                float t, a0, a1, dt = 1.0f;
                bool h;
                if (ISECT == WBVH_ISECT_WATERTIGHT) h = wb_tri_watertight(tri, r, best_t, t, a0, a1, dt);
                else h = wb_tri_triaccel(tri, r, best_t, t, a0, a1);
                if (h && (t < best_t || id < best_id)) {
                    best_t = t; best_id = id; h0 = a0; h1 = a1; hdet = dt;
                }
            }
        }
    }
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
    if ((e = hipMalloc(&a->d_counters, 8 * sizeof(uint32_t))) != hipSuccess) {
        (void)hipFree(a->d_nodes); (void)hipFree(a->d_tris); a->d_nodes = nullptr; a->d_tris = nullptr;
        return hip_error(e, "lmx_accel_upload: hipMalloc counters");
    }
    if ((e = hipMemcpy(a->d_nodes, a->host.nodes.data(), nb, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(a->d_tris, a->host.tris.data(), tb, hipMemcpyHostToDevice)) != hipSuccess) {
        (void)hipFree(a->d_nodes); (void)hipFree(a->d_tris); (void)hipFree(a->d_counters);
        a->d_nodes = nullptr; a->d_tris = nullptr; a->d_counters = nullptr;
        return hip_error(e, "lmx_accel_upload: hipMemcpy");
    }
    int cus = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
        return hip_error(e, "lmx_accel_upload: device attribute");
    int per_cu = 0;
    if (a->host.isect == WBVH_ISECT_WATERTIGHT)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<WBVH_ISECT_WATERTIGHT>, TRACE_WAVE, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<WBVH_ISECT_TRIACCEL>, TRACE_WAVE, 0);
    if (e != hipSuccess) return hip_error(e, "lmx_accel_upload: occupancy query");
    a->grid = std::max(8, cus * std::max(1, per_cu));
    a->device = dev;
    return LMX_OK;
}

int lmx_accel_intersect(lmx_accel* a, const lmx_ray* rays, uint32_t n, lmx_hit* hits, void* stream)
{
    if (!a) return set_error(LMX_EINVAL, "lmx_accel_intersect: NULL accel");
    if (n == 0) return LMX_OK;
    if (!rays || !hits) return set_error(LMX_EINVAL, "lmx_accel_intersect: NULL ray or hit array");
    if (n > 0x7fffffffu) return set_error(LMX_EINVAL, "lmx_accel_intersect: too many rays");
    int rc = lmx_accel_upload(a);
    if (rc != LMX_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int grid = std::max(1, std::min(a->grid, (int)((n + TRACE_WAVE - 1) / TRACE_WAVE)));
    k_init_cursors<<<1, 64, 0, s>>>(a->d_counters, n);
    if (a->host.isect == WBVH_ISECT_WATERTIGHT)
        k_trace<WBVH_ISECT_WATERTIGHT><<<grid, TRACE_WAVE, 0, s>>>(
            a->d_nodes, a->d_tris, reinterpret_cast<const WRay*>(rays), reinterpret_cast<WHit*>(hits), n, a->d_counters);
    else
        k_trace<WBVH_ISECT_TRIACCEL><<<grid, TRACE_WAVE, 0, s>>>(
            a->d_nodes, a->d_tris, reinterpret_cast<const WRay*>(rays), reinterpret_cast<WHit*>(hits), n, a->d_counters);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "lmx_accel_intersect: launch");
    return LMX_OK;
}

int lmx_accel_intersect_host(const lmx_accel* a, const lmx_ray* rays, uint32_t n, lmx_hit* hits,
                             int32_t nthreads)
{
    if (!a) return set_error(LMX_EINVAL, "lmx_accel_intersect_host: NULL accel");
    if (n == 0) return LMX_OK;
    if (!rays || !hits) return set_error(LMX_EINVAL, "lmx_accel_intersect_host: NULL array");
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    auto work = [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; i++)
            wbvh_walk(a->host, reinterpret_cast<const WRay*>(rays)[i], reinterpret_cast<WHit*>(hits)[i], nullptr, nullptr);
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++) {
        const uint32_t b = (uint32_t)((uint64_t)n * t / nthreads), e = (uint32_t)((uint64_t)n * (t + 1) / nthreads);
        th.emplace_back(work, b, e);
    }
    for (auto& x : th) x.join();
    return LMX_OK;
}

int lmx_accel_count(const lmx_accel* a, const lmx_ray* rays, uint32_t n, uint64_t* node_visits,
                    uint64_t* tri_tests, int32_t nthreads)
{
    if (!a || !node_visits || !tri_tests) return set_error(LMX_EINVAL, "lmx_accel_count: NULL argument");
    *node_visits = 0; *tri_tests = 0;
    if (n == 0) return LMX_OK;
    if (!rays) return set_error(LMX_EINVAL, "lmx_accel_count: NULL ray array");
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<uint64_t> nv(nthreads, 0), nt(nthreads, 0);
    auto work = [&](int t, uint32_t b, uint32_t e) {
        WHit h;
        for (uint32_t i = b; i < e; i++)
            wbvh_walk(a->host, reinterpret_cast<const WRay*>(rays)[i], h, &nv[t], &nt[t]);
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

}  // extern "C"
