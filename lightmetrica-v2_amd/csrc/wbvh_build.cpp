// wbvh_build.cpp -- host construction of the compressed 8-wide BVH.
//
// Replaces Accel_BVHSAHBin::Build / Accel_QBVH::Build
// (src/liblightmetrica/accel/accel_bvh_sahbin.cpp:70-232,
//  src/liblightmetrica/accel/accel_qbvh.cpp:152-389).  The reference builds
// a binned-SAH binary tree single-threaded with a heap allocation per node;
// here:
//   1. binned SAH over all three axes (32 bins), leaves of at most 3
//      triangles, built in parallel with deterministic node numbering (the
//      subtree over primitive range [b,e) owns node slots [2b, 2e-1) and its
//      root sits at slot 2m-1 for split point m), so the result does not
//      depend on thread timing;
//   2. collapse into 8-wide nodes by the cost-optimal dynamic program over
//      the binary tree (Ylitie, Karras and Laine, HPG 2017);
//   3. breadth-first layout (internal children of a node contiguous, leaf
//      triangles of a node contiguous), octant slot assignment, outward
//      quantisation of child boxes to a 255-step grid (stored as bytes).
// Triangle ids are the caller's enumeration order (primitive order x face
// order, accel_bvh_sahbin.cpp:79-111).
#include "wbvh_build.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace lmx {

namespace {

struct PrimRef {
    float lo[3], hi[3], c[3];
};

struct BNode {
    float lo[3], hi[3];
    int32_t left, right;      // -1 for a leaf
    uint32_t first, count;    // leaf range in the index array
};

struct Builder {
    const std::vector<PrimRef>& prims;
    std::vector<uint32_t>& idx;
    std::vector<BNode>& nodes;
    std::atomic<int> threads_left;

    Builder(const std::vector<PrimRef>& p, std::vector<uint32_t>& i, std::vector<BNode>& n, int nthreads)
        : prims(p), idx(i), nodes(n), threads_left(nthreads - 1) {}

    static constexpr int kBins = 32;
    static constexpr float kTraversalCost = 1.0f;

    static float area(const float lo[3], const float hi[3])
    {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }

    struct Bin {
        float lo[3], hi[3];
        uint32_t n;
        void reset()
        {
            for (int a = 0; a < 3; a++) { lo[a] = FLT_MAX; hi[a] = -FLT_MAX; }
            n = 0;
        }
        void add(const float l[3], const float h[3])
        {
            for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], l[a]); hi[a] = std::max(hi[a], h[a]); }
        }
    };

    // Builds the subtree over idx[b,e); returns the slot of its root.
    int32_t build(uint32_t b, uint32_t e)
    {
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        float clo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, chi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (uint32_t i = b; i < e; i++) {
            const PrimRef& p = prims[idx[i]];
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], p.lo[a]); hi[a] = std::max(hi[a], p.hi[a]);
                clo[a] = std::min(clo[a], p.c[a]); chi[a] = std::max(chi[a], p.c[a]);
            }
        }
        const uint32_t n = e - b;
        auto make_leaf = [&](int32_t slot) {
            BNode& nd = nodes[slot];
            std::memcpy(nd.lo, lo, sizeof lo); std::memcpy(nd.hi, hi, sizeof hi);
            nd.left = nd.right = -1; nd.first = b; nd.count = n;
            return slot;
        };
        if (n == 1) return make_leaf((int32_t)(2 * b));

        // binned SAH over the three axes
        int best_axis = -1, best_split = -1;
        float best_cost = FLT_MAX;
        const float node_area = area(lo, hi);
        for (int a = 0; a < 3; a++) {
            const float ext = chi[a] - clo[a];
            if (!(ext > 0.0f)) continue;
            const float k = (float)kBins / ext;
            Bin bins[kBins];
            for (auto& bn : bins) bn.reset();
            for (uint32_t i = b; i < e; i++) {
                const PrimRef& p = prims[idx[i]];
                int bi = (int)((p.c[a] - clo[a]) * k);
                bi = std::min(std::max(bi, 0), kBins - 1);
                bins[bi].add(p.lo, p.hi);
                bins[bi].n++;
            }
            float rarea[kBins];
            uint32_t rcount[kBins];
            Bin acc; acc.reset();
            for (int i = kBins - 1; i > 0; i--) {
                acc.add(bins[i].lo, bins[i].hi); acc.n += bins[i].n;
                rarea[i] = acc.n ? area(acc.lo, acc.hi) : 0.0f;
                rcount[i] = acc.n;
            }
            Bin l; l.reset();
            for (int s = 0; s < kBins - 1; s++) {
                l.add(bins[s].lo, bins[s].hi); l.n += bins[s].n;
                if (l.n == 0 || rcount[s + 1] == 0) continue;
                const float cost = kTraversalCost +
                    (area(l.lo, l.hi) * (float)l.n + rarea[s + 1] * (float)rcount[s + 1]) / node_area;
                if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = s; }
            }
        }
        if (n <= WBVH_MAX_LEAF_TRIS && (best_axis < 0 || (float)n <= best_cost))
            return make_leaf((int32_t)(2 * b));

        uint32_t m;
        if (best_axis < 0) {
            m = b + n / 2;      // all centroids coincide: split by index
        } else {
            const int a = best_axis;
            const float k = (float)kBins / (chi[a] - clo[a]);
            const float c0 = clo[a];
            auto it = std::partition(idx.begin() + b, idx.begin() + e, [&](uint32_t pi) {
                int bi = (int)((prims[pi].c[a] - c0) * k);
                bi = std::min(std::max(bi, 0), kBins - 1);
                return bi <= best_split;
            });
            m = (uint32_t)(it - idx.begin());
            if (m == b || m == e) m = b + n / 2;
        }

        int32_t l, r;
        const bool spawn = n > 65536 && threads_left.fetch_sub(1) > 0;
        if (spawn) {
            std::thread th([&] { l = build(b, m); });
            r = build(m, e);
            th.join();
            threads_left.fetch_add(1);
        } else {
            if (n > 65536) threads_left.fetch_add(1);
            l = build(b, m);
            r = build(m, e);
        }
        const int32_t slot = (int32_t)(2 * m - 1);
        BNode& nd = nodes[slot];
        std::memcpy(nd.lo, lo, sizeof lo); std::memcpy(nd.hi, hi, sizeof hi);
        nd.left = l; nd.right = r; nd.first = 0; nd.count = 0;
        return slot;
    }
};

inline float bn_area(const BNode& n)
{
    const float dx = n.hi[0] - n.lo[0], dy = n.hi[1] - n.lo[1], dz = n.hi[2] - n.lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Cost-optimal collapse of the binary tree into 8-wide nodes (the dynamic
// program of Ylitie, Karras and Laine, HPG 2017, with the binary leaves kept
// as they are).  cost[n][i-1] is the least surface-area cost of the subtree
// at n represented as a forest of at most i trees, a tree being a binary
// leaf (area x triangles x kTri) or a wide node (area x kNode plus the best
// forest of at most 8 trees for its slots).  Because a forest of at most i
// trees is also one of at most i + 1, the cost of splitting i slots between
// the two children is best with all i of them, so choice[n][i-1] is 0 when
// the subtree is best kept as one tree, else the number of the i slots its
// left child gets.  kNode / kTri is the cost of one wide-node visit (8 slot
// boxes) over one triangle test in the traversal kernel.
struct Collapse {
    static constexpr float kNode = 3.0f, kTri = 1.0f;
    const std::vector<BNode>& bn;
    std::vector<float> cost;       // [node][8]
    std::vector<uint8_t> choice;   // [node][8]
    std::vector<uint8_t> split8;   // [node]: left slots of the node's own 8

    explicit Collapse(const std::vector<BNode>& b)
        : bn(b), cost(b.size() * 8, 0.0f), choice(b.size() * 8, 0), split8(b.size(), 1) {}

    void run(int32_t root)
    {
        // post-order over the reachable nodes
        std::vector<std::pair<int32_t, bool>> st;
        st.push_back({root, false});
        while (!st.empty()) {
            const auto [n, expanded] = st.back();
            st.pop_back();
            const BNode& b = bn[n];
            if (b.left >= 0 && !expanded) {
                st.push_back({n, true});
                st.push_back({b.left, false});
                st.push_back({b.right, false});
                continue;
            }
            float* c = &cost[(size_t)n * 8];
            uint8_t* ch = &choice[(size_t)n * 8];
            const float a = bn_area(b);
            if (b.left < 0) {
                for (int i = 0; i < 8; i++) { c[i] = a * (float)b.count * kTri; ch[i] = 0; }
                continue;
            }
            const float* cl = &cost[(size_t)b.left * 8];
            const float* cr = &cost[(size_t)b.right * 8];
            // dist[j]: j slots split between the children, k of them to the left
            float dist[9];
            uint8_t dk[9];
            for (int j = 2; j <= 8; j++) {
                dist[j] = FLT_MAX; dk[j] = 1;
                for (int k = 1; k < j; k++) {
                    const float v = cl[k - 1] + cr[j - k - 1];
                    if (v < dist[j]) { dist[j] = v; dk[j] = (uint8_t)k; }
                }
            }
            const float one = a * kNode + dist[8];
            split8[n] = dk[8];
            c[0] = one; ch[0] = 0;
            for (int i = 2; i <= 8; i++) {
                if (dist[i] < one) { c[i - 1] = dist[i]; ch[i - 1] = dk[i]; }
                else { c[i - 1] = one; ch[i - 1] = 0; }
            }
        }
    }

    // the trees of the best forest of at most i trees for the subtree at n
    void forest(int32_t n, int i, int32_t* out, int& nk) const
    {
        const uint8_t k = bn[n].left < 0 ? 0 : choice[(size_t)n * 8 + (i - 1)];
        if (k == 0) { out[nk++] = n; return; }
        forest(bn[n].left, k, out, nk);
        forest(bn[n].right, i - k, out, nk);
    }

    // the slots of the wide node made from binary node n (at most 8)
    int children(int32_t n, int32_t* out) const
    {
        int nk = 0;
        if (bn[n].left < 0) { out[nk++] = n; return nk; }
        forest(bn[n].left, split8[n], out, nk);
        forest(bn[n].right, 8 - split8[n], out, nk);
        return nk;
    }
};

}  // namespace

const char* wbvh_build(const float* verts, uint32_t nvert, const uint32_t* tris, uint32_t ntri,
                       int isect, int nthreads, WbvhHost& out)
{
    out = WbvhHost();
    out.isect = isect;
    if (isect != WBVH_ISECT_TRIACCEL && isect != WBVH_ISECT_WATERTIGHT) return "unknown triangle test";
    if (ntri == 0) return "empty mesh";
    if (ntri >= (1u << 30)) return "too many triangles";
    for (uint64_t i = 0; i < 3ull * ntri; i++)
        if (tris[i] >= nvert) return "triangle index out of range";

    // Scene extent sets the world-space padding of leaf boxes.
    float slo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, shi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = 0; i < nvert; i++)
        for (int a = 0; a < 3; a++) {
            const float v = verts[3 * i + a];
            if (!std::isfinite(v)) return "non-finite vertex";
            slo[a] = std::min(slo[a], v); shi[a] = std::max(shi[a], v);
        }
    float mag = 0.0f;
    for (int a = 0; a < 3; a++) mag = std::max(mag, std::max(shi[a] - slo[a], std::max(std::fabs(slo[a]), std::fabs(shi[a]))));
    const float pad = std::max(mag * 3.8e-6f, 1e-30f);   // ~2^-18 of the scene magnitude

    std::vector<PrimRef> prims(ntri);
    for (uint32_t t = 0; t < ntri; t++) {
        PrimRef& p = prims[t];
        for (int a = 0; a < 3; a++) {
            const float v0 = verts[3 * tris[3 * t] + a], v1 = verts[3 * tris[3 * t + 1] + a], v2 = verts[3 * tris[3 * t + 2] + a];
            p.lo[a] = std::min(v0, std::min(v1, v2)) - pad;
            p.hi[a] = std::max(v0, std::max(v1, v2)) + pad;
            p.c[a] = 0.5f * (p.lo[a] + p.hi[a]);
        }
    }

    std::vector<uint32_t> idx(ntri);
    for (uint32_t i = 0; i < ntri; i++) idx[i] = i;
    std::vector<BNode> bn((size_t)2 * ntri - 1);
    if (nthreads < 1) nthreads = 1;
    Builder bld(prims, idx, bn, nthreads);
    const int32_t root = bld.build(0, ntri);

    // Collapse into wide nodes, breadth first, the slots of each chosen by
    // the cost-optimal dynamic program.
    Collapse col(bn);
    col.run(root);
    struct Pending { int32_t bnode; uint32_t depth; };
    std::vector<Pending> queue;
    queue.push_back({root, 1});
    out.nodes.clear();
    out.nodes.emplace_back();
    out.tris.clear();
    out.tris.reserve(ntri);
    size_t head = 0;
    uint32_t max_depth = 1;
    while (head < queue.size()) {
        const Pending pd = queue[head];
        const uint32_t wi = (uint32_t)head;
        head++;
        max_depth = std::max(max_depth, pd.depth);

        // children of this wide node (binary node ids)
        int32_t kids[WBVH_WIDTH];
        const int nk = col.children(pd.bnode, kids);   // a leaf at the root stays one slot

        // octant slot assignment: slot s holds the child that comes first
        // along direction octant s (bit a of s set: negative direction on axis a)
        float pc[3] = {0, 0, 0};
        {
            float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            for (int i = 0; i < nk; i++)
                for (int a = 0; a < 3; a++) {
                    lo[a] = std::min(lo[a], bn[kids[i]].lo[a]);
                    hi[a] = std::max(hi[a], bn[kids[i]].hi[a]);
                }
            for (int a = 0; a < 3; a++) pc[a] = 0.5f * (lo[a] + hi[a]);
        }
        float cost[WBVH_WIDTH][8];
        for (int i = 0; i < nk; i++) {
            const BNode& c = bn[kids[i]];
            for (int s = 0; s < 8; s++) {
                float dsum = 0.0f;
                for (int a = 0; a < 3; a++) {
                    const float dir = (s >> a) & 1 ? -1.0f : 1.0f;
                    dsum += (0.5f * (c.lo[a] + c.hi[a]) - pc[a]) * dir;
                }
                cost[i][s] = dsum;
            }
        }
        int slot_of[WBVH_WIDTH];
        int kid_in_slot[8];
        for (int s = 0; s < 8; s++) kid_in_slot[s] = -1;
        bool kid_done[WBVH_WIDTH] = {false};
        for (int round = 0; round < nk; round++) {
            int bi = -1, bs = -1;
            float bc = FLT_MAX;
            for (int i = 0; i < nk; i++) {
                if (kid_done[i]) continue;
                for (int s = 0; s < 8; s++) {
                    if (kid_in_slot[s] >= 0) continue;
                    if (cost[i][s] < bc) { bc = cost[i][s]; bi = i; bs = s; }
                }
            }
            kid_done[bi] = true;
            kid_in_slot[bs] = bi;
            slot_of[bi] = bs;
        }
        (void)slot_of;

        // quantisation frame
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int i = 0; i < nk; i++)
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], bn[kids[i]].lo[a]);
                hi[a] = std::max(hi[a], bn[kids[i]].hi[a]);
            }
        WNode wn;
        std::memset(&wn, 0, sizeof wn);
        wn.ox = lo[0]; wn.oy = lo[1]; wn.oz = lo[2];
        int ex[3];
        for (int a = 0; a < 3; a++) {
            const double ext = (double)hi[a] - (double)lo[a];
            int e = -100;
            if (ext > 0.0) {
                e = (int)std::ceil(std::log2(ext / (double)WBVH_QMAX));
                while (std::ldexp((double)WBVH_QMAX, e) < ext) e++;
                if (e < -100) e = -100;
            }
            if (e + WBVH_QBIAS > 254) return "scene extent too large to quantise";
            ex[a] = e;
        }
        wn.ex = (uint8_t)(ex[0] + WBVH_QBIAS);
        wn.ey = (uint8_t)(ex[1] + WBVH_QBIAS);
        wn.ez = (uint8_t)(ex[2] + WBVH_QBIAS);
        const float org[3] = {lo[0], lo[1], lo[2]};

        // children in slot order: internal ones get consecutive wide slots,
        // leaf triangles are appended contiguously
        wn.child_base = (uint32_t)queue.size();
        wn.tri_base = (uint32_t)out.tris.size();
        for (int s = 0; s < 8; s++) {
            const int i = kid_in_slot[s];
            if (i < 0) continue;
            const BNode& c = bn[kids[i]];
            for (int a = 0; a < 3; a++) {
                const double sc = std::ldexp(1.0, ex[a]);
                double ql = std::floor(((double)c.lo[a] - (double)org[a]) / sc);
                double qh = std::ceil(((double)c.hi[a] - (double)org[a]) / sc);
                ql = std::max(0.0, std::min((double)WBVH_QMAX, ql));
                qh = std::max(0.0, std::min((double)WBVH_QMAX, qh));
                wn.q[a][s][0] = (uint8_t)ql;
                wn.q[a][s][1] = (uint8_t)qh;
            }
            if (c.left >= 0) {
                wn.imask |= (uint8_t)(1u << s);
                queue.push_back({kids[i], pd.depth + 1});
                out.nodes.emplace_back();
            } else {
                wn.tvalid |= ((1u << c.count) - 1u) << (3 * s);
                for (uint32_t k = 0; k < c.count; k++) {
                    const uint32_t t = idx[c.first + k];
                    const float* p0 = &verts[3 * tris[3 * t]];
                    const float* p1 = &verts[3 * tris[3 * t + 1]];
                    const float* p2 = &verts[3 * tris[3 * t + 2]];
                    out.tris.push_back(make_tri(p0, p1, p2, t, isect));
                }
            }
        }
        out.nodes[wi] = wn;
    }
    out.max_depth = max_depth;
    out.ntri = ntri;
    if (max_depth > WBVH_STACK) return "BVH deeper than the traversal stack";
    // the traversal kernel reads nodes through 32-bit byte offsets
    if ((uint64_t)out.nodes.size() * sizeof(WNode) > 0xffffffffull) return "BVH nodes exceed 4 GB";
    return nullptr;
}

// TriAccel precompute with the reference's operation order
// (include/lightmetrica/triaccel.h:50-91): b = C - A, c = B - A, N = c x b,
// k = largest |N| axis, values divided by N[k] and by the projected
// determinant.  A zero determinant marks the triangle unhittable (k = 3).
WTri make_tri(const float* A, const float* B, const float* C, uint32_t id, int isect)
{
    WTri t;
    std::memset(&t, 0, sizeof t);
    auto bits = [](uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; };
    if (isect == WBVH_ISECT_WATERTIGHT) {
        t.f[0] = A[0]; t.f[1] = A[1]; t.f[2] = A[2]; t.f[3] = bits(id);
        t.f[4] = B[0]; t.f[5] = B[1]; t.f[6] = B[2];
        t.f[8] = C[0]; t.f[9] = C[1]; t.f[10] = C[2];
        return t;
    }
    static const int wald_modulo[4] = {1, 2, 0, 1};
    const float b[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    const float c[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    const float N[3] = {c[1] * b[2] - b[1] * c[2], c[2] * b[0] - b[2] * c[0], c[0] * b[1] - b[0] * c[1]};
    uint32_t k = 0;
    for (int j = 0; j < 3; j++)
        if (std::fabs(N[j]) > std::fabs(N[k])) k = (uint32_t)j;
    const int u = wald_modulo[k], v = wald_modulo[k + 1];
    const float n_k = N[k];
    const float denom = b[u] * c[v] - b[v] * c[u];
    t.f[10] = bits(id);
    if (denom == 0.0f) {
        t.f[0] = bits(3u);
        return t;
    }
    t.f[0] = bits(k);
    t.f[1] = N[u] / n_k;
    t.f[2] = N[v] / n_k;
    const float dotAN = (A[0] * N[0] + A[1] * N[1]) + A[2] * N[2];
    t.f[3] = dotAN / n_k;
    t.f[4] = A[u];
    t.f[5] = A[v];
    t.f[6] = b[u] / denom;
    t.f[7] = -b[v] / denom;
    t.f[8] = c[v] / denom;
    t.f[9] = -c[u] / denom;
    return t;
}

}  // namespace lmx
