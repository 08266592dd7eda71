// wbvh_math.h -- ray setup, child-box test and triangle tests, shared by the
// host walker and the gfx950 traversal kernel so both take exactly the same
// decisions (compiled with -ffp-contract=off; every FMA below is explicit).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include "wbvh.h"

#define WB_HD __host__ __device__ __forceinline__

// Relative widening of the far slab distance; with the world-space padding the
// builder applies to child boxes it keeps the quantised box test conservative
// under FP32 rounding.
#define WBVH_FAR_SCALE 1.00000048f   // 1 + 2^-21

WB_HD float wb_as_float(uint32_t u)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(u);
#else
    union { uint32_t u; float f; } c; c.u = u; return c.f;
#endif
}
WB_HD uint32_t wb_as_uint(float f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(f);
#else
    union { uint32_t u; float f; } c; c.f = f; return c.u;
#endif
}
WB_HD int wb_popc(uint32_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __popc(x);
#else
    return __builtin_popcount(x);
#endif
}
WB_HD int wb_ctz(uint32_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __ffs(x) - 1;
#else
    return __builtin_ctz(x);
#endif
}

// two binary16 grid coordinates (children 2w and 2w + 1 of one axis)
typedef _Float16 wb_h2 __attribute__((ext_vector_type(2)));

struct WRayPre {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
    float ix, iy, iz;          // clamped reciprocal direction (box test)
    uint32_t oct;              // direction octant: bit a set when d[a] < 0
    // watertight shear constants (Woop, Benthin, Wald, JCGT 2013)
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

WB_HD float wb_sel3(int k, float x, float y, float z)
{
    return k == 0 ? x : (k == 1 ? y : z);
}

WB_HD float wb_safe_inv(float d)
{
    const float tiny = 1e-20f;
    const float dd = fabsf(d) < tiny ? copysignf(tiny, d) : d;
    return 1.0f / dd;
}

WB_HD void wb_ray_setup(WRayPre& r, float ox, float oy, float oz, float tmin,
                        float dx, float dy, float dz, float tmax)
{
    r.ox = ox; r.oy = oy; r.oz = oz; r.tmin = tmin;
    r.dx = dx; r.dy = dy; r.dz = dz; r.tmax = tmax;
    r.ix = wb_safe_inv(dx); r.iy = wb_safe_inv(dy); r.iz = wb_safe_inv(dz);
    r.oct = (dx < 0.0f ? 1u : 0u) | (dy < 0.0f ? 2u : 0u) | (dz < 0.0f ? 4u : 0u);
    int kz = 0;
    if (fabsf(dy) > fabsf(dx)) kz = 1;
    if (fabsf(dz) > fabsf(wb_sel3(kz, dx, dy, dz))) kz = 2;
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    const float dkz = wb_sel3(kz, dx, dy, dz);
    if (dkz < 0.0f) { const int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = wb_sel3(kx, dx, dy, dz) / dkz;
    r.Sy = wb_sel3(ky, dx, dy, dz) / dkz;
    r.Sz = 1.0f / dkz;
}

// Two grid coordinates of a node word (bytes: lower and upper bound of
// slot 2w, then of slot 2w + 1, on one axis) as a pair of binary16
// subnormals q * 2^-24: the selector takes byte 0 or 1 and byte 2 or 3,
// each below a zero byte (selector byte 0x0c), one v_perm_b32 on the
// device.  The conversion to FP32 is exact.
WB_HD wb_h2 wb_q_pair(uint32_t word, uint32_t sel)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t v = __builtin_amdgcn_perm(0u, word, sel);
#else
    uint32_t v = 0;
    for (int i = 0; i < 4; i++) {
        const uint32_t s = (sel >> (8 * i)) & 0xffu;
        if (s < 4u) v |= ((word >> (8 * s)) & 0xffu) << (8 * i);
    }
#endif
    return __builtin_bit_cast(wb_h2, v);
}

// Box test of all eight child slots of a node against [tmin, tmax]: bit i
// of the result set when slot i's box is hit (empty slots included; they
// are neither internal nor hold triangles).
WB_HD uint32_t wb_slot_hits(const WNode& n, const WRayPre& r, float tmax)
{
    const float sx = wb_as_float((uint32_t)n.ex << 23) * r.ix;
    const float sy = wb_as_float((uint32_t)n.ey << 23) * r.iy;
    const float sz = wb_as_float((uint32_t)n.ez << 23) * r.iz;
    const float bx = (n.ox - r.ox) * r.ix;
    const float by = (n.oy - r.oy) * r.iy;
    const float bz = (n.oz - r.oz) * r.iz;
    // Near and far planes per axis chosen once per node by the sign of the
    // slab scale (for lo <= hi, fma(lo, s, b) <= fma(hi, s, b) exactly when
    // s >= 0), instead of a min and a max per plane pair and child: the
    // choice is the byte permute's selector.
    const uint32_t* qw = reinterpret_cast<const uint32_t*>(&n.q[0][0][0]);
    const bool neg[3] = {sx < 0.0f, sy < 0.0f, sz < 0.0f};
    wb_h2 nq[3][4], fq[3][4];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int a = 0; a < 3; a++) {
        const uint32_t nsel = neg[a] ? 0x0c030c01u : 0x0c020c00u;
        const uint32_t fsel = nsel ^ 0x00010001u;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
        for (int w = 0; w < 4; w++) {
            nq[a][w] = wb_q_pair(qw[4 * a + w], nsel);
            fq[a][w] = wb_q_pair(qw[4 * a + w], fsel);
        }
    }
    uint32_t hb = 0;                   // hit slots, bit i <-> slot i
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int i = 0; i < 8; i++) {
        // every slot is tested (no branch per slot); empty slots are masked
        // out below (neither internal nor holding triangles)
        const int w = i >> 1;
        const bool odd = (i & 1) != 0;
        const float xn = fmaf((float)(odd ? nq[0][w].y : nq[0][w].x), sx, bx);
        const float xf = fmaf((float)(odd ? fq[0][w].y : fq[0][w].x), sx, bx);
        const float yn = fmaf((float)(odd ? nq[1][w].y : nq[1][w].x), sy, by);
        const float yf = fmaf((float)(odd ? fq[1][w].y : fq[1][w].x), sy, by);
        const float zn = fmaf((float)(odd ? nq[2][w].y : nq[2][w].x), sz, bz);
        const float zf = fmaf((float)(odd ? fq[2][w].y : fq[2][w].x), sz, bz);
        const float tn = fmaxf(fmaxf(xn, yn), fmaxf(zn, r.tmin));
        const float tf = fminf(fminf(xf, yf), fminf(zf, tmax));
        hb |= (tn <= tf * WBVH_FAR_SCALE) ? 1u << i : 0u;
    }
    return hb;
}

// Internal children of a slot hit mask in traversal order: bit p <-> slot
// p ^ oct, by swapping bits, bit pairs and nibbles as the octant's bits say.
WB_HD constexpr uint32_t wb_octant_order(uint32_t inner, uint32_t oct)
{
    if (oct & 1u) inner = ((inner & 0x55u) << 1) | ((inner >> 1) & 0x55u);
    if (oct & 2u) inner = ((inner & 0x33u) << 2) | ((inner >> 2) & 0x33u);
    if (oct & 4u) inner = ((inner & 0x0fu) << 4) | ((inner >> 4) & 0x0fu);
    return inner;
}

// Leaf slots hit -> their triangle bits: slot bit i spread to bits 3 i ..
// 3 i + 2 (to be restricted to the node's valid triangles).
WB_HD constexpr uint32_t wb_leaf_spread(uint32_t leaf)
{
    leaf = (leaf | (leaf << 8)) & 0x00f00fu;
    leaf = (leaf | (leaf << 4)) & 0x0c30c3u;
    leaf = (leaf | (leaf << 2)) & 0x249249u;
    return leaf * 7u;
}

// Returns the hit mask of internal children permuted by the ray octant
// (bit p <-> slot p ^ oct) and, through tri_mask, the triangles of hit leaf
// children (bit b <-> triangle wb_tri_index(node, b)).
WB_HD uint32_t wb_node_hits(const WNode& n, const WRayPre& r, float tmax, uint32_t& tri_mask)
{
    const uint32_t hb = wb_slot_hits(n, r, tmax);
    tri_mask = wb_leaf_spread(hb & ~(uint32_t)n.imask) & n.tvalid;
    return wb_octant_order(hb & n.imask, r.oct);
}
// Triangle of bit b of a node's triangle mask: the node's leaf triangles are
// stored contiguously in slot order, so it is preceded by the valid bits
// below b.
WB_HD uint32_t wb_tri_index(uint32_t tri_base, uint32_t tvalid, uint32_t b)
{
    return tri_base + (uint32_t)wb_popc(tvalid & ((1u << b) - 1u));
}

// Kahan's difference of products a*b - c*d with two FMAs: exactly zero iff
// the exact value is zero, otherwise of the exact value's sign.
WB_HD float wb_dop(float a, float b, float c, float d)
{
    const float w = c * d;
    const float e = fmaf(-c, d, w);
    const float f = fmaf(a, b, -w);
    return f + e;
}

// Watertight test.  On a hit with t in [tmin, tmax] writes t and the
// unnormalised (V, W, det); b0 = V/det (2nd vertex), b1 = W/det (3rd vertex).
WB_HD bool wb_tri_watertight(const WTri& tri, const WRayPre& r, float tmax,
                             float& t, float& V, float& W, float& det)
{
    const float Ax0 = tri.f[0] - r.ox, Ay0 = tri.f[1] - r.oy, Az0 = tri.f[2] - r.oz;
    const float Bx0 = tri.f[4] - r.ox, By0 = tri.f[5] - r.oy, Bz0 = tri.f[6] - r.oz;
    const float Cx0 = tri.f[8] - r.ox, Cy0 = tri.f[9] - r.oy, Cz0 = tri.f[10] - r.oz;
    const float Akz = wb_sel3(r.kz, Ax0, Ay0, Az0), Bkz = wb_sel3(r.kz, Bx0, By0, Bz0), Ckz = wb_sel3(r.kz, Cx0, Cy0, Cz0);
    const float Ax = fmaf(-r.Sx, Akz, wb_sel3(r.kx, Ax0, Ay0, Az0));
    const float Ay = fmaf(-r.Sy, Akz, wb_sel3(r.ky, Ax0, Ay0, Az0));
    const float Bx = fmaf(-r.Sx, Bkz, wb_sel3(r.kx, Bx0, By0, Bz0));
    const float By = fmaf(-r.Sy, Bkz, wb_sel3(r.ky, Bx0, By0, Bz0));
    const float Cx = fmaf(-r.Sx, Ckz, wb_sel3(r.kx, Cx0, Cy0, Cz0));
    const float Cy = fmaf(-r.Sy, Ckz, wb_sel3(r.ky, Cx0, Cy0, Cz0));
    const float U = wb_dop(Cx, By, Cy, Bx);
    const float Vv = wb_dop(Ax, Cy, Ay, Cx);
    const float Ww = wb_dop(Bx, Ay, By, Ax);
    if ((U < 0.0f || Vv < 0.0f || Ww < 0.0f) && (U > 0.0f || Vv > 0.0f || Ww > 0.0f)) return false;
    const float dt = (U + Vv) + Ww;
    if (dt == 0.0f) return false;
    const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    const float T = fmaf(U, Az, fmaf(Vv, Bz, Ww * Cz));
    const float tt = T / dt;
    if (!(tt >= r.tmin && tt <= tmax)) return false;
    t = tt; V = Vv; W = Ww; det = dt;
    return true;
}

// TriAccel test with the reference's operation order
// (include/lightmetrica/triaccel.h:93-151); u = weight of the 2nd vertex,
// v = weight of the 3rd vertex.
WB_HD bool wb_tri_triaccel(const WTri& tri, const WRayPre& r, float tmax,
                           float& t, float& u, float& v)
{
    const uint32_t k = wb_as_uint(tri.f[0]);
    float o_u, o_v, o_k, d_u, d_v, d_k;
    if (k == 0) { o_u = r.oy; o_v = r.oz; o_k = r.ox; d_u = r.dy; d_v = r.dz; d_k = r.dx; }
    else if (k == 1) { o_u = r.oz; o_v = r.ox; o_k = r.oy; d_u = r.dz; d_v = r.dx; d_k = r.dy; }
    else if (k == 2) { o_u = r.ox; o_v = r.oy; o_k = r.oz; d_u = r.dx; d_v = r.dy; d_k = r.dz; }
    else return false;
    const float n_u = tri.f[1], n_v = tri.f[2], n_d = tri.f[3];
    const float a_u = tri.f[4], a_v = tri.f[5], b_nu = tri.f[6], b_nv = tri.f[7];
    const float c_nu = tri.f[8], c_nv = tri.f[9];
    const float demon = d_u * n_u + d_v * n_v + d_k;
    if (demon == 0.0f) return false;
    const float tt = (n_d - o_u * n_u - o_v * n_v - o_k) / demon;
    if (tt < r.tmin || tt > tmax) return false;
    const float hu = o_u + tt * d_u - a_u;
    const float hv = o_v + tt * d_v - a_v;
    const float uu = hv * b_nu + hu * b_nv;
    const float vv = hu * c_nu + hv * c_nv;
    if (!(uu >= 0.0f && vv >= 0.0f && uu + vv <= 1.0f)) return false;
    t = tt; u = uu; v = vv;
    return true;
}

WB_HD uint32_t wb_tri_id(const WTri& tri, int isect)
{
    return isect == WBVH_ISECT_TRIACCEL ? wb_as_uint(tri.f[10]) : wb_as_uint(tri.f[3]);
}
