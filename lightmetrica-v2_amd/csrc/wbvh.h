// wbvh.h -- compressed 8-wide BVH layout shared by the host builder, the
// host walker and the gfx950 traversal kernel.
//
// Replaces the reference's binary/4-wide BVH node types
// (src/liblightmetrica/accel/accel_bvh_sahbin.cpp:36-55,
//  src/liblightmetrica/accel/accel_qbvh.cpp:54-137).
//
// Node: 80 bytes = five 16-byte loads.  Child boxes are quantised per axis
// to a grid of 255 steps from the node origin, the step a power of two,
// rounded outward, and stored as bytes, a slot's lower and upper bound
// side by side, so one 4-byte word holds both bounds of two slots on one
// axis.  The box test turns the near (or far) bounds of such a word into a
// pair of binary16 subnormals q * 2^-24 with one byte permute (v_perm_b32:
// each chosen byte below a zero byte; the choice of lower or upper bound by
// the ray's direction sign is in the permute's selector), which its FMAs
// take as operands directly (v_fma_mix_f32, exact conversion); the node's
// scale exponent carries the 2^24.  Up to eight children per node; internal
// children of a node are stored contiguously (child_base + rank), and the
// triangles of all leaf children of a node are stored contiguously in slot
// order.  Each leaf holds 1..3
// triangles; leaf slot s owns bits 3 s .. 3 s + 2 of a 24-bit triangle mask,
// tvalid marks the bits in use, and the triangle of bit b is stored at
// tri_base + popcount(tvalid below b), so a node's hit leaves become a
// triangle mask by spreading the 8 hit bits to 3-bit groups.
//
// Child slots are assigned per node by octant (after Ylitie, Karras and
// Laine, "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide
// BVHs", HPG 2017): slot s holds the child that comes first along the
// direction octant s, so visiting hit slots in the order (p ^ octant) for
// p = 0..7 is approximately front to back for every ray.
#pragma once
#include <stdint.h>

#define WBVH_WIDTH 8
#define WBVH_MAX_LEAF_TRIS 3
#define WBVH_STACK 32         // traversal stack entries per ray (node groups)
#define WBVH_QMAX 255         // quantisation grid steps per axis (one byte)
#define WBVH_QBIAS 151        // scale exponent bias: grid step = 2^(e - 151), q * 2^-24 * 2^(e - 127)

// Triangle test selector (stored in the accel; picks the kernel variant).
#define WBVH_ISECT_TRIACCEL 0     // Wald projection test, reference semantics
#define WBVH_ISECT_WATERTIGHT 1   // Woop/Benthin/Wald watertight test

struct alignas(16) WNode {
    float ox, oy, oz;          // quantisation origin (lower corner)
    uint8_t ex, ey, ez;        // per-axis grid step = 2^(e - WBVH_QBIAS)
    uint8_t imask;             // bit s set: slot s is an internal child
    uint32_t child_base;       // index of the first internal child
    uint32_t tri_base;         // index of the first triangle of leaf children
    uint32_t tvalid;           // leaf triangles: bit 3 s + k set for the k-th triangle of leaf slot s
    uint32_t reserved;
    uint8_t q[3][8][2];        // quantised child bounds per axis and slot: lower, upper (grid units)
};
static_assert(sizeof(WNode) == 80, "WNode must be 80 bytes");

// Triangle record, 48 bytes = 3 x float4.
//   watertight: {v0.xyz, id} {v1.xyz, 0} {v2.xyz, 0}
//   triaccel:   {k, n_u, n_v, n_d} {a_u, a_v, b_nu, b_nv} {c_nu, c_nv, id, 0}
//   (the TriAccel fields follow include/lightmetrica/triaccel.h:35-48)
struct alignas(16) WTri {
    float f[12];
};
static_assert(sizeof(WTri) == 48, "WTri must be 48 bytes");

// Per-ray record at the C boundary (32 B) and hit record (16 B).
struct WRay { float o[3]; float tmin; float d[3]; float tmax; };
struct WHit { float t, b0, b1; uint32_t id; };

#define WBVH_MISS 0xffffffffu
