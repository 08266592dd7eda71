// wbvh_build.h -- host side of the compressed wide BVH: builder output and
// the host walker (used for the scalar Accel3::Intersect path and as the
// counting walker that supplies the traversal kernel's algorithmic bytes).
#pragma once
#include <stdint.h>
#include <vector>
#include "wbvh.h"

namespace lmx {

struct WbvhHost {
    std::vector<WNode> nodes;
    std::vector<WTri> tris;
    uint32_t ntri = 0;
    uint32_t max_depth = 0;
    int isect = WBVH_ISECT_WATERTIGHT;
};

// Returns nullptr on success, otherwise a static error string.
const char* wbvh_build(const float* verts, uint32_t nvert, const uint32_t* tris, uint32_t ntri,
                       int isect, int nthreads, WbvhHost& out);

WTri make_tri(const float* A, const float* B, const float* C, uint32_t id, int isect);

// Closest hit of one ray by walking the wide BVH on the host in exactly the
// order the traversal kernel uses.  Optionally counts node visits and
// triangle tests.  Returns false (and a miss) if a node group did not fit
// the WBVH_STACK-entry stack, which a tree the builder accepts never needs.
bool wbvh_walk(const WbvhHost& b, const WRay& ray, WHit& hit, uint64_t* nodes_visited,
               uint64_t* tris_tested);
// The any-hit walk of the occlusion kernel: stops after the node in which it
// found a triangle in range (whose other triangles it still tests).
bool wbvh_walk_anyhit(const WbvhHost& b, const WRay& ray, WHit& hit, uint64_t* nodes_visited,
                      uint64_t* tris_tested);

}  // namespace lmx
