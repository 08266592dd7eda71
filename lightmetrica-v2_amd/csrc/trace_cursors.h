// trace_cursors.h -- the work split of the traversal kernel's launches:
// the rays are cut into 8 regions (one per XCD), every wave takes a fixed
// first chunk of its region, and the rest of each region is claimed through
// an atomic cursor.  Shared by the kernel, the kernel that sets the cursors
// before a launch, and wavefront loops whose own commit kernel sets them
// for the next launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef LMX_CLAIM_CHUNK
#define LMX_CLAIM_CHUNK 64     // rays claimed from a region cursor per atomic
#endif
// words between two region cursors: each cursor sits in a 256-byte line of
// its own, so the atomics of different regions do not serialise on one line
#ifndef LMX_CURSOR_STRIDE
#define LMX_CURSOR_STRIDE 64
#endif
#define CURSOR_STRIDE LMX_CURSOR_STRIDE
// counter block of one trace launch: 8 region cursors, the ray count, and a
// sticky stack-overflow flag, each on a line of its own
#define LMX_TRACE_COUNTER_WORDS (10 * CURSOR_STRIDE)

__device__ __forceinline__ uint32_t region_start(uint32_t j, uint32_t n)
{
    return (uint32_t)(((uint64_t)n * j) / 8u);
}

__device__ __forceinline__ uint32_t region_end(uint32_t j, uint32_t n)
{
    return (uint32_t)(((uint64_t)n * (j + 1u)) / 8u);
}

// Rays per claim: LMX_CLAIM_CHUNK for full queues; short queues (the tail of
// a wavefront render) are spread over every wave of the grid, so that their
// time is that of a few rays per wave, not of whole chunks in turn.
__device__ __forceinline__ uint32_t trace_claim(uint32_t n, uint32_t grid)
{
    return min((uint32_t)LMX_CLAIM_CHUNK, max(1u, (n + grid - 1u) / grid));
}

// Every wave's first chunk is fixed by its index (wave b, of XCD region
// r = b mod 8, takes chunk b / 8 of its region); the region's cursor starts
// after those chunks, so a launch claims with atomics only what is left,
// and one with few rays (the tail of a wavefront render) with none.
__device__ __forceinline__ uint32_t region_dyn_start(uint32_t r, uint32_t n, uint32_t grid, uint32_t claim)
{
    const uint64_t waves = (grid + 7u - r) / 8u;          // waves b < grid with b mod 8 = r
    const uint64_t s = (uint64_t)region_start(r, n) + waves * claim;
    return (uint32_t)min(s, (uint64_t)region_end(r, n));
}

// Thread j of one wave: j < 8 sets region j's cursor, j == 8 the ray count,
// for a launch over n rays with `grid` workgroups.
__device__ __forceinline__ void trace_cursors_set(uint32_t* cursors, uint32_t n, uint32_t grid, uint32_t j)
{
    if (j < 8) cursors[j * CURSOR_STRIDE] = region_dyn_start(j, n, grid, trace_claim(n, grid));
    if (j == 8) cursors[8 * CURSOR_STRIDE] = n;
}
