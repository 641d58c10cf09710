// rowmask.hpp -- the tile hysteresis building blocks shared by the fused vision graph and the Canny
// edge trace: 64-bit row masks (lane = row), their 8-connected growth, and a lock-free union-find over
// global memory for the components that cross tile borders.
#pragma once
#include "mvx_common.hpp"

namespace mvx {

// union-find with atomicMin linking (the larger root hangs under the smaller): the final root of a
// component is its smallest index, whatever the order of the unions
MVX_DEV uint32_t gfind(gptr_t<const uint32_t> par, uint32_t p)
{
    uint32_t l = par[p];
    while (l != p) { p = l; l = par[p]; }
    return p;
}

MVX_DEV void gunion(gptr_t<uint32_t> par, uint32_t a, uint32_t b)
{
    while (true) {
        a = gfind(par, a);
        b = gfind(par, b);
        if (a == b) return;
        if (a < b) { uint32_t t = a; a = b; b = t; }
        const uint32_t old = __hip_atomic_fetch_min(par + a, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == a) return;
        a = old;
    }
}

// the 64-bit value of lane l (l wave-uniform)
MVX_DEV uint64_t readlane64(uint64_t v, int l)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// the 64-bit value of lane - 1 (DPP wave_shr:1) and of lane + 1 (wave_shl:1), 0 past the wave's ends;
// the whole wave must be active
MVX_DEV uint64_t lane_prev64(uint64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, 0x138, 0xf, 0xf, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), 0x138, 0xf, 0xf, true);
    return ((uint64_t)hi << 32) | lo;
}

MVX_DEV uint64_t lane_next64(uint64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, 0x130, 0xf, 0xf, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), 0x130, 0xf, 0xf, true);
    return ((uint64_t)hi << 32) | lo;
}

// Row masks, lane = row (rows 0..31 of one problem in lanes 0..31, of a second in lanes 32..63):
// grow m inside e by 8-connectivity until nothing changes.  Called by the whole wave.
MVX_DEV uint64_t grow_rows(uint64_t m, uint64_t e, int lane)
{
    const int r = lane & 31;
    while (true) {
        uint64_t up = lane_prev64(m), dn = lane_next64(m);
        if (r == 0) up = 0ull;
        if (r == 31) dn = 0ull;
        uint64_t v = m | up | dn;
        v |= (v << 1) | (v >> 1);
        const uint64_t n = (v & e) | m;
        const bool changed = n != m;
        m = n;
        if (!__any(changed)) break;
    }
    return m;
}

// The unions across one tile edge: lane i holds a = my id at border position i and b = the neighbour's
// id at position i (-1: no edge pixel, or past the edge's length); a joins b(i-1), b(i), b(i+1).  A pair
// that a lower lane also forms is left to it (same a in lane i-1 or i+1 forming the pair with its own b,
// or b(i +- 1) = b(i)), so a run of equal ids across the edge costs one union instead of three per pixel;
// every distinct pair is still joined, so the components -- and their minimum ids, the roots -- are
// the same.  Called by the whole wave.
MVX_DEV void border_unions(gptr_t<uint32_t> par, int32_t a, int32_t b, int lane)
{
    const int32_t aU = __shfl_up(a, 1), bU = __shfl_up(b, 1);
    const int32_t aD = __shfl_down(a, 1), bD = __shfl_down(b, 1);
    const bool hasU = lane >= 1, hasD = lane + 1 < 64;
    if (a < 0) return;
    if (b >= 0 && b != a && !(hasU && aU == a && bU == b)) gunion(par, (uint32_t)a, (uint32_t)b);
    if (hasU && bU >= 0 && bU != a && bU != b && aU != a) gunion(par, (uint32_t)a, (uint32_t)bU);
    if (hasD && bD >= 0 && bD != a && bD != b && aD != a) gunion(par, (uint32_t)a, (uint32_t)bD);
}

}  // namespace mvx
