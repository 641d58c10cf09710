// pick.hpp -- Harris sort-and-pick as a one-workgroup device routine (used by the per-node kernel and
// by the frame-batched graph).
//
// HafCpu_HarrisMergeSortAndPick_XY_XYS (ago_haf_cpu_harris.cpp:923-1005), grid set-up
// ago_kernel_api.cpp:20808-20822.  The reference sorts the XYS items as int64 keys (float bits, y, x)
// descending and then visits them in order, keeping a candidate unless a kept point lies within
// min_distance (dx^2 + dy^2 < ceil(d^2)); its grid of cellSize = floor(d / sqrt 2) is only an index
// for that test (same-cell points are always closer than d, and the +-2 cell window covers every
// point closer than d).  The result is the lexicographically-first maximal independent set of the
// "closer than d" graph in key order; parallel rounds reproduce it exactly: a candidate is kept once
// every closer higher-priority candidate is rejected, rejected once one of them is kept.
// One grid quirk is reproduced: HafCpu_MemSet_U32 clears only cells & ~3 cells to -1, so the last
// cells % 4 cells keep (0,0) from the zeroed scratch and behave as occupied by a point at (0,0).
#pragma once
#include "mvx_common.hpp"

namespace mvx {

struct XYS { int16_t x, y; float s; };                                                   // ago_keypoint_xys_t
struct Keypoint { int32_t x, y; float strength, scale, orientation; int32_t tracking_status; float error; };  // vx_keypoint_t

constexpr int PICK_THREADS = 1024;
constexpr uint32_t PICK_LDS_KEYS = 8192;          // keys sorted in LDS up to this many candidates
constexpr uint32_t PICK_NB = 32;                  // closer higher-priority candidates listed per candidate in LDS
constexpr uint8_t ST_UNDECIDED = 0, ST_IN = 1, ST_OUT = 2;
constexpr uint32_t HASH_EMPTY = 0xffffffffu;

MVX_DEV int64_t xys_key(const XYS &it)
{
    return (int64_t)(((uint64_t)__float_as_uint(it.s) << 32) | ((uint64_t)(uint16_t)it.y << 16) | (uint64_t)(uint16_t)it.x);
}

// bitonic sort, descending, of n (power of two) keys held in `k` (LDS or global), whole workgroup
MVX_DEV void bitonic_desc(int64_t *k, uint32_t n)
{
    for (uint32_t size = 2; size <= n; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < n / 2; t += PICK_THREADS) {
                const uint32_t i = 2 * t - (t & (stride - 1));
                const uint32_t j = i + stride;
                const bool desc = ((i & size) == 0);
                const int64_t a = k[i], b = k[j];
                if (desc ? (a < b) : (a > b)) { k[i] = b; k[j] = a; }
            }
            __syncthreads();
        }
    }
}

MVX_DEV uint32_t hash_u32(uint32_t v) { v ^= v >> 16; v *= 0x7feb352du; v ^= v >> 15; v *= 0x846ca68bu; v ^= v >> 16; return v; }

struct PickWork {
    int64_t *keys;        // pow2 >= maxList
    uint32_t *hkey;       // hash slots: pixel index, HASH_EMPTY when free
    uint32_t *hval;       // hash slots: rank
    uint32_t *slotOf;     // rank -> slot (to clear the table afterwards)
    uint8_t *state;       // rank -> ST_*
    uint32_t hmask;       // table size - 1
};

MVX_DEV uint32_t hash_find(const PickWork &W, uint32_t key)
{
    uint32_t s = hash_u32(key) & W.hmask;
    while (true) {
        const uint32_t k = W.hkey[s];
        if (k == key) return W.hval[s];
        if (k == HASH_EMPTY) return HASH_EMPTY;
        s = (s + 1) & W.hmask;
    }
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct PickLayout {
    uint32_t np, hsize;
    size_t keys, hkey, hval, slot, state, total;
};

inline PickLayout pick_layout(uint32_t maxList)
{
    PickLayout L;
    L.np = 1;
    while (L.np < maxList) L.np <<= 1;
    L.hsize = 1024;
    while (L.hsize < 2u * maxList) L.hsize <<= 1;
    L.keys = 0;
    L.hkey = align256(L.keys + (size_t)L.np * 8);
    L.hval = align256(L.hkey + (size_t)L.hsize * 4);
    L.slot = align256(L.hval + (size_t)L.hsize * 4);
    L.state = align256(L.slot + (size_t)maxList * 4);
    L.total = align256(L.state + (size_t)maxList);
    return L;
}

inline PickWork pick_work(void *workspace, uint32_t maxList)
{
    const PickLayout L = pick_layout(maxList ? maxList : 1);
    uint8_t *base = (uint8_t *)workspace;
    PickWork W;
    W.keys = (int64_t *)(base + L.keys);
    W.hkey = (uint32_t *)(base + L.hkey);
    W.hval = (uint32_t *)(base + L.hval);
    W.slotOf = (uint32_t *)(base + L.slot);
    W.state = base + L.state;
    W.hmask = L.hsize - 1;
    return W;
}

// A candidate's state before the rounds: the last cells % 4 grid cells keep (0,0) from the zeroed
// scratch, so a candidate in one of them, or within min_distance of the origin with one of them in its
// +-2 cell window, is never kept.
MVX_DEV uint8_t pick_initial_state(uint32_t px, uint32_t py, uint32_t cell, int gw, long cells, long firstPhantom, int md2)
{
    const int cx = (int)(px / cell), cy = (int)(py / cell);
    const long own = (long)cy * gw + cx;
    if (own >= firstPhantom) return ST_OUT;
    if ((int)(px * px + py * py) < md2) {
        for (long c = firstPhantom; c < cells; c++) {
            const int tx = (int)(c % gw), ty = (int)(c / gw);
            if (abs(tx - cx) <= 2 && abs(ty - cy) <= 2) return ST_OUT;
        }
    }
    return ST_UNDECIDED;
}

// The whole pick for one list, run by one workgroup of PICK_THREADS threads.  When resetCount is set,
// *listCount is cleared at the end (the list producer appends with atomics).
// Up to PICK_THREADS candidates (one per thread), the rounds run in LDS: each candidate first lists the
// higher-priority candidates closer than min_distance by a scan of the sorted keys (every lane of a
// wave reads the same key: an LDS broadcast), and a round then reads only those lists.  Longer lists
// (or more candidates) take the hash-table path, which probes the pixels around each candidate.
MVX_DEV void harris_pick_block(const XYS *list, uint32_t *listCount, bool resetCount, uint32_t maxList, float minDistance,
                               uint32_t width, uint32_t height, Keypoint *out, uint32_t capacity,
                               uint32_t *outNumItems, uint64_t *outCount, const PickWork &Wfull)
{
    __shared__ int64_t lkeys[PICK_LDS_KEYS];
    __shared__ uint32_t sflag, soverflow, ssum[PICK_THREADS];
    __shared__ uint16_t lnb[PICK_NB * PICK_THREADS];     // [k][candidate]: a wave's appends hit distinct banks
    __shared__ uint8_t lstate[PICK_THREADS];
    if (threadIdx.x == 0) soverflow = 0;
    const uint32_t n = min(*listCount, maxList);
    // use only a table of about 4n slots (>= 1024) so that the probes stay cache resident; the slots
    // beyond it stay empty, and the ones used here are emptied again at the end
    PickWork W = Wfull;
    {
        uint32_t hs = 1024;
        while (hs < 4u * n && hs <= Wfull.hmask) hs <<= 1;
        W.hmask = min(hs, Wfull.hmask + 1) - 1;
    }
    uint32_t np = 1;
    while (np < n) np <<= 1;
    const bool inLds = np <= PICK_LDS_KEYS;
    int64_t *keys = inLds ? lkeys : W.keys;
    for (uint32_t i = threadIdx.x; i < np; i += PICK_THREADS) keys[i] = (i < n) ? xys_key(list[i]) : INT64_MIN;
    __syncthreads();
    if (np > 1) bitonic_desc(keys, np);

    const bool useGrid = minDistance > 2.0f;
    uint32_t cell = 1, md2 = 0;
    int gw = 1;
    long cells = 0, firstPhantom = 0;
    if (useGrid) {
        cell = (uint32_t)floor((double)minDistance / 1.41421356237309504880);
        gw = (int)((width + cell - 1) / cell);
        const int gh = (int)((height + cell - 1) / cell);
        cells = (long)gw * gh;
        firstPhantom = cells & ~3L;
        md2 = (uint32_t)(int)ceilf(minDistance * minDistance);
    }
    bool lds = useGrid && inLds && n <= (uint32_t)PICK_THREADS;
    uint32_t nbCount = 0;
    if (lds) {
        const uint32_t r = threadIdx.x;
        if (r < n) {
            const int64_t key = keys[r];
            const int px = (int)(key & 0xffff), py = (int)((key >> 16) & 0xffff);
            lstate[r] = pick_initial_state((uint32_t)px, (uint32_t)py, cell, gw, cells, firstPhantom, (int)md2);
            for (uint32_t j = 0; j < r; j++) {
                const int64_t kj = keys[j];
                const int dx = (int)(kj & 0xffff) - px, dy = (int)((kj >> 16) & 0xffff) - py;
                if ((uint32_t)(dx * dx + dy * dy) < md2) {
                    if (nbCount < PICK_NB) lnb[nbCount * PICK_THREADS + r] = (uint16_t)j;
                    nbCount++;
                }
            }
            if (nbCount > PICK_NB) soverflow = 1;
        }
        __syncthreads();
        lds = soverflow == 0;
    }
    if (lds) {
        const uint32_t r = threadIdx.x;
        while (true) {
            if (threadIdx.x == 0) sflag = 0;
            __syncthreads();
            if (r < n && lstate[r] == ST_UNDECIDED) {
                bool pending = false, rejected = false;
                for (uint32_t k = 0; k < nbCount; k++) {
                    const uint8_t sj = lstate[lnb[k * PICK_THREADS + r]];
                    if (sj == ST_IN) { rejected = true; break; }
                    if (sj == ST_UNDECIDED) pending = true;
                }
                if (rejected) lstate[r] = ST_OUT;
                else if (!pending) lstate[r] = ST_IN;
                sflag = 1;
            }
            __syncthreads();
            const uint32_t again = sflag;
            __syncthreads();
            if (!again) break;
        }
    } else if (useGrid) {
        for (uint32_t r = threadIdx.x; r < n; r += PICK_THREADS) {
            const int64_t key = keys[r];
            const uint32_t px = (uint32_t)(key & 0xffff), py = (uint32_t)((key >> 16) & 0xffff);
            const uint32_t hk = py * width + px;
            uint32_t s = hash_u32(hk) & W.hmask;
            while (atomicCAS(&W.hkey[s], HASH_EMPTY, hk) != HASH_EMPTY) s = (s + 1) & W.hmask;
            W.hval[s] = r;
            W.slotOf[r] = s;
            W.state[r] = pick_initial_state(px, py, cell, gw, cells, firstPhantom, (int)md2);
        }
        __syncthreads();
        const int rad = (int)ceilf(minDistance);
        while (true) {
            if (threadIdx.x == 0) sflag = 0;
            __syncthreads();
            for (uint32_t r = threadIdx.x; r < n; r += PICK_THREADS) {
                if (W.state[r] != ST_UNDECIDED) continue;
                const int64_t key = keys[r];
                const int px = (int)(key & 0xffff), py = (int)((key >> 16) & 0xffff);
                bool pending = false, rejected = false;
                for (int dy = -rad; dy <= rad && !rejected; dy++) {
                    const int yy = py + dy;
                    if (yy < 0 || yy >= (int)height) continue;
                    for (int dx = -rad; dx <= rad; dx++) {
                        const int xx = px + dx;
                        if ((dx | dy) == 0 || xx < 0 || xx >= (int)width || dx * dx + dy * dy >= (int)md2) continue;
                        const uint32_t j = hash_find(W, (uint32_t)yy * width + (uint32_t)xx);
                        if (j == HASH_EMPTY || j >= r) continue;
                        const uint8_t sj = W.state[j];
                        if (sj == ST_IN) { rejected = true; break; }
                        if (sj == ST_UNDECIDED) pending = true;
                    }
                }
                if (rejected) W.state[r] = ST_OUT;
                else if (!pending) W.state[r] = ST_IN;
                sflag = 1;
            }
            __syncthreads();
            const uint32_t again = sflag;
            __syncthreads();
            if (!again) break;
        }
    } else {
        for (uint32_t r = threadIdx.x; r < n; r += PICK_THREADS) W.state[r] = ST_IN;
        __syncthreads();
    }
    const uint8_t *state = lds ? lstate : W.state;

    const uint32_t per = (n + PICK_THREADS - 1) / PICK_THREADS;
    const uint32_t r0 = threadIdx.x * per, r1 = min(r0 + per, n);
    uint32_t mine = 0;
    for (uint32_t r = r0; r < r1; r++) mine += (state[r] == ST_IN);
    ssum[threadIdx.x] = mine;
    __syncthreads();
    for (uint32_t off = 1; off < PICK_THREADS; off <<= 1) {
        const uint32_t v = (threadIdx.x >= off) ? ssum[threadIdx.x - off] : 0u;
        __syncthreads();
        ssum[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t pos = ssum[threadIdx.x] - mine;
    const uint32_t total = ssum[PICK_THREADS - 1];
    for (uint32_t r = r0; r < r1; r++) {
        if (state[r] != ST_IN) continue;
        if (pos < capacity) {
            const int64_t key = keys[r];
            Keypoint kp;
            kp.x = (int32_t)(key & 0xffff);
            kp.y = (int32_t)((key >> 16) & 0xffff);
            kp.strength = __uint_as_float((uint32_t)((uint64_t)key >> 32));
            kp.scale = 0.0f;
            kp.orientation = 0.0f;
            kp.tracking_status = 1;
            kp.error = 0.0f;
            out[pos] = kp;
        }
        pos++;
    }
    if (useGrid && !lds)
        for (uint32_t r = threadIdx.x; r < n; r += PICK_THREADS) W.hkey[W.slotOf[r]] = HASH_EMPTY;
    __syncthreads();
    if (threadIdx.x == 0) {
        *outNumItems = min(total, capacity);
        // the grid path counts every kept point; the copy path (min_distance <= 2) reports the copied
        // number (ago_haf_cpu_harris.cpp:971-975 vs :988-989)
        *outCount = useGrid ? total : min(total, capacity);
        if (resetCount) *listCount = 0;
    }
}

}  // namespace mvx
