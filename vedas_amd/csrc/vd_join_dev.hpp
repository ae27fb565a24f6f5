// Device side of the join's bounds step, shared by the single join
// (vd_join.hip) and the batched join (vd_batch.hip).
//
// For a tile of consecutive left rows (keys sorted) and the sorted right
// key range they can touch, every left row needs lower = first right row
// with key >= its key and cnt = number of right rows equal to its key.
// Each thread takes ITEMS consecutive left rows: one binary search for the
// first, then the bounds of the next rows are found by galloping forward
// from the previous answer (the keys only grow), and a repeated key reuses
// the previous answer.  The left tile is staged in LDS so the thread-
// contiguous reads cost nothing, and the results are written back through
// LDS so the global stores stay coalesced.
#pragma once

#include "vd_internal.hpp"

namespace vd {
namespace jdev {

// First i in [start, end) with at(i) >= v (GE) or at(i) > v (!GE), given
// that every index before start fails the predicate: exponential steps
// from start, then a binary search inside the last step.
template <bool GE, typename At, typename T>
__device__ __forceinline__ uint64_t gallop(const At &at, uint64_t start, uint64_t end, T v) {
    auto pass = [&](uint64_t i) { return GE ? !(at(i) < v) : (v < at(i)); };
    if (start >= end || pass(start)) return start;
    uint64_t prev = start, step = 1;  // at(prev) fails
    uint64_t hi;
    for (;;) {
        const uint64_t nxt = prev + step;
        if (nxt >= end) {
            hi = end;
            break;
        }
        if (pass(nxt)) {
            hi = nxt;
            break;
        }
        prev = nxt;
        step <<= 1;
    }
    uint64_t lo = prev + 1;
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        if (pass(mid)) hi = mid; else lo = mid + 1;
    }
    return lo;
}

// Bounds of the ITEMS rows [r0, r0 + ITEMS) of a staged left tile (s_l,
// nvalid rows, read as s_l[r]) against the right range at(0 .. m); results
// relative to the range start.
template <int ITEMS, typename T, typename LA, typename At>
__device__ __forceinline__ void tile_bounds(const LA &s_l, uint32_t nvalid, uint32_t r0,
                                            const At &at, uint64_t m, uint32_t *lo_out,
                                            uint32_t *cnt_out) {
    uint64_t plo = 0, phi = 0;
    T pkey = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        lo_out[j] = 0;
        cnt_out[j] = 0;
        if (r0 + j >= nvalid) continue;
        const T key = s_l[r0 + j];
        uint64_t lo, hi;
        if (j > 0 && key == pkey) {
            lo = plo;
            hi = phi;
        } else {
            if (j == 0) {
                uint64_t a = 0, b = m;  // plain binary search for the first row
                while (a < b) {
                    const uint64_t mid = a + ((b - a) >> 1);
                    if (at(mid) < key) a = mid + 1; else b = mid;
                }
                lo = a;
            } else {
                lo = gallop<true>(at, phi, m, key);  // key > pkey: lower >= phi
            }
            hi = gallop<false>(at, lo, m, key);
        }
        lo_out[j] = (uint32_t)lo;
        cnt_out[j] = (uint32_t)(hi - lo);
        plo = lo;
        phi = hi;
        pkey = key;
    }
}

}  // namespace jdev
}  // namespace vd
