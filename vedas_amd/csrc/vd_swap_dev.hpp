// Digit plans of the stable sorts on one 32-bit value (IndexSwap on column
// 0, and the index permutations derived from a sibling permutation's
// order): which digit width and which digits, and the LDS histograms the
// producing kernel counts for them.
#pragma once

#include "vd_internal.hpp"

namespace vd {
namespace swap {

// Digit plan from the value's maximum mx: digits below its bit length can
// vary, the ones above are 0 in every row.  `bits` is the radix_bits
// tunable: 0 = 9-bit digits (bits 0..26) when they need fewer passes than
// bytes, else 8-bit; 8 = always bytes; 9 = 9-bit digits whenever the value
// fits 27 bits.
struct Plan {
    bool use9;
    int ndig;  // digits that can vary
};
__host__ __device__ inline Plan plan_for(uint32_t mx, int bits) {
    int nb = 0;
    while (nb < 32 && (mx >> nb) != 0) ++nb;
    const int n8 = (nb + 7) / 8, n9 = (nb + 8) / 9;
    const bool use9 = nb <= 27 && (bits == 9 || (bits == 0 && n9 < n8));
    return Plan{use9, use9 ? n9 : n8};
}

// Histograms of either plan in one buffer: [8-bit: 4 x 256 | 9-bit: 3 x 512].
constexpr int H8 = 4 * 256;
constexpr int H9 = 3 * 512;
constexpr int HW = H9 > H8 ? H9 : H8;  // LDS words of one wave's copy
constexpr int HBUF = H8 + H9;          // global histogram / offsets words

// Counts one digit per calling lane into a wave's LDS bins; when every
// calling lane of the wave holds the same digit (the top digits of small
// IDs), one lane adds the lane count (LDS atomics on one address from many
// lanes are serialised).
__device__ __forceinline__ void count_digit(uint32_t *bins, uint32_t d) {
    const uint64_t act = __ballot(1);
    const int leader = __ffsll((unsigned long long)act) - 1;
    const uint32_t d0 = __shfl(d, leader, 64);
    if (__ballot(d == d0) == act) {
        if ((int)(threadIdx.x & 63) == leader) atomicAdd(&bins[d0], (uint32_t)__popcll(act));
    } else {
        atomicAdd(&bins[d], 1u);
    }
}

// Adds value v's digits under plan p to one wave's LDS copy h (HW words).
// The top digit counts through count_digit (its values are the most
// concentrated), the others with plain LDS atomics.
__device__ __forceinline__ void hist_add(uint32_t *h, uint32_t v, Plan p) {
    const int W = p.use9 ? 9 : 8, bins = 1 << W;
    const uint32_t mask = (uint32_t)bins - 1u;
    for (int d = 0; d + 1 < p.ndig; ++d) atomicAdd(&h[d * bins + ((v >> (W * d)) & mask)], 1u);
    if (p.ndig > 0) count_digit(&h[(p.ndig - 1) * bins], (v >> (W * (p.ndig - 1))) & mask);
}

// Folds a block's per-wave LDS copies (nwaves x HW words, preceded by a
// __syncthreads()) into the plan's global histogram (g8 or g9).
__device__ __forceinline__ void hist_flush(const uint32_t *s_h, int nwaves, Plan p,
                                           uint32_t *g8, uint32_t *g9) {
    const int bins = p.use9 ? 512 : 256;
    uint32_t *g = p.use9 ? g9 : g8;
    for (int i = threadIdx.x; i < p.ndig * bins; i += blockDim.x) {
        uint32_t c = 0;
        for (int ww = 0; ww < nwaves; ++ww) c += s_h[ww * HW + i];
        if (c) atomicAdd(&g[i], c);
    }
}

}  // namespace swap
}  // namespace vd
