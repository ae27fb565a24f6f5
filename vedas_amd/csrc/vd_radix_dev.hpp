// Device-side digit histogram used by the radix sort and fused into the
// kernels that produce sort keys (index packing, row-id packing), so the
// keys are read once for all passes' histograms.
#pragma once

#include "vd_internal.hpp"

namespace vd {
namespace radix {

constexpr int RBITS = 8;
constexpr int RBINS = 1 << RBITS;
constexpr int MAX_PASSES = 8;
// Threads of every kernel that uses the LDS histogram below.
constexpr int HIST_BLOCK = 256;
// Kernels that count histograms while they pack keys run at most this many
// blocks, each striding over tiles: every block ends with one global atomic
// per histogram bin, so a grid of one block per tile would issue tens of
// millions of same-address atomics on large inputs.
constexpr uint64_t HIST_GRID_MAX = 1024;
inline unsigned hist_grid(uint64_t ntiles) {
    return (unsigned)(ntiles < HIST_GRID_MAX ? (ntiles ? ntiles : 1) : HIST_GRID_MAX);
}
constexpr int HIST_WAVES = HIST_BLOCK / 64;

// LDS histogram: one copy per wave (fewer same-address atomics across
// waves), NP passes x 256 bins each.
template <int NP>
struct LdsHistN {
    uint32_t h[HIST_WAVES][NP][RBINS];
};
using LdsHist = LdsHistN<MAX_PASSES>;

template <int NP>
__device__ __forceinline__ void hist_zero(LdsHistN<NP> *s) {
    uint32_t *p = &s->h[0][0][0];
    for (int i = threadIdx.x; i < HIST_WAVES * NP * RBINS; i += HIST_BLOCK)
        p[i] = 0;
}

// Count the digits of one key for passes [0, npass), digit p covering bits
// [lo_bit + 8p, min(lo_bit + 8p + 8, hi_bit)).  With UNIFORM (for keys
// whose top digits are often equal across a wave, such as a segment id in
// the top bits), a digit held by every calling lane of the wave is added
// once with the lane count: LDS atomics on one address are serialised.
// Without it, one atomic per digit (cheaper for random digits).
template <bool UNIFORM = false, int NP>
__device__ __forceinline__ void hist_add(LdsHistN<NP> *s, uint64_t key, int lo_bit,
                                         int hi_bit, int npass) {
    const int w = threadIdx.x / 64;
    uint64_t act = 0;
    int leader = 0;
    if (UNIFORM) {
        act = __ballot(1);
        leader = __ffsll((unsigned long long)act) - 1;
    }
    for (int p = 0; p < npass && p < NP; ++p) {
        const int sh = lo_bit + p * RBITS;
        const int width = min(RBITS, hi_bit - sh);
        const uint32_t d = (uint32_t)(key >> sh) & ((1u << width) - 1u);
        if (UNIFORM) {
            const uint32_t d0 = __shfl(d, leader, 64);
            if (__ballot(d == d0) == act) {
                if ((int)(threadIdx.x & 63) == leader)
                    atomicAdd(&s->h[w][p][d0], (uint32_t)__popcll(act));
                continue;
            }
        }
        atomicAdd(&s->h[w][p][d], 1u);
    }
}

// Fold the block's LDS histogram into the global one (npass x 256 words).
// Must be preceded by a __syncthreads().
template <int NP>
__device__ __forceinline__ void hist_flush(LdsHistN<NP> *s, uint32_t *ghist,
                                           int npass) {
    for (int i = threadIdx.x; i < min(npass, NP) * RBINS; i += HIST_BLOCK) {
        const int p = i / RBINS, d = i % RBINS;
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < HIST_WAVES; ++w) c += s->h[w][p][d];
        if (c) atomicAdd(&ghist[i], c);
    }
}

}  // namespace radix
}  // namespace vd
