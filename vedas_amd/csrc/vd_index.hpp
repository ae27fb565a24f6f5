// Index structures shared by the index, select and batched-operator units.
#pragma once

#include "vd_internal.hpp"

struct vd_perm_arrays {
    bool built = false;
    uint64_t n = 0, n_l1 = 0, n_l2 = 0;
    uint32_t *l1v = nullptr, *l1o = nullptr, *l2v = nullptr, *l2o = nullptr;
    uint32_t *y = nullptr, *z = nullptr;
    // Maxima of the Y and Z columns (the input columns' maxima), when known.
    bool bounded = false;
    uint32_t ymax = 0, zmax = 0;
    // Rows equal to the row before them (repeated triples), when counted.
    bool dups_known = false;
    uint64_t dups = 0;
};

struct vd_index {
    vd_ctx *owner = nullptr;
    uint64_t n = 0;
    uint32_t mask = 0;
    vd_perm_arrays perm[VD_NUM_PERMS];
    // Owns the device arrays once the build succeeded; relations that read
    // ranges of them without a copy share it, so the arrays outlive
    // vd_index_free until the last such relation is destroyed.
    std::shared_ptr<void> hold;
    // Every triple occurs once (a built permutation counted no repeated
    // row): selects then return rows that are pairwise distinct.
    bool distinct() const {
        for (const auto &pa : perm)
            if (pa.built && pa.dups_known) return pa.dups == 0;
        return false;
    }
};

namespace vd {

struct PermDev {
    uint64_t n, n_l1;
    const uint32_t *l1v, *l1o, *y;
};

__device__ __forceinline__ void lookup_one(const PermDev &pd, uint32_t k1,
                                           uint32_t k2, bool has_k2,
                                           uint64_t *b, uint64_t *e) {
    const uint64_t i = dev::lower_bound(pd.l1v, 0, pd.n_l1, k1);
    if (i >= pd.n_l1 || pd.l1v[i] != k1) {
        const uint64_t pos = i < pd.n_l1 ? pd.l1o[i] : pd.n;
        *b = *e = pos;
        return;
    }
    uint64_t lo = pd.l1o[i];
    uint64_t hi = (i + 1 < pd.n_l1) ? pd.l1o[i + 1] : pd.n;
    if (has_k2) {
        lo = dev::lower_bound(pd.y, lo, hi, k2);
        hi = dev::upper_bound(pd.y, lo, hi, k2);
    }
    *b = lo;
    *e = hi;
}

// The key lookup of lookup_one, by one wave.
__device__ __forceinline__ void wave_lookup(const PermDev &pd, uint32_t k1, uint32_t k2,
                                            bool has_k2, uint64_t *b, uint64_t *e) {
    const uint64_t i = dev::wave_bound<false>(pd.l1v, 0, pd.n_l1, k1);
    if (i >= pd.n_l1 || pd.l1v[i] != k1) {
        const uint64_t pos = i < pd.n_l1 ? pd.l1o[i] : pd.n;
        *b = *e = pos;
        return;
    }
    uint64_t lo = pd.l1o[i];
    uint64_t hi = (i + 1 < pd.n_l1) ? pd.l1o[i + 1] : pd.n;
    if (has_k2) {
        lo = dev::wave_bound<false>(pd.y, lo, hi, k2);
        hi = dev::wave_bound<true>(pd.y, lo, hi, k2);
    }
    *b = lo;
    *e = hi;
}

PermDev perm_dev(const vd_perm_arrays &pa);
const vd_perm_arrays &get_perm(const vd_index *idx, int perm);

}  // namespace vd
