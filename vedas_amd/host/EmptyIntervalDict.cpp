#include "EmptyIntervalDict.h"

#include <algorithm>

bool EmptyIntervalDict::hasBoundFor(const std::string &var) const { return dict.count(var) > 0; }

VAR_BOUND EmptyIntervalDict::getBound(const std::string &var) const {
    auto it = dict.find(var);
    return it == dict.end() ? VAR_BOUND(0, 0, 0, 0) : it->second;
}

size_t EmptyIntervalDict::emptySize(const std::string &var) const {
    VAR_BOUND b = getBound(var);
    return (VAR_UB(b) - VAR_LB(b)) + (VAR_UB2(b) - VAR_LB2(b));
}

void EmptyIntervalDict::updateBound(const std::string &var, size_t lb1, size_t ub1,
                                    size_t lb2, size_t ub2) {
    if (!hasBoundFor(var)) {
        dict[var] = VAR_BOUND(lb1, ub1, lb2, ub2);
        return;
    }
    const VAR_BOUND old = dict.at(var);
    std::vector<std::pair<size_t, size_t>> iv = {
        {VAR_LB(old), VAR_UB(old)}, {VAR_LB2(old), VAR_UB2(old)}, {lb1, ub1}, {lb2, ub2}};
    std::sort(iv.begin(), iv.end());
    // Merge overlapping neighbours, right to left; the merged interval keeps
    // the later interval's upper end, as the reference does.
    for (size_t p = iv.size() - 1; p >= 1; --p) {
        if (iv[p].first <= iv[p - 1].second) {
            iv[p - 1].second = iv[p].second;
            iv.erase(iv.begin() + p);
        }
    }
    if (iv.empty()) {
        dict[var] = VAR_BOUND(0, 0, 0, 0);
    } else if (iv.size() == 1) {
        dict[var] = VAR_BOUND(0, 0, iv[0].first, iv[0].second);
    } else if (iv.size() == 2) {
        dict[var] = VAR_BOUND(iv[0].first, iv[0].second, iv[1].first, iv[1].second);
    } else {
        // Keep the two widest, in value order.
        size_t maxD = 0, secD = 0, maxI = 0, secI = 0;
        for (size_t i = 0; i < iv.size(); ++i) {
            const size_t d = iv[i].second - iv[i].first;
            if (d > maxD) {
                secD = maxD; secI = maxI; maxD = d; maxI = i;
            } else if (d > secD) {
                secD = d; secI = i;
            }
        }
        const size_t a = std::min(maxI, secI), b = std::max(maxI, secI);
        dict[var] = VAR_BOUND(iv[a].first, iv[a].second, iv[b].first, iv[b].second);
    }
}

std::vector<vd_interval_t> EmptyIntervalDict::skipIntervals(const std::string &var) const {
    std::vector<vd_interval_t> out;
    auto it = dict.find(var);
    if (it == dict.end()) return out;
    const VAR_BOUND &b = it->second;
    auto add = [&](size_t lo, size_t hi) {
        // Only well-formed 32-bit intervals; an empty one removes nothing.
        if (lo <= hi && hi <= 0xFFFFFFFFull) out.push_back({(uint32_t)lo, (uint32_t)hi});
    };
    if (VAR_UB(b) == 0 && VAR_UB2(b) == 0) return out;
    if (VAR_UB(b) != 0) add(VAR_LB(b), VAR_UB(b));
    add(VAR_LB2(b), VAR_UB2(b));
    return out;
}

void eifFromSortedColumn(const TYPEID_HOST_VEC &v0, size_t out[4]) {
    // Position of the widest gap, then the two widest gaps among the values
    // from two before it to three after it (the reference copies that window
    // back from the device).
    const size_t n = v0.size();
    size_t maxPos = 0;
    TYPEID maxDiff = 0;
    for (size_t i = 1; i < n; ++i) {
        const TYPEID d = v0[i] - v0[i - 1];
        if (d > maxDiff) { maxDiff = d; maxPos = i - 1; }
    }
    const size_t start = maxPos >= 2 ? maxPos - 2 : 0;
    const size_t end = std::min(maxPos + 4, n);
    TYPEID b[4] = {0, 0, 0, 0};
    size_t firstMax = 0, secondMax = 0;
    for (size_t i = start + 1; i < end; ++i) {
        const size_t gap = v0[i] - v0[i - 1];
        if (gap >= firstMax) {
            secondMax = firstMax;
            b[0] = b[2]; b[1] = b[3];
            firstMax = gap;
            b[2] = v0[i - 1]; b[3] = v0[i];
        } else if (gap > secondMax) {
            secondMax = gap;
            b[0] = v0[i - 1]; b[1] = v0[i];
        }
    }
    if (b[2] < b[0]) { std::swap(b[0], b[2]); std::swap(b[1], b[3]); }
    // A gap of fewer than two values holds no missing value; it (and an
    // unset gap) becomes the empty entry (0, 0) instead of the wrapped
    // interval the reference's unsigned arithmetic would produce.
    for (int k = 0; k < 2; ++k) {
        const TYPEID lo = b[2 * k], hi = b[2 * k + 1];
        if (hi > lo + 1) {
            out[2 * k] = (size_t)lo + 1;
            out[2 * k + 1] = (size_t)hi - 1;
        } else {
            out[2 * k] = out[2 * k + 1] = 0;
        }
    }
}
