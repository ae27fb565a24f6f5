// Seeded synthetic (S,P,O) triple generator (host code).
//
// Triples are drawn in chunks of 2^20; chunk k uses its own splitmix64
// stream seeded from (seed, k), so the output does not depend on the number
// of threads.  Zipf draws invert the CDF of w_k = k^-s over k = 1..M by
// binary search (ID = rank, ID 1 the most frequent).  With `distinct`, the
// first occurrence of each triple is kept and further chunks are drawn
// until n distinct triples exist.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "vedas_hip.h"

namespace {

inline uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// 64-bit finaliser of MurmurHash3 (a bijection that mixes all bits).
inline uint64_t fmix64(uint64_t h) {
    h ^= h >> 33;
    h *= 0xFF51AFD7ED558CCDull;
    h ^= h >> 33;
    h *= 0xC4CEB9FE1A85EC53ull;
    h ^= h >> 33;
    return h;
}

inline double unit(uint64_t &st) {
    return (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
}

struct Zipf {
    std::vector<double> cdf;  // cdf[k] = sum_{j<=k} w_j, normalised
    uint32_t m = 0;
    bool uniform = true;
    void init(uint32_t items, double s) {
        m = items;
        uniform = (s == 0.0);
        if (uniform) return;
        cdf.resize(items);
        double acc = 0.0;
        for (uint32_t k = 0; k < items; ++k) {
            acc += std::pow((double)(k + 1), -s);
            cdf[k] = acc;
        }
        for (auto &c : cdf) c /= acc;
        cdf.back() = 1.0;
    }
    // Returns an ID in [1, m].
    uint32_t draw(uint64_t &st) const {
        const double u = unit(st);
        if (uniform) {
            uint64_t k = (uint64_t)(u * m);
            if (k >= m) k = m - 1;
            return (uint32_t)k + 1;
        }
        size_t k = std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
        if (k >= m) k = m - 1;
        return (uint32_t)k + 1;
    }
};

constexpr uint64_t CHUNK = 1ull << 20;

struct Gen {
    const vd_gen_params_t &p;
    Zipf subj, pred;
    explicit Gen(const vd_gen_params_t &prm) : p(prm) {
        subj.init(prm.n_entities, prm.subject_zipf);
        pred.init(prm.n_predicates, prm.pred_zipf);
    }
    void chunk(uint64_t k, uint64_t count, uint32_t *s, uint32_t *pp,
               uint32_t *o) const {
        // Start of chunk k's stream: a hash of (seed, k).  (An offset of
        // k times the splitmix increment would only shift chunk 0's stream
        // by k draws.)
        uint64_t st = fmix64(fmix64(p.seed + 0x632BE59BD9B4E019ull) ^ (k * 0xD1B54A32D192ED03ull + 1));
        for (uint64_t i = 0; i < count; ++i) {
            s[i] = subj.draw(st);
            pp[i] = pred.draw(st);
            if (unit(st) < p.literal_frac) {
                uint64_t l = (uint64_t)(unit(st) * p.n_literals);
                if (l >= p.n_literals) l = p.n_literals - 1;
                o[i] = p.literal_base + (uint32_t)l;
            } else {
                uint64_t e = (uint64_t)(unit(st) * p.n_entities);
                if (e >= p.n_entities) e = p.n_entities - 1;
                o[i] = (uint32_t)e + 1;
            }
        }
    }
};

void parallel_chunks(const Gen &g, uint64_t first_chunk, uint64_t n,
                     uint32_t *s, uint32_t *p, uint32_t *o) {
    const uint64_t nchunks = (n + CHUNK - 1) / CHUNK;
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (nt > nchunks) nt = (unsigned)nchunks;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t]() {
            for (uint64_t c = t; c < nchunks; c += nt) {
                const uint64_t off = c * CHUNK;
                const uint64_t cnt = std::min(CHUNK, n - off);
                g.chunk(first_chunk + c, cnt, s + off, p + off, o + off);
            }
        });
    }
    for (auto &x : th) x.join();
}

inline uint64_t triple_hash(uint32_t s, uint32_t p, uint32_t o) {
    uint64_t h = (uint64_t)s * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)p * 0xC2B2AE3D27D4EB4Full + (h >> 29);
    h ^= (uint64_t)o * 0x165667B19E3779F9ull + (h >> 31);
    h ^= h >> 33;
    h *= 0xFF51AFD7ED558CCDull;
    h ^= h >> 33;
    return h;
}

}  // namespace

extern "C" int vd_gen_triples(const vd_gen_params_t *prm, uint32_t *s,
                              uint32_t *p, uint32_t *o) {
    if (!prm || (prm->n && (!s || !p || !o))) return VD_EINVAL;
    if (prm->n_entities == 0 || prm->n_predicates == 0) return VD_EINVAL;
    if (prm->literal_frac > 0.0 && prm->n_literals == 0) return VD_EINVAL;
    const uint64_t n = prm->n;
    if (n == 0) return VD_OK;
    try {
        Gen g(*prm);
        if (!prm->distinct) {
            parallel_chunks(g, 0, n, s, p, o);
            return VD_OK;
        }
        // Open addressing over row ids; 0 = empty, else row + 1.
        uint64_t cap = 1;
        while (cap < 2 * n) cap <<= 1;
        if (cap >> 32) return VD_ERANGE;
        std::vector<uint32_t> table(cap, 0);
        const uint64_t mask = cap - 1;
        uint64_t have = 0, next_chunk = 0;
        std::vector<uint32_t> bs, bp, bo;
        while (have < n) {
            const uint64_t want = std::max<uint64_t>(n - have, CHUNK);
            const uint64_t cnt = (want + CHUNK - 1) / CHUNK * CHUNK;
            bs.resize(cnt); bp.resize(cnt); bo.resize(cnt);
            parallel_chunks(g, next_chunk, cnt, bs.data(), bp.data(), bo.data());
            next_chunk += cnt / CHUNK;
            for (uint64_t i = 0; i < cnt && have < n; ++i) {
                uint64_t h = triple_hash(bs[i], bp[i], bo[i]) & mask;
                bool dup = false;
                for (;;) {
                    const uint32_t e = table[h];
                    if (e == 0) break;
                    const uint64_t r = e - 1;
                    if (s[r] == bs[i] && p[r] == bp[i] && o[r] == bo[i]) {
                        dup = true;
                        break;
                    }
                    h = (h + 1) & mask;
                }
                if (dup) continue;
                s[have] = bs[i]; p[have] = bp[i]; o[have] = bo[i];
                table[h] = (uint32_t)(have + 1);
                ++have;
            }
            if (next_chunk > 64 + 4 * (n / CHUNK + 1)) return VD_ERANGE;  // space too small
        }
    } catch (const std::bad_alloc &) {
        return VD_ENOMEM;
    }
    return VD_OK;
}
