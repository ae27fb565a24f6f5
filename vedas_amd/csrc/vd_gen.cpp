// Seeded synthetic (S,P,O) triple generator (host code).
//
// Triples are drawn in chunks of 2^20; chunk k uses its own splitmix64
// stream seeded from (seed, k), so the output does not depend on the number
// of threads.  Zipf draws invert the CDF of w_k = k^-s over k = 1..M by
// binary search inside a guide-table bucket (ID = rank, ID 1 the most
// frequent).  With `distinct`, the output is the first n distinct triples
// of the chunk stream, in stream order (first occurrences found per hash
// bucket in parallel).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
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

unsigned worker_count(uint64_t items) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (nt > items) nt = (unsigned)std::max<uint64_t>(items, 1);
    return nt;
}

template <class F>
void run_threads(unsigned nt, F &&f) {
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) th.emplace_back(f, t);
    f(0u);
    for (auto &x : th) x.join();
}

struct Zipf {
    std::vector<double> cdf;     // cdf[k] = sum_{j<=k} w_j, normalised
    std::vector<uint32_t> guide; // guide[j] = first k with cdf[k] > j / G
    uint32_t m = 0, G = 0;
    bool uniform = true;
    void init(uint32_t items, double s) {
        m = items;
        uniform = (s == 0.0);
        if (uniform) return;
        cdf.resize(items);
        // weights in parallel, then the running sum in order (the same sums
        // whatever the thread count)
        const unsigned nt = worker_count((items + 65535) / 65536);
        run_threads(nt, [&](unsigned t) {
            for (uint64_t k = (uint64_t)items * t / nt; k < (uint64_t)items * (t + 1) / nt; ++k)
                cdf[k] = std::pow((double)(k + 1), -s);
        });
        double acc = 0.0;
        for (uint32_t k = 0; k < items; ++k) {
            acc += cdf[k];
            cdf[k] = acc;
        }
        run_threads(nt, [&](unsigned t) {
            for (uint64_t k = (uint64_t)items * t / nt; k < (uint64_t)items * (t + 1) / nt; ++k)
                cdf[k] /= acc;
        });
        cdf.back() = 1.0;
        // Guide table: the draw for u in bucket j = floor(u G) lies in
        // [guide[j], guide[j + 1]], so the search covers only a few CDF
        // entries (the whole CDF is up to 400 MB at 50M entities).
        G = std::min<uint32_t>(items, 1u << 22);
        guide.resize((size_t)G + 1);
        uint32_t k = 0;
        for (uint32_t j = 0; j <= G; ++j) {
            const double x = (double)j / (double)G;
            while (k < m && cdf[k] <= x) ++k;
            guide[j] = k;
        }
    }
    // Returns an ID in [1, m]: the first k with cdf[k] > u, plus one (the
    // inverse CDF).  The bucket index may be off by one from rounding, so
    // the search spans the neighbouring buckets too; the result is the
    // same as an upper_bound over the whole CDF.
    uint32_t draw(uint64_t &st) const {
        const double u = unit(st);
        if (uniform) {
            uint64_t k = (uint64_t)(u * m);
            if (k >= m) k = m - 1;
            return (uint32_t)k + 1;
        }
        uint64_t j = (uint64_t)(u * (double)G);
        if (j >= G) j = G - 1;
        const size_t lo = guide[j > 0 ? j - 1 : 0];
        size_t hi = (size_t)guide[std::min<uint64_t>(j + 2, G)] + 1;
        if (hi > m) hi = m;
        size_t k = std::upper_bound(cdf.begin() + lo, cdf.begin() + hi, u) - cdf.begin();
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
    const unsigned nt = worker_count(nchunks);
    run_threads(nt, [&](unsigned t) {
        for (uint64_t c = t; c < nchunks; c += nt) {
            const uint64_t off = c * CHUNK;
            const uint64_t cnt = std::min(CHUNK, n - off);
            g.chunk(first_chunk + c, cnt, s + off, p + off, o + off);
        }
    });
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

// keep[i] = 1 for the first occurrence of each triple among candidates
// [0, m), else 0.  Candidates are grouped into buckets by the top bits of
// their hash (a counting sort that keeps ascending candidate order inside a
// bucket); each bucket is then deduplicated on its own with an open-address
// table, visiting its candidates in ascending order.  The flags equal those
// of one sequential pass, whatever the thread count.
void first_occurrences(const uint32_t *s, const uint32_t *p, const uint32_t *o,
                       uint64_t m, uint8_t *keep) {
    unsigned bbits = 6;
    while (bbits < 16 && (m >> (bbits + 16)) > 0) ++bbits;
    const uint64_t nb = 1ull << bbits;
    const unsigned nt = worker_count((m + 65535) / 65536);
    std::vector<uint64_t> cnt((size_t)nt * nb, 0);
    auto lo_of = [&](unsigned t) { return m * t / nt; };
    run_threads(nt, [&](unsigned t) {
        uint64_t *c = &cnt[(size_t)t * nb];
        for (uint64_t i = lo_of(t); i < lo_of(t + 1); ++i)
            ++c[triple_hash(s[i], p[i], o[i]) >> (64 - bbits)];
    });
    std::vector<uint64_t> bstart(nb + 1, 0);
    uint64_t acc = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        bstart[b] = acc;
        for (unsigned t = 0; t < nt; ++t) {
            const uint64_t v = cnt[(size_t)t * nb + b];
            cnt[(size_t)t * nb + b] = acc;
            acc += v;
        }
    }
    bstart[nb] = acc;
    std::unique_ptr<uint32_t[]> order(new uint32_t[m]);  // first touched by the workers
    run_threads(nt, [&](unsigned t) {
        uint64_t *c = &cnt[(size_t)t * nb];
        for (uint64_t i = lo_of(t); i < lo_of(t + 1); ++i)
            order[c[triple_hash(s[i], p[i], o[i]) >> (64 - bbits)]++] = (uint32_t)i;
    });
    std::atomic<uint64_t> next{0};
    run_threads(nt, [&](unsigned) {
        std::vector<uint32_t> table;
        for (;;) {
            const uint64_t b = next.fetch_add(1);
            if (b >= nb) break;
            const uint64_t lo = bstart[b], hi = bstart[b + 1];
            uint64_t cap = 16;
            while (cap < 2 * (hi - lo)) cap <<= 1;
            table.assign(cap, 0);  // 0 = empty, else candidate + 1
            const uint64_t mask = cap - 1;
            for (uint64_t k = lo; k < hi; ++k) {
                const uint32_t i = order[k];
                uint64_t h = triple_hash(s[i], p[i], o[i]) & mask;
                bool dup = false;
                for (;;) {
                    const uint32_t e = table[h];
                    if (e == 0) break;
                    const uint32_t r = e - 1;
                    if (s[r] == s[i] && p[r] == p[i] && o[r] == o[i]) {
                        dup = true;
                        break;
                    }
                    h = (h + 1) & mask;
                }
                keep[i] = !dup;  // every candidate is in exactly one bucket
                if (!dup) table[h] = i + 1;
            }
        }
    });
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
        // The candidate stream is chunks 0, 1, 2, ... of the generator; the
        // output is its first n distinct triples, in stream order.  Enough
        // chunks are drawn for a few percent of repeats; if the stream is
        // short of n distinct triples, more chunks are added and the flags
        // recomputed.
        uint64_t nchunks = (n + n / 16 + CHUNK - 1) / CHUNK + 1;
        const uint64_t max_chunks = 64 + 4 * (n / CHUNK + 1);
        for (;;) {
            const uint64_t m = nchunks * CHUNK;
            if (m >> 32) return VD_ERANGE;
            // Uninitialised buffers: the workers touch their pages first.
            // A longer stream redraws every chunk (chunk k is the same
            // draws each time).
            std::unique_ptr<uint32_t[]> bs(new uint32_t[m]), bp(new uint32_t[m]),
                bo(new uint32_t[m]);
            std::unique_ptr<uint8_t[]> keep(new uint8_t[m]);
            parallel_chunks(g, 0, m, bs.get(), bp.get(), bo.get());
            first_occurrences(bs.get(), bp.get(), bo.get(), m, keep.get());
            // first n kept candidates, in order: per-thread counts, then copies
            const unsigned nt = worker_count((m + 65535) / 65536);
            std::vector<uint64_t> kc(nt + 1, 0);
            run_threads(nt, [&](unsigned t) {
                uint64_t c = 0;
                for (uint64_t i = m * t / nt; i < m * (t + 1) / nt; ++i) c += keep[i];
                kc[t + 1] = c;
            });
            for (unsigned t = 0; t < nt; ++t) kc[t + 1] += kc[t];
            if (kc[nt] >= n) {
                run_threads(nt, [&](unsigned t) {
                    uint64_t w = kc[t];
                    for (uint64_t i = m * t / nt; i < m * (t + 1) / nt && w < n; ++i) {
                        if (!keep[i]) continue;
                        s[w] = bs[i]; p[w] = bp[i]; o[w] = bo[i];
                        ++w;
                    }
                });
                break;
            }
            if (nchunks >= max_chunks) return VD_ERANGE;  // ID space too small
            nchunks = std::min(max_chunks, nchunks + std::max<uint64_t>(nchunks / 2, 1));
        }
    } catch (const std::bad_alloc &) {
        return VD_ENOMEM;
    }
    return VD_OK;
}
