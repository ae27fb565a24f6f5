// C entry points of libvedas.so over the C++ operator API.
#include <chrono>
#include <cstring>
#include <memory>
#include <cctype>
#include <set>
#include <sstream>

#include "BatchExecutor.h"
#include "Dictionary.h"
#include "QueryExecutor.h"
#include "QueryGraph.h"
#include "Server.h"
#include "Statistics.h"
#include "vedas_host.h"

struct vdq_engine {
    std::unique_ptr<VedasContext> ctx;
    std::unique_ptr<VedasStorage> storage;
    std::unique_ptr<Statistics> stats;  // built on first use
    // Serving slots: contexts (streams + pools) that hold served results,
    // kept for the engine's lifetime.
    std::vector<std::unique_ptr<VedasContext>> slot_ctx;
    Statistics &statistics() {
        if (!stats) stats.reset(new Statistics(storage.get()));
        return *stats;
    }
    const vdq_dict *dict = nullptr;  // terms of the loaded store (caller owned)
    // Tunable values chosen by vdq_autotune; every context gets them.
    std::vector<std::pair<std::string, int>> tuned;
    void apply_tuned(VedasContext *c) {
        for (auto &kv : tuned) vd_ctx_tune(c->raw(), kv.first.c_str(), kv.second);
    }
    std::vector<VedasContext *> slots(int n) {
        while ((int)slot_ctx.size() < n) {
            slot_ctx.emplace_back(new VedasContext(ctx->device()));
            apply_tuned(slot_ctx.back().get());
        }
        std::vector<VedasContext *> v;
        for (int i = 0; i < n; ++i) v.push_back(slot_ctx[i].get());
        return v;
    }
};

struct vdq_result {
    SparqlResult result;
    std::vector<std::string> headers;
    std::vector<int> is_predicate;
    vdq_stats_t stats{};
};

struct vdq_dict {
    Dictionary d;
    explicit vdq_dict(bool ldict) : d(ldict) {}
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return VD_OK;
    } catch (const VedasError &e) {
        return fail(e.code(), e.what());
    } catch (const std::bad_alloc &) {
        return fail(VD_ENOMEM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(VD_EINVAL, e.what());
    }
}

// ---- query text -------------------------------------------------------------------
struct Parser {
    std::vector<std::string> tok;
    size_t pos = 0;
    explicit Parser(const std::string &s) {
        std::string cur;
        auto flush = [&]() { if (!cur.empty()) { tok.push_back(cur); cur.clear(); } };
        for (char c : s) {
            if (std::isspace((unsigned char)c)) { flush(); continue; }
            if (c == '{' || c == '}' || c == '(' || c == ')' || c == '.') {
                flush();
                tok.push_back(std::string(1, c));
                continue;
            }
            cur += c;
        }
        flush();
    }
    bool more() const { return pos < tok.size(); }
    const std::string &peek() const {
        if (!more()) throw VedasError(VD_EINVAL, "query text ends early");
        return tok[pos];
    }
    std::string next() { std::string t = peek(); ++pos; return t; }
    void expect(const std::string &t) {
        if (next() != t) throw VedasError(VD_EINVAL, "query text: expected " + t);
    }
};

struct PlanArena {
    std::vector<std::unique_ptr<PlanTreeNode>> nodes;
    PlanTreeNode *make() {
        nodes.emplace_back(new PlanTreeNode());
        return nodes.back().get();
    }
};

PlanTreeNode *parse_node(Parser &p, PlanArena &a, std::vector<TriplePattern> &pats) {
    p.expect("(");
    const std::string op = p.next();
    PlanTreeNode *n = a.make();
    if (op == "U") {
        const size_t i = std::stoul(p.next());
        if (i >= pats.size()) throw VedasError(VD_EINVAL, "plan names a missing pattern");
        n->op = UPLOAD;
        n->tp = &pats[i];
        n->var = p.next();
    } else if (op == "J") {
        n->op = JOIN;
        n->var = p.next();
        n->child1 = parse_node(p, a, pats);
        n->child2 = parse_node(p, a, pats);
    } else if (op == "S") {
        n->op = INDEXSWAP;
        n->var = p.next();
        n->child1 = parse_node(p, a, pats);
    } else {
        throw VedasError(VD_EINVAL, "unknown plan operator " + op);
    }
    p.expect(")");
    return n;
}

// One parsed query: patterns (stable storage), SELECT set, plan tree
// (from the text, or from the query graph planner).
struct ParsedQuery {
    std::vector<TriplePattern> pats;
    std::set<std::string> selected;
    PlanArena arena;
    std::unique_ptr<QueryGraph> graph;
    PlanTreeNode *root = nullptr;
};

void parse_query(const char *text, ParsedQuery &pq) {
    Parser p(text);
    p.expect("SELECT");
    bool star = false;
    while (p.peek() != "WHERE") {
        std::string v = p.next();
        if (v == "*") star = true; else pq.selected.insert(v);
    }
    p.expect("WHERE");
    p.expect("{");
    pq.pats.reserve(64);
    while (p.peek() != "}") {
        std::string s = p.next(), pr = p.next(), o = p.next();
        if (pq.pats.size() == 64) throw VedasError(VD_EINVAL, "at most 64 patterns");
        pq.pats.emplace_back(pq.pats.size(), s, pr, o);
        if (p.peek() == ".") p.next();
    }
    p.expect("}");
    if (pq.pats.empty()) throw VedasError(VD_EINVAL, "no patterns");
    if (star)
        for (auto &tp : pq.pats)
            for (auto t : {tp.getSubject(), tp.getPredicate(), tp.getObject()})
                if (!t.empty() && t[0] == '?') pq.selected.insert(t);
    if (p.more()) {
        p.expect("PLAN");
        pq.root = parse_node(p, pq.arena, pq.pats);
        if (p.more()) throw VedasError(VD_EINVAL, "text after the plan");
    }
}

// Rows of each pattern's index range (the reference's estimate_rows,
// QueryExecutor.cu:666,758): PSO for a bound or free subject, POS for a
// bound object.
void estimate_rows(vdq_engine *e, std::vector<TriplePattern> &pats) {
    VedasContext *ctx = e->ctx.get();
    vd_index *idx = e->storage->index();
    for (auto &tp : pats) {
        if (tp.predicateIsVariable()) {
            tp.estimate_rows = e->storage->getTripleSize();
            continue;
        }
        uint64_t b = 0, en = 0;
        int perm = VD_PERM_PSO, has2 = 0;
        uint32_t k2 = 0;
        if (!tp.subjectIsVariable()) {
            has2 = 1;
            k2 = tp.getSubjectId();
        } else if (!tp.objectIsVariable()) {
            perm = VD_PERM_POS;
            has2 = 1;
            k2 = tp.getObjectId();
        }
        ctx->check(vd_index_lookup(ctx->raw(), idx, perm, tp.getPredicateId(), has2, k2, &b, &en),
                   "vd_index_lookup");
        tp.estimate_rows = en - b;
    }
}

std::string plan_text(const PlanTreeNode *n) {
    switch (n->op) {
        case UPLOAD: return "(U " + std::to_string(n->tp->getId()) + " " + n->var + ")";
        case JOIN: return "(J " + n->var + " " + plan_text(n->child1) + " " + plan_text(n->child2) + ")";
        default: return "(S " + n->var + " " + plan_text(n->child1) + ")";
    }
}

// A query without PLAN is planned by the query graph (QueryExecutor.cu:
// 325-326), over the rows of each pattern's range.
void ensure_plan(vdq_engine *e, ParsedQuery &pq) {
    if (pq.root || pq.pats.size() == 1) return;
    estimate_rows(e, pq.pats);
    pq.graph.reset(new QueryGraph(pq.pats));
    pq.root = pq.graph->generateQueryPlan();
    if (!pq.root) throw VedasError(VD_EINVAL, "the planner produced no plan");
}

void fill_stats(vdq_result *res, VedasContext *ctx) {
    FullRelationIR *ir = res->result.getResultIR();
    for (size_t i = 0; ir && i < ir->getColumnNum(); ++i) {
        res->headers.push_back(ir->getHeader(i));
        res->is_predicate.push_back(ir->getIsPredicate(i) ? 1 : 0);
    }
    res->stats.rows = ir ? ir->size() : 0;
    if (!ctx) return;
    res->stats.join_rows = ctx->join_rows;
    res->stats.select_rows = ctx->select_rows;
    res->stats.swap_rows = ctx->swap_rows;
    res->stats.join_ms = ctx->join_ns / 1e6;
    res->stats.select_ms = ctx->upload_ns / 1e6;
    res->stats.swap_ms = ctx->swap_index_ns / 1e6;
    res->stats.jobs = (int)ctx->exe_log.size();
    res->stats.driven_joins = (int)ctx->driven_joins;
    res->stats.filtered_rows = ctx->filtered_rows;
}

}  // namespace

extern "C" {

int vdq_engine_create(int device, vdq_engine **out) {
    if (!out) return fail(VD_EINVAL, "null out");
    return guarded([&] {
        auto e = std::make_unique<vdq_engine>();
        e->ctx.reset(new VedasContext(device));
        *out = e.release();
    });
}

int vdq_engine_free(vdq_engine *e) {
    if (!e) return fail(VD_EINVAL, "null engine");
    return guarded([&] {
        e->storage.reset();
        e->ctx.reset();
        delete e;
    });
}

int vdq_engine_load(vdq_engine *e, const uint32_t *s, const uint32_t *p, const uint32_t *o,
                    uint64_t n, uint32_t perm_mask) {
    if (!e) return fail(VD_EINVAL, "null engine");
    return guarded([&] {
        vd_ctx *c = e->ctx->raw();
        void *d[3] = {nullptr, nullptr, nullptr};
        const uint32_t *h[3] = {s, p, o};
        try {
            for (int i = 0; i < 3; ++i) {
                e->ctx->check(vd_dev_alloc(c, (n ? n : 1) * 4, &d[i]), "vd_dev_alloc");
                if (n) e->ctx->check(vd_copy_h2d(c, d[i], h[i], n * 4), "vd_copy_h2d");
            }
            e->stats.reset();
            e->storage.reset();
            e->storage.reset(new VedasStorage(e->ctx.get(), (uint32_t *)d[0], (uint32_t *)d[1],
                                              (uint32_t *)d[2], n, perm_mask));
        } catch (...) {
            for (auto x : d) if (x) vd_dev_free(c, x);
            throw;
        }
        for (auto x : d) vd_dev_free(c, x);
    });
}

int vdq_dict_create(int literal_dict, vdq_dict **out) {
    if (!out) return fail(VD_EINVAL, "null argument");
    *out = nullptr;
    return guarded([&] { *out = new vdq_dict(literal_dict != 0); });
}

int vdq_dict_free(vdq_dict *d) {
    delete d;
    return VD_OK;
}

int vdq_dict_load_ntriples(vdq_dict *d, const char *path, uint64_t *n_triples) {
    if (!d || !path) return fail(VD_EINVAL, "null argument");
    return guarded([&] {
        d->d.loadNTriples(path);
        if (n_triples) *n_triples = d->d.subjects().size();
    });
}

int vdq_dict_load_ntriples_text(vdq_dict *d, const char *text, uint64_t *n_triples) {
    if (!d || !text) return fail(VD_EINVAL, "null argument");
    return guarded([&] {
        d->d.loadNTriplesText(text);
        if (n_triples) *n_triples = d->d.subjects().size();
    });
}

int vdq_dict_triples(const vdq_dict *d, uint32_t *s, uint32_t *p, uint32_t *o) {
    if (!d || !s || !p || !o) return fail(VD_EINVAL, "null argument");
    const size_t n = d->d.subjects().size();
    std::memcpy(s, d->d.subjects().data(), n * 4);
    std::memcpy(p, d->d.predicates().data(), n * 4);
    std::memcpy(o, d->d.objects().data(), n * 4);
    return VD_OK;
}

int vdq_dict_term(const vdq_dict *d, int is_predicate, uint32_t id, char *out, size_t cap) {
    if (!d || !out || cap == 0) return fail(VD_EINVAL, "null argument");
    const std::string t = d->d.term(id, is_predicate != 0);
    if (t.empty()) return fail(VD_EINVAL, "unknown ID " + std::to_string(id));
    if (t.size() + 1 > cap) return fail(VD_ERANGE, "term longer than the buffer");
    std::memcpy(out, t.c_str(), t.size() + 1);
    return VD_OK;
}

int vdq_dict_translate(const vdq_dict *d, const char *sparql, char *out, size_t cap) {
    if (!d || !sparql || !out || cap == 0) return fail(VD_EINVAL, "null argument");
    return guarded([&] {
        const std::string t = translateSparql(parseSparql(sparql), d->d);
        if (t.size() + 1 > cap) throw VedasError(VD_ERANGE, "query text longer than the buffer");
        std::memcpy(out, t.c_str(), t.size() + 1);
    });
}

int vdq_dict_save(const vdq_dict *d, const char *path) {
    if (!d || !path) return fail(VD_EINVAL, "null argument");
    return guarded([&] { d->d.save(path); });
}

int vdq_dict_open(vdq_dict *d, const char *path) {
    if (!d || !path) return fail(VD_EINVAL, "null argument");
    return guarded([&] { d->d.open(path); });
}

int vdq_engine_set_dict(vdq_engine *e, const vdq_dict *d) {
    if (!e || !d) return fail(VD_EINVAL, "null argument");
    e->dict = d;
    return VD_OK;
}

int vdq_engine_load_dict(vdq_engine *e, const vdq_dict *d, uint32_t perm_mask) {
    if (!e || !d) return fail(VD_EINVAL, "null argument");
    const size_t n = d->d.subjects().size();
    const int rc = vdq_engine_load(e, d->d.subjects().data(), d->d.predicates().data(),
                                   d->d.objects().data(), n, perm_mask);
    if (rc == VD_OK) e->dict = d;
    return rc;
}

int vdq_sparql(vdq_engine *e, const char *sparql, vdq_result **out) {
    if (!e || !sparql || !out) return fail(VD_EINVAL, "null argument");
    if (!e->dict) return fail(VD_EINVAL, "engine store has no dictionary (load it with vdq_engine_load_dict)");
    std::string text;
    const int rc = guarded([&] { text = translateSparql(parseSparql(sparql), e->dict->d); });
    if (rc != VD_OK) return rc;
    return vdq_query(e, text.c_str(), out);
}

int vdq_result_is_predicate(const vdq_result *r, int col) {
    if (!r || col < 0 || col >= (int)r->is_predicate.size()) return 0;
    return r->is_predicate[col];
}

int vdq_engine_save(vdq_engine *e, const char *path) {
    if (!e || !path) return fail(VD_EINVAL, "null argument");
    if (!e->storage) return fail(VD_EINVAL, "engine has no store loaded");
    return guarded([&] { e->storage->write(path); });
}

int vdq_engine_open(vdq_engine *e, const char *path) {
    if (!e || !path) return fail(VD_EINVAL, "null argument");
    return guarded([&] {
        e->stats.reset();
        e->storage.reset();
        e->storage.reset(new VedasStorage(e->ctx.get(), path));
    });
}

int vdq_engine_load_device(vdq_engine *e, const uint32_t *d_s, const uint32_t *d_p,
                           const uint32_t *d_o, uint64_t n, uint32_t perm_mask) {
    if (!e) return fail(VD_EINVAL, "null engine");
    return guarded([&] {
        e->stats.reset();
        e->storage.reset();
        e->storage.reset(new VedasStorage(e->ctx.get(), d_s, d_p, d_o, n, perm_mask));
    });
}

vd_ctx *vdq_engine_ctx(vdq_engine *e) { return e ? e->ctx->raw() : nullptr; }
vd_index *vdq_engine_index(vdq_engine *e) {
    return (e && e->storage) ? e->storage->index() : nullptr;
}

int vdq_query(vdq_engine *e, const char *text, vdq_result **out) {
    if (!e || !text || !out) return fail(VD_EINVAL, "null argument");
    if (!e->storage) return fail(VD_EINVAL, "engine has no store loaded");
    *out = nullptr;
    return guarded([&] {
        auto t0 = std::chrono::steady_clock::now();
        ParsedQuery pq;
        parse_query(text, pq);
        ensure_plan(e, pq);
        auto res = std::make_unique<vdq_result>();
        VedasContext *ctx = e->ctx.get();
        ctx->resetStats();
        QueryExecutor ex(e->storage.get(), ctx);
        ex.query(pq.pats, pq.selected, pq.root, res->result);
        ctx->sync();
        fill_stats(res.get(), ctx);
        res->stats.elapsed_ms = std::chrono::duration<double, std::milli>(
            std::chrono::steady_clock::now() - t0).count();
        *out = res.release();
    });
}

int vdq_batch(vdq_engine *e, const char *const *texts, int n, vdq_result **outs,
              vdq_batch_stats_t *stats) {
    if (!e || n < 0 || (n && (!texts || !outs))) return fail(VD_EINVAL, "null argument");
    if (!e->storage) return fail(VD_EINVAL, "engine has no store loaded");
    for (int i = 0; i < n; ++i) outs[i] = nullptr;
    return guarded([&] {
        auto t0 = std::chrono::steady_clock::now();
        std::vector<std::unique_ptr<ParsedQuery>> pqs;
        std::vector<BatchQuery> bq(n);
        for (int i = 0; i < n; ++i) {
            pqs.emplace_back(new ParsedQuery());
            parse_query(texts[i], *pqs.back());
            ensure_plan(e, *pqs.back());
            bq[i].patterns = &pqs.back()->pats;
            bq[i].selected = pqs.back()->selected;
            bq[i].root = pqs.back()->root;
        }
        std::vector<SparqlResult> results;
        BatchStats st;
        BatchExecutor bx(e->storage.get(), e->ctx.get());
        bx.run(bq, results, &st);
        e->ctx->sync();
        std::vector<std::unique_ptr<vdq_result>> rs;
        for (int i = 0; i < n; ++i) {
            rs.emplace_back(new vdq_result());
            rs.back()->result.setResult(results[i].release());
            fill_stats(rs.back().get(), nullptr);
        }
        for (int i = 0; i < n; ++i) outs[i] = rs[i].release();
        if (stats) {
            stats->rounds = st.rounds;
            stats->calls = st.launches_calls;
            stats->join_rows = st.join_rows;
            stats->select_rows = st.select_rows;
            stats->swap_rows = st.swap_rows;
            stats->elapsed_ms = std::chrono::duration<double, std::milli>(
                std::chrono::steady_clock::now() - t0).count();
        }
    });
}

// A batch captured as one HIP graph over the engine's context.
struct vdq_graph {
    vdq_engine *e = nullptr;
    vd_graph *g = nullptr;
    std::vector<std::unique_ptr<vdq_result>> results;  // rewritten by every launch
    ~vdq_graph() {
        results.clear();  // relations first: their columns are the graph's
        if (g) vd_graph_free(e->ctx->raw(), g);
    }
};

int vdq_graph_create(vdq_engine *e, const char *const *texts, int n, vdq_graph **out) {
    if (!e || !out || n < 0 || (n && !texts)) return fail(VD_EINVAL, "null argument");
    if (!e->storage) return fail(VD_EINVAL, "engine has no store loaded");
    *out = nullptr;
    return guarded([&] {
        std::vector<std::unique_ptr<ParsedQuery>> pqs;
        std::vector<BatchQuery> bq(n);
        for (int i = 0; i < n; ++i) {
            pqs.emplace_back(new ParsedQuery());
            parse_query(texts[i], *pqs.back());
            ensure_plan(e, *pqs.back());
            bq[i].patterns = &pqs.back()->pats;
            bq[i].selected = pqs.back()->selected;
            bq[i].root = pqs.back()->root;
        }
        VedasContext *ctx = e->ctx.get();
        vd_ctx *c = ctx->raw();
        // 1. the batch once, every read-back logged
        ctx->check(vd_ctx_record(c, 1), "vd_ctx_record");
        try {
            std::vector<SparqlResult> first;
            BatchExecutor bx(e->storage.get(), ctx);
            bx.run(bq, first, nullptr);
            ctx->sync();
        } catch (...) {
            vd_ctx_record(c, 0);
            throw;
        }
        ctx->check(vd_ctx_record(c, 0), "vd_ctx_record");
        // 2. the same batch captured (read-backs answered from the log)
        std::unique_ptr<vdq_graph> g(new vdq_graph());
        g->e = e;
        std::vector<SparqlResult> results;
        ctx->check(vd_ctx_capture_begin(c), "vd_ctx_capture_begin");
        try {
            BatchExecutor bx(e->storage.get(), ctx);
            bx.run(bq, results, nullptr);
        } catch (...) {
            results.clear();
            vd_ctx_capture_end(c, nullptr);
            throw;
        }
        const int rc = vd_ctx_capture_end(c, &g->g);
        if (rc != VD_OK) {
            results.clear();
            ctx->check(rc, "vd_ctx_capture_end");
        }
        for (int i = 0; i < n; ++i) {
            g->results.emplace_back(new vdq_result());
            g->results.back()->result.setResult(results[i].release());
            fill_stats(g->results.back().get(), nullptr);
        }
        // 3. the first launch computes the results
        ctx->check(vd_graph_launch(c, g->g), "vd_graph_launch");
        ctx->sync();
        *out = g.release();
    });
}

int vdq_graph_replay(vdq_graph *g) {
    if (!g || !g->g) return fail(VD_EINVAL, "null graph");
    return guarded([&] {
        VedasContext *ctx = g->e->ctx.get();
        ctx->check(vd_graph_launch(ctx->raw(), g->g), "vd_graph_launch");
        ctx->sync();
    });
}

vdq_result *vdq_graph_result(vdq_graph *g, int i) {
    if (!g || i < 0 || i >= (int)g->results.size()) return nullptr;
    return g->results[i].get();
}

int vdq_graph_free(vdq_graph *g) {
    if (!g) return fail(VD_EINVAL, "null graph");
    return guarded([&] { delete g; });
}

int vdq_serve(vdq_engine *e, const char *const *texts, const double *arrival_ms, int n,
              const vdq_serve_cfg_t *cfg, vdq_serve_rec_t *recs, vdq_result **outs) {
    if (!e || n < 0 || !cfg || (n && (!texts || !arrival_ms || !recs)))
        return fail(VD_EINVAL, "null argument");
    if (!e->storage) return fail(VD_EINVAL, "engine has no store loaded");
    if (outs)
        for (int i = 0; i < n; ++i) outs[i] = nullptr;
    return guarded([&] {
        std::vector<std::unique_ptr<ParsedQuery>> pqs;
        std::vector<ServeRequest> reqs(n);
        for (int i = 0; i < n; ++i) {
            pqs.emplace_back(new ParsedQuery());
            parse_query(texts[i], *pqs.back());
            ensure_plan(e, *pqs.back());
            reqs[i].patterns = pqs.back()->pats;
            reqs[i].selected = pqs.back()->selected;
            reqs[i].root = pqs.back()->root;
            reqs[i].arrival_ms = arrival_ms[i];
        }
        // plan nodes of request i must point at reqs[i].patterns
        for (int i = 0; i < n; ++i)
            for (auto &node : pqs[i]->arena.nodes)
                if (node->tp) node->tp = &reqs[i].patterns[node->tp->getId()];
        ServeConfig sc;
        sc.slots = cfg->slots;
        sc.max_batch = cfg->max_batch;
        sc.mem_budget = cfg->mem_budget;
        sc.plain = cfg->plain != 0;
        sc.min_fill = cfg->min_fill;
        sc.max_wait_ms = cfg->max_wait_ms;
        sc.keep_results = outs != nullptr;
        Statistics &st = e->statistics();
        Server srv(e->storage.get(), &st, sc, e->slots(sc.plain ? 1 : std::max(1, sc.slots)));
        std::vector<ServeRecord> rr;
        std::vector<SparqlResult> results;
        srv.run(reqs, rr, outs ? &results : nullptr);
        for (int i = 0; i < n; ++i) {
            recs[i].arrival_ms = rr[i].arrival_ms;
            recs[i].start_ms = rr[i].start_ms;
            recs[i].done_ms = rr[i].done_ms;
            recs[i].est_bytes = rr[i].est_bytes;
            recs[i].mem_held = rr[i].mem_held;
            recs[i].rows = rr[i].rows;
            recs[i].batch = rr[i].batch;
            recs[i].slot = rr[i].slot;
        }
        if (outs) {
            std::vector<std::unique_ptr<vdq_result>> rs;
            for (int i = 0; i < n; ++i) {
                rs.emplace_back(new vdq_result());
                rs.back()->result.setResult(results[i].release());
                fill_stats(rs.back().get(), nullptr);
            }
            for (int i = 0; i < n; ++i) outs[i] = rs[i].release();
        }
    });
}

// Coordinate descent over the kernel tunables: each candidate value of one
// tunable is timed on the caller's queries (best of `reps` runs of all of
// them) with the others held; a value is kept when it is at least 2%
// faster.  Two sweeps; the defaults are the starting point.
int vdq_autotune(vdq_engine *e, const char *const *texts, int n, int reps, vdq_tune_t *out) {
    if (!e || !texts || n <= 0 || !out) return fail(VD_EINVAL, "bad autotune arguments");
    if (!e->storage) return fail(VD_EINVAL, "engine has no store loaded");
    return guarded([&] {
        std::vector<std::unique_ptr<ParsedQuery>> qs;
        for (int i = 0; i < n; ++i) {
            qs.emplace_back(new ParsedQuery());
            parse_query(texts[i], *qs.back());
            ensure_plan(e, *qs.back());
        }
        VedasContext *ctx = e->ctx.get();
        auto set = [&](const char *key, int v) {
            ctx->check(vd_ctx_tune(ctx->raw(), key, v), "vd_ctx_tune");
        };
        auto run_all = [&]() {
            double best = 1e300;
            for (int r = 0; r < std::max(1, reps); ++r) {
                auto t0 = std::chrono::steady_clock::now();
                for (auto &q : qs) {
                    SparqlResult res;
                    QueryExecutor ex(e->storage.get(), ctx);
                    ex.query(q->pats, q->selected, q->root, res);
                }
                ctx->sync();
                best = std::min(best, std::chrono::duration<double, std::milli>(
                                          std::chrono::steady_clock::now() - t0).count());
            }
            return best;
        };
        struct Knob {
            const char *key;
            std::vector<int> values;  // values[0] = default
            int cur;
        };
        std::vector<Knob> knobs = {{"radix_items", {16, 8, 24}, 16},
                                   {"join_nv", {2048, 1024, 4096}, 2048},
                                   {"join_rcap", {4096, 2048, 8192}, 4096},
                                   {"radix_bits", {0, 8, 9}, 0},
                                   {"radix_group", {8, 0, 16}, 8}};
        for (auto &k : knobs) set(k.key, k.cur);
        run_all();  // warm-up
        const double base = run_all();
        double best = base;
        int evals = 1;
        for (int sweep = 0; sweep < 2; ++sweep) {
            for (auto &k : knobs) {
                for (int v : k.values) {
                    if (v == k.cur) continue;
                    set(k.key, v);
                    const double t = run_all();
                    ++evals;
                    if (t < 0.98 * best) {
                        best = t;
                        k.cur = v;
                    }
                }
                set(k.key, k.cur);
            }
        }
        e->tuned.clear();
        for (auto &k : knobs) e->tuned.emplace_back(k.key, k.cur);
        for (auto &sc : e->slot_ctx) e->apply_tuned(sc.get());
        out->radix_items = knobs[0].cur;
        out->join_nv = knobs[1].cur;
        out->join_rcap = knobs[2].cur;
        out->radix_bits = knobs[3].cur;
        out->radix_group = knobs[4].cur;
        out->base_ms = base;
        out->tuned_ms = best;
        out->evaluations = evals;
    });
}

int vdq_plan(const char *text, const uint64_t *est_rows, char *out, size_t cap) {
    if (!text || !out || cap == 0) return fail(VD_EINVAL, "null argument");
    return guarded([&] {
        ParsedQuery pq;
        parse_query(text, pq);
        std::string plan;
        if (pq.root) {
            plan = plan_text(pq.root);
        } else if (pq.pats.size() == 1) {
            const TriplePattern &tp = pq.pats[0];
            plan = "(U 0 " + (tp.subjectIsVariable() ? tp.getSubject() : tp.getObject()) + ")";
        } else {
            if (!est_rows) throw VedasError(VD_EINVAL, "estimates needed to plan");
            for (size_t i = 0; i < pq.pats.size(); ++i) pq.pats[i].estimate_rows = est_rows[i];
            QueryGraph g(pq.pats);
            PlanTreeNode *root = g.generateQueryPlan();
            if (!root) throw VedasError(VD_EINVAL, "the planner produced no plan");
            plan = plan_text(root);
        }
        if (plan.size() + 1 > cap) throw VedasError(VD_ERANGE, "plan text longer than the buffer");
        std::memcpy(out, plan.c_str(), plan.size() + 1);
    });
}

int vdq_estimate(vdq_engine *e, const char *text, double *rows, double *peak_bytes) {
    if (!e || !text) return fail(VD_EINVAL, "null argument");
    if (!e->storage) return fail(VD_EINVAL, "engine has no store loaded");
    return guarded([&] {
        ParsedQuery pq;
        parse_query(text, pq);
        ensure_plan(e, pq);
        PlanEstimator est(e->statistics());
        PlanEstimate r = est.estimate(pq.pats, pq.root);
        if (rows) *rows = r.rows;
        if (peak_bytes) *peak_bytes = r.peak_bytes;
    });
}

int vdq_result_shape(const vdq_result *r, int *ncols, uint64_t *nrows) {
    if (!r) return fail(VD_EINVAL, "null result");
    FullRelationIR *ir = r->result.getResultIR();
    if (ncols) *ncols = ir ? (int)ir->getColumnNum() : 0;
    if (nrows) *nrows = ir ? ir->size() : 0;
    return VD_OK;
}

const char *vdq_result_header(const vdq_result *r, int col) {
    if (!r || col < 0 || col >= (int)r->headers.size()) return nullptr;
    return r->headers[col].c_str();
}

int vdq_result_column(const vdq_result *r, int col, uint32_t *host_out) {
    if (!r || !host_out) return fail(VD_EINVAL, "null argument");
    return guarded([&] {
        FullRelationIR *ir = r->result.getResultIR();
        if (!ir || col < 0 || col >= (int)ir->getColumnNum())
            throw VedasError(VD_EINVAL, "no such column");
        TYPEID_HOST_VEC v = ir->download(col);
        if (!v.empty()) std::memcpy(host_out, v.data(), v.size() * 4);
    });
}

int vdq_result_stats(const vdq_result *r, vdq_stats_t *out) {
    if (!r || !out) return fail(VD_EINVAL, "null argument");
    *out = r->stats;
    return VD_OK;
}

int vdq_result_free(vdq_result *r) {
    if (!r) return fail(VD_EINVAL, "null result");
    return guarded([&] { delete r; });
}

int vdq_set_option(const char *key, int value) {
    if (!key) return fail(VD_EINVAL, "null option name");
    const std::string k(key);
    if (k == "bound_prescan") {
        QueryExecutor::ENABLE_BOUND_PRESCAN = value != 0;
    } else if (k == "bound_after_join") {
        QueryExecutor::ENABLE_UPDATE_BOUND_AFTER_JOIN = value != 0;
    } else if (k == "driven_swap") {
        QueryExecutor::ENABLE_DRIVEN_SWAP = value != 0;
    } else if (k == "filtered_swap") {
        if (value < 0) return fail(VD_EINVAL, "filtered_swap takes a row count >= 0");
        QueryExecutor::FILTERED_SWAP_MIN_ROWS = (uint64_t)value;
    } else {
        return fail(VD_EINVAL, "unknown option " + k);
    }
    return VD_OK;
}

int vdq_get_option(const char *key, int *value) {
    if (!key || !value) return fail(VD_EINVAL, "null argument");
    const std::string k(key);
    if (k == "bound_prescan") {
        *value = QueryExecutor::ENABLE_BOUND_PRESCAN;
    } else if (k == "bound_after_join") {
        *value = QueryExecutor::ENABLE_UPDATE_BOUND_AFTER_JOIN;
    } else if (k == "driven_swap") {
        *value = QueryExecutor::ENABLE_DRIVEN_SWAP;
    } else if (k == "filtered_swap") {
        *value = (int)std::min<uint64_t>(QueryExecutor::FILTERED_SWAP_MIN_ROWS, INT32_MAX);
    } else {
        return fail(VD_EINVAL, "unknown option " + k);
    }
    return VD_OK;
}

const char *vdq_last_error(void) { return g_err.c_str(); }

}  // extern "C"
