#include "BatchExecutor.h"

#include <climits>
#include <functional>

#include "FullRelationIR.h"
#include "QueryExecutor.h"

struct BatchExecutor::Rel {
    vd_rel *rel = nullptr;
    std::vector<std::string> headers;
    bool is_pred[VD_MAX_COLS] = {};
    size_t col(const std::string &v) const {
        for (size_t i = 0; i < headers.size(); ++i)
            if (headers[i] == v) return i;
        throw VedasError(VD_EINVAL, "no column for " + v);
    }
    uint64_t rows() const { return rel ? vd_rel_nrows(rel) : 0; }
    // moved to column 0 by an index swap, sorted by the join that takes it
    bool pending = false;
};

struct BatchExecutor::Node {
    PlanOperation op = UPLOAD;
    int c1 = -1, c2 = -1;
    std::string var;
    TriplePattern *tp = nullptr;
    bool done = false;
    int owner = -1;  // node holding the relation this node's result lives in
    int parent = -1;
    Rel out;
};

struct BatchExecutor::Query {
    BatchQuery *q = nullptr;
    std::vector<Node> nodes;
    int root = -1;
    std::map<std::string, size_t> counter;
    std::set<std::string> join_vars;
};

BatchExecutor::BatchExecutor(VedasStorage *s, VedasContext *c) : storage(s), context(c) {}

namespace {

// Mirrors QueryExecutor::create1VarSelectQueryJob / create2VarSelectQueryJob
// (QueryExecutor.cu:781-971): the permutation, constants and headers of the
// select for a pattern, given the variable its index is chosen for.
struct SelSpec {
    vd_select_job_t job;
    std::vector<std::string> headers;
    bool is_pred[2] = {false, false};
    bool distinct = false;  // second variable unused: distinct first column
};

SelSpec select_spec(const TriplePattern &tp, const std::string &var, bool rootOnly,
                    const std::function<bool(const std::string &)> &used) {
    SelSpec s{};
    s.job.k2 = VD_UNBOUND;
    if (tp.getVariableNum() == 1) {
        s.job.cols = 1;
        if (tp.subjectIsVariable()) {
            s.job.perm = VD_PERM_POS; s.job.k1 = tp.getPredicateId(); s.job.k2 = tp.getObjectId();
            s.headers = {tp.getSubject()};
        } else if (tp.predicateIsVariable()) {
            s.job.perm = VD_PERM_SOP; s.job.k1 = tp.getSubjectId(); s.job.k2 = tp.getObjectId();
            s.headers = {tp.getPredicate()};
            s.is_pred[0] = true;
        } else {
            s.job.perm = VD_PERM_PSO; s.job.k1 = tp.getPredicateId(); s.job.k2 = tp.getSubjectId();
            s.headers = {tp.getObject()};
        }
        return s;
    }
    if (tp.getVariableNum() != 2) throw VedasError(VD_EINVAL, "select needs 1 or 2 variables");
    s.job.cols = 3;
    // postorderTraversal passes "PSO" when the index variable is the
    // subject, else "POS"; a single-pattern query passes nothing (PSO).
    const bool subjIdx = rootOnly || tp.getSubject() == var;
    std::string second;
    switch (tp.getVariableBitmap()) {
        case 5:
            if (!subjIdx) {
                s.job.perm = VD_PERM_POS; s.job.k1 = tp.getPredicateId();
                s.headers = {tp.getObject(), tp.getSubject()};
            } else {
                s.job.perm = VD_PERM_PSO; s.job.k1 = tp.getPredicateId();
                s.headers = {tp.getSubject(), tp.getObject()};
            }
            break;
        case 3:  // the "PSO"/"POS" names never select SOP here
            s.job.perm = VD_PERM_SPO; s.job.k1 = tp.getSubjectId();
            s.headers = {tp.getPredicate(), tp.getObject()};
            s.is_pred[0] = true;
            break;
        case 6:
            s.job.perm = VD_PERM_OSP; s.job.k1 = tp.getObjectId();
            s.headers = {tp.getSubject(), tp.getPredicate()};
            s.is_pred[1] = true;
            break;
        default:
            throw VedasError(VD_EINVAL, "unsupported 2-variable pattern shape");
    }
    s.distinct = !used(s.headers[1]);
    return s;
}

}  // namespace

void BatchExecutor::selectRound(std::vector<std::pair<Query *, int>> &ready, BatchStats &st) {
    std::vector<vd_select_job_t> jobs;
    std::vector<SelSpec> specs;
    for (auto &qn : ready) {
        Query *Q = qn.first;
        Node &n = Q->nodes[qn.second];
        auto used = [Q](const std::string &v) {
            return Q->q->selected.count(v) > 0 || Q->join_vars.count(v) > 0;
        };
        SelSpec sp = select_spec(*n.tp, n.var, Q->nodes.size() == 1, used);
        if (!storage->hasPermutation(sp.job.perm))
            throw VedasError(VD_EINVAL, "pattern needs a permutation that was not built");
        jobs.push_back(sp.job);
        specs.push_back(sp);
    }
    std::vector<vd_rel *> outs(jobs.size(), nullptr);
    context->check(vd_select_multi(context->raw(), storage->index(), jobs.data(), (int)jobs.size(),
                                   outs.data()),
                   "vd_select_multi");
    st.launches_calls++;
    std::vector<vd_rel *> distinct;
    for (size_t i = 0; i < ready.size(); ++i) {
        Node &n = ready[i].first->nodes[ready[i].second];
        n.out.rel = outs[i];
        n.out.headers = specs[i].headers;
        n.out.is_pred[0] = specs[i].is_pred[0];
        n.out.is_pred[1] = specs[i].is_pred[1];
        if (specs[i].distinct) {
            // the select of SelectQueryJob.cu:151-158: first column only, distinct
            context->check(vd_rel_drop_col(context->raw(), n.out.rel, 1), "vd_rel_drop_col");
            n.out.headers.resize(1);
            distinct.push_back(n.out.rel);
        }
        n.owner = ready[i].second;
        n.done = true;
        st.select_rows += vd_rel_nrows(outs[i]);
    }
    if (!distinct.empty()) {
        context->check(vd_unique_adjacent_multi(context->raw(), distinct.data(), (int)distinct.size()),
                       "vd_unique_adjacent_multi");
        st.launches_calls++;
    }
}

namespace {

void swap_cols(BatchExecutor *, vd_rel *rel, std::vector<std::string> &h, bool *isp, size_t i,
               size_t j) {
    if (i == j) return;
    vd_rel_swap_cols(rel, (int)i, (int)j);
    std::swap(h[i], h[j]);
    std::swap(isp[i], isp[j]);
}

}  // namespace

void BatchExecutor::joinRound(std::vector<std::pair<Query *, int>> &joins,
                              std::vector<std::pair<Query *, int>> &swaps, BatchStats &st) {
    vd_ctx *c = context->raw();
    std::vector<vd_rel *> to_sort;
    // Index swaps: move the variable to column 0, sort below.  A large
    // relation whose parent joins on that variable is left to the join,
    // which first drops the rows the other input cannot match (as
    // IndexSwapJob and JoinQueryJob::join do on the one-query path).
    for (auto &qn : swaps) {
        Query *Q = qn.first;
        Node &n = Q->nodes[qn.second];
        Rel &r = Q->nodes[Q->nodes[n.c1].owner].out;
        swap_cols(this, r.rel, r.headers, r.is_pred, 0, r.col(n.var));
        const Node *par = n.parent >= 0 ? &Q->nodes[n.parent] : nullptr;
        const uint64_t min_rows = QueryExecutor::FILTERED_SWAP_MIN_ROWS;
        if (par && par->op == JOIN && par->var == n.var && min_rows > 0 && r.rows() >= min_rows) {
            r.pending = true;
        } else {
            to_sort.push_back(r.rel);
            st.swap_rows += r.rows();
        }
        n.owner = Q->nodes[n.c1].owner;
    }
    // Joins: operand order, projection push-down, key to column 0
    // (JoinQueryJob.cu:125-180).
    struct Pending { Query *Q; int node; Rel *L; Rel *R; };
    std::vector<Pending> pend;
    for (auto &qn : joins) {
        Query *Q = qn.first;
        Node &n = Q->nodes[qn.second];
        Rel *L = &Q->nodes[Q->nodes[n.c1].owner].out;
        Rel *R = &Q->nodes[Q->nodes[n.c2].owner].out;
        if (L->rows() > R->rows()) std::swap(L, R);
        for (Rel *X : {L, R}) {
            bool resort = false;
            for (int i = (int)X->headers.size() - 1; i >= 0; --i) {
                const std::string h = X->headers[i];
                if (Q->counter[h] == 0 && h != n.var) {
                    context->check(vd_rel_drop_col(c, X->rel, i), "vd_rel_drop_col");
                    X->headers.erase(X->headers.begin() + i);
                    for (size_t k = i; k + 1 < VD_MAX_COLS; ++k) X->is_pred[k] = X->is_pred[k + 1];
                    if (i == 0) resort = true;
                }
            }
            const size_t k = X->col(n.var);
            if (k != 0) {
                swap_cols(this, X->rel, X->headers, X->is_pred, 0, k);
                resort = true;
            }
            if (X->pending) {
                // keep the rows whose key the other input holds (when that
                // is at most half of them), then sort
                Rel *Y = X == L ? R : L;
                vd_rel *f = nullptr;
                uint64_t kept = 0;
                context->check(vd_semi_filter(c, X->rel, 0, Y->rel, 0, 0xFFFFFFFFu,
                                              X->rows() / 2, &f, &kept),
                               "vd_semi_filter");
                if (f) {
                    vd_rel_destroy(c, X->rel);
                    X->rel = f;
                }
                X->pending = false;
                resort = true;
            }
            if (resort) {
                to_sort.push_back(X->rel);
                st.swap_rows += X->rows();
            }
        }
        Q->counter[n.var] -= 1;
        pend.push_back({Q, qn.second, L, R});
    }
    if (!to_sort.empty()) {
        context->check(vd_sort_col0_multi(c, to_sort.data(), (int)to_sort.size()),
                       "vd_sort_col0_multi");
        st.launches_calls++;
    }
    for (auto &qn : swaps) qn.first->nodes[qn.second].done = true;
    if (pend.empty()) return;
    std::vector<const vd_rel *> Ls, Rs;
    for (auto &p : pend) {
        Ls.push_back(p.L->rel);
        Rs.push_back(p.R->rel);
    }
    std::vector<vd_rel *> outs(pend.size(), nullptr);
    context->check(vd_join_multi(c, Ls.data(), Rs.data(), (int)pend.size(), outs.data()),
                   "vd_join_multi");
    st.launches_calls++;
    for (size_t i = 0; i < pend.size(); ++i) {
        Pending &p = pend[i];
        Node &n = p.Q->nodes[p.node];
        n.out.rel = outs[i];
        n.out.headers.clear();
        size_t k = 0;
        n.out.headers.push_back(p.L->headers[0]);
        n.out.is_pred[k++] = p.L->is_pred[0];
        for (size_t j = 1; j < p.L->headers.size(); ++j) {
            n.out.headers.push_back(p.L->headers[j]);
            n.out.is_pred[k++] = p.L->is_pred[j];
        }
        for (size_t j = 1; j < p.R->headers.size(); ++j) {
            n.out.headers.push_back(p.R->headers[j]);
            n.out.is_pred[k++] = p.R->is_pred[j];
        }
        st.join_rows += vd_rel_nrows(outs[i]);
        // the inputs are consumed
        vd_rel_destroy(c, p.L->rel);
        vd_rel_destroy(c, p.R->rel);
        p.L->rel = p.R->rel = nullptr;
        n.owner = p.node;
        n.done = true;
    }
}

void BatchExecutor::run(std::vector<BatchQuery> &queries, std::vector<SparqlResult> &results,
                        BatchStats *stats) {
    ContextGuard guard(context);
    BatchStats st;
    std::vector<Query> Qs(queries.size());
    // Flatten each plan in postorder, as postorderTraversal creates jobs.
    for (size_t qi = 0; qi < queries.size(); ++qi) {
        Query &Q = Qs[qi];
        Q.q = &queries[qi];
        std::function<int(PlanTreeNode *)> walk = [&](PlanTreeNode *pn) -> int {
            int a = pn->child1 ? walk(pn->child1) : -1;
            int b = pn->child2 ? walk(pn->child2) : -1;
            Node n;
            n.op = pn->op;
            n.c1 = a;
            n.c2 = b;
            n.var = pn->var;
            n.tp = pn->tp;
            if ((n.op == JOIN && (a < 0 || b < 0)) || (n.op == INDEXSWAP && a < 0) ||
                (n.op == UPLOAD && !n.tp))
                throw VedasError(VD_EINVAL, "malformed plan tree");
            Q.nodes.push_back(n);
            if (n.op == JOIN) Q.counter[n.var] += 1;  // QueryPlan::pushJob
            const int self = (int)Q.nodes.size() - 1;
            if (a >= 0) Q.nodes[a].parent = self;
            if (b >= 0) Q.nodes[b].parent = self;
            return self;
        };
        if (Q.q->root) {
            Q.root = walk(Q.q->root);
        } else {
            if (Q.q->patterns->size() != 1) throw VedasError(VD_EINVAL, "plan required");
            Node n;
            n.op = UPLOAD;
            n.tp = const_cast<TriplePattern *>(&(*Q.q->patterns)[0]);
            n.var = n.tp->getSubject();
            Q.nodes.push_back(n);
            Q.root = 0;
        }
        std::map<std::string, int> cnt;
        for (auto &tp : *Q.q->patterns)
            for (auto t : {tp.getSubject(), tp.getPredicate(), tp.getObject()})
                if (!t.empty() && t[0] == '?') cnt[t]++;
        for (auto &kv : cnt)
            if (kv.second > 1) Q.join_vars.insert(kv.first);
        for (auto &v : Q.q->selected) Q.counter[v] = ULONG_MAX;
    }
    try {
        for (;;) {
            std::vector<std::pair<Query *, int>> sel, join, swp;
            bool pending = false;
            for (auto &Q : Qs)
                for (int i = 0; i < (int)Q.nodes.size(); ++i) {
                    Node &n = Q.nodes[i];
                    if (n.done) continue;
                    pending = true;
                    const bool r1 = n.c1 < 0 || Q.nodes[n.c1].done;
                    const bool r2 = n.c2 < 0 || Q.nodes[n.c2].done;
                    if (!(r1 && r2)) continue;
                    if (n.op == UPLOAD) sel.emplace_back(&Q, i);
                    else if (n.op == JOIN) join.emplace_back(&Q, i);
                    else swp.emplace_back(&Q, i);
                }
            if (!pending) break;
            st.rounds++;
            if (!sel.empty()) {
                selectRound(sel, st);
                continue;
            }
            joinRound(join, swp, st);
        }
        // Finalise (QueryPlan.cu:188-204): drop unused columns i >= 1, then
        // adjacent duplicates.
        std::vector<vd_rel *> finals;
        for (auto &Q : Qs) {
            Rel &r = Q.nodes[Q.nodes[Q.root].owner].out;
            for (int i = (int)r.headers.size() - 1; i >= 1; --i)
                if (Q.counter[r.headers[i]] == 0) {
                    context->check(vd_rel_drop_col(context->raw(), r.rel, i), "vd_rel_drop_col");
                    r.headers.erase(r.headers.begin() + i);
                    for (size_t k = i; k + 1 < VD_MAX_COLS; ++k) r.is_pred[k] = r.is_pred[k + 1];
                }
            finals.push_back(r.rel);
        }
        context->check(vd_unique_adjacent_multi(context->raw(), finals.data(), (int)finals.size()),
                       "vd_unique_adjacent_multi");
        st.launches_calls++;
        results.clear();
        results.resize(Qs.size());
        for (size_t qi = 0; qi < Qs.size(); ++qi) {
            Query &Q = Qs[qi];
            Rel &r = Q.nodes[Q.nodes[Q.root].owner].out;
            auto *ir = new FullRelationIR(context, r.rel);
            r.rel = nullptr;
            for (size_t i = 0; i < r.headers.size(); ++i) ir->setHeader(i, r.headers[i], r.is_pred[i]);
            results[qi].setResult(ir);
        }
    } catch (...) {
        for (auto &Q : Qs)
            for (auto &n : Q.nodes)
                if (n.out.rel) vd_rel_destroy(context->raw(), n.out.rel), n.out.rel = nullptr;
        throw;
    }
    if (stats) *stats = st;
}
