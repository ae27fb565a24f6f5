#include "QueryPlan.h"

#include <climits>
#include <iostream>
#include <map>
#include <vector>

#include "FullRelationIR.h"
#include "IndexSwapJob.h"
#include "JoinQueryJob.h"
#include "SelectQueryJob.h"

QueryPlan::QueryPlan(const std::set<std::string> &sel) {
    for (auto &v : sel) select_variables.push_back(v);
}

QueryPlan::~QueryPlan() {
    for (auto *j : jobs) delete j;
}

// QueryPlan::pushJob, QueryPlan.cu:34-43: every join counts its variable.
void QueryPlan::pushJob(QueryJob *job) {
    jobs.push_back(job);
    if (auto *jj = dynamic_cast<JoinQueryJob *>(job))
        query_variable_counter[jj->getJoinVariable()] += 1;
}

QueryJob *QueryPlan::getJob(size_t i) { return jobs.at(i); }

// Resolves the row range of every select of the plan with one batched
// lookup per store (the ranges depend only on the pattern constants), so
// that each select job then only filters its range: an unfiltered range is
// a view and needs no device round trip.
static void resolveSelectRanges(const std::vector<QueryJob *> &jobs) {
    std::map<vd_index *, std::vector<SelectQueryJob *>> groups;
    for (auto *job : jobs)
        if (auto *sj = dynamic_cast<SelectQueryJob *>(job)) groups[sj->indexHandle()].push_back(sj);
    for (auto &g : groups) {
        const auto &sel = g.second;
        if (sel.size() < 2) continue;  // one lookup: vd_select_key costs the same
        std::vector<vd_select_job_t> lk;
        for (auto *sj : sel) lk.push_back(sj->lookupJob());
        std::vector<uint64_t> b(sel.size()), e(sel.size());
        VedasContext *ctx = sel[0]->getContext();
        ctx->check(vd_index_lookup_multi(ctx->raw(), g.first, lk.data(), (int)lk.size(),
                                         b.data(), e.data()),
                   "vd_index_lookup_multi");
        for (size_t i = 0; i < sel.size(); ++i) sel[i]->setRange(b[i], e[i]);
    }
}

// QueryPlan::execute, single-GPU branch (QueryPlan.cu:160-205).
void QueryPlan::execute(SparqlResult &result) {
    resolveSelectRanges(jobs);
    // A join whose output an IndexSwapJob re-sorts next learns the swap
    // variable (it may deliver that order itself).
    for (auto *job : jobs)
        if (auto *sw = dynamic_cast<IndexSwapJob *>(job))
            if (auto *jj = dynamic_cast<JoinQueryJob *>(sw->childJob()))
                jj->setSwapHint(sw->swapVariable());
    // An IndexSwapJob whose output feeds a join on its variable leaves a
    // large relation's sort to that join.
    for (auto *job : jobs)
        if (auto *jj = dynamic_cast<JoinQueryJob *>(job))
            for (QueryJob *in : {jj->leftOperand(), jj->rightOperand()})
                if (auto *sw = dynamic_cast<IndexSwapJob *>(in))
                    sw->setJoinParent(sw->swapVariable() == jj->getJoinVariable());
    for (auto &v : select_variables) query_variable_counter[v] = ULONG_MAX;
    for (auto *job : jobs) {
        if (auto *jj = dynamic_cast<JoinQueryJob *>(job)) jj->setQueryVarCounter(&query_variable_counter);
        job->startJob(0);
    }
    if (jobs.empty()) return;
    auto *ir = dynamic_cast<FullRelationIR *>(jobs.back()->releaseIR());
    if (!ir) throw VedasError(VD_EINVAL, "plan produced no relation");
    result.setResult(ir);
    // Drop columns (never column 0) no selected variable needs, then remove
    // adjacent duplicate rows.
    for (int i = (int)ir->getColumnNum() - 1; i >= 1; --i)
        if (query_variable_counter[ir->getHeader(i)] == 0) ir->removeColumn(i);
    ir->removeDuplicate();
}

void QueryPlan::print() const {
    std::cout << "=========================== Query Plan ============================\n";
    for (auto *j : jobs) j->print();
}
