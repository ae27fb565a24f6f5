#include "QueryPlan.h"

#include <climits>
#include <iostream>

#include "FullRelationIR.h"
#include "JoinQueryJob.h"

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

// QueryPlan::execute, single-GPU branch (QueryPlan.cu:160-205).
void QueryPlan::execute(SparqlResult &result) {
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
