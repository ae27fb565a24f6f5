// Job list of one query and its single-GPU execution (reference
// QueryPlan.h / QueryPlan.cu:20-205).  The 2-GPU branch is not built here.
#ifndef VEDAS_HOST_QUERYPLAN_H
#define VEDAS_HOST_QUERYPLAN_H

#include <map>
#include <set>
#include <string>
#include <vector>

#include "QueryJob.h"
#include "SparqlResult.h"

class QueryPlan {
  public:
    explicit QueryPlan(const std::set<std::string> &select_variable_set);
    ~QueryPlan();
    void pushJob(QueryJob *job);
    QueryJob *getJob(size_t i);
    size_t size() const { return jobs.size(); }
    void execute(SparqlResult &sparqlResult);
    void print() const;

  private:
    std::map<std::string, size_t> query_variable_counter;
    std::vector<QueryJob *> jobs;
    std::vector<std::string> select_variables;
};

#endif
