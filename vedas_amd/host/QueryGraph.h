// Query graph and greedy star planner (reference QueryGraph.h / QueryGraph.cu:
// 11-279).  Vertices are the terms of the patterns (variables and
// constants), edges the patterns; the planner repeatedly takes the vertex
// whose incident patterns share a variable most often (fewest estimated
// rows on ties), joins those patterns on it (smallest estimate first,
// IndexSwap where a pattern is indexed on another variable) and merges the
// joined edges into one, until two vertices are left.
//
// Deterministic where the reference iterates unordered_maps: vertices and
// variables are visited in name order, and equal estimates keep pattern
// order.  Any valid plan gives the same result set, so this only fixes
// which of the reference's possible plans is produced.
#ifndef VEDAS_HOST_QUERYGRAPH_H
#define VEDAS_HOST_QUERYGRAPH_H

#include <map>
#include <memory>
#include <set>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "PlanTreeNode.h"
#include "TriplePattern.h"

struct GraphEdge {
    std::string id;         // predicate ID as text
    std::string tmpTriple;  // "s p o" for debug names
    std::set<std::string> vars;
    std::string indexVar1, indexVar2;  // variables the edge's relation is sorted on
    TriplePattern *tp = nullptr;
    unsigned long long estimateCardinality = 0;
    PlanTreeNode *treeNode = nullptr;
    bool hasVariable(const std::string &v) const { return vars.count(v) > 0; }
};

class QueryGraph {
  public:
    // Patterns must carry estimate_rows (rows of their index range).
    explicit QueryGraph(std::vector<TriplePattern> &patterns);
    // The plan root (nullptr for a single pattern); nodes stay owned by
    // the graph.
    PlanTreeNode *generateQueryPlan();
    void print() const;

  private:
    using Link = std::pair<std::string, GraphEdge *>;
    void addEdge(const std::string &source, const std::string &destination, GraphEdge *edge);
    bool edgeExist(const std::string &v1, const std::string &v2);
    void mergeEdge(const std::string joinVar, const std::string end1, const std::string end2);
    std::tuple<size_t, unsigned long long, std::string> sameMaxDegAndTotalCard(
        const std::vector<Link> &links) const;
    std::pair<std::string, std::string> findBestStarNode() const;
    PlanTreeNode *generateJoinPlanForNode(const std::string &vertex, const std::string &joinVar);
    PlanTreeNode *newNode();

    std::map<std::string, std::vector<Link>> adjacencyList;
    std::vector<std::unique_ptr<GraphEdge>> edges_;
    std::vector<std::unique_ptr<PlanTreeNode>> nodes_;
    size_t npatterns_ = 0;
};

#endif
