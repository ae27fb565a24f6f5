#include "QueryGraph.h"

#include <algorithm>
#include <cmath>
#include <iostream>

#include "vedas.h"

// QueryGraph::QueryGraph, QueryGraph.cu:38-62: one edge per pattern between
// its subject and object terms (constants by their ID), with the pattern's
// variables and estimate rows x variables.
QueryGraph::QueryGraph(std::vector<TriplePattern> &patterns) : npatterns_(patterns.size()) {
    for (auto &tp : patterns) {
        const std::string s = tp.subjectIsVariable() ? tp.getSubject() : std::to_string(tp.getSubjectId());
        const std::string o = tp.objectIsVariable() ? tp.getObject() : std::to_string(tp.getObjectId());
        edges_.emplace_back(new GraphEdge());
        GraphEdge *e = edges_.back().get();
        // GraphEdge::initialize, QueryGraph.cu:19-31
        e->id = std::to_string(tp.getPredicateId());
        e->tmpTriple = s + " " + e->id + " " + o;
        if (s[0] == '?') {
            e->vars.insert(s);
            e->indexVar1 = s;
        }
        if (o[0] == '?') {
            e->vars.insert(o);
            if (e->indexVar1.empty()) e->indexVar1 = o; else e->indexVar2 = o;
        }
        e->tp = &tp;
        e->estimateCardinality = (unsigned long long)tp.estimate_rows * tp.getVariableNum();
        // Two patterns over the same two terms need a 2-variable join,
        // which this planner (like the reference's) does not produce.
        if (edgeExist(s, o))
            throw VedasError(VD_EINVAL, "patterns sharing two terms need an explicit PLAN");
        addEdge(s, o, e);
    }
}

PlanTreeNode *QueryGraph::newNode() {
    nodes_.emplace_back(new PlanTreeNode());
    return nodes_.back().get();
}

// QueryGraph::generateQueryPlan, :64-74.
PlanTreeNode *QueryGraph::generateQueryPlan() {
    PlanTreeNode *root = nullptr;
    size_t rounds = 0;
    while (adjacencyList.size() > 2) {
        // Every round merges at least one edge into another; more rounds
        // than patterns means the graph has a shape the greedy rule cannot
        // reduce (a cycle or a disconnected part).
        if (++rounds > npatterns_ + 1)
            throw VedasError(VD_EINVAL, "query graph cannot be planned (cycle or disconnected)");
        auto star = findBestStarNode();
        root = generateJoinPlanForNode(star.first, star.second);
    }
    return root;
}

void QueryGraph::print() const {
    std::cout << "Query Graph\n";
    for (auto &src : adjacencyList) {
        std::cout << '\t' << src.first << " => ";
        for (auto &dst : src.second) std::cout << dst.first << '(' << dst.second->id << ") ";
        std::cout << '\n';
    }
}

// :87-90
void QueryGraph::addEdge(const std::string &source, const std::string &destination,
                         GraphEdge *edge) {
    adjacencyList[source].push_back(std::make_pair(destination, edge));
    adjacencyList[destination].push_back(std::make_pair(source, edge));
}

// :92-96
bool QueryGraph::edgeExist(const std::string &v1, const std::string &v2) {
    for (auto &edge : adjacencyList[v1])
        if (edge.first == v2) return true;
    return false;
}

// QueryGraph::mergeEdge, :98-118: move the edges of end2 to end1, drop the
// links back to end2, then end2 itself.
// The names are taken by value: end2 may name a link that this erases.
void QueryGraph::mergeEdge(const std::string joinVar, const std::string end1,
                           const std::string end2) {
    (void)joinVar;
    const std::vector<Link> links = adjacencyList[end2];
    for (auto &edge : links) {
        if (!edgeExist(end1, edge.first)) addEdge(end1, edge.first, edge.second);
        auto &other = adjacencyList[edge.first];
        for (auto it = other.rbegin(); it != other.rend(); ++it) {
            if (it->first == end2) {
                other.erase(std::next(it).base());
                break;
            }
        }
    }
    adjacencyList.erase(end2);
}

// QueryGraph::sameMaxDegAndTotalCard, :127-155: the variable shared by the
// most links (fewest summed estimates among equals) and that count.
std::tuple<size_t, unsigned long long, std::string> QueryGraph::sameMaxDegAndTotalCard(
    const std::vector<Link> &links) const {
    std::map<std::string, size_t> varCount;
    std::map<std::string, unsigned long long> varEstCard;
    for (auto &node : links)
        for (auto &var : node.second->vars) {
            varCount[var] += 1;
            varEstCard[var] += node.second->estimateCardinality;
        }
    size_t maxDeg = 0;
    unsigned long long maxDegMinCard = 0;
    std::string bestVar;
    for (auto &v : varCount) {
        const unsigned long long estCard = varEstCard[v.first];
        if (v.second > maxDeg || (v.second == maxDeg && estCard < maxDegMinCard)) {
            maxDeg = v.second;
            maxDegMinCard = estCard;
            bestVar = v.first;
        }
    }
    return std::make_tuple(maxDeg, maxDegMinCard, bestVar);
}

// QueryGraph::findBestStarNode, :157-177.
std::pair<std::string, std::string> QueryGraph::findBestStarNode() const {
    std::string bestNodeName, bestNodeJoinVar;
    size_t maxDegree = 0;
    unsigned long long maxDegCardSum = 0;
    for (auto &node : adjacencyList) {
        auto vdc = sameMaxDegAndTotalCard(node.second);
        const size_t degree = std::get<0>(vdc);
        const unsigned long long totalEstCard = std::get<1>(vdc);
        if (degree > maxDegree || (degree == maxDegree && totalEstCard < maxDegCardSum)) {
            maxDegree = degree;
            bestNodeName = node.first;
            maxDegCardSum = totalEstCard;
            bestNodeJoinVar = std::get<2>(vdc);
        }
    }
    return std::make_pair(bestNodeName, bestNodeJoinVar);
}

// QueryGraph::generateJoinPlanForNode, :219-279: join the vertex's links
// that carry joinVar, smallest estimate first, into the first of them.
PlanTreeNode *QueryGraph::generateJoinPlanForNode(const std::string &vertex,
                                                  const std::string &joinVar) {
    auto &adjList = adjacencyList[vertex];
    std::stable_sort(adjList.begin(), adjList.end(), [](const Link &a, const Link &b) {
        return a.second->estimateCardinality < b.second->estimateCardinality;
    });
    size_t firstIndex = 0;
    while (firstIndex < adjList.size() && !adjList[firstIndex].second->hasVariable(joinVar))
        ++firstIndex;
    if (firstIndex == adjList.size()) throw VedasError(VD_EINVAL, "no pattern carries " + joinVar);
    const std::string indexVar = joinVar;
    size_t lastJoinIdx = 0;
    for (size_t i = firstIndex; i < adjList.size(); ++i)
        if (adjList[i].second->hasVariable(joinVar)) lastJoinIdx = i;

    for (size_t i = firstIndex; i < adjList.size(); ++i) {
        if (!adjList[i].second->hasVariable(joinVar)) continue;
        if (!adjList[i].second->treeNode) {  // createUploadJobNode, :187-194
            PlanTreeNode *u = newNode();
            u->op = UPLOAD;
            u->tp = adjList[i].second->tp;
            u->debugName = adjList[i].second->tmpTriple;
            u->var = indexVar;
            adjList[i].second->treeNode = u;
        }
        if (i > firstIndex) {
            GraphEdge *e1 = adjList[firstIndex].second;
            GraphEdge *e2 = adjList[i].second;
            if (e1->indexVar1 != indexVar && e1->indexVar2 != indexVar) {
                PlanTreeNode *sw = newNode();  // createIndexSwapJobNode, :208-215
                sw->op = INDEXSWAP;
                sw->child1 = e1->treeNode;
                sw->debugName = indexVar;
                sw->var = indexVar;
                e1->treeNode = sw;
                e1->indexVar1 = indexVar;
                e1->indexVar2 = "";
            }
            if (e2->indexVar1 != indexVar && e2->indexVar2 != indexVar) {
                PlanTreeNode *sw = newNode();
                sw->op = INDEXSWAP;
                sw->child1 = e2->treeNode;
                sw->debugName = indexVar;
                sw->var = indexVar;
                e2->treeNode = sw;
            }
            PlanTreeNode *j = newNode();  // createJoinJobNode, :196-206
            j->op = JOIN;
            j->child1 = e1->treeNode;
            j->child2 = e2->treeNode;
            j->debugName = indexVar;
            j->var = indexVar;
            j->reuseVar = i != lastJoinIdx;
            e1->treeNode = j;
            double sigma = 0.000001;  // the reference's join selectivity guess
            if (e1->estimateCardinality < 1000 && e2->estimateCardinality < 1000)
                sigma = 0.001;
            else if (e1->estimateCardinality < 1000 || e2->estimateCardinality < 1000)
                sigma = 0.0001;
            e1->estimateCardinality *=
                (unsigned long long)std::ceil((double)e2->estimateCardinality * sigma);
            for (auto &v : e2->vars) e1->vars.insert(v);
        }
    }
    const std::string keep = adjList[firstIndex].first;
    GraphEdge *kept = adjList[firstIndex].second;
    for (size_t i = adjList.size(); i-- > firstIndex + 1;) {
        if (!adjList[i].second->hasVariable(joinVar)) continue;
        mergeEdge(vertex, keep, adjList[i].first);
    }
    kept->indexVar1 = indexVar;
    kept->indexVar2 = "";
    return kept->treeNode;
}
