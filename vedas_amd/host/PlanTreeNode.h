// Plan tree node (reference PlanTreeNode.h:9-20).
#ifndef VEDAS_HOST_PLANTREENODE_H
#define VEDAS_HOST_PLANTREENODE_H

#include <string>

#include "TriplePattern.h"

enum PlanOperation { UPLOAD, JOIN, INDEXSWAP };

struct PlanTreeNode {
    PlanOperation op;
    PlanTreeNode *child1{nullptr};
    PlanTreeNode *child2{nullptr};
    std::string debugName = "";
    bool reuseVar = false;        // JOIN only
    TriplePattern *tp = nullptr;  // UPLOAD only
    std::string var;              // join var (JOIN), swap var (INDEXSWAP), index var (UPLOAD)
};

#endif
