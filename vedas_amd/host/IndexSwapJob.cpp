#include "IndexSwapJob.h"

#include <chrono>
#include <iostream>

#include "QueryExecutor.h"

IndexSwapJob::IndexSwapJob(QueryJob *beforeJob, std::string swapVar, VedasContext *context,
                           VariableBoundMap *variables_bound, EmptyIntervalDict *ei_dict)
    : context(context), variables_bound(variables_bound), swapVar(swapVar),
      beforeJob(beforeJob), ei_dict(ei_dict) {}

// The relation belongs to the child job; this job only aliases it.
IndexSwapJob::~IndexSwapJob() = default;

IR *IndexSwapJob::getIR() { return intermediateResult; }

IR *IndexSwapJob::releaseIR() { return beforeJob->releaseIR(); }

// IndexSwapJob::startJob, IndexSwapJob.cu:28-95.
int IndexSwapJob::startJob(int gpuId) {
    this->gpuId = gpuId;
    ContextGuard guard(context);
    auto t0 = std::chrono::steady_clock::now();
    auto *ir = dynamic_cast<FullRelationIR *>(beforeJob->getIR());
    if (!ir) throw VedasError(VD_EINVAL, "index swap on a job without a relation");
    const size_t col = ir->getColumnId(swapVar);
    // A driven join already delivered this order (column 0 = swapVar).
    const bool done = col == 0 && ir->sortedOnColumn0();
    ir->swapColumn(0, col);
    // The join that consumes a large relation sorts it after dropping the
    // rows it cannot match (JoinQueryJob::join); such a relation is too
    // large for the empty-interval update below either way.
    const bool defer = !done && joinParent && QueryExecutor::FILTERED_SWAP_MIN_ROWS > 0 &&
                       ir->size() >= QueryExecutor::FILTERED_SWAP_MIN_ROWS &&
                       ir->size() >= EIF_ALPHA;
    if (defer) ir->markPendingSort();
    else ir->sortByFirstColumn();
    context->exe_log.emplace_back(gpuId, SWAP_OP, ir->getHeader(0) + " <-> " + ir->getHeader(col),
                                  ir->size(), ir->getColumnNum());
    if (!done && !defer) context->swap_rows += ir->size();
    context->swap_index_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(
        std::chrono::steady_clock::now() - t0).count();
    if (ir->size() > 2 && ir->size() < EIF_ALPHA && ei_dict) {
        size_t b[4];
        eifFromSortedColumn(ir->download(0), b);
        context->eif_count++;
        ei_dict->updateBound(ir->getHeader(0), b[0], b[1], b[2], b[3]);
    }
    intermediateResult = ir;
    return 0;
}

void IndexSwapJob::print() const { std::cout << "\tINDEX SWAP JOB - to " << swapVar << "\n"; }

std::string IndexSwapJob::jobTypeName() const { return "Index Swap Job"; }
