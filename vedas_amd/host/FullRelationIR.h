// Column-store relation of uint32 IDs (reference FullRelationIR.h:10-47).
// Columns live in HBM behind a vd_rel handle; the header list, predicate
// flags and variable -> column map are host state, as in the reference.
#ifndef VEDAS_HOST_FULLRELATIONIR_H
#define VEDAS_HOST_FULLRELATIONIR_H

#include <ostream>
#include <string>
#include <unordered_map>
#include <vector>

#include "IR.h"
#include "VedasContext.h"
#include "vedas.h"

class FullRelationIR : public IR {
  public:
    FullRelationIR();
    // Allocates on VedasContext::current() (FullRelationIR.cu:19-31).
    FullRelationIR(size_t columnNum, size_t relationSize);
    // Takes ownership of a relation produced by the C ABI.
    FullRelationIR(VedasContext *ctx, vd_rel *rel);
    ~FullRelationIR() override;
    FullRelationIR(const FullRelationIR &) = delete;
    FullRelationIR &operator=(const FullRelationIR &) = delete;

    void sort();              // lexicographic, all columns (:218-336)
    void removeDuplicate();   // adjacent duplicates (:338-522)
    TYPEID *getRelationRawPointer(size_t i);  // device pointer
    void getRelationPointers(TYPEID **relations);
    size_t getColumnNum() const;
    size_t getColumnId(std::string var) const;
    size_t getRelationSize(size_t i) const;
    void setHeader(size_t i, std::string h, bool is_predicate);
    std::string getHeader(size_t i) const;
    std::string getHeaders(std::string delimitor = " ") const;
    bool getIsPredicate(size_t i) const;
    // Host -> device copies into column i (FullRelationIR.cu:578-592).
    void setRelation(size_t i, size_t offset, TYPEID_HOST_VEC::iterator bit,
                     TYPEID_HOST_VEC::iterator eit);
    void setRelation(size_t i, TYPEID_HOST_VEC::iterator bit, TYPEID_HOST_VEC::iterator eit);
    // Device -> device copies into column i.
    void setRelation(size_t i, size_t offset, const TYPEID *dbegin, const TYPEID *dend);
    void sortByFirstColumn();                 // (:594-713)
    void swapColumn(size_t i, size_t j);      // (:715-727)
    void removeColumn(size_t i);              // (:729-737)
    void removeColumn(size_t i, std::string &maintain_var);  // (:739-745)
    void print() const;
    void print(bool full_version, std::ostream &out) const;
    size_t size() const;

    TYPEID_HOST_VEC download(size_t i) const;  // one column to the host
    vd_rel *raw() const { return rel_; }
    VedasContext *context() const { return ctx_; }
    // The producer delivered the rows already stably sorted on column 0 (a
    // join driven in IndexSwap order): sortByFirstColumn has nothing to do
    // until column 0 changes.
    void markSortedOnColumn0() { sorted0_ = true; }
    bool sortedOnColumn0() const { return sorted0_; }
    // An IndexSwapJob moved its variable to column 0 and left the sort to
    // the join that consumes the relation (which may first drop the rows
    // the join cannot match): sortByFirstColumn clears it.
    void markPendingSort() { pending_ = true; }
    bool pendingSort() const { return pending_; }
    // Takes ownership of `rel` (same columns, e.g. a subset of the rows)
    // in place of the current relation; headers stay.
    void replaceRelation(vd_rel *rel);

  private:
    bool sorted0_ = false;
    bool pending_ = false;
    VedasContext *ctx_ = nullptr;
    vd_rel *rel_ = nullptr;
    std::vector<bool> is_predicates;
    std::vector<std::string> headers;
    std::unordered_map<std::string, size_t> var_column_map;
};

#endif
