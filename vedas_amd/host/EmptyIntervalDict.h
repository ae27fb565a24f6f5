// Per-variable "known empty" value intervals (reference
// EmptyIntervalDict.h:6-22).  Pruning only: a select drops rows whose
// leading value lies in an interval that no final binding can take.
#ifndef VEDAS_HOST_EMPTYINTERVALDICT_H
#define VEDAS_HOST_EMPTYINTERVALDICT_H

#include <string>
#include <unordered_map>
#include <vector>

#include "vedas.h"

class EmptyIntervalDict {
  public:
    VAR_BOUND getBound(const std::string &var) const;
    bool hasBoundFor(const std::string &var) const;
    // Merge two new intervals into var's entry (EmptyIntervalDict.cu:33-106).
    void updateBound(const std::string &var, size_t lb1, size_t ub1, size_t lb2, size_t ub2);
    // The value intervals a select on var must skip (the value form of
    // getUploadRelationData, EmptyIntervalDict.cu:108-170).
    std::vector<vd_interval_t> skipIntervals(const std::string &var) const;
    size_t emptySize(const std::string &var) const;

  private:
    std::unordered_map<std::string, VAR_BOUND> dict;
};

// The two widest gaps between consecutive values of a small sorted key
// column, as intervals of missing values (JoinQueryJob.cu:398-433,
// IndexSwapJob.cu:53-90).  col0 holds the whole (sorted) column.
void eifFromSortedColumn(const TYPEID_HOST_VEC &col0, size_t out[4]);

#endif
