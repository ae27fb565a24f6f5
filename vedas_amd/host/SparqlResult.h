// Final relation of a query (reference SparqlResult.h); owns it.
#ifndef VEDAS_HOST_SPARQLRESULT_H
#define VEDAS_HOST_SPARQLRESULT_H

#include <memory>

#include "FullRelationIR.h"

class SparqlResult {
  public:
    void setResult(FullRelationIR *ir) { result.reset(ir); }
    FullRelationIR *getResultIR() const { return result.get(); }
    FullRelationIR *release() { return result.release(); }
  private:
    std::unique_ptr<FullRelationIR> result;
};

#endif
