// Host triple container fed to the index build (reference RdfData.h:7-20).
#ifndef VEDAS_HOST_RDFDATA_H
#define VEDAS_HOST_RDFDATA_H

#include "vedas.h"

class RdfData {
  public:
    RdfData(TYPEID_HOST_VEC &subjects, TYPEID_HOST_VEC &predicates, TYPEID_HOST_VEC &objects);
    size_t size() const;
    TYPEID_HOST_VEC &getSubject();
    TYPEID_HOST_VEC &getPredicate();
    TYPEID_HOST_VEC &getObject();
  private:
    TYPEID_HOST_VEC subjects, predicates, objects;
};

#endif
