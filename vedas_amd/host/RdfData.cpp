#include "RdfData.h"

RdfData::RdfData(TYPEID_HOST_VEC &s, TYPEID_HOST_VEC &p, TYPEID_HOST_VEC &o)
    : subjects(s), predicates(p), objects(o) {
    if (s.size() != p.size() || s.size() != o.size())
        throw VedasError(VD_EINVAL, "RdfData: column lengths differ");
}
size_t RdfData::size() const { return subjects.size(); }
TYPEID_HOST_VEC &RdfData::getSubject() { return subjects; }
TYPEID_HOST_VEC &RdfData::getPredicate() { return predicates; }
TYPEID_HOST_VEC &RdfData::getObject() { return objects; }
