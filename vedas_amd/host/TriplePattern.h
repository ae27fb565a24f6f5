// Triple pattern of a basic graph pattern (reference TriplePattern.h:7-36).
// Terms starting with '?' are variables; other terms are decimal IDs (the
// reference maps IRIs to IDs through its dictionaries first).
#ifndef VEDAS_HOST_TRIPLEPATTERN_H
#define VEDAS_HOST_TRIPLEPATTERN_H

#include <string>

#include "vedas.h"

class TriplePattern {
  public:
    TriplePattern(size_t id, std::string s, std::string p, std::string o);
    size_t getId() const { return id; }
    TYPEID getSubjectId() const { return ids[0]; }
    TYPEID getPredicateId() const { return ids[1]; }
    TYPEID getObjectId() const { return ids[2]; }
    std::string getSubject() const { return terms[0]; }
    std::string getPredicate() const { return terms[1]; }
    std::string getObject() const { return terms[2]; }
    size_t getVariableNum() const { return isVar[0] + isVar[1] + isVar[2]; }
    // isVar[0]*4 + isVar[1]*2 + isVar[2] (TriplePattern.cu)
    size_t getVariableBitmap() const { return isVar[0] * 4 + isVar[1] * 2 + isVar[2]; }
    bool subjectIsVariable() const { return isVar[0]; }
    bool predicateIsVariable() const { return isVar[1]; }
    bool objectIsVariable() const { return isVar[2]; }
    bool hasVariable(const std::string &v) const {
        return (isVar[0] && terms[0] == v) || (isVar[1] && terms[1] == v) ||
               (isVar[2] && terms[2] == v);
    }
    std::string toString() const { return terms[0] + " " + terms[1] + " " + terms[2]; }
    size_t estimate_rows = 0;

  private:
    size_t id;
    std::string terms[3];
    TYPEID ids[3] = {0, 0, 0};
    bool isVar[3] = {false, false, false};
};

#endif
