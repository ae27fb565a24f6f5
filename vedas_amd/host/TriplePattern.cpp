#include "TriplePattern.h"

#include <cstdlib>

TriplePattern::TriplePattern(size_t id, std::string s, std::string p, std::string o) : id(id) {
    terms[0] = s;
    terms[1] = p;
    terms[2] = o;
    for (int i = 0; i < 3; ++i) {
        if (terms[i].empty()) throw VedasError(VD_EINVAL, "empty pattern term");
        isVar[i] = terms[i][0] == '?';
        if (!isVar[i]) {
            char *end = nullptr;
            unsigned long v = std::strtoul(terms[i].c_str(), &end, 10);
            if (!end || *end != '\0' || v > 0xFFFFFFFFul)
                throw VedasError(VD_EINVAL, "constant term is not a 32-bit ID: " + terms[i]);
            ids[i] = (TYPEID)v;
        }
    }
}
