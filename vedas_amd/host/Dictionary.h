// Term dictionaries and the N-Triples / SPARQL front-ends of the query
// path (reference loader.cu:12-93, SparqlQuery.cu:29-124, TriplePattern.cu:
// 11-47, SparqlResult.cu:96-104).  raptor2 and rasqal are not used: both
// formats are read by small parsers of their own here.
//
// IDs: subjects and objects share one space from 1, predicates have their
// own from 1, and -- with the literal dictionary on -- literal objects have
// their own from LITERAL_START_ID (5e7).  IDs follow first appearance in the
// file; within a triple the subject is numbered before the object and the
// object before the predicate (the order of the reference's statement
// handler).
#ifndef VEDAS_HOST_DICTIONARY_H
#define VEDAS_HOST_DICTIONARY_H

#include <array>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "vedas.h"

class Dictionary {
  public:
    explicit Dictionary(bool literal_dict) : ldict_(literal_dict) {}
    bool literalDict() const { return ldict_; }

    // Number a triple's terms (loading).
    void addTriple(const std::string &s, const std::string &p, const std::string &o);
    // ID of a query constant; 0 when the term is not in the store (it then
    // matches nothing, as the reference's dictionary lookup gives 0).
    TYPEID subjectObjectId(const std::string &t) const;
    TYPEID predicateId(const std::string &t) const;
    // Term of an ID of a result column (predicate columns use the predicate
    // dictionary); empty when the ID is unknown.
    std::string term(TYPEID id, bool is_predicate) const;

    // The .vdd dictionary file (save_dict, VedasBuild.cu:145-160; read by
    // load_dict, VedasQuery.cu:213-268): for the subject/object, predicate
    // and literal dictionaries in turn, a line with the term count, then one
    // "id term" line per term (here in ID order).  open replaces the terms
    // (the literal dictionary is on when the file has literal terms) and
    // clears the triples; the literal section may be missing.
    void save(const char *path) const;
    void open(const char *path);

    // Parse an N-Triples file and number its terms; the triples are appended.
    void loadNTriples(const char *path);
    void loadNTriplesText(const std::string &text);
    const TYPEID_HOST_VEC &subjects() const { return s_; }
    const TYPEID_HOST_VEC &predicates() const { return pv_; }
    const TYPEID_HOST_VEC &objects() const { return o_; }
    size_t soCount() const { return so_rev_.size(); }
    size_t predicateCount() const { return p_rev_.size(); }
    size_t literalCount() const { return l_rev_.size(); }

  private:
    static TYPEID number(std::unordered_map<std::string, TYPEID> &map,
                         std::vector<std::string> &rev, TYPEID first, const std::string &t);
    bool ldict_;
    std::unordered_map<std::string, TYPEID> so_map_, p_map_, l_map_;
    std::vector<std::string> so_rev_, p_rev_, l_rev_;
    TYPEID_HOST_VEC s_, pv_, o_;
};

// A SPARQL SELECT query over a basic graph pattern: PREFIX declarations,
// SELECT [DISTINCT] (variables | *), WHERE { triples }, with ';' and ','
// abbreviations, 'a' for rdf:type, IRIs, prefixed names, literals (with
// @lang or ^^type) and variables.  Terms come out in N-Triples form (IRIs
// as <...>, literals as written) and variables as ?name.
struct SparqlBgp {
    std::vector<std::string> select;  // empty = SELECT *
    std::vector<std::array<std::string, 3>> patterns;
};
SparqlBgp parseSparql(const std::string &text);

// The query in the engine's own text form (decimal IDs), terms mapped by
// the dictionary; no PLAN, so the query graph planner plans it.
std::string translateSparql(const SparqlBgp &q, const Dictionary &d);

#endif
