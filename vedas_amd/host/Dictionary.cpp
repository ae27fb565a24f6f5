#include "Dictionary.h"

#include <array>
#include <cctype>
#include <fstream>
#include <map>
#include <sstream>

// ---- dictionary -------------------------------------------------------------
TYPEID Dictionary::number(std::unordered_map<std::string, TYPEID> &map,
                          std::vector<std::string> &rev, TYPEID first, const std::string &t) {
    auto it = map.find(t);
    if (it != map.end()) return it->second;
    const TYPEID id = first + (TYPEID)rev.size();
    map.emplace(t, id);
    rev.push_back(t);
    return id;
}

// triple_statement_handler, loader.cu:18-53: subject, then object (literal
// dictionary or the shared one), then predicate.
void Dictionary::addTriple(const std::string &s, const std::string &p, const std::string &o) {
    if (s.empty() || s[0] == '"') throw VedasError(VD_EINVAL, "literal subject: " + s);
    const TYPEID sid = number(so_map_, so_rev_, 1, s);
    TYPEID oid;
    if (o[0] == '"' && ldict_)
        oid = number(l_map_, l_rev_, (TYPEID)LITERAL_START_ID, o);
    else
        oid = number(so_map_, so_rev_, 1, o);
    const TYPEID pid = number(p_map_, p_rev_, 1, p);
    s_.push_back(sid);
    pv_.push_back(pid);
    o_.push_back(oid);
}

TYPEID Dictionary::subjectObjectId(const std::string &t) const {
    const auto &m = (ldict_ && !t.empty() && t[0] == '"') ? l_map_ : so_map_;
    auto it = m.find(t);
    return it == m.end() ? 0 : it->second;
}

TYPEID Dictionary::predicateId(const std::string &t) const {
    auto it = p_map_.find(t);
    return it == p_map_.end() ? 0 : it->second;
}

// SparqlResult::print's decoding, SparqlResult.cu:96-104.
std::string Dictionary::term(TYPEID id, bool is_predicate) const {
    if (is_predicate) return (id >= 1 && id <= p_rev_.size()) ? p_rev_[id - 1] : std::string();
    if (ldict_ && id >= (TYPEID)LITERAL_START_ID) {
        const size_t k = id - (TYPEID)LITERAL_START_ID;
        return k < l_rev_.size() ? l_rev_[k] : std::string();
    }
    return (id >= 1 && id <= so_rev_.size()) ? so_rev_[id - 1] : std::string();
}

// ---- .vdd file ----------------------------------------------------------------
namespace {

void save_section(std::ofstream &out, const std::vector<std::string> &rev, TYPEID first) {
    size_t count = 0;
    for (auto &t : rev) count += !t.empty();
    out << count << "\n";
    for (size_t i = 0; i < rev.size(); ++i)
        if (!rev[i].empty()) out << (first + (TYPEID)i) << " " << rev[i] << "\n";
}

// One section: the count line, then "id term" lines; an ID at or above
// `first` goes to rev[id - first].  Returns false at the end of the file
// before the count (an optional section).
bool load_section(std::ifstream &in, std::unordered_map<std::string, TYPEID> &map,
                  std::vector<std::string> &rev, TYPEID first, bool optional, const char *what) {
    std::string line;
    if (!std::getline(in, line)) {
        if (optional) return false;
        throw VedasError(VD_EINVAL, std::string(".vdd: missing ") + what + " count");
    }
    size_t count = 0;
    try {
        count = std::stoull(line);
    } catch (const std::exception &) {
        throw VedasError(VD_EINVAL, std::string(".vdd: bad ") + what + " count: " + line);
    }
    for (size_t k = 0; k < count; ++k) {
        if (!std::getline(in, line))
            throw VedasError(VD_EINVAL, std::string(".vdd: ") + what + " section truncated");
        const size_t sp = line.find(' ');
        unsigned long long id = 0;
        try {
            id = std::stoull(line.substr(0, sp));
        } catch (const std::exception &) {
            throw VedasError(VD_EINVAL, std::string(".vdd: bad ") + what + " line: " + line);
        }
        if (sp == std::string::npos || id < first || id > 0xFFFFFFFFull ||
            id - first > (64ull << 20) + count)
            throw VedasError(VD_EINVAL, std::string(".vdd: bad ") + what + " line: " + line);
        const std::string term = line.substr(sp + 1);
        const size_t slot = (size_t)(id - first);
        if (slot >= rev.size()) rev.resize(slot + 1);
        rev[slot] = term;
        map[term] = (TYPEID)id;
    }
    return true;
}

}  // namespace

void Dictionary::save(const char *path) const {
    std::ofstream out(path, std::ios::out | std::ios::trunc);
    if (!out) throw VedasError(VD_EINVAL, std::string("cannot open ") + path);
    save_section(out, so_rev_, 1);
    save_section(out, p_rev_, 1);
    save_section(out, l_rev_, (TYPEID)LITERAL_START_ID);
    if (!out) throw VedasError(VD_EINVAL, std::string("write failed: ") + path);
}

void Dictionary::open(const char *path) {
    std::ifstream in(path);
    if (!in) throw VedasError(VD_EINVAL, std::string("cannot open ") + path);
    Dictionary d(false);
    load_section(in, d.so_map_, d.so_rev_, 1, false, "subject/object");
    load_section(in, d.p_map_, d.p_rev_, 1, false, "predicate");
    load_section(in, d.l_map_, d.l_rev_, (TYPEID)LITERAL_START_ID, true, "literal");
    d.ldict_ = !d.l_map_.empty() || ldict_;
    *this = std::move(d);
}

// ---- N-Triples ---------------------------------------------------------------
namespace {

bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

// One N-Triples term starting at line[i]: <iri>, _:label or "literal" with
// an optional @lang or ^^<iri>; returned as written.
std::string nt_term(const std::string &line, size_t &i, size_t lineno) {
    auto bad = [&](const char *what) {
        throw VedasError(VD_EINVAL, "N-Triples line " + std::to_string(lineno) + ": " + what);
    };
    const size_t b = i;
    if (i >= line.size()) bad("missing term");
    if (line[i] == '<') {
        const size_t e = line.find('>', i + 1);
        if (e == std::string::npos) bad("unterminated IRI");
        i = e + 1;
    } else if (line.compare(i, 2, "_:") == 0) {
        i += 2;
        while (i < line.size() && !is_ws(line[i]) && line[i] != '<' && line[i] != '"') ++i;
        // a label does not end with '.': that dot ends the statement
        while (i > b + 2 && line[i - 1] == '.') --i;
    } else if (line[i] == '"') {
        ++i;
        while (i < line.size() && line[i] != '"') i += (line[i] == '\\') ? 2 : 1;
        if (i >= line.size()) bad("unterminated literal");
        ++i;
        if (i < line.size() && line[i] == '@') {
            ++i;
            while (i < line.size() && (std::isalnum((unsigned char)line[i]) || line[i] == '-')) ++i;
        } else if (line.compare(i, 2, "^^") == 0) {
            i += 2;
            if (i >= line.size() || line[i] != '<') bad("datatype IRI expected");
            const size_t e = line.find('>', i + 1);
            if (e == std::string::npos) bad("unterminated datatype IRI");
            i = e + 1;
        }
    } else {
        bad("unexpected character");
    }
    return line.substr(b, i - b);
}

void skip_ws(const std::string &line, size_t &i) {
    while (i < line.size() && is_ws(line[i])) ++i;
}

}  // namespace

void Dictionary::loadNTriplesText(const std::string &text) {
    std::istringstream in(text);
    std::string line;
    size_t lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        size_t i = 0;
        skip_ws(line, i);
        if (i >= line.size() || line[i] == '#') continue;
        const std::string s = nt_term(line, i, lineno);
        skip_ws(line, i);
        const std::string p = nt_term(line, i, lineno);
        skip_ws(line, i);
        const std::string o = nt_term(line, i, lineno);
        skip_ws(line, i);
        if (i >= line.size() || line[i] != '.')
            throw VedasError(VD_EINVAL, "N-Triples line " + std::to_string(lineno) + ": '.' expected");
        ++i;
        skip_ws(line, i);
        if (i < line.size() && line[i] != '#')
            throw VedasError(VD_EINVAL, "N-Triples line " + std::to_string(lineno) + ": text after '.'");
        if (p[0] != '<') throw VedasError(VD_EINVAL, "N-Triples line " + std::to_string(lineno) + ": predicate must be an IRI");
        addTriple(s, p, o);
    }
}

void Dictionary::loadNTriples(const char *path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw VedasError(VD_EINVAL, std::string("cannot open ") + path);
    std::stringstream ss;
    ss << in.rdbuf();
    loadNTriplesText(ss.str());
}

// ---- SPARQL -----------------------------------------------------------------
namespace {

struct Lexer {
    const std::string &s;
    size_t i = 0;
    explicit Lexer(const std::string &t) : s(t) {}

    void skip() {
        for (;;) {
            while (i < s.size() && is_ws(s[i])) ++i;
            if (i < s.size() && s[i] == '#') {
                while (i < s.size() && s[i] != '\n') ++i;
                continue;
            }
            return;
        }
    }
    bool done() {
        skip();
        return i >= s.size();
    }
    [[noreturn]] void bad(const std::string &what) const {
        throw VedasError(VD_EINVAL, "SPARQL: " + what + " at offset " + std::to_string(i));
    }
    // Next token: punctuation, <iri>, "literal"(@lang|^^type)?, ?var, or a
    // word (keyword, prefixed name, 'a').
    std::string next() {
        skip();
        if (i >= s.size()) bad("unexpected end");
        const size_t b = i;
        const char c = s[i];
        if (c == '{' || c == '}' || c == '.' || c == ';' || c == ',' || c == '*') {
            ++i;
            return s.substr(b, 1);
        }
        if (c == '<') {
            const size_t e = s.find('>', i + 1);
            if (e == std::string::npos) bad("unterminated IRI");
            i = e + 1;
            return s.substr(b, i - b);
        }
        if (c == '"') {
            ++i;
            while (i < s.size() && s[i] != '"') i += (s[i] == '\\') ? 2 : 1;
            if (i >= s.size()) bad("unterminated literal");
            ++i;
            if (i < s.size() && s[i] == '@') {
                ++i;
                while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '-')) ++i;
            } else if (s.compare(i, 2, "^^") == 0) {
                i += 2;
                const std::string dt = next();  // <iri> or prefixed name, resolved later
                return s.substr(b, i - b - dt.size()) + dt;
            }
            return s.substr(b, i - b);
        }
        if (c == '?' || c == '$') {
            ++i;
            while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_')) ++i;
            if (i == b + 1) bad("empty variable name");
            return "?" + s.substr(b + 1, i - b - 1);
        }
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == ':' ||
                                s[i] == '-' || s[i] == '.'))
            ++i;
        while (i > b && s[i - 1] == '.') --i;  // a name does not end with '.'
        if (i == b) bad(std::string("unexpected character '") + c + "'");
        return s.substr(b, i - b);
    }
    std::string peek() {
        const size_t save = i;
        std::string t = next();
        i = save;
        return t;
    }
};

std::string upper(std::string t) {
    for (auto &ch : t) ch = (char)std::toupper((unsigned char)ch);
    return t;
}

}  // namespace

SparqlBgp parseSparql(const std::string &text) {
    Lexer lx(text);
    std::map<std::string, std::string> prefixes;
    SparqlBgp q;
    auto resolve = [&](const std::string &t) -> std::string {
        if (t.empty()) lx.bad("empty term");
        if (t[0] == '<' || t[0] == '?') return t;
        if (t == "a") return "<http://www.w3.org/1999/02/22-rdf-syntax-ns#type>";
        if (t[0] == '"') {  // resolve a prefixed datatype
            const size_t k = t.rfind("^^");
            if (k != std::string::npos && t.size() > k + 2 && t[k + 2] != '<') {
                const std::string dt = t.substr(k + 2);
                const size_t c = dt.find(':');
                auto it = prefixes.find(dt.substr(0, c));
                if (c == std::string::npos || it == prefixes.end()) lx.bad("unknown prefix in " + dt);
                return t.substr(0, k + 2) + "<" + it->second + dt.substr(c + 1) + ">";
            }
            return t;
        }
        const size_t c = t.find(':');
        if (c == std::string::npos) lx.bad("unexpected word '" + t + "'");
        auto it = prefixes.find(t.substr(0, c));
        if (it == prefixes.end()) lx.bad("unknown prefix '" + t.substr(0, c) + "'");
        return "<" + it->second + t.substr(c + 1) + ">";
    };
    while (upper(lx.peek()) == "PREFIX") {
        lx.next();
        std::string name = lx.next();
        if (name.empty() || name.back() != ':') lx.bad("prefix name expected");
        name.pop_back();
        const std::string iri = lx.next();
        if (iri.size() < 2 || iri[0] != '<') lx.bad("prefix IRI expected");
        prefixes[name] = iri.substr(1, iri.size() - 2);
    }
    if (upper(lx.next()) != "SELECT") lx.bad("SELECT expected");
    std::string t = lx.next();
    if (upper(t) == "DISTINCT" || upper(t) == "REDUCED") t = lx.next();
    if (t == "*") {
        t = lx.next();
    } else {
        while (!t.empty() && t[0] == '?') {
            q.select.push_back(t);
            t = lx.next();
        }
        if (q.select.empty()) lx.bad("variables or * expected after SELECT");
    }
    if (upper(t) == "WHERE") t = lx.next();
    if (t != "{") lx.bad("'{' expected");
    for (;;) {
        t = lx.next();
        if (t == "}") break;
        if (t == ".") continue;
        const std::string subj = resolve(t);
        for (;;) {  // predicate-object list
            const std::string pred = resolve(lx.next());
            for (;;) {  // object list
                const std::string obj = resolve(lx.next());
                q.patterns.push_back({subj, pred, obj});
                if (lx.peek() != ",") break;
                lx.next();
            }
            if (lx.peek() != ";") break;
            lx.next();
            const std::string after = lx.peek();
            if (after == "." || after == "}") break;
        }
    }
    if (!lx.done()) lx.bad("text after the graph pattern (only basic graph patterns are read)");
    if (q.patterns.empty()) lx.bad("empty graph pattern");
    return q;
}

std::string translateSparql(const SparqlBgp &q, const Dictionary &d) {
    std::string out = "SELECT";
    if (q.select.empty()) out += " *";
    for (auto &v : q.select) out += " " + v;
    out += " WHERE {";
    for (size_t k = 0; k < q.patterns.size(); ++k) {
        const auto &tp = q.patterns[k];
        auto so = [&](const std::string &t) {
            return t[0] == '?' ? t : std::to_string(d.subjectObjectId(t));
        };
        const std::string p = tp[1][0] == '?' ? tp[1] : std::to_string(d.predicateId(tp[1]));
        out += (k ? " . " : " ") + so(tp[0]) + " " + p + " " + so(tp[2]);
    }
    out += " }";
    return out;
}
