#include "FullRelationIR.h"

#include <chrono>
#include <iostream>

namespace {
VedasContext *need_current() {
    VedasContext *c = VedasContext::current();
    if (!c) throw VedasError(VD_EINVAL, "no current VedasContext");
    return c;
}
}  // namespace

FullRelationIR::FullRelationIR() = default;

FullRelationIR::FullRelationIR(size_t columnNum, size_t relationSize)
    : ctx_(need_current()) {
    headers.resize(columnNum);
    is_predicates.resize(columnNum);
    if (columnNum > 0)
        ctx_->check(vd_rel_create(ctx_->raw(), (int)columnNum, relationSize, &rel_),
                    "vd_rel_create");
}

FullRelationIR::FullRelationIR(VedasContext *ctx, vd_rel *rel) : ctx_(ctx), rel_(rel) {
    const size_t nc = rel ? (size_t)vd_rel_ncols(rel) : 0;
    headers.resize(nc);
    is_predicates.resize(nc);
}

FullRelationIR::~FullRelationIR() {
    if (rel_ && ctx_) vd_rel_destroy(ctx_->raw(), rel_);
}

void FullRelationIR::sort() {
    if (!rel_) return;
    ctx_->check(vd_sort_lex(ctx_->raw(), rel_), "vd_sort_lex");
}

void FullRelationIR::removeDuplicate() {
    if (!rel_) return;
    auto t0 = std::chrono::steady_clock::now();
    ctx_->check(vd_unique_adjacent(ctx_->raw(), rel_), "vd_unique_adjacent");
    ctx_->eliminate_duplicate_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(
        std::chrono::steady_clock::now() - t0).count();
}

TYPEID *FullRelationIR::getRelationRawPointer(size_t i) {
    return rel_ ? vd_rel_col(rel_, (int)i) : nullptr;
}

void FullRelationIR::getRelationPointers(TYPEID **relations) {
    for (size_t i = 0; i < getColumnNum(); ++i) relations[i] = getRelationRawPointer(i);
}

size_t FullRelationIR::getColumnNum() const { return headers.size(); }

size_t FullRelationIR::getColumnId(std::string var) const {
    auto it = var_column_map.find(var);
    if (it == var_column_map.end()) throw VedasError(VD_EINVAL, "no column for " + var);
    return it->second;
}

size_t FullRelationIR::getRelationSize(size_t) const { return size(); }

void FullRelationIR::setHeader(size_t i, std::string h, bool is_predicate) {
    headers.at(i) = h;
    var_column_map[h] = i;
    is_predicates.at(i) = is_predicate;
}

std::string FullRelationIR::getHeader(size_t i) const { return headers.at(i); }

std::string FullRelationIR::getHeaders(std::string delimitor) const {
    if (headers.empty()) return "";
    std::string s = headers[0];
    for (size_t i = 1; i < headers.size(); ++i) s += delimitor + headers[i];
    return s;
}

bool FullRelationIR::getIsPredicate(size_t i) const { return is_predicates.at(i); }

void FullRelationIR::setRelation(size_t i, size_t offset, TYPEID_HOST_VEC::iterator bit,
                                 TYPEID_HOST_VEC::iterator eit) {
    const size_t n = eit - bit;
    if (n == 0) return;
    if (offset + n > size()) throw VedasError(VD_EINVAL, "setRelation out of range");
    if (i == 0) sorted0_ = false;
    ctx_->check(vd_rel_materialize(ctx_->raw(), rel_), "vd_rel_materialize");
    ctx_->check(vd_copy_h2d(ctx_->raw(), getRelationRawPointer(i) + offset, &*bit,
                            n * sizeof(TYPEID)), "vd_copy_h2d");
}

void FullRelationIR::setRelation(size_t i, TYPEID_HOST_VEC::iterator bit,
                                 TYPEID_HOST_VEC::iterator eit) {
    setRelation(i, 0, bit, eit);
}

void FullRelationIR::setRelation(size_t i, size_t offset, const TYPEID *dbegin,
                                 const TYPEID *dend) {
    const size_t n = dend - dbegin;
    if (n == 0) return;
    if (offset + n > size()) throw VedasError(VD_EINVAL, "setRelation out of range");
    if (i == 0) sorted0_ = false;
    ctx_->check(vd_rel_materialize(ctx_->raw(), rel_), "vd_rel_materialize");
    ctx_->check(vd_copy_d2d(ctx_->raw(), getRelationRawPointer(i) + offset, dbegin,
                            n * sizeof(TYPEID)), "vd_copy_d2d");
}

void FullRelationIR::sortByFirstColumn() {
    pending_ = false;
    if (!rel_ || sorted0_) return;
    ctx_->check(vd_sort_col0(ctx_->raw(), rel_), "vd_sort_col0");
}

void FullRelationIR::replaceRelation(vd_rel *rel) {
    if (!rel || vd_rel_ncols(rel) != (int)headers.size())
        throw VedasError(VD_EINVAL, "replacement relation has other columns");
    if (rel_ && ctx_) vd_rel_destroy(ctx_->raw(), rel_);
    rel_ = rel;
}

void FullRelationIR::swapColumn(size_t i, size_t j) {
    if (i == j) return;
    if (i == 0 || j == 0) sorted0_ = false;
    ctx_->check(vd_rel_swap_cols(rel_, (int)i, (int)j), "vd_rel_swap_cols");
    std::swap(headers[i], headers[j]);
    bool t = is_predicates[i];
    is_predicates[i] = is_predicates[j];
    is_predicates[j] = t;
    var_column_map[headers[i]] = i;
    var_column_map[headers[j]] = j;
}

void FullRelationIR::removeColumn(size_t i) {
    if (i == 0) sorted0_ = false;
    if (getColumnNum() == 1) {
        vd_rel_destroy(ctx_->raw(), rel_);
        rel_ = nullptr;
    } else {
        ctx_->check(vd_rel_drop_col(ctx_->raw(), rel_, (int)i), "vd_rel_drop_col");
    }
    var_column_map.erase(headers[i]);
    headers.erase(headers.begin() + i);
    is_predicates.erase(is_predicates.begin() + i);
    for (size_t k = i; k < headers.size(); ++k) var_column_map[headers[k]] = k;
}

void FullRelationIR::removeColumn(size_t i, std::string &maintain_var) {
    removeColumn(i);
    if (i == 0) {
        swapColumn(0, getColumnId(maintain_var));
        sortByFirstColumn();
    }
}

void FullRelationIR::print() const { print(false, std::cout); }

void FullRelationIR::print(bool full_version, std::ostream &out) const {
    if (getColumnNum() == 0) {
        out << "Empty Relation\n";
        return;
    }
    std::vector<TYPEID_HOST_VEC> cols;
    for (size_t c = 0; c < getColumnNum(); ++c) cols.push_back(download(c));
    for (auto &h : headers) out << h << " ";
    out << "\n";
    const size_t n = size();
    auto row = [&](size_t r) {
        for (auto &c : cols) out << c[r] << " ";
        out << "\n";
    };
    if (n > 50 && !full_version) {
        for (size_t r = 0; r < 3; ++r) row(r);
        out << "... Other " << (n - 6) << " rows\n";
        for (size_t r = n - 3; r < n; ++r) row(r);
    } else {
        for (size_t r = 0; r < n; ++r) row(r);
    }
}

size_t FullRelationIR::size() const { return rel_ ? vd_rel_nrows(rel_) : 0; }

TYPEID_HOST_VEC FullRelationIR::download(size_t i) const {
    TYPEID_HOST_VEC v(size());
    if (!v.empty())
        ctx_->check(vd_copy_d2h(ctx_->raw(), v.data(), vd_rel_col(rel_, (int)i),
                                v.size() * sizeof(TYPEID)), "vd_copy_d2h");
    return v;
}
