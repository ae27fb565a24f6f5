#include "VedasStorage.h"

#include <cstdio>
#include <fstream>
#include <vector>

VedasStorage::VedasStorage(RdfData &rdf, bool /*preload*/, bool fullIndex, VedasContext *ctx) {
    ctx_ = ctx ? ctx : VedasContext::current();
    if (!ctx_) {
        own_ctx_.reset(new VedasContext(0));
        ctx_ = own_ctx_.get();
    }
    n_ = rdf.size();
    uint32_t mask = VD_PERM_BIT(VD_PERM_PSO) | VD_PERM_BIT(VD_PERM_POS);
    if (fullIndex)
        mask |= VD_PERM_BIT(VD_PERM_SPO) | VD_PERM_BIT(VD_PERM_SOP) |
                VD_PERM_BIT(VD_PERM_OSP) | VD_PERM_BIT(VD_PERM_OPS);
    void *d[3] = {nullptr, nullptr, nullptr};
    TYPEID_HOST_VEC *h[3] = {&rdf.getSubject(), &rdf.getPredicate(), &rdf.getObject()};
    try {
        for (int c = 0; c < 3; ++c) {
            ctx_->check(vd_dev_alloc(ctx_->raw(), (n_ ? n_ : 1) * 4, &d[c]), "vd_dev_alloc");
            if (n_) ctx_->check(vd_copy_h2d(ctx_->raw(), d[c], h[c]->data(), n_ * 4), "vd_copy_h2d");
        }
        build((uint32_t *)d[0], (uint32_t *)d[1], (uint32_t *)d[2], mask);
    } catch (...) {
        for (auto p : d) if (p) vd_dev_free(ctx_->raw(), p);
        throw;
    }
    for (auto p : d) vd_dev_free(ctx_->raw(), p);
}

VedasStorage::VedasStorage(VedasContext *ctx, const uint32_t *d_s, const uint32_t *d_p,
                           const uint32_t *d_o, uint64_t n, uint32_t perm_mask)
    : ctx_(ctx), n_(n) {
    build(d_s, d_p, d_o, perm_mask);
}

void VedasStorage::build(const uint32_t *d_s, const uint32_t *d_p, const uint32_t *d_o,
                         uint32_t mask) {
    ctx_->check(vd_index_build(ctx_->raw(), d_s, d_p, d_o, n_, mask, &index_), "vd_index_build");
}

namespace {

// .vds array: a size_t element count, then the uint32 elements.
void write_array(std::ofstream &out, const TYPEID_HOST_VEC &a) {
    const size_t n = a.size();
    out.write(reinterpret_cast<const char *>(&n), sizeof(size_t));
    out.write(reinterpret_cast<const char *>(a.data()), n * sizeof(TYPEID));
}

// Reads one array; false at a clean end of file before its size.
bool read_array(std::ifstream &in, TYPEID_HOST_VEC &a, bool may_end = false) {
    size_t n = 0;
    in.read(reinterpret_cast<char *>(&n), sizeof(size_t));
    if (may_end && in.gcount() == 0 && in.eof()) return false;
    if (!in) throw VedasError(VD_EINVAL, ".vds file truncated (array size)");
    if (n > (size_t(1) << 34)) throw VedasError(VD_EINVAL, ".vds array size out of range");
    a.resize(n);
    in.read(reinterpret_cast<char *>(a.data()), n * sizeof(TYPEID));
    if (!in) throw VedasError(VD_EINVAL, ".vds file truncated (array data)");
    return true;
}

// Run arrays (L1 or L2) of n rows: equal lengths, offsets strictly
// increasing from 0 and below n, and for an L1 strictly increasing values.
void check_runs(const TYPEID_HOST_VEC &v, const TYPEID_HOST_VEC &o, size_t n, bool l1,
                const char *what) {
    if (v.size() != o.size()) throw VedasError(VD_EINVAL, std::string(what) + ": length mismatch");
    if (n == 0 ? !o.empty() : (o.empty() || o[0] != 0))
        throw VedasError(VD_EINVAL, std::string(what) + ": runs do not start at row 0");
    for (size_t i = 1; i < o.size(); ++i)
        if (o[i] <= o[i - 1] || o[i] >= n || (l1 && v[i] <= v[i - 1]))
            throw VedasError(VD_EINVAL, std::string(what) + ": runs out of order");
}

}  // namespace

// File order (VedasStorage.cu:492-516): the default index -- the predicate
// L1 shared by PSO and POS, the PSO and POS L2, then the Y and Z columns --
// and, for a full index (SPO, SOP, OSP and OPS built as well), the
// subject L1 (which the reference leaves empty; written filled here), the
// SPO and SOP L2, the object L1, the OSP Y column, the OPS L2, and the Z
// columns of SPO, SOP, OPS and OSP.
void VedasStorage::write(const char *fname) {
    if (!hasPermutation(VD_PERM_PSO) || !hasPermutation(VD_PERM_POS))
        throw VedasError(VD_EINVAL, ".vds needs the PSO and POS permutations");
    std::ofstream out(fname, std::ios::binary);
    if (!out) throw VedasError(VD_EINVAL, std::string("cannot open ") + fname);
    write_array(out, *host(VD_PERM_PSO, L1V));  // p_idx_values
    write_array(out, *host(VD_PERM_PSO, L1O));  // p_idx_offsets
    write_array(out, *host(VD_PERM_PSO, L2V));  // ps_s_idx_values
    write_array(out, *host(VD_PERM_PSO, L2O));  // ps_idx_offsets
    write_array(out, *host(VD_PERM_POS, L2V));  // po_o_idx_values
    write_array(out, *host(VD_PERM_POS, L2O));  // po_idx_offsets
    write_array(out, *host(VD_PERM_PSO, COLY)); // ps_data  (S in PSO order)
    write_array(out, *host(VD_PERM_POS, COLY)); // po_data  (O in POS order)
    write_array(out, *host(VD_PERM_PSO, COLZ)); // pso_data (O in PSO order)
    write_array(out, *host(VD_PERM_POS, COLZ)); // pos_data (S in POS order)
    if (isFullIndex()) {
        write_array(out, *host(VD_PERM_SPO, L1V));  // s_idx_values
        write_array(out, *host(VD_PERM_SPO, L1O));  // s_idx_offsets
        write_array(out, *host(VD_PERM_SPO, L2V));  // sp_p_idx_values
        write_array(out, *host(VD_PERM_SPO, L2O));  // sp_idx_offsets
        write_array(out, *host(VD_PERM_SOP, L2V));  // so_o_idx_values
        write_array(out, *host(VD_PERM_SOP, L2O));  // so_idx_offsets
        write_array(out, *host(VD_PERM_OSP, L1V));  // o_idx_values
        write_array(out, *host(VD_PERM_OSP, L1O));  // o_idx_offsets
        write_array(out, *host(VD_PERM_OSP, COLY)); // os_data  (S in OSP order)
        write_array(out, *host(VD_PERM_OPS, L2V));  // op_p_idx_values
        write_array(out, *host(VD_PERM_OPS, L2O));  // op_idx_offsets
        write_array(out, *host(VD_PERM_SPO, COLZ)); // spo_data (O in SPO order)
        write_array(out, *host(VD_PERM_SOP, COLZ)); // sop_data (P in SOP order)
        write_array(out, *host(VD_PERM_OPS, COLZ)); // ops_data (S in OPS order)
        write_array(out, *host(VD_PERM_OSP, COLZ)); // osp_data (P in OSP order)
    }
    if (!out) throw VedasError(VD_EINVAL, std::string("write failed: ") + fname);
}

bool VedasStorage::isFullIndex() const {
    for (int p : {VD_PERM_SPO, VD_PERM_SOP, VD_PERM_OSP, VD_PERM_OPS})
        if (!hasPermutation(p)) return false;
    return true;
}

// VedasStorage::open (VedasStorage.cu:459-484): the arrays go to the device
// as they are (no sort).  The full-index arrays lack the SPO, SOP and OPS Y
// columns and the OSP L2; the device rebuilds them from the runs.  A file
// whose subject L1 is empty (the reference never fills it,
// VedasStorage.cu:319-339) gets SPO and SOP by an index build over the
// triples of PSO instead.
VedasStorage::VedasStorage(VedasContext *ctx, const char *fname) : ctx_(ctx) {
    std::ifstream in(fname, std::ios::binary);
    if (!in) throw VedasError(VD_EINVAL, std::string("cannot open ") + fname);
    TYPEID_HOST_VEC a[25];
    for (int i = 0; i < 10; ++i) read_array(in, a[i]);
    const bool full = read_array(in, a[10], true);
    if (full)
        for (int i = 11; i < 25; ++i) read_array(in, a[i]);
    const size_t n = a[6].size();
    if (a[7].size() != n || a[8].size() != n || a[9].size() != n)
        throw VedasError(VD_EINVAL, ".vds arrays have inconsistent sizes");
    check_runs(a[0], a[1], n, true, "predicate L1");
    check_runs(a[2], a[3], n, false, "PSO L2");
    check_runs(a[4], a[5], n, false, "POS L2");
    const bool s_l1 = full && !a[10].empty();
    if (full) {
        if (s_l1) {
            check_runs(a[10], a[11], n, true, "subject L1");
            check_runs(a[12], a[13], n, false, "SPO L2");
            check_runs(a[14], a[15], n, false, "SOP L2");
            if (a[21].size() != n || a[22].size() != n)
                throw VedasError(VD_EINVAL, ".vds SPO / SOP columns have inconsistent sizes");
        }
        check_runs(a[16], a[17], n, true, "object L1");
        check_runs(a[19], a[20], n, false, "OPS L2");
        if (a[18].size() != n || a[23].size() != n || a[24].size() != n)
            throw VedasError(VD_EINVAL, ".vds object columns have inconsistent sizes");
    }
    n_ = n;
    const int na = full ? 25 : 10;
    std::vector<void *> d(na, nullptr);
    vd_index *tmp = nullptr;  // SPO / SOP built from the triples
    void *tri[3] = {nullptr, nullptr, nullptr};
    auto release = [&]() {
        for (auto p : d) if (p) vd_dev_free(ctx_->raw(), p);
        for (auto p : tri) if (p) vd_dev_free(ctx_->raw(), p);
        if (tmp) vd_index_free(ctx_->raw(), tmp);
    };
    try {
        for (int i = 0; i < na; ++i) {
            if (full && !s_l1 && i >= 10 && i <= 15) continue;  // empty subject-first runs
            if (full && !s_l1 && (i == 21 || i == 22)) continue;
            ctx_->check(vd_dev_alloc(ctx_->raw(), (a[i].size() ? a[i].size() : 1) * 4, &d[i]),
                        "vd_dev_alloc");
            if (!a[i].empty())
                ctx_->check(vd_copy_h2d(ctx_->raw(), d[i], a[i].data(), a[i].size() * 4),
                            "vd_copy_h2d");
        }
        vd_perm_view_t views[VD_NUM_PERMS] = {};
        auto u = [&](int i) { return static_cast<const uint32_t *>(d[i]); };
        views[VD_PERM_PSO] = vd_perm_view_t{n, a[0].size(), a[2].size(), u(0), u(1), u(2), u(3),
                                            u(6), u(8)};
        views[VD_PERM_POS] = vd_perm_view_t{n, a[0].size(), a[4].size(), u(0), u(1), u(4), u(5),
                                            u(7), u(9)};
        uint32_t mask = VD_PERM_BIT(VD_PERM_PSO) | VD_PERM_BIT(VD_PERM_POS);
        if (full) {
            views[VD_PERM_OPS] = vd_perm_view_t{n, a[16].size(), a[19].size(), u(16), u(17),
                                                u(19), u(20), nullptr, u(23)};
            views[VD_PERM_OSP] = vd_perm_view_t{n, a[16].size(), 0, u(16), u(17), nullptr,
                                                nullptr, u(18), u(24)};
            if (s_l1) {
                views[VD_PERM_SPO] = vd_perm_view_t{n, a[10].size(), a[12].size(), u(10), u(11),
                                                    u(12), u(13), nullptr, u(21)};
                views[VD_PERM_SOP] = vd_perm_view_t{n, a[10].size(), a[14].size(), u(10), u(11),
                                                    u(14), u(15), nullptr, u(22)};
            } else if (n) {
                // triples of PSO: P from its L1 runs, S and O its columns
                ctx_->check(vd_dev_alloc(ctx_->raw(), n * 4, &tri[1]), "vd_dev_alloc");
                ctx_->check(vd_runs_expand(ctx_->raw(), u(0), u(1), a[0].size(), n,
                                           (uint32_t *)tri[1]),
                            "vd_runs_expand");
                ctx_->check(vd_index_build(ctx_->raw(), u(6), (const uint32_t *)tri[1], u(8), n,
                                           VD_PERM_BIT(VD_PERM_SPO) | VD_PERM_BIT(VD_PERM_SOP),
                                           &tmp),
                            "vd_index_build");
                ctx_->check(vd_index_view(tmp, VD_PERM_SPO, &views[VD_PERM_SPO]), "vd_index_view");
                ctx_->check(vd_index_view(tmp, VD_PERM_SOP, &views[VD_PERM_SOP]), "vd_index_view");
            } else {
                views[VD_PERM_SPO] = views[VD_PERM_OSP];
                views[VD_PERM_SOP] = views[VD_PERM_OSP];
            }
            mask |= VD_PERM_BIT(VD_PERM_SPO) | VD_PERM_BIT(VD_PERM_SOP) |
                    VD_PERM_BIT(VD_PERM_OSP) | VD_PERM_BIT(VD_PERM_OPS);
        }
        ctx_->check(vd_index_from_arrays(ctx_->raw(), mask, views, &index_), "vd_index_from_arrays");
    } catch (...) {
        release();
        throw;
    }
    release();
}

VedasStorage::~VedasStorage() {
    if (index_) vd_index_free(ctx_->raw(), index_);
}

vd_perm_view_t VedasStorage::view(int perm) const {
    vd_perm_view_t v;
    if (vd_index_view(index_, perm, &v) != VD_OK)
        throw VedasError(VD_EINVAL, "permutation not built");
    return v;
}

size_t VedasStorage::getPredicateIndexSize() const { return view(VD_PERM_PSO).n_l1; }
size_t VedasStorage::getSubjectIndexSize() const { return view(VD_PERM_SPO).n_l1; }
size_t VedasStorage::getObjectIndexSize() const { return view(VD_PERM_OSP).n_l1; }

TYPEID_HOST_VEC *VedasStorage::host(int perm, Part part) {
    auto key = std::make_pair(perm, (int)part);
    auto it = host_cache_.find(key);
    if (it != host_cache_.end()) return &it->second;
    vd_perm_view_t v = view(perm);
    const uint32_t *src = nullptr;
    uint64_t cnt = 0;
    switch (part) {
        case L1V: src = v.l1_values; cnt = v.n_l1; break;
        case L1O: src = v.l1_offsets; cnt = v.n_l1; break;
        case L2V: src = v.l2_values; cnt = v.n_l2; break;
        case L2O: src = v.l2_offsets; cnt = v.n_l2; break;
        case COLY: src = v.col_y; cnt = v.n; break;
        case COLZ: src = v.col_z; cnt = v.n; break;
    }
    TYPEID_HOST_VEC &dst = host_cache_[key];
    dst.resize(cnt);
    if (cnt) ctx_->check(vd_copy_d2h(ctx_->raw(), dst.data(), src, cnt * 4), "vd_copy_d2h");
    return &dst;
}

// Name -> (permutation, part), following VedasStorage.cu:25-73.
TYPEID_HOST_VEC *VedasStorage::getSubjectIndexValues() { return host(VD_PERM_SPO, L1V); }
TYPEID_HOST_VEC *VedasStorage::getSubjectIndexOffsets() { return host(VD_PERM_SPO, L1O); }
TYPEID_HOST_VEC *VedasStorage::getSubjectPredicateIndexValues() { return host(VD_PERM_SPO, L2V); }
TYPEID_HOST_VEC *VedasStorage::getSubjectPredicateIndexOffsets() { return host(VD_PERM_SPO, L2O); }
TYPEID_HOST_VEC *VedasStorage::getSubjectObjectIndexValues() { return host(VD_PERM_SOP, L2V); }
TYPEID_HOST_VEC *VedasStorage::getSubjectObjectIndexOffsets() { return host(VD_PERM_SOP, L2O); }
TYPEID_HOST_VEC *VedasStorage::getPredicateIndexValues() { return host(VD_PERM_PSO, L1V); }
TYPEID_HOST_VEC *VedasStorage::getPredicateIndexOffsets() { return host(VD_PERM_PSO, L1O); }
TYPEID_HOST_VEC *VedasStorage::getPredicateSubjectIndexValues() { return host(VD_PERM_PSO, L2V); }
TYPEID_HOST_VEC *VedasStorage::getPredicateSubjectIndexOffsets() { return host(VD_PERM_PSO, L2O); }
TYPEID_HOST_VEC *VedasStorage::getPredicateObjectIndexValues() { return host(VD_PERM_POS, L2V); }
TYPEID_HOST_VEC *VedasStorage::getPredicateObjectIndexOffsets() { return host(VD_PERM_POS, L2O); }
TYPEID_HOST_VEC *VedasStorage::getObjectIndexValues() { return host(VD_PERM_OSP, L1V); }
TYPEID_HOST_VEC *VedasStorage::getObjectIndexOffsets() { return host(VD_PERM_OSP, L1O); }
TYPEID_HOST_VEC *VedasStorage::getObjectPredicateIndexValues() { return host(VD_PERM_OPS, L2V); }
TYPEID_HOST_VEC *VedasStorage::getObjectPredicateIndexOffsets() { return host(VD_PERM_OPS, L2O); }
TYPEID_HOST_VEC *VedasStorage::getPSdata() { return host(VD_PERM_PSO, COLY); }
TYPEID_HOST_VEC *VedasStorage::getPOdata() { return host(VD_PERM_POS, COLY); }
TYPEID_HOST_VEC *VedasStorage::getOSdata() { return host(VD_PERM_OSP, COLY); }
TYPEID_HOST_VEC *VedasStorage::getSPOdata() { return host(VD_PERM_SPO, COLZ); }
TYPEID_HOST_VEC *VedasStorage::getSOPdata() { return host(VD_PERM_SOP, COLZ); }
TYPEID_HOST_VEC *VedasStorage::getPSOdata() { return host(VD_PERM_PSO, COLZ); }
TYPEID_HOST_VEC *VedasStorage::getPOSdata() { return host(VD_PERM_POS, COLZ); }
TYPEID_HOST_VEC *VedasStorage::getOPSdata() { return host(VD_PERM_OPS, COLZ); }
TYPEID_HOST_VEC *VedasStorage::getOSPdata() { return host(VD_PERM_OSP, COLZ); }
