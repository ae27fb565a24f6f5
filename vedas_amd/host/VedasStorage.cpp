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

void read_array(std::ifstream &in, TYPEID_HOST_VEC &a) {
    size_t n = 0;
    in.read(reinterpret_cast<char *>(&n), sizeof(size_t));
    if (!in) throw VedasError(VD_EINVAL, ".vds file truncated (array size)");
    if (n > (size_t(1) << 34)) throw VedasError(VD_EINVAL, ".vds array size out of range");
    a.resize(n);
    in.read(reinterpret_cast<char *>(a.data()), n * sizeof(TYPEID));
    if (!in) throw VedasError(VD_EINVAL, ".vds file truncated (array data)");
}

}  // namespace

// File order of the default index (VedasStorage.cu:486-496): the predicate
// L1 shared by PSO and POS, the PSO and POS L2, then the Y and Z columns.
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
    if (!out) throw VedasError(VD_EINVAL, std::string("write failed: ") + fname);
}

VedasStorage::VedasStorage(VedasContext *ctx, const char *fname) : ctx_(ctx) {
    std::ifstream in(fname, std::ios::binary);
    if (!in) throw VedasError(VD_EINVAL, std::string("cannot open ") + fname);
    TYPEID_HOST_VEC a[10];
    for (auto &x : a) read_array(in, x);
    const size_t n = a[6].size();
    if (a[0].size() != a[1].size() || a[2].size() != a[3].size() ||
        a[4].size() != a[5].size() || a[7].size() != n || a[8].size() != n ||
        a[9].size() != n)
        throw VedasError(VD_EINVAL, ".vds arrays have inconsistent sizes");
    n_ = n;
    // Device copies of the ten arrays, then the index adopts them.
    std::vector<void *> d(10, nullptr);
    try {
        for (int i = 0; i < 10; ++i) {
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
        ctx_->check(vd_index_from_arrays(ctx_->raw(),
                                         VD_PERM_BIT(VD_PERM_PSO) | VD_PERM_BIT(VD_PERM_POS),
                                         views, &index_),
                    "vd_index_from_arrays");
    } catch (...) {
        for (auto p : d) if (p) vd_dev_free(ctx_->raw(), p);
        throw;
    }
    for (auto p : d) vd_dev_free(ctx_->raw(), p);
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
