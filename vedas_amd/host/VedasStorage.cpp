#include "VedasStorage.h"

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
