// Triple store with HBM-resident permutations (reference VedasStorage.h:8-132).
//
// The reference sorts on the GPU, copies everything back to host vectors and
// uploads ranges per select (or preloads the data columns with -preload).
// Here the permutations stay in HBM after the build (always "preloaded");
// the host getters download an array on first use, for callers that want
// host copies (printing, statistics, the .vds writer).
#ifndef VEDAS_HOST_VEDASSTORAGE_H
#define VEDAS_HOST_VEDASSTORAGE_H

#include <map>
#include <memory>

#include "RdfData.h"
#include "VedasContext.h"
#include "vedas.h"

class VedasStorage {
  public:
    // Default index: PSO + POS; fullIndex adds SPO, SOP, OSP, OPS
    // (VedasStorage.cu:25-73).
    VedasStorage(RdfData &rdfData, bool preload, bool fullIndex,
                 VedasContext *ctx = nullptr);
    // Explicit permutation set (mask of VD_PERM_BIT) from device triples.
    VedasStorage(VedasContext *ctx, const uint32_t *d_s, const uint32_t *d_p,
                 const uint32_t *d_o, uint64_t n, uint32_t perm_mask);
    // A store read from a .vds file (VedasStorage::open, VedasStorage.cu:
    // 452-484): the PSO and POS permutations of the default index, and all
    // six of a full-index file, uploaded as they are (no sort; the columns
    // the file leaves out are rebuilt from its runs on the device).
    VedasStorage(VedasContext *ctx, const char *fname);
    ~VedasStorage();
    // Write the .vds format (VedasStorage::write, VedasStorage.cu:492-516):
    // arrays, each a size_t count followed by the uint32 values -- the ten
    // of the default index (needs PSO and POS), then, when SPO, SOP, OSP and
    // OPS are built too, the fifteen of the full index.
    void write(const char *fname);
    VedasStorage(const VedasStorage &) = delete;
    VedasStorage &operator=(const VedasStorage &) = delete;

    size_t getTripleSize() const { return n_; }
    size_t getPredicateIndexSize() const;
    size_t getSubjectIndexSize() const;
    size_t getObjectIndexSize() const;
    bool isPreload() const { return true; }
    bool hasPermutation(int perm) const { return (vd_index_perms(index_) >> perm) & 1u; }
    // SPO, SOP, OSP and OPS built as well as PSO and POS (the reference's
    // -full-index store): write() then adds the full-index arrays.
    bool isFullIndex() const;
    vd_index *index() const { return index_; }
    VedasContext *context() const { return ctx_; }
    vd_perm_view_t view(int perm) const;

    // Host copies, names as in the reference.
    TYPEID_HOST_VEC *getSubjectIndexValues();
    TYPEID_HOST_VEC *getSubjectIndexOffsets();
    TYPEID_HOST_VEC *getSubjectPredicateIndexValues();
    TYPEID_HOST_VEC *getSubjectPredicateIndexOffsets();
    TYPEID_HOST_VEC *getSubjectObjectIndexValues();
    TYPEID_HOST_VEC *getSubjectObjectIndexOffsets();
    TYPEID_HOST_VEC *getPredicateIndexValues();
    TYPEID_HOST_VEC *getPredicateIndexOffsets();
    TYPEID_HOST_VEC *getPredicateSubjectIndexValues();
    TYPEID_HOST_VEC *getPredicateSubjectIndexOffsets();
    TYPEID_HOST_VEC *getPredicateObjectIndexValues();
    TYPEID_HOST_VEC *getPredicateObjectIndexOffsets();
    TYPEID_HOST_VEC *getObjectIndexValues();
    TYPEID_HOST_VEC *getObjectIndexOffsets();
    TYPEID_HOST_VEC *getObjectPredicateIndexValues();
    TYPEID_HOST_VEC *getObjectPredicateIndexOffsets();
    TYPEID_HOST_VEC *getPSdata();
    TYPEID_HOST_VEC *getPOdata();
    TYPEID_HOST_VEC *getOSdata();
    TYPEID_HOST_VEC *getSPOdata();
    TYPEID_HOST_VEC *getSOPdata();
    TYPEID_HOST_VEC *getPSOdata();
    TYPEID_HOST_VEC *getPOSdata();
    TYPEID_HOST_VEC *getOPSdata();
    TYPEID_HOST_VEC *getOSPdata();

  private:
    enum Part { L1V, L1O, L2V, L2O, COLY, COLZ };
    TYPEID_HOST_VEC *host(int perm, Part part);
    void build(const uint32_t *d_s, const uint32_t *d_p, const uint32_t *d_o, uint32_t mask);

    VedasContext *ctx_ = nullptr;
    std::unique_ptr<VedasContext> own_ctx_;
    vd_index *index_ = nullptr;
    uint64_t n_ = 0;
    std::map<std::pair<int, int>, TYPEID_HOST_VEC> host_cache_;
};

#endif
