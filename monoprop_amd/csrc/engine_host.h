// Host side of the engine: device memory cache, device arrays and the engine state (mpa_engine).
// Part of the monoprop_amd engine; included by engine.hip (one translation unit).

#pragma once

namespace {

// ------------------------------------------------------------------------------------------------
// host side

// Device memory blocks are cached per process instead of returned: a regrowing store (capacities are
// powers of two) and the next propagator get the same sizes back without mapping new memory, which
// for blocks of many GB costs far more than the work on them.  A block freed on one stream and taken
// on another waits for the freeing stream's work first (an event recorded at free time).  When HIP
// cannot allocate, every cached block goes back to HIP and the allocation is retried once.
class BlockCache {
   public:
    static BlockCache &get() {
        static BlockCache c;
        return c;
    }
    void *take(size_t &bytes, hipStream_t s) {
        int dev = 0;
        HIP_TRY(hipGetDevice(&dev));
        {
            std::lock_guard<std::mutex> lk(mu_);
            size_t best = free_.size();
            for (size_t i = 0; i < free_.size(); ++i) {
                const Blk &b = free_[i];
                if (b.dev != dev || b.bytes < bytes || b.bytes > 2 * bytes + (size_t{64} << 20)) continue;
                if (best == free_.size() || b.bytes < free_[best].bytes) best = i;
            }
            if (best != free_.size()) {
                Blk b = free_[best];
                free_.erase(free_.begin() + static_cast<std::ptrdiff_t>(best));
                if (b.ev) {
                    if (b.stream != s) HIP_TRY(hipStreamWaitEvent(s, b.ev, 0));
                    (void)hipEventDestroy(b.ev);
                }
                bytes = b.bytes;
                return b.p;
            }
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            release_all();
            e = hipMalloc(&p, bytes);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();  // a failed allocation is reported here, not by the next launch check
            throw HipFailure(std::string("hipMalloc(") + std::to_string(bytes) + " bytes): " + hipGetErrorString(e));
        }
        return p;
    }
    void give(void *p, size_t bytes, hipStream_t s) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipEvent_t ev = nullptr;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(ev, s) != hipSuccess) {
            if (ev) (void)hipEventDestroy(ev);
            (void)hipStreamSynchronize(s);
            ev = nullptr;
        }
        std::lock_guard<std::mutex> lk(mu_);
        free_.push_back({p, bytes, dev, s, ev});
    }
    size_t cached_bytes(int dev) {
        std::lock_guard<std::mutex> lk(mu_);
        size_t t = 0;
        for (const Blk &b : free_)
            if (b.dev == dev) t += b.bytes;
        return t;
    }
    void release_all() {
        std::lock_guard<std::mutex> lk(mu_);
        for (Blk &b : free_) {
            if (b.ev) {
                (void)hipEventSynchronize(b.ev);
                (void)hipEventDestroy(b.ev);
            }
            (void)hipFree(b.p);
        }
        free_.clear();
    }

   private:
    struct Blk {
        void *p;
        size_t bytes;
        int dev;
        hipStream_t stream;
        hipEvent_t ev;
    };
    std::mutex mu_;
    std::vector<Blk> free_;
};

// Device address of `scratch` on each GPU (each GPU holds its own copy), looked up once per device
// with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

void *noalloc_scratch(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw ArgError("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (rc != hipSuccess || ptr == nullptr)
            throw HipFailure(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(rc));
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

// A device array owned by one engine: allocations and frees are ordered on the engine's stream (the
// null stream once the engine has synchronised and destroyed it).
struct DevArray {
    void *p = nullptr;
    size_t bytes = 0;
    const hipStream_t *sp = nullptr;
    DevArray() = default;
    explicit DevArray(const hipStream_t *s) : sp(s) {}
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    ~DevArray() {
        if (p) BlockCache::get().give(p, bytes, stream());
    }
    hipStream_t stream() const { return sp ? *sp : nullptr; }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
    // Back to the cache.
    void release() {
        if (p) BlockCache::get().give(p, bytes, stream());
        p = nullptr;
        bytes = 0;
    }
    // Grow to at least `need` bytes (contents not preserved).
    void reserve(size_t need) {
        if (need <= bytes) return;
        size_t nb = std::max(need, bytes + bytes / 2);
        nb = (nb + 4095) & ~size_t{4095};
        if (p) BlockCache::get().give(p, bytes, stream());
        p = nullptr;
        bytes = 0;
        p = BlockCache::get().take(nb, stream());
        bytes = nb;
    }
};

}  // namespace

struct mpa_engine {
    int basis = 0, n_modes = 0, L = 0, nw = 0, ncols = 0;
    Trunc tr{};
    double core = 0.0;
    std::vector<uint64_t> occ;
    int device = 0;
    hipStream_t stream = nullptr;

    uint64_t N = 0;     // live terms (exact whenever the host has synchronised)
    uint64_t cap = 0;   // row capacity (a power of two, at least 64 * kBlock)
    DevArray rows{&stream}, coef{&stream}, cols{&stream};
    DevArray AL{&stream}, lead{&stream}, foll{&stream}, blk_a{&stream}, blk_b{&stream}, rec{&stream}, lv{&stream},
        match{&stream}, miss{&stream}, ins{&stream}, table{&stream}, counters{&stream}, stats{&stream},
        gdesc{&stream}, fold{&stream}, occ_dev{&stream};
    // ordered sums (ordered_sum): block counts + total, compacted terms, their members, results
    DevArray sum_blk{&stream}, sum_v{&stream}, sum_m{&stream}, sum_out{&stream};
    // Cutoff basis (basis_change): 64 * nw rows of nw words, row s the image of slot s; empty: native.
    std::vector<uint64_t> bchg_host;
    DevArray bchg_dev{&stream};

    // Upload bchg_host and point the truncation at it (or at nothing for the native basis).
    void upload_basis() {
        if (bchg_host.empty()) {
            tr.bchg = nullptr;
            return;
        }
        bchg_dev.reserve(bchg_host.size() * 8);
        HIP_TRY(hipMemcpyAsync(bchg_dev.p, bchg_host.data(), bchg_host.size() * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        tr.bchg = bchg_dev.as<uint64_t>();
    }

    // Graph (build_graph): per layer the rotation records and what recomputes its cos-only rows (the
    // live count before it, G's fold columns, per 256-word block the cos-only rows before the block),
    // the coefficient vector the seeded build evolves, the operator coefficients the graph starts
    // from, and replay scratch.  cos_begin / cos_end number the layer's cos-only rows in build order
    // (they place its pre-layer values in a segment's snapshot).
    struct LayerInfo {
        uint64_t cos_begin, cos_end, rot_begin, rot_end;
        int64_t param_index;
        double gen_coeff;
        uint64_t n_before;
        int32_t fold_off, fold_cnt;  // in graph_folds / g_fold
        uint64_t cofs_off;           // first block offset in g_cofs
    };
    std::vector<LayerInfo> layers;
    std::vector<int32_t> graph_folds;
    int64_t graph_nparams = 0;
    DevArray g_rot{&stream}, g_fold{&stream}, g_cofs{&stream}, pmark{&stream}, pmark1{&stream}, seedc{&stream}, base{&stream},
        rep{&stream}, snap_c{&stream}, snap_r{&stream}, lam{&stream}, gpartial{&stream}, goffs{&stream}, gdl{&stream};
    uint64_t base_n = 0;
    uint64_t gcos_total = 0, grot_total = 0;  // cos-only rows counted, rotation records stored
    uint64_t cofs_total = 0;                  // block offsets stored
    // Checkpointed adjoint: coefficient vectors at segment starts; the byte budget for them plus one
    // segment's snapshots; with keep_ckpt, mpa_expectation_value runs the checkpointed forward and a
    // following gradient at the same parameters (ckpt_params, to the bit) starts from it.
    DevArray ckpt{&stream};
    int64_t grad_budget = int64_t{24} << 30;
    bool keep_ckpt = false, ckpt_valid = false;
    int ckpt_kind = 0;
    std::vector<double> ckpt_params;
    double ckpt_value = 0.0;
    bool building = false;                    // seedc tracks capacity while a graph is built

    // Fused pipeline state: global row index; per 1024-term sweep tile a region of emitter rows (t_item)
    // and their pre-gate values (t_val); tile, sweep-block and group counts; leader emit masks per word;
    // look-back descriptors.
    DevArray gidx{&stream}, f_elead{&stream}, t_val{&stream}, t_item{&stream}, t_bsum{&stream},
        t_gsum{&stream}, desc_m{&stream};
    uint64_t gsum_stride = 0;  // per gate parity
    int group_log2 = kMinGroupLog2;
    // launch parameters of the fused pipeline (mpa_set_launch): resolve blocks that take chunks, and
    // the smallest count group (2^min_group_log2 sweep tiles)
    uint32_t res_grid_cap = kMaxGrid;
    int min_group_log2 = kMinGroupLog2;
    uint64_t gidx_slots = 0;
    uint64_t gidx_rows = 0;     // rows [0, gidx_rows) are indexed
    uint64_t fused_cap = 0;     // capacity the per-row fused arrays were sized for
    uint64_t cap_items = 0;
    uint32_t desc_epoch = 0;    // per fused gate; descriptors of other epochs read as unpublished
    static constexpr size_t kChunk = 256;  // gates enqueued between host synchronisations
    bool fused_enabled = true;             // mpa_set_pipeline(e, 0) forces the per-gate pipeline
    uint64_t table_slots = 0;
    uint64_t epoch = 0;
    int64_t *h_counters = nullptr;  // pinned

    // Batched propagation: `nbatch` independent members (0: a lone propagator) share the store;
    // bmember[row] is the member of a row, btrig[gate * nbatch + member] its rotation constants, bcore
    // the members' identity coefficients.
    uint32_t nbatch = 0;
    std::vector<double> bcore;
    uint64_t cap_limit = 0;  // fused pipeline: stop before a gate that would grow the store past this (0: none)
    uint64_t store_limit = 0;  // lone propagation: rows the store may not outgrow (0: device memory decides)
    size_t fused_done = 0;   // gates the last fused propagate applied
    DevArray bmember{&stream}, btrig{&stream}, bout{&stream};

    // Schrodinger picture (schrodinger_cutoff set): the store holds the evolved state; the observable's
    // terms (core excluded) are matched against the rows when <O> is evaluated (obs_tab: their
    // open-addressing table, built on first use).
    bool schr = false;
    std::vector<uint64_t> obs_rows;
    std::vector<double> obs_coef;
    uint64_t obs_tmask = 0;
    DevArray obs_tab{&stream}, obs_c{&stream}, obs_hit{&stream}, obs_val{&stream};

    std::vector<double> init_coef;  // initial coefficients of rows [0, n_init)
    uint64_t n_init = 0;
    std::vector<int64_t> stat_n, stat_a, stat_e, stat_i, stat_l, stat_p;

    // Profiler: an event pair around every launch, resolved whenever the stream is idle.
    bool prof_on = false;
    std::vector<hipEvent_t> ev_pool;
    struct Pending {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    double prof_ms[K_KINDS] = {};
    double prof_bytes[K_KINDS] = {};
    int64_t prof_launches[K_KINDS] = {};
    double last_propagate_ms = 0.0;

    ~mpa_engine() {
        for (auto &p : pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (hipEvent_t ev : ev_pool) (void)hipEventDestroy(ev);
        if (h_counters) (void)hipHostFree(h_counters);
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
            stream = nullptr;  // the arrays below go back to the block cache on the null stream
        }
    }

    hipEvent_t ev_get() {
        if (!ev_pool.empty()) {
            hipEvent_t ev = ev_pool.back();
            ev_pool.pop_back();
            return ev;
        }
        hipEvent_t ev;
        HIP_TRY(hipEventCreate(&ev));
        return ev;
    }
    hipEvent_t prof_begin() {
        if (!prof_on) return nullptr;
        hipEvent_t a = ev_get();
        HIP_TRY(hipEventRecord(a, stream));
        return a;
    }
    void prof_end(int kind, hipEvent_t a) {
        if (!prof_on) return;
        hipEvent_t b = ev_get();
        HIP_TRY(hipEventRecord(b, stream));
        pending.push_back({kind, a, b});
    }
    // Call only when the stream is idle.
    void prof_resolve() {
        for (const Pending &p : pending) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
            prof_ms[p.kind] += ms;
            prof_launches[p.kind] += 1;
            ev_pool.push_back(p.a);
            ev_pool.push_back(p.b);
        }
        pending.clear();
    }
    // Algorithmic bytes of one propagate's kernels from the per-gate counters (the traffic the
    // algorithm must move, not what the caches see): NW-word rows, SW-word hash slots.
    void add_algorithmic_bytes(const std::vector<int32_t> &fold_plus_piv) {
        const double NW = nw, SW = nw + 1.0;
        for (size_t g = 0; g < stat_n.size(); ++g) {
            if (stat_a[g] == 0 && fold_plus_piv[g] == 0) continue;
            const double words = std::ceil(static_cast<double>(stat_n[g]) / 64.0);
            const double nA = stat_a[g], nL = stat_l[g], nF = nA - nL, pairs = stat_p[g], I = stat_i[g];
            prof_bytes[K_SCAN] += words * 8.0 * (fold_plus_piv[g] + 2);
            prof_bytes[K_SCANSUM] += 2.0 * 8.0 * std::ceil(words / kBlock);
            prof_bytes[K_EXPAND] += words * 16.0 + nA * 4.0;
            if (nA == 0) continue;
            prof_bytes[K_LEADERS] += nL * (4 + 8 * NW + 8 + 8 * SW + 1 + 8 + 1);
            prof_bytes[K_FOLLOWERS] += nF * (4 + 8 * NW + 8 + 8 * SW + 8 + 1 + 8) + pairs * (4 + 8 + 1 + 8 + 1);
            prof_bytes[K_UNMATCHED] += nL * 1 + (nL - pairs) * (1 + 8 + 4 + 8 + 1 + 8);
            prof_bytes[K_COUNT] += nA;
            prof_bytes[K_INSERT] += nA + I * (4 + 16 * NW + 8);
            prof_bytes[K_COLS] += I * (8 * NW + ncols / 8.0);
        }
    }

    uint64_t colstride() const { return cap / 64; }

    // Grow the store to hold at least `need` rows, preserving rows, coefficients and columns.
    void ensure_capacity(uint64_t need, bool exact = false) {
        if (need <= cap) return;
        try {
            grow_store(need, exact);
        } catch (const HipFailure &f) {
            throw RunError("out of device memory growing the term store to " + std::to_string(need) + " rows (" +
                           std::to_string(N) + " live terms): " + f.what);
        }
    }

    // Capacities are multiples of the 16384-row granule that grow by at least half each time, so every
    // per-row array of a store, and of the next propagator of the same workload, recurs at the same
    // sizes and the memory pool hands back cached blocks.
    // A fresh store of `c` rows of capacity (a multiple of 64; contents undefined, column index zero).
    void alloc_store(uint64_t c) {
        rows.reserve(c * nw * 8);
        coef.reserve(c * 8);
        cols.reserve(static_cast<size_t>(ncols) * (c / 64) * 8);
        cap = c;
        HIP_TRY(hipMemsetAsync(cols.p, 0, static_cast<size_t>(ncols) * (c / 64) * 8, stream));
        if (nbatch) bmember.reserve(c * 2);
    }

    // Grows the store to `need` rows (exact: the granule above it, else by at least half), first giving
    // back the capacity-sized work arrays when `release` (they are rebuilt); false when the device memory
    // does not hold it (the store is then unchanged).
    bool try_grow(uint64_t need, bool exact, bool release) {
        if (need <= cap) return true;
        if (release) release_work_arrays();
        try {
            ensure_capacity(need, exact);
        } catch (const RunError &) {
            return false;
        }
        return true;
    }

    // Gives back the arrays sized by the capacity other than the store itself (the fused pipeline's,
    // the row index, the per-gate pipeline's); they are sized again by the next propagate.
    void release_work_arrays() {
        for (DevArray *a : {&gidx, &f_elead, &t_val, &t_item, &t_bsum, &t_gsum, &desc_m, &AL, &lead, &foll,
                            &blk_a, &blk_b})
            a->release();
        fused_cap = 0;
        cap_items = 0;
        gidx_slots = 0;
        gidx_rows = static_cast<uint64_t>(-1);
    }

    static constexpr uint64_t kGranule = 64ull * kBlock;

    // Capacity grow_store(need) picks: a multiple of the granule, at least need and 1.5 times the current.
    uint64_t grown_capacity(uint64_t need) const {
        const uint64_t want = std::max<uint64_t>({need, cap + cap / 2, kGranule});
        return (want + kGranule - 1) / kGranule * kGranule;
    }

    // Device bytes per row of capacity the fused pipeline holds: rows, coefficient, column index, row
    // index slots, the emitter regions (t_item + t_val over half the capacity), leader masks.
    double bytes_per_row() const {
        return 8.0 * nw + 8 + ncols / 8.0 + (nbatch ? 2 : 0) + 8.0 * kIdxSlotsPerRow + 6 + 0.125;
    }

    // Scratch of the per-gate pipeline, sized by the capacity (the current gate's item list kept).
    void ensure_gate_arrays() {
        grow_keep(AL, cap * 4);
        lead.reserve((cap / 64) * 8);
        foll.reserve((cap / 64) * 8);
        const uint64_t nblk = cap / 64 / kBlock + 2;
        blk_a.reserve(std::max<uint64_t>(nblk, cap / kBlock + 2) * 4);
        blk_b.reserve(nblk * 4);
    }

    void grow_store(uint64_t need, bool exact = false) {
        const uint64_t nc = exact ? (need + kGranule - 1) / kGranule * kGranule : grown_capacity(need);
        DevArray r2{&stream}, c2{&stream}, k2{&stream};
        r2.reserve(nc * nw * 8);
        c2.reserve(nc * 8);
        k2.reserve(static_cast<size_t>(ncols) * (nc / 64) * 8);
        HIP_TRY(hipMemsetAsync(k2.p, 0, static_cast<size_t>(ncols) * (nc / 64) * 8, stream));
        if (N > 0) {
            HIP_TRY(hipMemcpyAsync(r2.p, rows.p, N * nw * 8, hipMemcpyDeviceToDevice, stream));
            HIP_TRY(hipMemcpyAsync(c2.p, coef.p, N * 8, hipMemcpyDeviceToDevice, stream));
            const uint64_t nwords = (N + 63) / 64;
            HIP_TRY(hipMemcpy2DAsync(k2.p, (nc / 64) * 8, cols.p, colstride() * 8, nwords * 8, ncols,
                                     hipMemcpyDeviceToDevice, stream));
        }
        HIP_TRY(hipStreamSynchronize(stream));
        std::swap(rows.p, r2.p);
        std::swap(rows.bytes, r2.bytes);
        std::swap(coef.p, c2.p);
        std::swap(coef.bytes, c2.bytes);
        std::swap(cols.p, k2.p);
        std::swap(cols.bytes, k2.bytes);
        if (nbatch) {
            DevArray b2{&stream};
            b2.reserve(nc * 2);
            if (N > 0) HIP_TRY(hipMemcpyAsync(b2.p, bmember.p, N * 2, hipMemcpyDeviceToDevice, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            std::swap(bmember.p, b2.p);
            std::swap(bmember.bytes, b2.bytes);
        }
        if (building) {
            DevArray s2{&stream};
            s2.reserve(nc * 8);
            if (N > 0) HIP_TRY(hipMemcpyAsync(s2.p, seedc.p, N * 8, hipMemcpyDeviceToDevice, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            std::swap(seedc.p, s2.p);
            std::swap(seedc.bytes, s2.bytes);
        }
        cap = nc;
        if (AL.p) ensure_gate_arrays();  // the per-gate pipeline is in use on this engine
    }

    void ensure_items(uint64_t nL, uint64_t nA) {
        rec.reserve(std::max<uint64_t>(nL, 1));
        lv.reserve(std::max<uint64_t>(nL, 1) * 8);
        match.reserve(std::max<uint64_t>(nL, 1));
        miss.reserve(std::max<uint64_t>(nA, 1));
        ins.reserve(std::max<uint64_t>(nA, 1) * 8);
        blk_a.reserve((nA / kBlock + 2) * 4);
        uint64_t want = 64;
        while (want < 2 * nL) want <<= 1;
        if (want > table_slots) {
            table.reserve(want * (nw + 1) * 8);
            HIP_TRY(hipMemsetAsync(table.p, 0, table.bytes, stream));
            table_slots = want;
            epoch = 0;  // fresh memory: every tag reads as empty for any epoch >= 1
        }
    }

    // Grow `a` to `need` bytes keeping its contents.
    void grow_keep(DevArray &a, size_t need) {
        if (need <= a.bytes) return;
        DevArray b{&stream};
        b.reserve(std::max(need, a.bytes + a.bytes / 2));
        if (a.p) HIP_TRY(hipMemcpyAsync(b.p, a.p, a.bytes, hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        std::swap(a.p, b.p);
        std::swap(a.bytes, b.bytes);
    }

    // Move `a` into an array of exactly `keep` bytes (its first `keep` bytes), when that saves a tenth.
    void shrink_keep(DevArray &a, size_t keep) {
        if (a.bytes <= keep + keep / 10 + 4096) return;
        DevArray b{&stream};
        size_t want = std::max<size_t>((keep + 4095) & ~size_t{4095}, 4096);
        b.p = BlockCache::get().take(want, stream);
        b.bytes = want;
        if (a.p && keep) HIP_TRY(hipMemcpyAsync(b.p, a.p, std::min(keep, a.bytes), hipMemcpyDeviceToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        std::swap(a.p, b.p);
        std::swap(a.bytes, b.bytes);
    }

    void sync_counters() {
        HIP_TRY(hipMemcpyAsync(h_counters, counters.p, C_COUNT * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
    }

    template <int NW, int PAULI>
    void launch_cols_update(uint64_t start, uint64_t count, bool from_counters, uint64_t count_ub) {
        const uint64_t words = (count_ub + 63) / 64 + 1;
        const uint64_t blocks = (words + kWaves - 1) / kWaves;
        if (blocks == 0) return;
        hipLaunchKernelGGL((k_cols_update<NW, PAULI>), dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0, stream,
                           rows.as<uint64_t>(), cols.as<uint64_t>(), colstride(), n_modes,
                           counters.as<uint64_t>(), start, count, from_counters ? 1 : 0);
        HIP_TRY(hipGetLastError());
    }

    template <int NW, int PAULI>
    void propagate_impl(const std::vector<GateDesc> &gds, const std::vector<int32_t> &folds, bool graph,
                        Trunc trl) {
        const size_t ng = gds.size();
        gdesc.reserve(std::max<size_t>(ng, 1) * sizeof(GateDesc));
        fold.reserve(std::max<size_t>(folds.size(), 1) * 4);
        stats.reserve(std::max<size_t>(ng, 1) * kStatW * 8);
        if (ng) HIP_TRY(hipMemcpyAsync(gdesc.p, gds.data(), ng * sizeof(GateDesc), hipMemcpyHostToDevice, stream));
        if (!folds.empty())
            HIP_TRY(hipMemcpyAsync(fold.p, folds.data(), folds.size() * 4, hipMemcpyHostToDevice, stream));
        ensure_capacity(std::max<uint64_t>(N, 1));
        ensure_gate_arrays();
        uint64_t n_ub = N;  // upper bound of the live count at the current gate
        hipEvent_t ev[2];
        HIP_TRY(hipEventCreate(&ev[0]));
        HIP_TRY(hipEventCreate(&ev[1]));
        HIP_TRY(hipEventRecord(ev[0], stream));
        std::vector<int32_t> fold_plus_piv(ng, 0);
        // in a graph build the seed vector is evolved and the operator's own coefficients get zeros
        uint64_t grot_ub = grot_total;
        if (graph) {
            const int64_t gc[2] = {static_cast<int64_t>(gcos_total), static_cast<int64_t>(grot_total)};
            HIP_TRY(hipMemcpyAsync(counters.as<uint64_t>() + C_GCOS, gc, 16, hipMemcpyHostToDevice, stream));
            HIP_TRY(hipStreamSynchronize(stream));
        }
        for (size_t gi = 0; gi < ng; ++gi) {
            const GateDesc &gd = gds[gi];
            const int igi = static_cast<int>(gi);
            if (gd.identity) {
                hipEvent_t pa = prof_begin();
                hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, stream, counters.as<uint64_t>(),
                                   stats.as<int64_t>(), igi, 1);
                HIP_TRY(hipGetLastError());
                prof_end(K_ADVANCE, pa);
                continue;
            }
            fold_plus_piv[gi] = gd.fold_cnt + (gd.piv0 >= 0) + (gd.piv1 >= 0);
            const uint64_t nwords_ub = (n_ub + 63) / 64;
            const uint32_t sblocks = static_cast<uint32_t>(std::max<uint64_t>((nwords_ub + kBlock - 1) / kBlock, 1));
            hipEvent_t pa = prof_begin();
            hipLaunchKernelGGL(k_scan, dim3(sblocks), dim3(kBlock), 0, stream, cols.as<uint64_t>(), colstride(),
                               fold.as<int32_t>() + gd.fold_off, gd.fold_cnt, gd.piv0, gd.piv1,
                               counters.as<uint64_t>(), lead.as<uint64_t>(), foll.as<uint64_t>(),
                               blk_a.as<uint32_t>(), blk_b.as<uint32_t>());
            HIP_TRY(hipGetLastError());
            prof_end(K_SCAN, pa);
            pa = prof_begin();
            hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, stream, blk_a.as<uint32_t>(),
                               blk_b.as<uint32_t>(), sblocks, counters.as<uint64_t>(), C_NL, C_NF, C_NINS,
                               C_NPAIR + 1);
            HIP_TRY(hipGetLastError());
            prof_end(K_SCANSUM, pa);
            pa = prof_begin();
            hipLaunchKernelGGL(k_expand, dim3(sblocks), dim3(kBlock), 0, stream, lead.as<uint64_t>(),
                               foll.as<uint64_t>(), blk_a.as<uint32_t>(), blk_b.as<uint32_t>(),
                               counters.as<uint64_t>(), AL.as<uint32_t>());
            HIP_TRY(hipGetLastError());
            prof_end(K_EXPAND, pa);
            sync_counters();
            prof_resolve();
            N = static_cast<uint64_t>(h_counters[C_N]);
            if (h_counters[C_ERR]) throw RunError("device hash table overflow");
            const uint64_t nL = static_cast<uint64_t>(h_counters[C_NL]);
            const uint64_t nF = static_cast<uint64_t>(h_counters[C_NF]);
            const uint64_t nA = nL + nF;
            n_ub = N + nA;
            // (graph records keep the phase in bit 31 of the target row)
            if (n_ub >= (uint64_t{1} << (graph ? 31 : 32))) throw RunError("term count exceeds the 32-bit row index");
            ensure_capacity(n_ub);
            ensure_items(nL, nA);
            double *cv = graph ? seedc.as<double>() : coef.as<double>();
            const uint64_t *occ_s = schr ? occ_dev.as<uint64_t>() : nullptr;  // Schrodinger: fresh rows scored
            GraphRec grec{nullptr};
            if (graph && nA > 0) {
                // every anticommuting term is cos-only or an endpoint of at most one rotation, on top
                // of the records written so far (counted on the device, read with the gate's counts)
                grot_ub = static_cast<uint64_t>(h_counters[C_GROT]) + nA;
                grow_keep(g_rot, grot_ub * 8);
                grec = GraphRec{g_rot.as<uint2>()};
            }
            if (nA > 0) {
                ++epoch;
                const uint64_t tmask = table_slots - 1;
                const uint32_t lb = static_cast<uint32_t>((nL + kBlock - 1) / kBlock);
                const uint32_t fb = static_cast<uint32_t>((nF + kBlock - 1) / kBlock);
                const uint32_t ab = static_cast<uint32_t>((nA + kBlock - 1) / kBlock);
                if (lb) {
                    pa = prof_begin();
                    hipLaunchKernelGGL((k_leaders<NW, PAULI>), dim3(lb), dim3(kBlock), 0, stream,
                                       gdesc.as<GateDesc>(), igi, trl, rows.as<uint64_t>(), cv,
                                       AL.as<uint32_t>(), static_cast<uint32_t>(nL), table.as<uint64_t>(), tmask,
                                       epoch, rec.as<uint8_t>(), lv.as<double>(), match.as<uint8_t>(),
                                       counters.as<uint64_t>());
                    HIP_TRY(hipGetLastError());
                    prof_end(K_LEADERS, pa);
                }
                if (fb) {
                    pa = prof_begin();
                    hipLaunchKernelGGL((k_followers<NW, PAULI>), dim3(fb), dim3(kBlock), 0, stream,
                                       gdesc.as<GateDesc>(), igi, trl, rows.as<uint64_t>(), cv,
                                       AL.as<uint32_t>(), static_cast<uint32_t>(nL), static_cast<uint32_t>(nF),
                                       table.as<uint64_t>(), tmask, epoch, rec.as<uint8_t>(), lv.as<double>(),
                                       match.as<uint8_t>(), miss.as<uint8_t>(), ins.as<double>(),
                                       counters.as<uint64_t>(), grec, occ_s);
                    HIP_TRY(hipGetLastError());
                    prof_end(K_FOLLOWERS, pa);
                }
                if (lb) {
                    pa = prof_begin();
                    hipLaunchKernelGGL((k_leaders_unmatched<NW, PAULI>), dim3(lb), dim3(kBlock), 0, stream,
                                       gdesc.as<GateDesc>(), igi, rows.as<uint64_t>(), occ_s, cv, AL.as<uint32_t>(),
                                       static_cast<uint32_t>(nL),
                                       rec.as<uint8_t>(), lv.as<double>(), match.as<uint8_t>(), miss.as<uint8_t>(),
                                       ins.as<double>(), counters.as<uint64_t>(), grec);
                    HIP_TRY(hipGetLastError());
                    prof_end(K_UNMATCHED, pa);
                }
                pa = prof_begin();
                hipLaunchKernelGGL(k_count_miss, dim3(ab), dim3(kBlock), 0, stream, miss.as<uint8_t>(),
                                   static_cast<uint32_t>(nA), blk_a.as<uint32_t>());
                HIP_TRY(hipGetLastError());
                prof_end(K_COUNT, pa);
                pa = prof_begin();
                hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, stream, blk_a.as<uint32_t>(),
                                   static_cast<uint32_t *>(nullptr), ab, counters.as<uint64_t>(), C_NINS, -1, 0, 0);
                HIP_TRY(hipGetLastError());
                prof_end(K_SCANSUM, pa);
                pa = prof_begin();
                hipLaunchKernelGGL((k_insert<NW, PAULI>), dim3(ab), dim3(kBlock), 0, stream, gdesc.as<GateDesc>(), igi,
                                   rows.as<uint64_t>(), cv, graph ? coef.as<double>() : static_cast<double *>(nullptr),
                                   AL.as<uint32_t>(), static_cast<uint32_t>(nA), miss.as<uint8_t>(), ins.as<double>(),
                                   blk_a.as<uint32_t>(), counters.as<uint64_t>(), grec, occ_s);
                HIP_TRY(hipGetLastError());
                prof_end(K_INSERT, pa);
                pa = prof_begin();
                launch_cols_update<NW, PAULI>(0, 0, true, nA);
                prof_end(K_COLS, pa);
            }
            pa = prof_begin();
            hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, stream, counters.as<uint64_t>(), stats.as<int64_t>(),
                               igi, 0);
            HIP_TRY(hipGetLastError());
            prof_end(K_ADVANCE, pa);
        }
        HIP_TRY(hipEventRecord(ev[1], stream));
        sync_counters();
        prof_resolve();
        N = static_cast<uint64_t>(h_counters[C_N]);
        if (h_counters[C_ERR]) throw RunError("device hash table overflow");
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        last_propagate_ms = ms;
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        std::vector<int64_t> st(ng * kStatW);
        if (ng) HIP_TRY(hipMemcpy(st.data(), stats.p, ng * kStatW * 8, hipMemcpyDeviceToHost));
        for (auto *v : {&stat_n, &stat_a, &stat_e, &stat_i, &stat_l, &stat_p}) v->resize(ng);
        for (size_t g = 0; g < ng; ++g) {
            stat_n[g] = st[kStatW * g];
            stat_a[g] = st[kStatW * g + 1];
            stat_e[g] = st[kStatW * g + 2];
            stat_i[g] = st[kStatW * g + 3];
            stat_l[g] = st[kStatW * g + 4];
            stat_p[g] = st[kStatW * g + 5];
        }
        if (prof_on) add_algorithmic_bytes(fold_plus_piv);
        if (graph) {
            uint64_t cb = gcos_total, rb = grot_total;
            for (size_t g = 0; g < ng; ++g) {
                const uint64_t ce = static_cast<uint64_t>(st[kStatW * g + 6]);
                const uint64_t re = static_cast<uint64_t>(st[kStatW * g + 7]);
                LayerInfo ly{cb, ce, rb, re, gds[g].param_index, gds[g].gen_coeff,
                             static_cast<uint64_t>(st[kStatW * g]), static_cast<int32_t>(graph_folds.size()), 0, 0};
                if (ce > cb) {  // the fold columns that recompute its cos-only rows
                    graph_folds.insert(graph_folds.end(), folds.begin() + gds[g].fold_off,
                                       folds.begin() + gds[g].fold_off + gds[g].fold_cnt);
                    ly.fold_cnt = gds[g].fold_cnt;
                }
                layers.push_back(ly);
                cb = ce;
                rb = re;
            }
            gcos_total = cb;
            grot_total = rb;
        }
    }

    // Replay the stored graph at `params` into `rep` (evolve_operator, Evolution.cpp:529-540).
    //   kOperator   (Heisenberg): the operator from the coefficients the graph started at, layers in
    //               processing order, +alpha.
    //   kState      (Schrodinger contract_partially, MonomialPropagator.inl:1069-1083): the state,
    //               oldest layer first, -alpha.
    //   kObservable (Schrodinger functionals, MonomialPropagator.inl:918-1000): the observable's
    //               coefficients, which the caller has put in `rep`, newest layer first (the order the
    //               reference keeps Schrodinger layers in, MPGraph.h:41-43), +alpha.
    enum ReplayKind { kOperator, kState, kObservable };
    void replay_init(ReplayKind kind) {
        rep.reserve(std::max<uint64_t>(N, 1) * 8);
        if (kind == kOperator) {
            if (base_n) HIP_TRY(hipMemcpyAsync(rep.p, base.p, base_n * 8, hipMemcpyDeviceToDevice, stream));
            if (N > base_n) HIP_TRY(hipMemsetAsync(rep.as<double>() + base_n, 0, (N - base_n) * 8, stream));
        } else if (kind == kState && N) {
            HIP_TRY(hipMemcpyAsync(rep.p, coef.p, N * 8, hipMemcpyDeviceToDevice, stream));
        }
    }
    size_t replay_layer_at(ReplayKind kind, size_t pos) const {
        return kind == kObservable ? layers.size() - 1 - pos : pos;
    }
    // The pair-row marks of the cosine recompute: one byte per row, all zero between layers.
    // (two buffers: a layer's launch reads and clears one while it marks the next layer's rows in the other)
    void ensure_pmark() {
        const uint64_t need = std::max<uint64_t>(N, 1);
        for (DevArray *m : {&pmark, &pmark1}) {
            if (m->bytes >= need) continue;
            m->release();
            m->reserve(need);
            HIP_TRY(hipMemsetAsync(m->p, 0, m->bytes, stream));  // all of it: later sizes reuse the block
        }
    }
    // 16-word chunks (one wavefront each) of a layer's cosine rows, their blocks, its rotation blocks.
    static uint64_t cos_chunks(const LayerInfo &ly) {
        return ly.cos_end > ly.cos_begin ? ((ly.n_before + 63) / 64 + kCosChunkWords - 1) / kCosChunkWords : 0;
    }
    static uint32_t cos_blocks(const LayerInfo &ly) {
        return static_cast<uint32_t>((cos_chunks(ly) + kWaves - 1) / kWaves);
    }
    LayerCols layer_cols(const LayerInfo &ly) const {
        LayerCols lc{};
        lc.n = ly.fold_cnt;
        for (int f = 0; f < std::min(ly.fold_cnt, kLayerCols); ++f) lc.col[f] = graph_folds[static_cast<size_t>(ly.fold_off + f)];
        return lc;
    }
    static uint32_t rot_blocks(const LayerInfo &ly) {
        return static_cast<uint32_t>((ly.rot_end - ly.rot_begin + kBlock - 1) / kBlock);
    }
    uint8_t *pm_buf(size_t i) { return (i & 1) ? pmark1.as<uint8_t>() : pmark.as<uint8_t>(); }
    static bool layer_live(const LayerInfo &ly) { return ly.cos_end > ly.cos_begin || ly.rot_end > ly.rot_begin; }
    // One pass over graph layers (`order`: layer indices in processing order), one k_layer launch per
    // layer with records: the replay on `rep` (snapshots of the pre-layer values when `snaps`, indexed
    // from the segment's first cos-only row sc0 and first rotation sr0), or with ADJ the adjoint on
    // `lam` (its derivative partials from gpartial + offs[layer]).  Each launch marks the pair rows of
    // the pass's next layer in the other mark buffer; the first layer's marks go first, on their own.
    template <int ADJ>
    void run_pass(ReplayKind kind, const double *params, const std::vector<size_t> &order, bool snaps, uint64_t sc0,
                  uint64_t sr0, const uint64_t *offs) {
        std::vector<size_t> live;
        for (size_t l : order)
            if (layer_live(layers[l])) live.push_back(l);
        if (live.empty()) return;
        ensure_pmark();
        {
            const LayerInfo &f = layers[live[0]];
            const uint64_t nr = f.rot_end - f.rot_begin;
            if (f.cos_end > f.cos_begin && nr) {
                hipLaunchKernelGGL(k_layer_mark, dim3(rot_blocks(f)), dim3(kBlock), 0, stream, g_rot.as<uint2>(),
                                   f.rot_begin, nr, pm_buf(0), f.n_before);
                HIP_TRY(hipGetLastError());
            }
        }
        for (size_t i = 0; i < live.size(); ++i) {
            const LayerInfo &ly = layers[live[i]];
            LayerArgs a{};
            a.rb = ly.rot_begin;
            a.nrot = ly.rot_end - ly.rot_begin;
            a.n_before = ly.n_before;
            a.nchunks = cos_chunks(ly);
            a.snap_base = ly.cos_begin - sc0;
            a.sr0 = sr0;
            sin_cos(2 * replay_alpha(kind, params, ly), &a.s, &a.k);
            a.lc = layer_cols(ly);
            a.fold = g_fold.as<int32_t>() + ly.fold_off;
            a.cofs = g_cofs.as<uint32_t>() + ly.cofs_off;
            a.snap_c = (ADJ || snaps) ? snap_c.as<double>() : nullptr;
            a.snap_r = (ADJ || snaps) ? snap_r.as<double>() : nullptr;
            a.rot_blocks = rot_blocks(ly);
            a.cos_blocks = cos_blocks(ly);
            uint32_t mark_blocks = 0;
            if (i + 1 < live.size()) {
                const LayerInfo &nx = layers[live[i + 1]];
                if (nx.cos_end > nx.cos_begin) {
                    a.next_rb = nx.rot_begin;
                    a.next_nrot = nx.rot_end - nx.rot_begin;
                    a.next_n_before = nx.n_before;
                    mark_blocks = rot_blocks(nx);
                }
            }
            const uint32_t blocks = a.rot_blocks + a.cos_blocks + mark_blocks;
            if (blocks == 0) continue;
            hipLaunchKernelGGL(k_layer<ADJ>, dim3(blocks), dim3(kBlock), 0, stream, g_rot.as<uint2>(),
                               cols.as<uint64_t>(), colstride(), ADJ ? lam.as<double>() : rep.as<double>(), pm_buf(i),
                               pm_buf(i + 1), a, ADJ ? gpartial.as<double>() + offs[live[i]] : nullptr);
            HIP_TRY(hipGetLastError());
        }
    }
    std::vector<size_t> pass_order(ReplayKind kind, size_t from, size_t to, bool reverse) const {
        std::vector<size_t> o;
        for (size_t i = from; i < to; ++i) o.push_back(replay_layer_at(kind, i));
        if (reverse) std::reverse(o.begin(), o.end());
        return o;
    }
    // After a build: the new layers' fold columns on the device and, per layer and 64-word chunk, the
    // cos-only rows before the chunk (the snapshot positions of the adjoint).  The recomputed count
    // of every layer must equal the count its build recorded.
    void graph_finalize(size_t from) {
        if (!graph_folds.empty()) {
            g_fold.release();
            g_fold.reserve(graph_folds.size() * 4);
            HIP_TRY(hipMemcpyAsync(g_fold.p, graph_folds.data(), graph_folds.size() * 4, hipMemcpyHostToDevice, stream));
        }
        uint64_t total = cofs_total;
        for (size_t l = from; l < layers.size(); ++l) {
            layers[l].cofs_off = total;
            total += cos_chunks(layers[l]);
        }
        if (total == cofs_total) return;
        grow_keep(g_cofs, total * 4);
        ensure_pmark();
        for (size_t l = from; l < layers.size(); ++l) {
            const LayerInfo &ly = layers[l];
            if (cos_chunks(ly) == 0) continue;
            const uint64_t nr = ly.rot_end - ly.rot_begin;
            if (nr) {
                hipLaunchKernelGGL(k_layer_mark, dim3(rot_blocks(ly)), dim3(kBlock), 0, stream, g_rot.as<uint2>(),
                                   ly.rot_begin, nr, pmark.as<uint8_t>(), ly.n_before);
                HIP_TRY(hipGetLastError());
            }
            hipLaunchKernelGGL(k_layer_cos_count, dim3(cos_blocks(ly)), dim3(kBlock), 0, stream, cols.as<uint64_t>(),
                               colstride(), layer_cols(ly), g_fold.as<int32_t>() + ly.fold_off, ly.n_before,
                               pmark.as<uint8_t>(), g_cofs.as<uint32_t>() + ly.cofs_off, cos_chunks(ly));
            HIP_TRY(hipGetLastError());
        }
        std::vector<uint32_t> cnt(total - cofs_total);
        HIP_TRY(hipMemcpyAsync(cnt.data(), g_cofs.as<uint32_t>() + cofs_total, cnt.size() * 4, hipMemcpyDeviceToHost,
                               stream));
        HIP_TRY(hipStreamSynchronize(stream));
        for (size_t l = from; l < layers.size(); ++l) {
            const LayerInfo &ly = layers[l];
            uint64_t acc = 0;
            for (uint64_t b = 0; b < cos_chunks(ly); ++b) {
                uint32_t &v = cnt[ly.cofs_off - cofs_total + b];
                const uint32_t c = v;
                v = static_cast<uint32_t>(acc);
                acc += c;
            }
            if (acc != ly.cos_end - ly.cos_begin)
                throw RunError("graph layer " + std::to_string(l) + ": " + std::to_string(acc) +
                               " recomputed cos-only rows, the build recorded " +
                               std::to_string(ly.cos_end - ly.cos_begin));
        }
        HIP_TRY(hipMemcpyAsync(g_cofs.as<uint32_t>() + cofs_total, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice,
                               stream));
        HIP_TRY(hipStreamSynchronize(stream));
        cofs_total = total;
    }
    double replay_alpha(ReplayKind kind, const double *params, const LayerInfo &ly) const {
        return (kind == kState ? -1.0 : 1.0) * params[ly.param_index] * ly.gen_coeff;
    }
    void replay(const double *params, ReplayKind kind = kOperator) {
        ckpt_valid = false;  // rep no longer holds a checkpointed forward
        replay_init(kind);
        run_pass<0>(kind, params, pass_order(kind, 0, layers.size(), false), false, 0, 0, nullptr);
    }

    // Checkpointed adjoint.  The replay order is cut into segments of about equal snapshot bytes
    // (8 B per cos-only record, 16 B per rotation).  The forward replay keeps the coefficient vector at
    // the start of every segment but the last (ckpt) and the last segment's pre-layer values; the
    // adjoint sweeps the segments last to first, replaying each earlier one again from its checkpoint
    // to regain its pre-layer values.  Every replayed value is the forward's to the bit, so the gradient
    // does not depend on the segmentation.  The segment count is the smallest whose checkpoints plus
    // one segment's snapshots fit grad_budget bytes (else the count that needs the least memory; a
    // zero budget asks for the most segments, one per layer up to 256).
    std::vector<size_t> seg_start;  // positions in the replay order, nseg + 1 entries
    std::vector<uint64_t> seg_sc0, seg_sr0;
    uint64_t seg_snap_c = 0, seg_snap_r = 0;  // largest segment's records
    void plan_segments(ReplayKind kind) {
        const size_t nl = layers.size();
        std::vector<double> cum(nl + 1, 0.0);
        for (size_t i = 0; i < nl; ++i) {
            const LayerInfo &ly = layers[replay_layer_at(kind, i)];
            cum[i + 1] = cum[i] + 8.0 * static_cast<double>(ly.cos_end - ly.cos_begin) +
                         16.0 * static_cast<double>(ly.rot_end - ly.rot_begin);
        }
        const double vec = 8.0 * static_cast<double>(std::max<uint64_t>(N, 1));
        auto cut = [&](size_t S, std::vector<size_t> &st) {  // greedy split, returns the largest segment
            st.assign(1, 0);
            double best = 0.0;
            size_t from = 0;
            for (size_t k = 1; k < S; ++k) {
                const double target = cum[nl] * static_cast<double>(k) / static_cast<double>(S);
                size_t p = static_cast<size_t>(std::lower_bound(cum.begin(), cum.end(), target) - cum.begin());
                p = std::min(std::max(p, from + 1), nl);
                if (p >= nl) break;
                best = std::max(best, cum[p] - cum[from]);
                st.push_back(p);
                from = p;
            }
            st.push_back(nl);
            return std::max(best, cum[nl] - cum[from]);
        };
        std::vector<size_t> st, best_st;
        double best_mem = -1.0;
        const size_t smax = std::max<size_t>(std::min<size_t>(nl, 256), 1);
        if (grad_budget == 0) {  // the most segments: one per layer, up to 256
            (void)cut(smax, st);
            best_st = st;
            best_mem = 0.0;
        }
        for (size_t S = 1; S <= smax && grad_budget != 0; ++S) {
            const double seg = cut(S, st);
            const double mem = static_cast<double>(st.size() - 2) * vec + seg;
            if (best_mem < 0 || mem < best_mem) {
                best_mem = mem;
                best_st = st;
            }
            if (mem <= static_cast<double>(grad_budget)) {
                best_st = st;
                break;
            }
        }
        seg_start = best_st;
        const size_t ns = seg_start.size() - 1;
        seg_sc0.assign(ns, 0);
        seg_sr0.assign(ns, 0);
        seg_snap_c = seg_snap_r = 0;
        for (size_t sg = 0; sg < ns; ++sg) {
            uint64_t c0 = UINT64_MAX, r0 = UINT64_MAX, c1 = 0, r1 = 0;
            for (size_t i = seg_start[sg]; i < seg_start[sg + 1]; ++i) {
                const LayerInfo &ly = layers[replay_layer_at(kind, i)];
                c0 = std::min(c0, ly.cos_begin);
                r0 = std::min(r0, ly.rot_begin);
                c1 = std::max(c1, ly.cos_end);
                r1 = std::max(r1, ly.rot_end);
            }
            if (c0 == UINT64_MAX) c0 = c1 = 0;
            if (r0 == UINT64_MAX) r0 = r1 = 0;
            seg_sc0[sg] = c0;
            seg_sr0[sg] = r0;
            seg_snap_c = std::max(seg_snap_c, c1 - c0);
            seg_snap_r = std::max(seg_snap_r, r1 - r0);
        }
    }
    void replay_segment(ReplayKind kind, const double *params, size_t sg, bool snaps) {
        run_pass<0>(kind, params, pass_order(kind, seg_start[sg], seg_start[sg + 1], false), snaps, seg_sc0[sg],
                    seg_sr0[sg], nullptr);
    }
    // Forward replay with checkpoints (rep must hold the starting vector); afterwards rep holds the
    // evolved vector and the last segment's snapshots are kept.
    void forward_checkpointed(ReplayKind kind, const double *params, int64_t n_params) {
        plan_segments(kind);
        const size_t ns = seg_start.size() - 1;
        snap_c.reserve(std::max<uint64_t>(seg_snap_c, 1) * 8);
        snap_r.reserve(std::max<uint64_t>(seg_snap_r, 1) * 16);
        if (ns > 1) ckpt.reserve((ns - 1) * std::max<uint64_t>(N, 1) * 8);
        for (size_t sg = 0; sg < ns; ++sg) {
            if (sg + 1 < ns && N)
                HIP_TRY(hipMemcpyAsync(ckpt.as<double>() + sg * N, rep.p, N * 8, hipMemcpyDeviceToDevice, stream));
            replay_segment(kind, params, sg, sg + 1 == ns);
        }
        ckpt_kind = kind;
        ckpt_params.assign(params, params + n_params);
        ckpt_valid = true;
    }

    // d<O>/d params: adjoint sweep over the layers in the reverse of the replay's order, after
    // forward_checkpointed (kOperator in the Heisenberg picture, kObservable in the Schrodinger
    // picture, whose adjoint starts from the state instead of the diagonal phases).  Consumes the
    // checkpointed forward (rep is overwritten by the segment replays).
    template <int NW, int PAULI>
    void gradient(const double *params, int64_t n_params, double *grad) {
        const ReplayKind kind = static_cast<ReplayKind>(ckpt_kind);
        ckpt_valid = false;
        lam.reserve(std::max<uint64_t>(N, 1) * 8);
        if (N && schr) {
            HIP_TRY(hipMemcpyAsync(lam.p, coef.p, N * 8, hipMemcpyDeviceToDevice, stream));
        } else if (N) {
            hipLaunchKernelGGL((k_init_lambda<NW, PAULI>), dim3(static_cast<uint32_t>((N + kBlock - 1) / kBlock)),
                               dim3(kBlock), 0, stream, rows.as<uint64_t>(), N, occ_dev.as<uint64_t>(),
                               lam.as<double>());
            HIP_TRY(hipGetLastError());
        }
        const size_t nl = layers.size();
        std::vector<uint64_t> offs(nl + 1, 0);  // per layer: its rotation blocks' partials, then its cosine blocks'
        for (size_t l = 0; l < nl; ++l) offs[l + 1] = offs[l] + rot_blocks(layers[l]) + cos_blocks(layers[l]);
        ensure_pmark();
        gpartial.reserve(std::max<uint64_t>(offs[nl], 1) * 8);
        goffs.reserve((nl + 1) * 8);
        gdl.reserve(std::max<size_t>(nl, 1) * 8);
        const size_t ns = seg_start.size() - 1;
        for (size_t sg = ns; sg-- > 0;) {
            if (sg + 1 < ns) {  // regain this segment's pre-layer values from its checkpoint
                if (N) HIP_TRY(hipMemcpyAsync(rep.p, ckpt.as<double>() + sg * N, N * 8, hipMemcpyDeviceToDevice, stream));
                replay_segment(kind, params, sg, true);
            }
            run_pass<1>(kind, params, pass_order(kind, seg_start[sg], seg_start[sg + 1], true), true, seg_sc0[sg],
                        seg_sr0[sg], offs.data());
        }
        std::vector<double> dl(nl, 0.0);
        if (nl) {
            HIP_TRY(hipMemcpyAsync(goffs.p, offs.data(), (nl + 1) * 8, hipMemcpyHostToDevice, stream));
            hipLaunchKernelGGL(k_reduce_partials, dim3(static_cast<uint32_t>(nl)), dim3(kBlock), 0, stream,
                               gpartial.as<double>(), goffs.as<uint64_t>(), gdl.as<double>());
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(dl.data(), gdl.p, nl * 8, hipMemcpyDeviceToHost, stream));
        }
        HIP_TRY(hipStreamSynchronize(stream));
        for (int64_t p = 0; p < n_params; ++p) grad[p] = 0.0;
        for (size_t l = nl; l-- > 0;) grad[layers[l].param_index] += layers[l].gen_coeff * dl[l];
    }

    // Open-addressing table of the observable's terms (Schrodinger picture), uploaded once.
    template <int NW>
    void ensure_obs_table() {
        if (obs_tmask) return;
        const size_t n = obs_coef.size();
        uint64_t slots = 16;
        while (slots < 2 * std::max<size_t>(n, 1)) slots <<= 1;
        std::vector<uint64_t> tab(slots * (NW + 1), 0);
        for (size_t j = 0; j < n; ++j) {
            const uint64_t *m = obs_rows.data() + j * NW;
            uint64_t h = mpa::hash_key<NW>(m) & (slots - 1);
            while (tab[h * (NW + 1) + NW]) h = (h + 1) & (slots - 1);
            std::copy(m, m + NW, tab.begin() + static_cast<ptrdiff_t>(h * (NW + 1)));
            tab[h * (NW + 1) + NW] = j + 1;
        }
        obs_tab.reserve(tab.size() * 8);
        obs_c.reserve(std::max<size_t>(n, 1) * 8);
        obs_hit.reserve(std::max<size_t>(n, 1) * 8);
        obs_val.reserve(std::max<size_t>(n, 1) * 8);
        HIP_TRY(hipMemcpyAsync(obs_tab.p, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, stream));
        if (n) HIP_TRY(hipMemcpyAsync(obs_c.p, obs_coef.data(), n * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        obs_tmask = slots - 1;
    }

    // Rows holding the observable's terms and the values `c` has there (hit_row -1: no row holds the
    // term); with `scatter`, the observable's coefficient vector over the rows is written there.
    template <int NW>
    void match_observable(const double *c, std::vector<int64_t> &hit_row, std::vector<double> &hit_val,
                          double *scatter) {
        ensure_obs_table<NW>();
        const size_t n = obs_coef.size();
        if (scatter && N) HIP_TRY(hipMemsetAsync(scatter, 0, N * 8, stream));
        if (n) HIP_TRY(hipMemsetAsync(obs_hit.p, 0xff, n * 8, stream));
        if (n && N) {
            hipLaunchKernelGGL((k_obs_match<NW>), dim3(static_cast<uint32_t>((N + kBlock - 1) / kBlock)), dim3(kBlock),
                               0, stream, rows.as<uint64_t>(), c, N, obs_tab.as<uint64_t>(), obs_tmask,
                               obs_hit.as<int64_t>(), obs_val.as<double>(), obs_c.as<double>(), scatter);
            HIP_TRY(hipGetLastError());
        }
        hit_row.assign(n, -1);
        hit_val.assign(n, 0.0);
        if (n) {
            HIP_TRY(hipMemcpyAsync(hit_row.data(), obs_hit.p, n * 8, hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipMemcpyAsync(hit_val.data(), obs_val.p, n * 8, hipMemcpyDeviceToHost, stream));
        }
        HIP_TRY(hipStreamSynchronize(stream));
    }

    // Schrodinger <O> with no graph: core + state . op over the rows in ascending order
    // (MPFunctions.cpp:86-94, 196-205); op is zero away from the observable's rows, so the sum runs
    // over those rows only and is the reference's to the bit.
    template <int NW>
    double observable_expect(const double *c) {
        std::vector<int64_t> hr;
        std::vector<double> hv;
        match_observable<NW>(c, hr, hv, nullptr);
        std::vector<std::pair<int64_t, size_t>> at;
        for (size_t j = 0; j < hr.size(); ++j)
            if (hr[j] >= 0) at.emplace_back(hr[j], j);
        std::sort(at.begin(), at.end());
        double r = 0.0;
        for (const auto &x : at) r += hv[x.second] * obs_coef[x.second];
        return core + r;
    }

    // The reference's ordered sum over rows [0, N) (k_terms): MODE TERM_PAIRED adds phase * a[i] over the
    // fully paired rows, TERM_DOT adds a[i] * b[i] over all rows.  out[0], or out[q] per member q of
    // `member` (nb members).
    // With `fixed` (a device buffer of kNoAllocBytes, one member), the sum allocates nothing: its counts,
    // terms and result live there, and more than kNoAllocTerms nonzero terms is an error.
    template <int NW, int PAULI, int MODE>
    void ordered_sum(const double *a, const double *b, const uint16_t *member, uint32_t nb, double *out,
                     unsigned char *fixed = nullptr) {
        for (uint32_t q = 0; q < nb; ++q) out[q] = 0.0;
        if (N == 0) return;
        const uint64_t nblk = std::min<uint64_t>((N + kBlock - 1) / kBlock, kNoAllocBlocks);
        const uint64_t seg = ((N + nblk - 1) / nblk + kBlock - 1) / kBlock * kBlock;
        const uint32_t grid = static_cast<uint32_t>((N + seg - 1) / seg);
        uint32_t *blk;
        double *vals, *res;
        if (fixed) {
            blk = reinterpret_cast<uint32_t *>(fixed);
            vals = reinterpret_cast<double *>(fixed + ((kNoAllocCountBytes + 255) / 256) * 256);
            res = vals + kNoAllocTerms;
        } else {
            sum_blk.reserve((static_cast<size_t>(grid) + 2) * 4 + 8);
            blk = sum_blk.as<uint32_t>();
        }
        uint64_t *total = reinterpret_cast<uint64_t *>(blk + ((grid + 1) & ~1u));
        const uint64_t *occ_s = occ_dev.as<uint64_t>();
        hipLaunchKernelGGL((k_terms<NW, PAULI, MODE, 0>), dim3(grid), dim3(kBlock), 0, stream, rows.as<uint64_t>(), a,
                           b, member, N, seg, occ_s, blk, static_cast<double *>(nullptr),
                           static_cast<uint16_t *>(nullptr));
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kScanBlock), 0, stream, blk, grid, total);
        HIP_TRY(hipGetLastError());
        // the count goes to the engine's pinned words (no staging buffer is needed for the copy)
        HIP_TRY(hipMemcpyAsync(h_counters + C_COUNT, total, 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        const uint64_t K = static_cast<uint64_t>(h_counters[C_COUNT]);
        if (K == 0) return;
        if (fixed) {
            if (K > kNoAllocTerms)
                throw ArgError("the operator's <O> has " + std::to_string(K) + " nonzero terms, more than the " +
                               std::to_string(kNoAllocTerms) + " the allocation-free entry point holds; use "
                               "mpa_expectation_value");
        } else {
            sum_v.reserve(K * 8);
            if (member) sum_m.reserve(K * 2);
            sum_out.reserve(static_cast<size_t>(nb) * 8);
            vals = sum_v.as<double>();
            res = sum_out.as<double>();
        }
        hipLaunchKernelGGL((k_terms<NW, PAULI, MODE, 1>), dim3(grid), dim3(kBlock), 0, stream, rows.as<uint64_t>(), a,
                           b, member, N, seg, occ_s, blk, vals,
                           member ? sum_m.as<uint16_t>() : static_cast<uint16_t *>(nullptr));
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_seq_sum, dim3(member ? (nb + kBlock - 1) / kBlock : 1), dim3(kBlock), 0, stream, vals,
                           member ? sum_m.as<uint16_t>() : static_cast<uint16_t *>(nullptr), K, nb, res);
        HIP_TRY(hipGetLastError());
        if (fixed) {
            HIP_TRY(hipMemcpyAsync(h_counters + C_COUNT, res, 8, hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            std::memcpy(out, h_counters + C_COUNT, 8);
            return;
        }
        HIP_TRY(hipMemcpyAsync(out, res, static_cast<size_t>(nb) * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
    }

    // core + sum_i a[i] b[i], the reference's state . op (MPFunctions.cpp:86-94), to the bit.
    template <int NW, int PAULI>
    double dot_core(const double *a, const double *b) {
        double r = 0.0;
        ordered_sum<NW, PAULI, TERM_DOT>(a, b, nullptr, 1, &r);
        return core + r;
    }

    // Fresh descriptor memory reads as epoch 0, which no gate uses.
    void clear_descriptors() {
        HIP_TRY(hipMemsetAsync(desc_m.p, 0, desc_m.bytes, stream));
        desc_epoch = 0;
    }

    // Arrays of the fused pipeline sized to the store (a gate runs only while N <= cap / 2), and a global
    // index with >= 2 slots per row of capacity covering every stored row.
    // Per-tile arrays cover the rows a gate can sweep: half the capacity normally (a gate starts at
    // N <= cap / 2), all of it on a tight store.
    template <int NW, int BATCH>
    void ensure_fused_arrays(bool tight = false) {
        if (fused_cap != cap || cap_items != (tight ? cap : cap / 2)) {
            cap_items = tight ? cap : cap / 2;
            const uint64_t tiles = (cap_items + kWTileRows - 1) / kWTileRows;
            t_item.reserve(tiles * kWTileRows * 4);
            t_bsum.reserve(((tiles + kWaves - 1) / kWaves + 1) * 8);
            group_log2 = min_group_log2;
            while (group_log2 < kMaxGroupLog2 && ((tiles + (uint64_t{1} << group_log2) - 1) >> group_log2) + 1 > kMaxGroups)
                ++group_log2;
            gsum_stride = ((tiles + (uint64_t{1} << group_log2) - 1) >> group_log2) + 1;
            if (gsum_stride > kMaxGroups)
                throw RunError("store of " + std::to_string(cap) + " rows exceeds the fused pipeline's limit; "
                               "use the per-gate pipeline (mpa_set_pipeline(e, 0))");
            t_gsum.reserve(2 * gsum_stride * 8);
            HIP_TRY(hipMemsetAsync(t_gsum.p, 0, t_gsum.bytes, stream));  // the sweep accumulates into it
            t_val.reserve(tiles * kWTileRows * 8);
            f_elead.reserve((cap / 64) * 8);
            desc_m.reserve((cap_items / kResChunk + 2) * 8);
            clear_descriptors();
            fused_cap = cap;
        }
        if (desc_epoch >= (1u << 30) - kChunk - 1) clear_descriptors();
        // a gate starts at N <= cap / 2 and ends at N <= cap: load <= 1 / kIdxSlotsPerRow after it
        const uint64_t want = (static_cast<uint64_t>(static_cast<double>(cap) * kIdxSlotsPerRow) + 63) / 64 * 64;
        if (want > gidx_slots) {
            gidx.reserve((want + kIdxPad + kIdxGroup) * 8);
            gidx_slots = want;
            gidx_rows = static_cast<uint64_t>(-1);  // force a rebuild
        }
        if (gidx_rows != N) {
            HIP_TRY(hipMemsetAsync(gidx.p, 0, (gidx_slots + kIdxPad + kIdxGroup) * 8, stream));
            if (N) {
                hipLaunchKernelGGL((k_gidx_build<NW>), dim3(static_cast<uint32_t>((N + kBlock - 1) / kBlock)),
                                   dim3(kBlock), 0, stream, rows.as<uint64_t>(), 0, N,
                                   BATCH ? bmember.as<uint16_t>() : nullptr, gidx.as<uint64_t>(), gidx_slots,
                                   counters.as<uint64_t>());
                HIP_TRY(hipGetLastError());
            }
            gidx_rows = N;
        }
    }

    // Algorithmic bytes of the fused kernels per gate g, the per-unit figure of the propagate path's
    // measurement definition: the anticommutation test reads, for every live term, the limbs that can
    // meet supp(G) (lb bytes per limb unit, w_g units); each anticommuting coefficient is read and
    // written; the anticommuting terms' remaining limbs form the partner keys; each emitted partner
    // record (row + value) is written by the sweep and read back by the resolve; each new term is
    // appended.
    //   k_sweep:   lb*w_g*N_g + 16*A_g + lb*(U - w_g)*A_g + (R + 8)*E_g
    //   k_resolve: (R + 8)*E_g + (R + 8)*I_g
    // lb = 16 (an x and a z limb) for Pauli, 8 (one slot word) for Majorana; U the row's limb units;
    // R = 8*NW the row bytes.  The column index reads far fewer bytes than the lb*w_g*N_g of the
    // test (the PMC traffic counters show what the kernels actually move).
    void add_algorithmic_bytes_fused(const std::vector<GateDesc> &gds) {
        const bool pauli = basis == MPA_BASIS_PAULI;
        const int units = pauli ? L : nw;
        const double lb = pauli ? 16.0 : 8.0, R = 8.0 * nw;
        for (size_t g = 0; g < stat_n.size() && g < gds.size(); ++g) {
            if (gds[g].identity) continue;
            int w = 0;
            for (int u = 0; u < units; ++u) w += pauli ? ((gds[g].g[u] | gds[g].g[L + u]) != 0) : (gds[g].g[u] != 0);
            const double N = static_cast<double>(stat_n[g]), A = static_cast<double>(stat_a[g]);
            const double E = static_cast<double>(stat_e[g]), I = static_cast<double>(stat_i[g]);
            prof_bytes[K_SWEEP] += lb * w * N + 16.0 * A + lb * (units - w) * A + (R + 8.0) * E;
            prof_bytes[K_RESOLVE] += (R + 8.0) * E + (R + 8.0) * I;
        }
    }

    // Live count of gate `g`'s parity, the work ticket and both parities' group counts, as the device
    // pipeline expects them before gate `g`; nothing stopped.
    void reset_fused_counters(size_t g) {
        if (t_gsum.p) HIP_TRY(hipMemsetAsync(t_gsum.p, 0, t_gsum.bytes, stream));
        uint64_t h[C_COUNT];
        HIP_TRY(hipMemcpyAsync(h, counters.p, C_COUNT * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        h[C_NB + (g & 1)] = N;
        h[C_N] = N;
        h[C_STOP] = 0;
        h[C_TK_RES] = 0;
        h[C_ACNT] = h[C_ACNT + 1] = 0;
        HIP_TRY(hipMemcpyAsync(counters.p, h, C_COUNT * 8, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipStreamSynchronize(stream));
    }

    // Contract-immediately propagation without host synchronisation between gates: two persistent
    // launches per gate, enqueued in chunks; the device stops before a gate whose worst case could
    // outgrow the store, the host grows it and resumes there.
    // With `graph` (build_graph, no orbital cap): the seed vector is evolved with the unfused graph
    // arithmetic, each gate's rotations are recorded and its cos-only rows scaled by k_build_cos, and
    // the gates become graph layers (GraphArgs).  `trc` is the truncation of this run.
    template <int NW, int PAULI, int BATCH>
    void propagate_fused(const std::vector<GateDesc> &gds, const std::vector<int32_t> &folds) {
        propagate_fused<NW, PAULI, BATCH>(gds, folds, tr, false);
    }
    template <int NW, int PAULI, int BATCH>
    void propagate_fused(const std::vector<GateDesc> &gds, const std::vector<int32_t> &folds, const Trunc &trc,
                         bool graph) {
        const size_t ng = gds.size();
        gdesc.reserve(std::max<size_t>(ng, 1) * sizeof(GateDesc));
        fold.reserve(std::max<size_t>(folds.size(), 1) * 4);
        stats.reserve(std::max<size_t>(ng, 1) * kStatW * 8);
        if (ng) HIP_TRY(hipMemcpyAsync(gdesc.p, gds.data(), ng * sizeof(GateDesc), hipMemcpyHostToDevice, stream));
        if (!folds.empty())
            HIP_TRY(hipMemcpyAsync(fold.p, folds.data(), folds.size() * 4, hipMemcpyHostToDevice, stream));
        HIP_TRY(hipMemsetAsync(stats.p, 0, std::max<size_t>(ng, 1) * kStatW * 8, stream));
        reset_fused_counters(0);
        if (graph) {  // records continue after the stored ones
            const uint64_t g0 = grot_total;
            HIP_TRY(hipMemcpyAsync(counters.as<uint64_t>() + C_GROT, &g0, 8, hipMemcpyHostToDevice, stream));
            HIP_TRY(hipStreamSynchronize(stream));
            h_counters[C_GROT] = static_cast<int64_t>(g0);
        }
        hipEvent_t ev[2];
        HIP_TRY(hipEventCreate(&ev[0]));
        HIP_TRY(hipEventCreate(&ev[1]));
        HIP_TRY(hipEventRecord(ev[0], stream));
        size_t gi = 0;
        double rate = 1.0;  // per-gate growth of the live count over the last chunk
        // A gate may at most double the live count, so the store is grown to more than 2N before a gate
        // whose worst case it could not take.  A batch past its cap_limit stops there (its driver
        // splits it).  A lone store that cannot have that (out of device memory, or past its
        // store_limit) turns tight: each gate's anticommuting rows are counted first (k_count_anti) and
        // bound what it can append, and the store grows only to what a stopped gate needs.
        bool tight = false;
        uint64_t tight_need = 0;   // live + anticommuting rows of the gate a tight store stopped before
        uint64_t no_grow_cap = 0;  // capacity at which growing past the 2N worst case failed
        // a batch's cap_limit stops the batch (its driver splits it); a lone store's limit is an error.
        // Rows are 32-bit indices (row + 1 in the index slots, 31 bits in graph records).
        const uint64_t row_cap = graph ? (uint64_t{1} << 31) : (uint64_t{1} << 32) - kGranule;
        const uint64_t user_limit = cap_limit ? cap_limit : store_limit;
        const uint64_t limit = user_limit ? std::min(user_limit, row_cap) : row_cap;
        while (gi < ng) {
            if (2 * N > cap && cap_limit) {
                // a batch splits rather than going tight: smaller batches at the 2N rule run faster than
                // a tight store with a count per gate (measured on the parameter sweep)
                const uint64_t need = 2 * N + N / 4;
                if (grown_capacity(need) > limit) break;  // stop before gate gi (the driver splits the batch)
                ensure_capacity(need);
                tight = false;
            } else if (2 * N > cap) {
                const uint64_t need = 2 * N + N / 4;
                const bool allowed = !(limit && grown_capacity(need) > limit);
                if (allowed && no_grow_cap != cap && try_grow(need, false, false)) {
                    tight = false;
                } else {
                    no_grow_cap = cap;
                    tight = true;
                    if (tight_need > cap) {
                        const uint64_t lim = limit ? std::max<uint64_t>(limit / kGranule * kGranule, cap) : UINT64_MAX;
                        if (tight_need > lim) {
                            if (tight_need > row_cap)
                                throw RunError("term count exceeds the 32-bit row index at gate " + std::to_string(gi) +
                                               " (" + std::to_string(N) + " live terms, at most " +
                                               std::to_string(tight_need - N) + " more)");
                            throw RunError("the term store limit of " + std::to_string(store_limit) +
                                           " rows cannot take gate " + std::to_string(gi) + " (" + std::to_string(N) +
                                           " live terms, at most " + std::to_string(tight_need - N) + " more)");
                        }
                        // the largest of five steps from half again the capacity down to the exact need
                        // that device memory holds (each failed try allocates nothing)
                        const uint64_t hi = std::max(std::min(grown_capacity(tight_need), lim), tight_need);
                        bool grown = false;
                        for (int q = 4; q >= 0 && !grown; --q)
                            grown = try_grow(tight_need + (hi - tight_need) / 4 * static_cast<uint64_t>(q), true, true);
                        if (!grown)
                            throw RunError("out of device memory growing the term store to " +
                                           std::to_string(tight_need) + " rows (" + std::to_string(N) +
                                           " live terms, at most " + std::to_string(tight_need - N) +
                                           " more this gate)");
                        no_grow_cap = cap;
                    }
                }
            } else {
                tight = false;
            }
            tight_need = 0;
            ensure_fused_arrays<NW, BATCH>(tight);
            GraphArgs ga{nullptr, nullptr, nullptr, 0};
            if (graph) {
                // room for at least this gate's records (one per live term at most), grown by half
                const uint64_t nrec = static_cast<uint64_t>(h_counters[C_GROT]);
                if (g_rot.bytes / 8 < nrec + 2 * N + 1024) grow_keep(g_rot, (nrec + 2 * N + 1024) * 8);
                const uint64_t need_mark = std::max<uint64_t>(cap, 1);
                for (DevArray *m : {&pmark}) {
                    if (m->bytes >= need_mark) continue;
                    m->release();
                    m->reserve(need_mark);
                    HIP_TRY(hipMemsetAsync(m->p, 0, m->bytes, stream));
                }
                ga = GraphArgs{g_rot.as<uint2>(), pmark.as<uint8_t>(), coef.as<double>(), g_rot.bytes / 8};
            }
            double *cv = graph ? seedc.as<double>() : coef.as<double>();
            // enqueue about as many gates as the store takes at the last chunk's growth rate, so that a
            // device stop leaves few empty launches behind
            size_t span = kChunk;
            if (rate > 1.0 && N > 0) {
                const double room = std::log(static_cast<double>(cap) / (2.0 * static_cast<double>(N)));
                span = static_cast<size_t>(std::clamp(room / std::log(rate) + 8.0, 16.0, static_cast<double>(kChunk)));
            }
            const size_t end = std::min(ng, gi + span);
            const uint64_t n_start = N;
            const uint64_t tiles_ub = (cap_items + kWTileRows - 1) / kWTileRows;
            const uint32_t sgrid = static_cast<uint32_t>(std::max<uint64_t>((tiles_ub + kWaves - 1) / kWaves, 1));
            // a graph build resolves every applied pair, typically thousands of chunks per gate: a wider grid
            const uint64_t rcap = graph ? std::max<uint64_t>(res_grid_cap, kGraphGrid) : res_grid_cap;
            const uint32_t rgrid =
                static_cast<uint32_t>(std::clamp<uint64_t>((cap_items + kResChunk - 1) / kResChunk, 1, rcap));
            const size_t rlds = 2 * (gsum_stride + 1) * sizeof(uint32_t);
            for (size_t g = gi; g < end; ++g) {
                const GateDesc &gd = gds[g];
                const int ig = static_cast<int>(g);
                hipEvent_t pa;
                if (gd.identity) {
                    pa = prof_begin();
                    hipLaunchKernelGGL(k_advance_fused, dim3(1), dim3(64), 0, stream, counters.as<uint64_t>(),
                                       stats.as<int64_t>(), ig);
                    HIP_TRY(hipGetLastError());
                    prof_end(K_ADVANCE, pa);
                    continue;
                }
                const uint32_t ep = ++desc_epoch;
                uint64_t *gsum_cur = t_gsum.as<uint64_t>() + (g & 1) * gsum_stride;
                uint64_t *gsum_nxt = t_gsum.as<uint64_t>() + ((g + 1) & 1) * gsum_stride;
                uint16_t *mem = BATCH ? bmember.as<uint16_t>() : nullptr;
                const Trig *gtrig = BATCH ? btrig.as<Trig>() + g * nbatch : nullptr;
                SweepCols sc{};
                sc.nfold = gd.fold_cnt;
                sc.piv0 = gd.piv0;
                sc.piv1 = gd.piv1;
                {
                    int c = 0;
                    for (int f = 0; f < gd.fold_cnt; ++f, ++c)
                        if (c < kColArgs) sc.col[c] = folds[static_cast<size_t>(gd.fold_off + f)];
                    for (int pv : {gd.piv0, gd.piv1})
                        if (pv >= 0) {
                            if (c < kColArgs) sc.col[c] = pv;
                            ++c;
                        }
                    sc.ncol = c;
                }
                pa = prof_begin();
                if (tight) {
                    hipLaunchKernelGGL(k_count_anti, dim3(std::min<uint32_t>(sgrid, 256)), dim3(kBlock), 0, stream,
                                       cols.as<uint64_t>(), colstride(), sc, fold.as<int32_t>() + gd.fold_off,
                                       counters.as<uint64_t>(), ig);
                    HIP_TRY(hipGetLastError());
                }
                auto *ksw = (!BATCH && graph) ? k_sweep<NW, PAULI, 0, 1> : k_sweep<NW, PAULI, BATCH, 0>;
                hipLaunchKernelGGL(ksw, dim3(sgrid), dim3(kBlock),
                                   BATCH ? static_cast<size_t>(nbatch) * 16 : 0, stream, gd, ig, trc,
                                   cols.as<uint64_t>(), colstride(), sc, fold.as<int32_t>() + gd.fold_off,
                                   rows.as<uint64_t>(), cv, t_val.as<double>(), t_item.as<uint32_t>(),
                                   f_elead.as<uint64_t>(), t_bsum.as<uint64_t>(), gsum_cur,
                                   group_log2, counters.as<uint64_t>(), mem, gtrig, nbatch, cap, ga, tight ? 1 : 0);
                HIP_TRY(hipGetLastError());
                prof_end(K_SWEEP, pa);
                pa = prof_begin();
                auto *kres = (!BATCH && schr) ? k_resolve<NW, PAULI, 0, 1> : k_resolve<NW, PAULI, BATCH, 0>;
                hipLaunchKernelGGL(kres, dim3(rgrid), dim3(kBlock), rlds, stream, gd, ig, ep, rows.as<uint64_t>(), cv, cols.as<uint64_t>(), colstride(),
                                   n_modes, t_item.as<uint32_t>(), t_bsum.as<uint64_t>(),
                                   gsum_cur, gsum_nxt,
                                   gsum_stride, t_val.as<double>(), f_elead.as<uint64_t>(), gidx.as<uint64_t>(), gidx_slots,
                                   desc_m.as<uint64_t>(), counters.as<uint64_t>(), stats.as<int64_t>(),
                                   static_cast<uint32_t>(gsum_stride), group_log2, mem, gtrig,
                                   schr ? occ_dev.as<uint64_t>() : static_cast<uint64_t *>(nullptr), ga);
                HIP_TRY(hipGetLastError());
                prof_end(K_RESOLVE, pa);
                if (graph) {
                    LayerCols lc{};
                    lc.n = gd.fold_cnt;
                    for (int f = 0; f < std::min(gd.fold_cnt, kLayerCols); ++f)
                        lc.col[f] = folds[static_cast<size_t>(gd.fold_off + f)];
                    const uint64_t cchunks = (cap / 64 + kCosChunkWords - 1) / kCosChunkWords;
                    hipLaunchKernelGGL(k_build_cos, dim3(static_cast<uint32_t>((cchunks + kWaves - 1) / kWaves)),
                                       dim3(kBlock), 0, stream, cols.as<uint64_t>(), colstride(), lc,
                                       fold.as<int32_t>() + gd.fold_off, gd.k, cv, pmark.as<uint8_t>(),
                                       counters.as<uint64_t>(), ig);
                    HIP_TRY(hipGetLastError());
                }
            }
            sync_counters();
            prof_resolve();
            if (h_counters[C_ERR]) {
                const int64_t err = h_counters[C_ERR];
                throw RunError(err & static_cast<int64_t>(kErrIndex) ? "device index overflow"
                                                                     : "device pipeline wait limit exceeded");
            }
            const uint64_t stop = static_cast<uint64_t>(h_counters[C_STOP]);
            const size_t next = stop ? static_cast<size_t>(stop - 1) : end;
            N = static_cast<uint64_t>(h_counters[C_NB + (next & 1)]);
            if (stop && tight) tight_need = N + static_cast<uint64_t>(h_counters[C_ACNT + (next & 1)]);
            if (next > gi && n_start > 0 && N > n_start)
                rate = std::pow(static_cast<double>(N) / static_cast<double>(n_start), 1.0 / static_cast<double>(next - gi));
            else
                rate = 1.0;
            gidx_rows = N;
            if (N > row_cap) throw RunError("term count exceeds the 32-bit row index");  // (the bound above prevents it)
            reset_fused_counters(next);
            gi = next;
        }
        fused_done = gi;
        HIP_TRY(hipEventRecord(ev[1], stream));
        HIP_TRY(hipStreamSynchronize(stream));
        prof_resolve();
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        last_propagate_ms = ms;
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        std::vector<int64_t> st(ng * kStatW);
        if (ng) HIP_TRY(hipMemcpy(st.data(), stats.p, ng * kStatW * 8, hipMemcpyDeviceToHost));
        for (auto *v : {&stat_n, &stat_a, &stat_e, &stat_i, &stat_l, &stat_p}) v->resize(ng);
        for (size_t g = 0; g < ng; ++g) {
            stat_n[g] = st[kStatW * g];
            stat_a[g] = st[kStatW * g + 1];
            stat_e[g] = st[kStatW * g + 2];
            stat_i[g] = st[kStatW * g + 3];
            stat_l[g] = st[kStatW * g + 4];
            stat_p[g] = st[kStatW * g + 5];
        }
        if (prof_on) add_algorithmic_bytes_fused(gds);
        if (graph) {
            // per gate: rotations = hits + inserts; cos-only rows = anticommuting - rows in rotations
            // below the live count (2 per hit, 1 per insert)
            uint64_t cb = gcos_total, rb = grot_total;
            for (size_t g = 0; g < ng; ++g) {
                const uint64_t a = static_cast<uint64_t>(stat_a[g]), h = static_cast<uint64_t>(stat_p[g]);
                const uint64_t ins = static_cast<uint64_t>(stat_i[g]);
                if (a < 2 * h + ins) throw RunError("graph build: inconsistent gate counts at gate " + std::to_string(g));
                const uint64_t ce = cb + (a - 2 * h - ins), re = rb + h + ins;
                LayerInfo ly{cb, ce, rb, re, gds[g].param_index, gds[g].gen_coeff, static_cast<uint64_t>(stat_n[g]),
                             static_cast<int32_t>(graph_folds.size()), 0, 0};
                if (ce > cb) {
                    graph_folds.insert(graph_folds.end(), folds.begin() + gds[g].fold_off,
                                       folds.begin() + gds[g].fold_off + gds[g].fold_cnt);
                    ly.fold_cnt = gds[g].fold_cnt;
                }
                layers.push_back(ly);
                cb = ce;
                rb = re;
            }
            sync_counters();
            if (static_cast<uint64_t>(h_counters[C_GROT]) != rb)
                throw RunError("graph build: " + std::to_string(h_counters[C_GROT]) + " rotation records written, " +
                               std::to_string(rb) + " counted");
            gcos_total = cb;
            grot_total = rb;
        }
    }

    // <O> of every batch member (its identity coefficient plus its paired rows), into out[nbatch].
    template <int NW, int PAULI>
    void expect_batch(double *out) {
        std::vector<double> h(nbatch, 0.0);
        ordered_sum<NW, PAULI, TERM_PAIRED>(coef.as<double>(), nullptr, bmember.as<uint16_t>(), nbatch, h.data());
        for (uint32_t q = 0; q < nbatch; ++q) out[q] = bcore[q] + h[q];
    }

    // Stored rows of every batch member, into out[nbatch].
    void member_sizes(int64_t *out) {
        bout.reserve(static_cast<size_t>(nbatch) * 8);
        HIP_TRY(hipMemsetAsync(bout.p, 0, static_cast<size_t>(nbatch) * 8, stream));
        if (N) {
            const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((N + kBlock - 1) / kBlock, kExpectBlocks));
            hipLaunchKernelGGL(k_count_members, dim3(grid), dim3(kBlock), static_cast<size_t>(nbatch) * 4, stream,
                               bmember.as<uint16_t>(), N, nbatch, bout.as<unsigned long long>());
            HIP_TRY(hipGetLastError());
        }
        std::vector<uint64_t> h(nbatch);
        HIP_TRY(hipMemcpyAsync(h.data(), bout.p, nbatch * 8, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        for (uint32_t q = 0; q < nbatch; ++q) out[q] = static_cast<int64_t>(h[q]);
    }

    template <int NW, int PAULI>
    double expect_impl(const double *c) {
        double r = 0.0;
        ordered_sum<NW, PAULI, TERM_PAIRED>(c, nullptr, nullptr, 1, &r);
        return core + r;
    }

    // Runtime word count / basis -> template instantiation.
    template <typename F>
    void dispatch(F &&f) {
        if (basis == MPA_BASIS_PAULI) {
            switch (nw) {
                case 2: f.template operator()<2, 1>(); return;
                case 4: f.template operator()<4, 1>(); return;
                case 6: f.template operator()<6, 1>(); return;
                case 8: f.template operator()<8, 1>(); return;
            }
        } else {
            switch (nw) {
                case 1: f.template operator()<1, 0>(); return;
                case 2: f.template operator()<2, 0>(); return;
                case 3: f.template operator()<3, 0>(); return;
                case 4: f.template operator()<4, 0>(); return;
                case 5: f.template operator()<5, 0>(); return;
                case 6: f.template operator()<6, 0>(); return;
                case 7: f.template operator()<7, 0>(); return;
                case 8: f.template operator()<8, 0>(); return;
            }
        }
        throw ArgError("unsupported row width");
    }

    // Reference slot tuple -> row words (conversion_utils.py:66-88 Pauli slots).
    void slots_to_words(const int32_t *s, int64_t n, uint64_t *out) const {
        std::fill(out, out + nw, 0);
        for (int64_t k = 0; k < n; ++k) {
            const int32_t x = s[k];
            if (x < 0 || x >= 2 * n_modes)
                throw ArgError("Majorana/Pauli index " + std::to_string(x) + " is out of range; must be less than " +
                               std::to_string(2 * n_modes) + ".");
            if (basis == MPA_BASIS_PAULI) {
                const int q = x / 2;
                const uint64_t b = uint64_t{1} << (q & 63);
                out[q >> 6] ^= b;                    // slot 2q (u) and slot 2q+1 (v) both flip x = u ^ v
                if (x & 1) out[L + (q >> 6)] ^= b;  // slot 2q+1 (v) is z
            } else {
                out[x >> 6] |= uint64_t{1} << (x & 63);
            }
        }
    }

    void words_to_slots(const uint64_t *m, std::vector<int32_t> &out) const {
        out.clear();
        if (basis == MPA_BASIS_PAULI) {
            for (int q = 0; q < n_modes; ++q) {
                const uint64_t b = uint64_t{1} << (q & 63);
                const bool x = m[q >> 6] & b, z = m[L + (q >> 6)] & b;
                if (x ^ z) out.push_back(2 * q);
                if (z) out.push_back(2 * q + 1);
            }
            return;
        }
        for (int s = 0; s < 2 * n_modes; ++s)
            if (m[s >> 6] & (uint64_t{1} << (s & 63))) out.push_back(s);
    }

    // Fold and pivot columns of a generator (Scan.h:265-271; Algebra.h:76-78, 108-110).
    void gate_columns(const uint64_t *g, std::vector<int32_t> &fold_out, int32_t &piv0, int32_t &piv1) const {
        piv0 = piv1 = -1;
        if (basis == MPA_BASIS_PAULI) {
            int qmax = -1;
            for (int q = 0; q < n_modes; ++q) {
                const uint64_t b = uint64_t{1} << (q & 63);
                const bool x = g[q >> 6] & b, z = g[L + (q >> 6)] & b;
                if (z) fold_out.push_back(q);            // x_M . z_G
                if (x) fold_out.push_back(n_modes + q);  // z_M . x_G
                if (x || z) qmax = q;
            }
            if (qmax >= 0) {
                const uint64_t b = uint64_t{1} << (qmax & 63);
                if (g[L + (qmax >> 6)] & b) {
                    piv0 = n_modes + qmax;  // Y or Z letter: the pivot is the z bit
                } else {
                    piv0 = qmax;  // X letter: the pivot is u = x ^ z
                    piv1 = n_modes + qmax;
                }
            }
            return;
        }
        int pop = 0, smax = -1;
        for (int s = 0; s < 2 * n_modes; ++s)
            if (g[s >> 6] & (uint64_t{1} << (s & 63))) {
                fold_out.push_back(s);
                ++pop;
                smax = s;
            }
        if (pop & 1) fold_out.push_back(2 * n_modes);  // odd |G|: parity(|M|) correction
        piv0 = smax;
    }
};
