// Kernels of the fused two-launch pipeline: row index, k_sweep, k_resolve.
// Part of the monoprop_amd engine; included by engine.hip (one translation unit).

#pragma once

namespace {

// ---------------------------------------------------------------------------------------------
// Fused contract-immediately pipeline (the default propagate: no orbital cap, cos != 0).
//
// A global open-addressing index over every stored row answers "is M^G stored?" for the emitting
// terms only.  It holds no keys: a slot is one u64, (fingerprint << 32) | (row + 1), 0 = empty, and a
// probe whose fingerprint matches confirms against the row itself (the store is the key array), with
// the row's coefficient loaded in the same round trip.  `member` is the batch member owning the row
// (0 for a lone propagator): it is mixed into the hash and compared with the row's member, so equal
// monomials of different members are different keys.  The home slot of hash h is the high word of
// h * gslots (any slot count works), the fingerprint h's low word; probes run forward from the home
// slot and never wrap (kIdxPad slots past the last home slot absorb the runs; a run past them is an
// index overflow, reported, never silent).  A slot is claimed whole by one CAS, so a reader never sees
// half an entry; lookups skip rows appended during the current gate (they are nobody's partner), so no
// row still being written is compared.
constexpr uint64_t kIdxPad = 1024;
constexpr int kIdxGroup = 4;  // slots read per probe round trip

template <int NW>
__device__ __forceinline__ uint64_t gidx_hash(const uint64_t *key, uint32_t member) {
    return mpa::hash_key<NW>(key) ^ mpa::mix64(static_cast<uint64_t>(member) * 0x9e3779b97f4a7c15ull);
}

__device__ __forceinline__ uint64_t gidx_home(uint64_t h, uint64_t gslots) { return __umul64hi(h, gslots); }

template <int NW>
__device__ __forceinline__ bool gidx_insert(uint64_t *__restrict__ gidx, uint64_t gslots, const uint64_t *key,
                                            uint32_t row, uint32_t member) {
    const uint64_t h = gidx_hash<NW>(key, member);
    const uint64_t val = (h << 32) | (static_cast<uint64_t>(row) + 1);
    const uint64_t end = gslots + kIdxPad;
    for (uint64_t pos = gidx_home(h, gslots); pos < end; ++pos)
        if (atomicCAS(reinterpret_cast<unsigned long long *>(gidx + pos), 0ull, val) == 0ull) return true;
    return false;
}

// Row of (`key`, `member`) among rows [0, limit), or -1; *val receives the row's coefficient and, when
// `flags` is given, *flag_word the word of `flags` (one bit per row) that holds the row's bit, loaded
// in the same round trip.  member_of null: every row is member 0.
template <int NW>
__device__ __forceinline__ int64_t gidx_find_below(const uint64_t *__restrict__ gidx, uint64_t gslots,
                                                   const uint64_t *__restrict__ rows,
                                                   const uint16_t *__restrict__ member_of,
                                                   const double *__restrict__ coef, const uint64_t *key,
                                                   uint64_t limit, uint32_t member, double *val,
                                                   const uint64_t *__restrict__ flags = nullptr,
                                                   uint64_t *flag_word = nullptr) {
    const uint64_t h = gidx_hash<NW>(key, member);
    const uint32_t fp = static_cast<uint32_t>(h);
    const uint64_t end = gslots + kIdxPad;
    for (uint64_t pos = gidx_home(h, gslots); pos < end; pos += kIdxGroup) {
        uint64_t e[kIdxGroup];
#pragma unroll
        for (int i = 0; i < kIdxGroup; ++i) e[i] = gidx[pos + i];  // the array has kIdxGroup slots past end
#pragma unroll
        for (int i = 0; i < kIdxGroup; ++i) {
            if (e[i] == 0 || pos + i >= end) return -1;
            const uint64_t r = (e[i] & 0xffffffffull) - 1;
            if (static_cast<uint32_t>(e[i] >> 32) != fp || r >= limit) continue;
            uint64_t m[NW];
            load_row<NW>(rows, static_cast<uint32_t>(r), m);
            const double c = coef[r];
            const uint32_t q = member_of ? member_of[r] : 0u;
            const uint64_t fw = flags ? flags[r >> 6] : 0;
            bool eq = q == member;
#pragma unroll
            for (int w = 0; w < NW; ++w) eq = eq && (m[w] == key[w]);
            if (eq) {
                *val = c;
                if (flags) *flag_word = fw;
                return static_cast<int64_t>(r);
            }
        }
    }
    return -1;
}

// Index rows [start, start + count); `member` of row r is member[r] (null: every row is member 0).
template <int NW>
__global__ __launch_bounds__(kBlock) void k_gidx_build(const uint64_t *__restrict__ rows, uint64_t start,
                                                       uint64_t count, const uint16_t *__restrict__ member,
                                                       uint64_t *__restrict__ gidx, uint64_t gslots,
                                                       uint64_t *__restrict__ counters) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= count) return;
    uint64_t m[NW];
    load_row<NW>(rows, static_cast<uint32_t>(start + i), m);
    const uint32_t q = member ? member[start + i] : 0u;
    if (!gidx_insert<NW>(gidx, gslots, m, static_cast<uint32_t>(start + i), q))
        atomicOr(reinterpret_cast<unsigned long long *>(counters + C_ERR), 2ull);
}

// Loads of data other blocks of the same kernel wrote: agent-scope atomics read the coherence point
// (the L2 caches of the XCDs are not coherent with each other).
__device__ __forceinline__ uint64_t load_agent(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return lane_u32(wave_incl_scan(v), 63); }

// Ordered prefix sums across the tiles of one launch (single-pass scan with decoupled look-back).
// Descriptor word: epoch (30 bits) | status (2 bits) | value (32 bits); a descriptor of another epoch
// (another gate) reads as not yet published, so the arrays are never cleared between gates.
constexpr uint64_t kDescAgg = 1, kDescIncl = 2;
constexpr uint32_t kSpinLimit = 1u << 22;  // bounded waits: a broken invariant flags an error, never hangs
constexpr uint64_t kErrIndex = 2, kErrLookBack = 4;

__device__ __forceinline__ uint64_t desc_pack(uint32_t ep, uint64_t status, uint32_t value) {
    return (static_cast<uint64_t>(ep) << 34) | (status << 32) | value;
}

__device__ __forceinline__ void desc_publish(uint64_t *d, uint32_t ep, uint64_t status, uint32_t value) {
    __hip_atomic_store(d, desc_pack(ep, status, value), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of tile t over tiles [0, t); called by all 64 lanes of one wave.  Each lane reads
// kLookSpan descriptors per round trip (512 per wave), so a tile whose predecessors have published
// only aggregates still finds its prefix in one or two round trips.
constexpr int kLookSpan = 8;

__device__ uint32_t look_back(const uint64_t *desc, uint32_t t, uint32_t ep, uint64_t *counters) {
    const uint32_t lane = lane_id();
    uint32_t sum = 0, spins = 0;
    int64_t pos = static_cast<int64_t>(t) - 1;
    while (pos >= 0) {
        // lane l covers tiles pos - kLookSpan l - i, i = 0 .. kLookSpan - 1, nearest first
        uint64_t d[kLookSpan];
#pragma unroll
        for (int i = 0; i < kLookSpan; ++i) {
            const int64_t q = pos - static_cast<int64_t>(kLookSpan * lane) - i;
            d[i] = q >= 0 ? load_agent(desc + q) : desc_pack(ep, kDescIncl, 0);  // before tile 0: prefix 0
        }
        uint32_t part = 0;
        int stop_i = kLookSpan;
        bool stop_incl = false;
#pragma unroll
        for (int i = 0; i < kLookSpan; ++i) {
            const uint32_t st = (d[i] >> 34) == ep ? static_cast<uint32_t>((d[i] >> 32) & 3u) : 0u;
            if (stop_i == kLookSpan) {
                if (st == kDescAgg) {
                    part += static_cast<uint32_t>(d[i]);
                } else {
                    stop_i = i;
                    stop_incl = st == kDescIncl;
                    if (stop_incl) part += static_cast<uint32_t>(d[i]);
                }
            }
        }
        const uint64_t has_stop = __ballot(stop_i < kLookSpan);
        if (has_stop == 0) {  // 64 kLookSpan aggregates
            sum += wave_sum(part);
            pos -= 64 * kLookSpan;
            continue;
        }
        const int first = __ffsll(static_cast<unsigned long long>(has_stop)) - 1;
        const int f_incl = static_cast<int>(lane_u32(stop_incl ? 1u : 0u, first));
        const int f_i = static_cast<int>(lane_u32(static_cast<uint32_t>(stop_i), first));
        sum += wave_sum(static_cast<int>(lane) <= first ? part : 0u);
        if (f_incl) break;
        pos -= static_cast<int64_t>(kLookSpan) * first + f_i;  // the first tile not yet published
        // wait on that one descriptor only, then read the window again from it
        int ok = 1;
        if (lane == 0) {
            for (;;) {
                const uint64_t dd = load_agent(desc + pos);
                if ((dd >> 34) == ep && ((dd >> 32) & 3u) != 0) break;
                if (++spins > kSpinLimit) {
                    atomicOr(reinterpret_cast<unsigned long long *>(counters + C_ERR), kErrLookBack);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        if (!lane_u32(static_cast<uint32_t>(ok), 0)) break;
    }
    return sum;
}

// Column bits of an appended term, gathered in LDS per (column, word) for the chunk's words
// [wlo, wlo + 5) (at most 256 consecutive new terms) before one global OR per nonzero word.
template <int NW, int PAULI>
constexpr int max_cols() {
    return PAULI ? 64 * NW : 64 * NW + 1;
}

template <int NW, int PAULI>
__device__ __forceinline__ void gather_column_bits(uint64_t *lcol, int ncols, int n_modes, const uint64_t *m,
                                                   uint64_t rel_word, uint64_t bit) {
    uint64_t *base = lcol + rel_word * ncols;
    if (PAULI) {
        constexpr int L = NW / 2;
#pragma unroll
        for (int l = 0; l < L; ++l) {
            for (uint64_t b = m[l]; b; b &= b - 1)
                atomicOr(reinterpret_cast<unsigned long long *>(base + l * 64 + __ffsll(static_cast<unsigned long long>(b)) - 1), bit);
            for (uint64_t b = m[L + l]; b; b &= b - 1)
                atomicOr(reinterpret_cast<unsigned long long *>(base + n_modes + l * 64 +
                                                                __ffsll(static_cast<unsigned long long>(b)) - 1),
                         bit);
        }
    } else {
        int pop = 0;
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            pop += __popcll(m[k]);
            for (uint64_t b = m[k]; b; b &= b - 1)
                atomicOr(reinterpret_cast<unsigned long long *>(base + k * 64 + __ffsll(static_cast<unsigned long long>(b)) - 1), bit);
        }
        if (pop & 1) atomicOr(reinterpret_cast<unsigned long long *>(base + 2 * n_modes), bit);
    }
}

// sweep: one wavefront per 16 column words.  (8 words, 512 terms, keeps more waves resident and takes
// 4-7% off the sweep when propagations run one at a time, but three propagations on concurrent streams
// ran 2% slower, also when only large stores used it.)
constexpr int kWTileWords = 16;
constexpr int kWTileRows = kWTileWords * 64;              // = 1024 terms
constexpr int kTileLog2 = kWTileWords == 16 ? 10 : kWTileWords == 8 ? 9 : 8;
// Count groups: 2^group_log2 tiles, at least one sweep block (4 tiles) and up to 2^20 rows for the
// largest stores, so that at most kMaxGroups prefixes sit in LDS.  A group of one block lets the resolve
// place an emitter from the prefixes alone.
constexpr int kMinGroupLog2 = 2, kMaxGroupLog2 = 20 - kTileLog2;
constexpr uint32_t kMaxGroups = 4096;
// per-group count word: leaders | followers << 21 | anticommuting << 42 (each <= 2^20)
constexpr uint32_t kGrpBits = 21;
constexpr uint64_t kGrpMask = (uint64_t{1} << kGrpBits) - 1;
constexpr int kResChunk = kBlock;                         // resolve: one emitter per thread
// resolve blocks: each draws one arrival ticket (one atomic on one address), so a grid much larger than
// a gate's chunk count costs more than it gives; chunks past the grid go by further tickets
constexpr uint32_t kMaxGrid = 128;
constexpr uint32_t kGraphGrid = 512;  // resolve blocks of a graph build (measured: 512 beats 128, 1024, 2048)
constexpr int kBsumBatch = 16;                            // resolve: sweep-block sums read per round

// Lane l's x as a wave-uniform (scalar) value; l must be wave-uniform.
__device__ __forceinline__ uint64_t lane_value(uint64_t x, int l) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x)), l));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x >> 32)), l));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Wave-level LDS hand-off between lanes of one wavefront.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Column ids of one gate for the sweep, passed by value: G's fold columns, then its pivot columns.
// Generators with more than kColArgs columns read the rest from the gate's fold list in memory.
constexpr int kColArgs = 8;
struct SweepCols {
    int32_t col[kColArgs];
    int32_t nfold, ncol;  // ncol = nfold + pivot columns
    int32_t piv0, piv1;
};

// Sweep lanes hold (column slot, word): 64 / kWTileWords column slots per pass over the tile's words,
// kColPasses passes for the kColArgs by-value columns.
constexpr int kColSlots = 64 / kWTileWords;
constexpr int kColPasses = kColArgs / kColSlots;

// Column id of slot pass * kColSlots + sl (< kColArgs) for lane slot sl (wave-uniform `pass`): a select
// over the by-value ids, so no memory access is needed to know which column to read.
__device__ __forceinline__ int sweep_col(const SweepCols &sc, int pass, int sl) {
    int c = 0;
    // the ids as values (scalar registers), so the select is over values, not over their addresses
#pragma unroll
    for (int i = 0; i < kColSlots; ++i) {
        const int ci = __builtin_amdgcn_readfirstlane(sc.col[pass * kColSlots + i]);
        c = sl == i ? ci : c;
    }
    return c;
}

// A graph build on the fused kernels (build_graph without an orbital cap): the seed vector is evolved
// with the unfused arithmetic of GraphSink (Engine.h:69-180): the sweep leaves the anticommuting
// coefficients unscaled, the resolve rotates every applied pair from the exact pre-gate values, records
// it in `rot` ({source row, target row | phase bit}, appended at the C_GROT counter; at most rot_cap
// records) and marks its rows below the live count in `pmark`, and k_build_cos then scales the unmarked
// anticommuting rows.  New rows get the operator's own starting value in `zero_fill` (0, or the state
// score in the Schrodinger picture).  rot == nullptr: a plain propagation.
struct GraphArgs {
    uint2 *rot;
    uint8_t *pmark;
    double *zero_fill;
    uint64_t rot_cap;
};

// Store bound of a gate on a tight store (the store could not grow to the 2N worst case): its
// anticommuting rows among the live ones, which bound the rows the gate can append, added into
// counters[C_ACNT + (gi & 1)] (zeroed by the previous gate).  Grid-stride over the column words.
__global__ __launch_bounds__(kBlock) void k_count_anti(const uint64_t *__restrict__ cols, uint64_t colstride,
                                                       SweepCols sc, const int32_t *__restrict__ fold,
                                                       uint64_t *__restrict__ counters, int gi) {
    __shared__ uint32_t lds[kWaves];
    if (counters[C_STOP]) return;
    const uint64_t N = counters[C_NB + (gi & 1)];
    const uint64_t nwords = (N + 63) / 64;
    uint32_t cnt = 0;
    for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; w < nwords;
         w += static_cast<uint64_t>(gridDim.x) * kBlock) {
        uint64_t y[kColArgs];
#pragma unroll
        for (int f = 0; f < kColArgs; ++f) y[f] = f < sc.nfold ? cols[static_cast<uint64_t>(sc.col[f]) * colstride + w] : 0;
        uint64_t a = 0;
#pragma unroll
        for (int f = 0; f < kColArgs; ++f) a ^= y[f];
        for (int f = kColArgs; f < sc.nfold; ++f) a ^= cols[static_cast<uint64_t>(fold[f]) * colstride + w];
        if (w == nwords - 1 && (N & 63)) a &= (uint64_t{1} << (N & 63)) - 1;
        cnt += static_cast<uint32_t>(__popcll(a));
    }
    const uint32_t tot = block_sum(cnt, lds);
    if (threadIdx.x == 0 && tot)
        atomicAdd(reinterpret_cast<unsigned long long *>(counters + C_ACNT + (gi & 1)), static_cast<unsigned long long>(tot));
}

// Anticommutation + fused cos sweep + emission gate (Scan.h:384-401) over one gate.  One wavefront
// (one 64-thread block) per 1024-term tile, no dependence between tiles:
//   A  G's fold and pivot columns of the tile's 16 words (lane = slot * 16 + word), loaded by the kernel
//      before anything waits (xc), XOR-reduced across lanes -> anticommuting mask a_j, pivot mask p_j
//   B  c <- c k on every anticommuting term (the reference's in-place sweep); lane b owns bit b of every
//      word, so at most 16 loads per lane, all in flight; the pre-gate values stay in the lane's
//      registers.  Terms past the lower_atol gate go to the tile's candidate list (LDS)
//   E  candidates dense over the lanes: partner key m ^ G against the structural cutoff (or the
//      upper_atol rescue) -> emit bit (LDS)
//   D  the tile's counts (returned) and leader emit masks per word; the emitters themselves are
//      written by sweep_emit once the block knows every tile's counts
//   BATCH  rows belong to batch members (member[row]); each term takes its member's k and |s| from the
//          block's copy of the gate's per-member constants (s_trig: k, |s| per member, LDS)
template <int NW, int PAULI, int BATCH, int GRAPH>
__device__ __forceinline__ uint64_t sweep_tile(const GateDesc &gd, const Trunc &tr, const uint64_t *__restrict__ cols,
                                               uint64_t colstride, const SweepCols &sc, const int32_t *__restrict__ fold,
                                               const uint64_t *xc, const uint64_t *__restrict__ rows,
                                               double *__restrict__ coef, uint64_t *__restrict__ elead, uint64_t N,
                                               uint64_t tile, uint64_t *__restrict__ sem, uint16_t *__restrict__ cidx,
                                               const uint16_t *__restrict__ member, const double *__restrict__ s_trig,
                                               double *v, uint64_t &p_out) {
    const uint64_t nwords = (N + 63) / 64;
    const uint64_t w0 = tile * kWTileWords;
    const uint32_t lane = lane_id();
    const uint64_t row0 = tile * kWTileRows;
    // ---- A: lanes hold (column slot s = lane / kWTileWords, word j = lane % kWTileWords)
    uint64_t a = 0, p = 0;
    {
        const int j = static_cast<int>(lane & (kWTileWords - 1)), sl = static_cast<int>(lane / kWTileWords);
        const uint64_t w = w0 + j;
#pragma unroll
        for (int pass = 0; pass < kColPasses; ++pass) {
            const int cc = pass * kColSlots + sl;
            if (cc < sc.nfold) a ^= xc[pass];
            else p ^= xc[pass];
        }
        for (int cc0 = kColArgs; cc0 < sc.ncol; cc0 += kColSlots) {  // generators wider than kColArgs columns
            const int cc = cc0 + sl;
            uint64_t y = 0;
            if (w < nwords && cc < sc.ncol) {
                const int col = cc < sc.nfold ? fold[cc] : (cc == sc.nfold ? sc.piv0 : sc.piv1);
                y = cols[static_cast<uint64_t>(col) * colstride + w];
            }
            if (cc < sc.nfold) a ^= y;
            else p ^= y;
        }
#pragma unroll
        for (int d = kWTileWords; d < 64; d <<= 1) {
            a ^= __shfl_xor(a, d, 64);
            p ^= __shfl_xor(p, d, 64);
        }
        if (w == nwords - 1 && (N & 63)) a &= (uint64_t{1} << (N & 63)) - 1;
        if (w >= nwords) a = 0;
    }
    if (lane < kWTileWords) sem[lane] = 0;
    // ---- B: lane owns bit `lane` of each word
    uint32_t cand = 0, resc = 0;  // candidates; those the upper_atol rescue emits regardless
    // v: pre-gate values (kWTileWords), kept for sweep_emit
    {
        // BATCH: the 32-bit word holding the row's 16-bit member (a whole-word load, so the loads of
        // all 16 words stay in flight together)
        uint32_t q[BATCH ? kWTileWords : 1];
        const uint32_t *mw = reinterpret_cast<const uint32_t *>(member);
        uint32_t act = 0;
#pragma unroll
        for (int j = 0; j < kWTileWords; ++j) {
            const uint64_t aj = lane_value(a, j);
            v[j] = 0.0;
            if (BATCH) q[j] = 0;
            if ((aj >> lane) & 1u) {
                const uint64_t r = row0 + static_cast<uint64_t>(j) * 64 + lane;
                v[j] = coef[r];
                if (BATCH) q[j] = mw[r >> 1];
                act |= 1u << j;
            }
        }
#pragma unroll
        for (int j = 0; j < kWTileWords; ++j) {
            if (!((act >> j) & 1u)) continue;
            const uint64_t r = row0 + static_cast<uint64_t>(j) * 64 + lane;
            const uint32_t qm = BATCH ? ((q[BATCH ? j : 0] >> (16 * (lane & 1u))) & 0xffffu) : 0u;  // r & 1 == lane & 1
            const double k = BATCH ? s_trig[2 * qm] : gd.k;
            const double abs_s = BATCH ? s_trig[2 * qm + 1] : gd.abs_s;
            if (!GRAPH) {  // a graph build scales after its resolve (k_build_cos)
                const double nv = v[j] * k;
                if (__double_as_longlong(nv) != __double_as_longlong(v[j])) coef[r] = nv;  // unchanged bits: no store
            }
            if (!(tr.check_lower && abs_s * fabs(v[j]) <= tr.lower)) {
                cand |= 1u << j;
                if (tr.check_upper && abs_s * fabs(v[j]) >= tr.upper) resc |= 1u << j;
            }
        }
    }
    const uint32_t mine = static_cast<uint32_t>(__popc(cand));
    const uint32_t incl = wave_incl_scan(mine);
    const uint32_t ncand = lane_u32(incl, 63);
    {
        uint32_t pos = incl - mine;
#pragma unroll
        for (int j = 0; j < kWTileWords; ++j)
            if ((cand >> j) & 1u) cidx[pos++] = static_cast<uint16_t>(j * 64 + lane) | (((resc >> j) & 1u) << 15);
    }
    wave_lds_sync();
    // ---- E
    for (uint32_t c0 = 0; c0 < ncand; c0 += 128) {
        uint64_t m[2][NW];
        uint32_t id[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint32_t c = c0 + q * 64 + lane;
            id[q] = c < ncand ? cidx[c] : 0xffffu;
            if (id[q] < 0x8000u) load_row<NW>(rows, static_cast<uint32_t>(row0 + id[q]), m[q]);
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (id[q] == 0xffffu) continue;
            bool e = true;  // rescued
            if (id[q] < 0x8000u) {
#pragma unroll
                for (int k = 0; k < NW; ++k) m[q][k] ^= gd.g[k];
                e = keep_cut<NW, PAULI>(m[q], tr);
            }
            const uint32_t i = id[q] & 0x7fffu;
            if (e) atomicOr(reinterpret_cast<unsigned long long *>(&sem[i >> 6]), uint64_t{1} << (i & 63));
        }
    }
    wave_lds_sync();
    // ---- D: per word j the emit mask sem[j] split by the pivot mask p_j (lanes j, j + 16, ... hold it)
    uint32_t nl = 0, nf = 0;
#pragma unroll
    for (int j = 0; j < kWTileWords; ++j) {
        const uint64_t em = sem[j], pj = lane_value(p, j);
        nl += static_cast<uint32_t>(__popcll(em & ~pj));
        nf += static_cast<uint32_t>(__popcll(em & pj));
    }
    const uint32_t na = wave_sum(lane < kWTileWords ? static_cast<uint32_t>(__popcll(a)) : 0u);
    if (lane < kWTileWords && w0 + lane < nwords) elead[w0 + lane] = sem[lane] & ~p;
    p_out = p;
    return static_cast<uint64_t>(nl) | (static_cast<uint64_t>(nf) << kGrpBits) |
           (static_cast<uint64_t>(na) << (2 * kGrpBits));
}

// The tile's emitters into its sweep block's list (t_item: row, t_val: pre-gate value), which holds the
// block's leaders ascending, then its followers ascending (pivot clear / set): lane b writes those of
// bit b of every word, with the values it kept from B, at lpre0 / fpre0 (the block's emitters of that
// kind in earlier tiles; fpre0 includes all the block's leaders) plus per-word popcounts.
__device__ __forceinline__ void sweep_emit(const uint64_t *__restrict__ sem, uint64_t p, const double *v,
                                           uint32_t *__restrict__ region, double *__restrict__ vregion,
                                           uint64_t row0, uint32_t lpre0, uint32_t fpre0) {
    const uint32_t lane = lane_id();
    const uint64_t below = (uint64_t{1} << lane) - 1;
    uint32_t lpre = lpre0, fpre = fpre0;  // list positions before word j: leaders, followers
#pragma unroll
    for (int j = 0; j < kWTileWords; ++j) {
        const uint64_t em = sem[j], pj = lane_value(p, j);
        const uint64_t sel = em & ~pj, sef = em & pj;
        if ((em >> lane) & 1u) {
            const uint32_t pos = ((pj >> lane) & 1u) ? fpre + static_cast<uint32_t>(__popcll(sef & below))
                                                     : lpre + static_cast<uint32_t>(__popcll(sel & below));
            region[pos] = static_cast<uint32_t>(row0 + static_cast<uint64_t>(j) * 64 + lane);
            vregion[pos] = v[j];
        }
        lpre += static_cast<uint32_t>(__popcll(sel));
        fpre += static_cast<uint32_t>(__popcll(sef));
    }
}

// Column words of G's first kColArgs columns for tile `tile` (lane = slot * kWTileWords + word); the
// tiles cover the capacity, so the words are in bounds for any live count.
__device__ __forceinline__ void sweep_load_cols(const uint64_t *__restrict__ cols, uint64_t colstride,
                                                const SweepCols &sc, uint64_t tile, uint64_t *xc) {
    const uint32_t lane = lane_id();
    const int j = static_cast<int>(lane & (kWTileWords - 1)), sl = static_cast<int>(lane / kWTileWords);
    const uint64_t w = tile * kWTileWords + j;
#pragma unroll
    for (int pass = 0; pass < kColPasses; ++pass) {
        const int cc = pass * kColSlots + sl;
        xc[pass] = cc < sc.ncol ? cols[static_cast<uint64_t>(sweep_col(sc, pass, sl)) * colstride + w] : 0;
    }
}

// One 1024-term tile per wavefront, four per block; the block stores its tiles' summed counts
// (bsum[block]) and adds them into their group word (a plain store when the group is this block), and
// lists its emitters in its region of t_item / t_val (block b: entries [4096 b, 4096 b + its
// emitters)), so that the resolve finds emitter f of a block from the group prefixes and the block sums
// alone.  The live count, the stop flag, the tile's column words and (BATCH) the per-member constants
// are all requested before anything waits on any of them.  (A persistent grid of resident blocks walking the tiles, with the next
// tiles' columns prefetched, was measured slower: more registers, and a barrier per tile set.)
template <int NW, int PAULI, int BATCH, int GRAPH>
__global__ __launch_bounds__(kBlock) void k_sweep(GateDesc gdv, int gi, Trunc tr, const uint64_t *__restrict__ cols,
                                                  uint64_t colstride, SweepCols sc, const int32_t *__restrict__ fold,
                                                  const uint64_t *__restrict__ rows, double *__restrict__ coef,
                                                  double *__restrict__ t_val, uint32_t *__restrict__ t_item,
                                                  uint64_t *__restrict__ elead, uint64_t *__restrict__ bsum,
                                                  uint64_t *__restrict__ gsum,
                                                  int group_log2, uint64_t *__restrict__ counters,
                                                  const uint16_t *__restrict__ member, const Trig *__restrict__ gtrig,
                                                  uint32_t nbatch, uint64_t cap_rows, GraphArgs ga, int tight) {
    __shared__ uint64_t sem[kWaves][kWTileWords];
    __shared__ uint16_t cidx[kWaves][kWTileRows];
    __shared__ uint64_t s_grp[kWaves];
    extern __shared__ double s_trig[];  // BATCH: k, |s| per member
    const uint64_t stop = counters[C_STOP];
    const uint64_t N = counters[C_NB + (gi & 1)];
    const uint64_t nrec = GRAPH ? counters[C_GROT] : 0;
    const uint64_t acnt = counters[C_ACNT + (gi & 1)];
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t tile0 = static_cast<uint64_t>(blockIdx.x) * kWaves;
    const uint64_t tile = tile0 + wave;
    uint64_t xc[kColPasses];
    sweep_load_cols(cols, colstride, sc, tile, xc);
    if (BATCH) {
        for (uint32_t q = threadIdx.x; q < nbatch; q += kBlock) {
            s_trig[2 * q] = gtrig[q].k;
            s_trig[2 * q + 1] = gtrig[q].abs_s;
        }
        __syncthreads();
    }
    if (stop) return;
    // a gate may double the live count (on a tight store: add at most its anticommuting rows, counted by
    // k_count_anti): stop before it when the store could not take that (the host grows the store and
    // resumes here); a graph build also stops when this gate's records (at most one per live term) might
    // not fit
    if ((tight ? N + acnt : 2 * N) > cap_rows || (GRAPH && nrec + N > ga.rot_cap)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) counters[C_STOP] = static_cast<uint64_t>(gi) + 1;
        return;
    }
    // the next gate's live count starts at N; the resolve's chunks add their misses to it
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[C_TK_RES] = 0;
        counters[C_NB + ((gi & 1) ^ 1)] = N;
        counters[C_ACNT + ((gi & 1) ^ 1)] = 0;
    }
    const uint64_t nwords = (N + 63) / 64;
    if (tile0 * kWTileWords >= nwords) return;  // whole block idle
    uint64_t grp = 0, p = 0;
    double v[kWTileWords];
    const bool live = tile * kWTileWords < nwords;
    if (live)
        grp = sweep_tile<NW, PAULI, BATCH, GRAPH>(gdv, tr, cols, colstride, sc, fold, xc, rows, coef, elead, N, tile,
                                                  sem[wave], cidx[wave], member, s_trig, v, p);
    if (lane_id() == 0) s_grp[wave] = grp;
    __syncthreads();
    uint64_t g = 0;
    uint32_t lpre = 0, fpre = 0;  // this tile's list offsets: the block's leaders / followers in earlier tiles
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint64_t gw = s_grp[w];
        g += gw;
        if (static_cast<uint32_t>(w) < wave) {
            lpre += static_cast<uint32_t>(gw & kGrpMask);
            fpre += static_cast<uint32_t>((gw >> kGrpBits) & kGrpMask);
        }
    }
    if (live && (grp & ((kGrpMask << kGrpBits) | kGrpMask)))
        sweep_emit(sem[wave], p, v, t_item + tile0 * kWTileRows, t_val + tile0 * kWTileRows, tile * kWTileRows, lpre,
                   static_cast<uint32_t>(g & kGrpMask) + fpre);
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = g;
        // a group of one block has this block as its only writer: a store; larger groups add up
        if (g) {
            if (group_log2 == 2) gsum[blockIdx.x] = g;
            else atomicAdd(reinterpret_cast<unsigned long long *>(gsum + (tile0 >> group_log2)), g);
        }
    }
}

// Partner resolution of every emitter (Engine.h:364-387, 465-513) and the ordered append of the
// partners to create (leader misses then follower misses, ascending, Engine.h:428-444).
//   positions  every block turns the sweep's per-group counts into prefixes (LDS); list position f
//              lies in group g (binary search) and in one of its sweep blocks (the group itself when it
//              is one block, else the blocks' count sums): that block's list holds its row
//   resolution one emitter per thread, 256-emitter chunks (leader chunks, then follower chunks) by work
//              ticket.  A follower whose partner leader emits is rotated by that leader (the
//              reference's matched-follower skip).  A hit applies the rotation against the swept
//              partner value recovered as (v k)(1/k) (Engine.h:209-211, FusedApply.h:61-72); a miss
//              writes the new term's coefficient k*0 + (s ph) v
//   append     a chunk's misses land at N + (misses of earlier chunks) by look-back; new terms enter
//              the row index, and their column bits are gathered per word in LDS before one OR per
//              word.  A block appends a chunk one iteration late: it resolves its next chunk first
//              (misses staged in LDS), so the look-back rarely waits on chunks still resolving
// The next gate's live count N + I is the sweep's N plus every chunk's misses (atomic adds: a chunk
// that creates no term neither looks back nor waits); the next gate's sweep stops the device when its
// worst case could outgrow the store or the index.  Every block also clears its share of the other
// gate parity's group counts.
// With BATCH, an emitter's member (member[S]) selects its rotation constants (gtrig[member]) and keys its
// index lookup; the partner it creates joins the same member.
template <int NW, int PAULI, int BATCH, int SCHR>
__global__ __launch_bounds__(kBlock) void k_resolve(GateDesc gd, int gi, uint32_t ep,
                                                    uint64_t *__restrict__ rows, double *__restrict__ coef,
                                                    uint64_t *__restrict__ cols, uint64_t colstride, int n_modes,
                                                    const uint32_t *__restrict__ t_item, const uint64_t *__restrict__ bsum,
                                                    const uint64_t *__restrict__ gsum, uint64_t *__restrict__ gsum_next,
                                                    uint64_t n_groups_next, const double *__restrict__ t_val,
                                                    const uint64_t *__restrict__ elead, uint64_t *__restrict__ gidx,
                                                    uint64_t gslots, uint64_t *__restrict__ desc_m,
                                                    uint64_t *__restrict__ counters, int64_t *__restrict__ stats,
                                                    uint32_t max_groups,
                                                    int group_log2, uint16_t *__restrict__ member,
                                                    const Trig *__restrict__ gtrig,
                                                    const uint64_t *__restrict__ occ_dev, GraphArgs ga) {
    constexpr int NC = max_cols<NW, PAULI>();
    constexpr int kColWords = (kResChunk + 63) / 64 + 1;  // words a chunk's consecutive new terms can touch
    __shared__ uint64_t lcol[kColWords * NC];
    extern __shared__ uint32_t dyn_lds[];  // group prefixes: leaders [0, max_groups], followers after
    __shared__ uint32_t lds[kWaves];
    __shared__ uint64_t lds64[kWaves];
    __shared__ uint32_t s_ticket, s_base, s_first;
    __shared__ uint32_t st_row[2][kResChunk];  // a chunk's misses: emitter row, new coefficient
    __shared__ double st_ins[2][kResChunk];
    __shared__ uint8_t st_ph[2][kResChunk];     // graph build: a miss's phase bit
    __shared__ uint64_t s_hbase, s_rbase;        // graph build: first record of a chunk's hits / misses
    uint32_t *gpl = dyn_lds, *gpf = dyn_lds + max_groups + 1;
    const int ncols = PAULI ? 2 * n_modes : 2 * n_modes + 1;
    // The first chunk goes by a ticket drawn on arrival: a chunk's predecessors then belong to blocks
    // that arrived earlier and are running, so the look-back always ends.  Block index order would not
    // do: blocks are dispatched per XCD, and an XCD busy with other work (kernels of other streams) can
    // hold back a lower-indexed block while a higher-indexed one already waits on it.  The ticket, the
    // stop flag, the live count and every group word of the capacity (groups past this gate's are zero)
    // are all requested before anything waits on them.
    uint64_t first_ticket = 0;
    if (threadIdx.x == 0)
        first_ticket = atomicAdd(reinterpret_cast<unsigned long long *>(counters + C_TK_RES), 1ull);
    const uint64_t stop = counters[C_STOP];
    const uint64_t N = counters[C_NB + (gi & 1)];
    const uint32_t gper = (max_groups + kBlock - 1) / kBlock;  // <= kMaxGroups / kBlock
    const uint32_t g0 = threadIdx.x * gper;
    uint64_t gv[kMaxGroups / kBlock];
#pragma unroll
    for (uint32_t i = 0; i < kMaxGroups / kBlock; ++i) gv[i] = (i < gper && g0 + i < max_groups) ? gsum[g0 + i] : 0;
    if (stop) return;
    for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; g < n_groups_next;
         g += static_cast<uint64_t>(gridDim.x) * kBlock)
        gsum_next[g] = 0;
    const int nxt = (gi & 1) ^ 1;
    const uint64_t ntiles = ((N + 63) / 64 + kWTileWords - 1) / kWTileWords;
    const uint32_t ng = static_cast<uint32_t>((ntiles + (uint64_t{1} << group_log2) - 1) >> group_log2);
    const uint64_t nblocks = (ntiles + kWaves - 1) / kWaves;  // sweep blocks of this gate
    // ---- group prefixes (every block): thread t sums its gper consecutive groups, one block scan of
    // (leaders | followers << 32) places them
    uint64_t mine = 0;
    uint32_t mine_a = 0;
#pragma unroll
    for (uint32_t i = 0; i < kMaxGroups / kBlock; ++i) {
        mine += (gv[i] & kGrpMask) | (((gv[i] >> kGrpBits) & kGrpMask) << 32);
        mine_a += static_cast<uint32_t>(gv[i] >> (2 * kGrpBits));
    }
    if (threadIdx.x == 0) s_first = static_cast<uint32_t>(first_ticket);
    uint64_t tot;
    uint32_t nA;
    uint64_t x = block_excl_scan64_sum(mine, mine_a, lds64, lds, &tot, &nA);  // (its barriers also publish s_first)
    const uint32_t nEL = static_cast<uint32_t>(tot), nEF = static_cast<uint32_t>(tot >> 32);
#pragma unroll
    for (uint32_t i = 0; i < kMaxGroups / kBlock; ++i) {
        if (i < gper && g0 + i <= ng) {
            gpl[g0 + i] = static_cast<uint32_t>(x);
            gpf[g0 + i] = static_cast<uint32_t>(x >> 32);
        }
        x += (gv[i] & kGrpMask) | (((gv[i] >> kGrpBits) & kGrpMask) << 32);
    }
    const uint32_t cL = (nEL + kResChunk - 1) / kResChunk;
    const uint32_t nchunks = cL + (nEF + kResChunk - 1) / kResChunk;
    int64_t *st = stats + static_cast<int64_t>(gi) * kStatW;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[C_NL] = nEL;
        counters[C_NF] = nEF;
        st[0] = static_cast<int64_t>(N);
        st[1] = static_cast<int64_t>(nA);
        st[2] = static_cast<int64_t>(nEL) + static_cast<int64_t>(nEF);
        st[4] = static_cast<int64_t>(nEL);
    }
    const uint32_t ticket0 = s_first;
    if (ticket0 >= nchunks) return;  // only blocks with work take chunks
    // Chunk c is the one of ticket c: a block's arrival ticket first, then (only when there are more
    // chunks than blocks: every block draws one arrival ticket, so otherwise those cover all chunks)
    // further tickets from the same counter.  Every chunk publishes its aggregate in the iteration that
    // draws it, before its block waits on anything.  One chunk in flight behind the current one: chunk
    // c is resolved and its misses staged (LDS) while the previous chunk of this block, whose look-back
    // has had chunk c's time to settle, is appended.
    const bool more = nchunks > gridDim.x;
    uint32_t pend = 0xffffffffu, pend_miss = 0, buf = 0;
    for (bool first = true;; first = false) {
        __syncthreads();  // LDS of the previous iteration fully consumed (and the group prefixes written)
        uint32_t chunk = ticket0;  // the first chunk's ticket is known to every thread
        if (!first) {
            if (threadIdx.x == 0)
                s_ticket = !more ? 0xffffffffu
                                 : static_cast<uint32_t>(atomicAdd(
                                       reinterpret_cast<unsigned long long *>(counters + C_TK_RES), 1ull));
            __syncthreads();
            chunk = s_ticket;
        }
        const bool have = chunk < nchunks;
        if (!have && pend == 0xffffffffu) break;
        uint32_t n_miss = 0;
        if (have) {
            // ---- positions: this thread's emitter f of side L (leaders) or F (followers)
            const bool leader = chunk < cL;
            const uint32_t *gp = leader ? gpl : gpf;
            const uint32_t total = leader ? nEL : nEF;
            const uint32_t first = (leader ? chunk : chunk - cL) * kResChunk;
            const uint32_t cnt = std::min<uint32_t>(kResChunk, total - first);
            const uint32_t f = first + threadIdx.x;
            const bool valid = threadIdx.x < cnt;
            // group of f: binary search over the group prefixes (LDS); a group of one sweep block is that
            // block, a larger one is searched through its blocks' count sums (a round of independent loads)
            uint32_t S = 0;
            uint64_t sidx = 0;
            double vS = 0.0;
            if (valid) {
                uint32_t lo = 0, hi = ng;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (gp[mid] <= f) lo = mid;
                    else hi = mid;
                }
                uint32_t rem = f - gp[lo];
                uint64_t blk = lo;
                // the block's leaders (its followers' list starts after them)
                uint32_t nl_b = gpl[lo + 1] - gpl[lo];
                if (group_log2 > 2) {
                    // groups of several sweep blocks: the block sums in rounds of independent loads
                    const uint32_t blocks_per_group = 1u << (group_log2 - 2);
                    const uint64_t b0 = static_cast<uint64_t>(lo) << (group_log2 - 2);
                    for (uint32_t base = 0; base < blocks_per_group; base += kBsumBatch) {
                        uint32_t cb[kBsumBatch], cl[kBsumBatch];
#pragma unroll
                        for (int k = 0; k < kBsumBatch; ++k) {
                            const uint64_t bb = b0 + base + k;
                            const uint64_t v = (base + k < blocks_per_group && bb < nblocks) ? bsum[bb] : 0;
                            cl[k] = static_cast<uint32_t>(v & kGrpMask);
                            cb[k] = leader ? cl[k] : static_cast<uint32_t>((v >> kGrpBits) & kGrpMask);
                        }
                        bool found = false;
#pragma unroll
                        for (int k = 0; k < kBsumBatch; ++k) {
                            if (found) continue;
                            if (rem < cb[k]) {
                                blk = b0 + base + k;
                                nl_b = cl[k];
                                found = true;
                            } else {
                                rem -= cb[k];
                            }
                        }
                        if (found) break;
                    }
                }
                sidx = blk * (kWaves * kWTileRows) + (leader ? 0u : nl_b) + rem;
                S = t_item[sidx];
                vS = t_val[sidx];
            }
            // ---- resolution
            double ins = 0.0;
            uint32_t miss = 0, hit = 0, mph = 0;
            uint2 hrec = make_uint2(0u, 0u);
            if (valid) {
                const uint32_t q = BATCH ? member[S] : 0u;
                uint64_t m[NW], t[NW];
                load_row<NW>(rows, S, m);
#pragma unroll
                for (int k = 0; k < NW; ++k) t[k] = m[k] ^ gd.g[k];
                double wT = 0.0;
                uint64_t lw = 0;  // followers: the partner's word of emitting-leader bits, with its coefficient
                const int64_t T = gidx_find_below<NW>(gidx, gslots, rows, BATCH ? member : nullptr, coef, t, N, q, &wT,
                                                      leader ? nullptr : elead, &lw);
                const bool skip = T >= 0 && !leader && ((lw >> (T & 63)) & 1u);
                if (!skip) {
                    const double gk = BATCH ? gtrig[q].k : gd.k;
                    const double gs = BATCH ? gtrig[q].s : gd.s;
                    const double ginv = BATCH ? gtrig[q].inv_k : gd.inv_k;
                    const int ph = mpa::emit_phase<NW, PAULI>(m, gd.g, t);
                    const double wS = vS * gk;
                    if (T >= 0) {
                        // fused: the stored partner value was swept (k v_T), recovered as (k v_T)(1/k);
                        // graph build: it is the pre-gate value (the gate's 1/k is 1)
                        const double vTu = wT * ginv;
                        coef[S] = wS + gs * static_cast<double>(-ph) * vTu;
                        coef[T] = (ga.rot ? wT * gk : wT) + gs * static_cast<double>(ph) * vS;
                        hit = 1;
                        if (ga.rot) {
                            ga.pmark[S] = 1;
                            ga.pmark[T] = 1;
                            hrec = make_uint2(S, static_cast<uint32_t>(T) | (ph < 0 ? kPhaseBit : 0u));
                        }
                    } else {  // the new partner starts at vT0: 0, or its state score (SCHR)
                        const double vT0 = SCHR ? fresh_score<NW, PAULI>(t, occ_dev) : 0.0;
                        coef[S] = wS + gs * static_cast<double>(-ph) * vT0;
                        ins = gk * vT0 + gs * static_cast<double>(ph) * vS;
                        miss = 1;
                        mph = ph < 0 ? 1u : 0u;
                        if (ga.rot) ga.pmark[S] = 1;
                    }
                }
            }
            const uint32_t rank = block_excl_scan(miss, lds, &n_miss);
            uint32_t n_hit;
            const uint32_t hrank = block_excl_scan(hit, lds, &n_hit);
            if (miss) {  // staged in the chunk's rank order
                st_row[buf][rank] = S;
                st_ins[buf][rank] = ins;
                st_ph[buf][rank] = static_cast<uint8_t>(mph);
            }
            if (ga.rot && n_hit) {  // the chunk's rotations of existing pairs, one counter add per chunk
                if (threadIdx.x == 0)
                    s_hbase = atomicAdd(reinterpret_cast<unsigned long long *>(counters + C_GROT),
                                        static_cast<unsigned long long>(n_hit));
                __syncthreads();
                if (hit) ga.rot[s_hbase + hrank] = hrec;
            }
            if (threadIdx.x == 0) {
                desc_publish(desc_m + chunk, ep, chunk == 0 ? kDescIncl : kDescAgg, n_miss);
                if (n_hit) atomicAdd(reinterpret_cast<unsigned long long *>(st + 5), static_cast<unsigned long long>(n_hit));
                if (n_miss) {  // the next gate's live count and this gate's insert count (no order needed)
                    atomicAdd(reinterpret_cast<unsigned long long *>(counters + C_NB + nxt),
                              static_cast<unsigned long long>(n_miss));
                    atomicAdd(reinterpret_cast<unsigned long long *>(st + 3), static_cast<unsigned long long>(n_miss));
                }
            }
        }
        // A chunk that creates no term has nothing to place: its aggregate (0) is all its successors need.
        if (pend != 0xffffffffu && pend_miss != 0) {
            // ---- append the previous chunk: its misses land at N + (misses of earlier chunks) + rank
            const uint32_t pb = buf ^ 1u;
            // the emitters' rows are requested before the look-back, so both round trips overlap
            uint64_t t[NW];
            uint32_t q = 0;
            if (threadIdx.x < pend_miss) {
                const uint32_t S = st_row[pb][threadIdx.x];
                q = BATCH ? member[S] : 0u;
                load_row<NW>(rows, S, t);
            }
            if (threadIdx.x < 64) {
                uint32_t excl = 0;
                if (pend != 0) {
                    excl = look_back(desc_m, pend, ep, counters);
                    if (threadIdx.x == 0) desc_publish(desc_m + pend, ep, kDescIncl, excl + pend_miss);
                }
                if (threadIdx.x == 0) {
                    s_base = excl;
                    if (ga.rot)  // the misses' rotation records
                        s_rbase = atomicAdd(reinterpret_cast<unsigned long long *>(counters + C_GROT),
                                            static_cast<unsigned long long>(pend_miss));
                }
            }
            for (int k = threadIdx.x; k < kColWords * NC; k += kBlock) lcol[k] = 0;
            __syncthreads();
            const uint64_t first_row = N + s_base, wlo = first_row >> 6;
            if (threadIdx.x < pend_miss) {
#pragma unroll
                for (int k = 0; k < NW; ++k) t[k] ^= gd.g[k];
                const uint64_t dest = first_row + threadIdx.x;
                uint64_t *pd = rows + dest * NW;
#pragma unroll
                for (int k = 0; k < NW; ++k) pd[k] = t[k];
                coef[dest] = st_ins[pb][threadIdx.x];
                if (BATCH) member[dest] = static_cast<uint16_t>(q);
                if (ga.rot) {
                    ga.zero_fill[dest] = SCHR ? fresh_score<NW, PAULI>(t, occ_dev) : 0.0;
                    ga.rot[s_rbase + threadIdx.x] =
                        make_uint2(st_row[pb][threadIdx.x],
                                   static_cast<uint32_t>(dest) | (st_ph[pb][threadIdx.x] ? kPhaseBit : 0u));
                }
                gather_column_bits<NW, PAULI>(lcol, ncols, n_modes, t, (dest >> 6) - wlo, uint64_t{1} << (dest & 63));
                if (!gidx_insert<NW>(gidx, gslots, t, static_cast<uint32_t>(dest), q))
                    atomicOr(reinterpret_cast<unsigned long long *>(counters + C_ERR), kErrIndex);
            }
            __syncthreads();
            if (pend_miss) {
                const int nwt = static_cast<int>(((first_row + pend_miss - 1) >> 6) - wlo + 1);
                for (int k = threadIdx.x; k < nwt * ncols; k += kBlock) {
                    const uint64_t bits = lcol[k];
                    if (bits)
                        atomicOr(reinterpret_cast<unsigned long long *>(cols + static_cast<uint64_t>(k % ncols) * colstride +
                                                                        wlo + k / ncols),
                                 bits);
                }
            }
        }
        pend = have ? chunk : 0xffffffffu;
        pend_miss = n_miss;
        buf ^= 1u;
    }
}

// Identity gate of the fused pipeline: the live count carries over to the next gate parity.
__global__ void k_advance_fused(uint64_t *__restrict__ counters, int64_t *__restrict__ stats, int gi) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (counters[C_STOP]) return;
    int64_t *s = stats + static_cast<int64_t>(gi) * kStatW;
    s[0] = static_cast<int64_t>(counters[C_NB + (gi & 1)]);
    counters[C_NB + ((gi & 1) ^ 1)] = counters[C_NB + (gi & 1)];
    counters[C_ACNT + ((gi & 1) ^ 1)] = 0;
}

}  // namespace
