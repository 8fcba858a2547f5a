// monoprop_amd engine: Heisenberg operator propagation on one MI355X (gfx950).
//
// Term store in HBM (append-only, insertion order = the reference's row order):
//   rows  : AoS, NW u64 words per term (layout in algebra.h)            -- gathered per anticommuting term
//   coef  : f64 per term
//   cols  : bit-sliced column index, one u64 per 64 terms per column     -- streamed by the anticommutation
//           Pauli: x-plane qubit q -> column q, z-plane -> column n+q;    scan (reference inverted index,
//           Majorana: slot s -> column s, popcount parity -> column 2n.   detail/operator/InvertedIndex.h)
//
// Per gate G (reference build_layer + apply_fused_contract, layer_build/Engine.h:522-637,
// FusedApply.h:32-83), all on one stream.  The default (fused) pipeline is two launches per gate with no
// host synchronisation between gates:
//   k_sweep      anticommutation over G's columns, c <- c cos 2a in place, the emission gate, emitters
//                listed per 4096-term sweep block (leaders, then followers) with their pre-gate values
//   k_resolve    every emitter's partner looked up in a keyless global row index, the rotation applied,
//                misses appended in the reference's order by a single-pass chained scan
// A graph build runs the same two (the sweep's graph instance leaves the coefficients unscaled, the
// resolve records the rotations) plus k_build_cos for the cos-only rows; graph replays and the adjoint
// run one k_layer launch per layer.  A store that cannot hold a gate's 2N worst case is bounded gate
// by gate by k_count_anti.  The per-gate pipeline (orbital cap, mpa_set_pipeline(e, 0)) is:
//   k_scan       XOR of G's fold columns -> anticommuting masks, split by G's pivot bit into leaders
//                (pivot clear) and followers (pivot set) (Scan.h:125-185, 265-269)
//   k_scan_sums  exclusive scan of per-block counts
//   k_expand     masks -> ascending item list AL = [leaders..., followers...]
//   (host reads nL, nF)
//   k_leaders    classify leaders (emission gate, partner phase), insert their monomial into a per-gate
//                open-addressing hash table keyed by the full monomial
//   k_followers  classify followers, look up their partner M^G among the leaders: a term's partner
//                exists in the store iff it is an anticommuting leader, so no global index is needed.
//                The follower thread applies the whole 2x2 update of a pair (Engine.h:465-513 net rule:
//                rotated iff either endpoint emits, the leader's emission taking precedence).
//   k_leaders_unmatched  cos-only / insert arm of leaders without a partner
//   k_count_miss, k_scan_sums, k_insert   stable compaction of partners to create (leader misses then
//                follower misses, Engine.h:428-444) appended at N..N+I
//   k_cols_update  wave64 ballots transpose the new rows into the column index
//   k_advance    N += I and the per-gate counters
// Wave-level scans and sums use DPP row shifts and broadcasts (no LDS permutes).  Floating point follows
// the reference's operation order; the file is compiled with -ffp-contract=off so nothing is fused into
// FMA.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/monoprop_amd.h"
#include "algebra.h"

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kScanBlock = 1024;
constexpr int kExpectBlocks = 1024;
constexpr uint64_t kMaxSchrodingerRows = uint64_t{1} << 28;  // initial Schrodinger state terms
constexpr double kIdxSlotsPerRow = 1.6;  // row index slots per row of store capacity
constexpr int kMemFields = 11;  // mpa_memory_breakdown entries

// counters[] on the device
enum : int {
    C_N = 0, C_NL = 1, C_NF = 2, C_NINS = 3, C_NEMIT = 4, C_NPAIR = 5, C_ERR = 6,
    C_STOP = 7,              // fused pipeline: 1 + index of the gate to resume at (0 = running)
    C_GCOS = 8, C_GROT = 9,  // graph record counts (cumulative over a build)
    // fused pipeline: work ticket of the resolve kernel (reset by the sweep) and the live count per
    // gate parity (C_NB + (gate & 1))
    C_TK_RES = 11,
    C_NB = 16,
    // fused pipeline on a tight store: anticommuting rows of the gate, per gate parity (C_ACNT + (gate & 1))
    C_ACNT = 18,
    C_COUNT = 24
};
// per-gate stats: N, nA, emitted, inserted, leaders, pairs, graph cos-record end, graph rotation end
constexpr int kStatW = 8;
constexpr uint32_t kPhaseBit = 0x80000000u;  // rotation record: target row | (phase < 0) << 31

// Kernel classes timed by the profiler (mpa_kernel_profile).
enum KernelKind : int { K_SCAN, K_SCANSUM, K_EXPAND, K_LEADERS, K_FOLLOWERS, K_UNMATCHED, K_COUNT, K_INSERT,
                        K_COLS, K_ADVANCE, K_SWEEP, K_RESOLVE, K_KINDS };
const char *const kKindNames[K_KINDS] = {"k_scan", "k_scan_sums", "k_expand", "k_leaders", "k_followers",
                                         "k_leaders_unmatched", "k_count_miss", "k_insert", "k_cols_update",
                                         "k_advance", "k_sweep", "k_resolve"};

thread_local std::string g_error;

// sin and cos of one angle from one glibc sincos call, on the host.  The reference evaluates std::sin
// and std::cos of the same angle side by side (FusedApply.h:47-48, Evolution.cpp:475-476, Scan.h:44
// next to Engine.h:552), which GCC, the compiler of its Linux builds, turns into a single sincos
// call; glibc's sincos can differ from its sin or cos in the last bit, so every angle goes through it.
void sin_cos(double x, double *s, double *c) { ::sincos(x, s, c); }

int set_error(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            throw HipFailure(std::string(#expr) + ": " + hipGetErrorString(_e));                   \
    } while (0)

struct HipFailure {
    std::string what;
    explicit HipFailure(std::string w) : what(std::move(w)) {}
};
struct ArgError {
    std::string what;
    explicit ArgError(std::string w) : what(std::move(w)) {}
};
struct RunError {
    std::string what;
    explicit RunError(std::string w) : what(std::move(w)) {}
};

// ------------------------------------------------------------------------------------------------
// device helpers

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

// Inclusive scan across one wavefront of 64 lanes.
// v of lane l - shift (DPP_CTRL) for the lanes ROWS enables, 0 where the source lies outside the row or
// the row is not enabled: a cross-lane move of the VALU, no LDS round trip.
template <int DPP_CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_shifted(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), DPP_CTRL, ROWS, 0xf, false));
}

// Inclusive scan across one wavefront of 64 lanes by DPP: row_shr 1, 2, 4, 8 scan each row of 16 lanes,
// then row_bcast:15 adds the last lane of rows 0 and 2 to rows 1 and 3, and row_bcast:31 the last lane of
// row 1 to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += dpp_shifted<0x111, 0xf>(v);
    v += dpp_shifted<0x112, 0xf>(v);
    v += dpp_shifted<0x114, 0xf>(v);
    v += dpp_shifted<0x118, 0xf>(v);
    v += dpp_shifted<0x142, 0xa>(v);
    v += dpp_shifted<0x143, 0xc>(v);
    return v;
}

// Lane l's 32-bit value as a wave-uniform one; l must be wave-uniform.
__device__ __forceinline__ uint32_t lane_u32(uint32_t x, int l) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), l));
}

// Exclusive scan across a block of WAVES wavefronts; returns the exclusive prefix, total in *total.
template <int WAVES = kWaves>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *lds, uint32_t *total) {
    const uint32_t incl = wave_incl_scan(v);
    const uint32_t wave = threadIdx.x >> 6;
    if (lane_id() == 63) lds[wave] = incl;
    __syncthreads();
    uint32_t wave_base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        const uint32_t s = lds[w];
        if (static_cast<uint32_t>(w) < wave) wave_base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return wave_base + incl - v;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t *lds) {
    uint32_t tot;
    (void)block_excl_scan(v, lds, &tot);
    return tot;
}

// Exclusive scan of 64-bit values across a block of kWaves wavefronts (two packed 32-bit counts that
// never carry into each other), total in *total, and with the same two barriers the block's sum of a
// (in *total_a).
__device__ __forceinline__ uint64_t block_excl_scan64_sum(uint64_t v, uint32_t a, uint64_t *lds, uint32_t *lds_a,
                                                          uint64_t *total, uint32_t *total_a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t incl = static_cast<uint64_t>(wave_incl_scan(static_cast<uint32_t>(v))) |
                          (static_cast<uint64_t>(wave_incl_scan(static_cast<uint32_t>(v >> 32))) << 32);
    const uint32_t wa = lane_u32(wave_incl_scan(a), 63);
    const uint32_t wave = threadIdx.x >> 6;
    if (lane == 63) {
        lds[wave] = incl;
        lds_a[wave] = wa;
    }
    __syncthreads();
    uint64_t wave_base = 0, tot = 0;
    uint32_t tot_a = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint64_t s = lds[w];
        if (static_cast<uint32_t>(w) < wave) wave_base += s;
        tot += s;
        tot_a += lds_a[w];
    }
    __syncthreads();
    *total = tot;
    *total_a = tot_a;
    return wave_base + incl - v;
}

// Wave-aggregated append: every lane of the wave must call it; lanes with `pred` get consecutive slots
// of the list whose length lives at *counter (one atomic per wave).
__device__ __forceinline__ uint64_t wave_append(bool pred, uint64_t *counter) {
    const uint64_t ball = __ballot(pred);
    unsigned long long base = 0;
    if (lane_id() == 0 && ball)
        base = atomicAdd(reinterpret_cast<unsigned long long *>(counter), static_cast<unsigned long long>(__popcll(ball)));
    base = (static_cast<uint64_t>(lane_u32(static_cast<uint32_t>(base >> 32), 0)) << 32) |
           lane_u32(static_cast<uint32_t>(base), 0);
    const uint32_t lane = lane_id();
    const uint64_t below = lane == 0 ? 0 : (ball & ((uint64_t{1} << lane) - 1));
    return base + static_cast<uint64_t>(__popcll(below));
}

// ------------------------------------------------------------------------------------------------
// per-gate constants

struct GateDesc {
    uint64_t g[mpa::kMaxWords];
    double k;       // cos(2 alpha)
    double s;       // sin(2 alpha)
    double inv_k;   // 1 / cos(2 alpha), fused mode
    double abs_s;   // |sin(2 alpha)| of the emission gate
    int32_t fused;  // 1: cos folded into the scan, partner value recovered as (v k) (1/k)
    int32_t fold_off, fold_cnt;
    int32_t piv0, piv1;
    int32_t identity;
    int64_t param_index;  // graph layer bookkeeping (host only)
    double gen_coeff;
};

struct Trunc {
    uint32_t cutoff;
    int32_t cutoff_type;
    int32_t check_lower, check_upper;
    double lower, upper;
    int32_t only_rot;  // -1: none
    // Cutoff basis (Majorana only; null: native): row s (NW words) is the image of slot s, for every
    // slot of the row width, rows past 2n zero (MonomialPropagator.inl:484-495).
    const uint64_t *bchg;
};

// Structural keep test of a partner t, in the cutoff basis when one is set: the test runs on the XOR
// of the images of t's slots (change_basis, MajoranaAlgebra.h:178-189, applied by the cutoff function
// of MonomialPropagatorCommon.h:44-63), the fully-paired exemption included.
template <int NW, int PAULI>
__device__ __forceinline__ bool keep_cut(const uint64_t *t, const Trunc &tr) {
    if (!PAULI && tr.bchg) {
        uint64_t mp[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) mp[k] = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w)
            for (uint64_t b = t[w]; b; b &= b - 1) {
                const uint64_t *r = tr.bchg + (static_cast<uint64_t>(w) * 64 + static_cast<uint64_t>(__builtin_ctzll(b))) * NW;
#pragma unroll
                for (int k = 0; k < NW; ++k) mp[k] ^= r[k];
            }
        return mpa::keep_partner<NW, PAULI>(mp, tr.cutoff, tr.cutoff_type);
    }
    return mpa::keep_partner<NW, PAULI>(t, tr.cutoff, tr.cutoff_type);
}

// Rotation constants of one (gate, batch member) of a batched propagation: the member's own angle
// alpha = theta[param_map] * g gives k = cos 2a, s = sin 2a, inv_k = 1 / cos 2a, abs_s = |sin 2a|.
struct Trig {
    double k, s, inv_k, abs_s;
};

// ------------------------------------------------------------------------------------------------
// kernels

// Anticommutation + pivot split over the column index: one thread per 64-term word.
__global__ __launch_bounds__(kBlock) void k_scan(const uint64_t *__restrict__ cols, uint64_t colstride,
                                                 const int32_t *__restrict__ fold, int nfold, int piv0,
                                                 int piv1, const uint64_t *__restrict__ counters,
                                                 uint64_t *__restrict__ lead, uint64_t *__restrict__ foll,
                                                 uint32_t *__restrict__ blk_l, uint32_t *__restrict__ blk_f) {
    __shared__ uint32_t lds[kWaves];
    const uint64_t N = counters[C_N];
    const uint64_t nwords = (N + 63) / 64;
    const uint64_t w = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    uint64_t lw = 0, fw = 0;
    if (w < nwords) {
        uint64_t a = 0, p = 0;
        for (int j = 0; j < nfold; ++j) a ^= cols[static_cast<uint64_t>(fold[j]) * colstride + w];
        if (piv0 >= 0) p = cols[static_cast<uint64_t>(piv0) * colstride + w];
        if (piv1 >= 0) p ^= cols[static_cast<uint64_t>(piv1) * colstride + w];
        if (w == nwords - 1 && (N & 63)) a &= (uint64_t{1} << (N & 63)) - 1;
        lw = a & ~p;
        fw = a & p;
        lead[w] = lw;
        foll[w] = fw;
    }
    const uint32_t cl = block_sum(static_cast<uint32_t>(__popcll(lw)), lds);
    const uint32_t cf = block_sum(static_cast<uint32_t>(__popcll(fw)), lds);
    if (threadIdx.x == 0) {
        blk_l[blockIdx.x] = cl;
        blk_f[blockIdx.x] = cf;
    }
}

// Exclusive scan of one or two per-block count arrays in a single block; totals to counters.
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(uint32_t *__restrict__ a, uint32_t *__restrict__ b,
                                                          uint32_t nb, uint64_t *__restrict__ counters,
                                                          int ta, int tb, int zero_from, int zero_to) {
    __shared__ uint32_t lds[kScanBlock / 64];
    __shared__ uint32_t carry_a, carry_b;
    if (counters[C_STOP]) return;
    if (threadIdx.x == 0) {
        carry_a = 0;
        carry_b = 0;
    }
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += kScanBlock) {
        const uint32_t i = base + threadIdx.x;
        uint32_t va = (i < nb) ? a[i] : 0;
        uint32_t vb = (b != nullptr && i < nb) ? b[i] : 0;
        for (int which = 0; which < 2; ++which) {
            uint32_t v = which == 0 ? va : vb;
            if (which == 1 && b == nullptr) break;
            const uint32_t incl = wave_incl_scan(v);
            const uint32_t wave = threadIdx.x >> 6;
            if (lane_id() == 63) lds[wave] = incl;
            __syncthreads();
            uint32_t wave_base = 0, tot = 0;
            for (int w = 0; w < kScanBlock / 64; ++w) {
                if (static_cast<uint32_t>(w) < wave) wave_base += lds[w];
                tot += lds[w];
            }
            uint32_t &carry = which == 0 ? carry_a : carry_b;
            const uint32_t excl = carry + wave_base + incl - v;
            if (i < nb) (which == 0 ? a : b)[i] = excl;
            __syncthreads();
            if (threadIdx.x == 0) carry += tot;
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        if (ta >= 0) counters[ta] = carry_a;
        if (tb >= 0) counters[tb] = carry_b;
        for (int z = zero_from; z < zero_to; ++z) counters[z] = 0;
    }
}

// Masks -> ascending item list: leaders at [0, nL), followers at [nL, nL + nF).
__global__ __launch_bounds__(kBlock) void k_expand(const uint64_t *__restrict__ lead,
                                                   const uint64_t *__restrict__ foll,
                                                   const uint32_t *__restrict__ blk_l,
                                                   const uint32_t *__restrict__ blk_f,
                                                   const uint64_t *__restrict__ counters,
                                                   uint32_t *__restrict__ AL) {
    __shared__ uint32_t lds[kWaves];
    if (counters[C_STOP]) return;
    const uint64_t N = counters[C_N];
    const uint64_t nwords = (N + 63) / 64;
    const uint64_t nL = counters[C_NL];
    const uint64_t w = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    uint64_t lw = 0, fw = 0;
    if (w < nwords) {
        lw = lead[w];
        fw = foll[w];
    }
    uint32_t tot;
    const uint32_t ol = block_excl_scan(static_cast<uint32_t>(__popcll(lw)), lds, &tot);
    const uint32_t of = block_excl_scan(static_cast<uint32_t>(__popcll(fw)), lds, &tot);
    uint64_t pos = blk_l[blockIdx.x] + ol;
    for (uint64_t m = lw; m; m &= m - 1) AL[pos++] = static_cast<uint32_t>(w * 64 + __ffsll(static_cast<unsigned long long>(m)) - 1);
    pos = nL + blk_f[blockIdx.x] + of;
    for (uint64_t m = fw; m; m &= m - 1) AL[pos++] = static_cast<uint32_t>(w * 64 + __ffsll(static_cast<unsigned long long>(m)) - 1);
}

template <int NW>
__device__ __forceinline__ void load_row(const uint64_t *__restrict__ rows, uint32_t r, uint64_t *m) {
    const uint64_t *p = rows + static_cast<uint64_t>(r) * NW;
#pragma unroll
    for (int w = 0; w < NW; ++w) m[w] = p[w];
}

// Value a new row takes in the Schrodinger picture before the gate rotates it: its diagonal phase
// against the initial product state if fully paired, else 0 (MPOperator.h:158-181 dense_state()
// scoring, read back as the insert's v_tgt by FusedApply.h:38-44).  The Heisenberg value is 0.
template <int NW, int PAULI>
__device__ __forceinline__ double fresh_score(const uint64_t *t, const uint64_t *__restrict__ occ_dev) {
    if (!mpa::is_paired<NW, PAULI>(t)) return 0.0;
    uint64_t occ[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) occ[w] = occ_dev[w];
    return mpa::state_phase<NW, PAULI>(t, occ);
}

// Emission decision of one anticommuting term (Scan.h:146-157, 315-347; CutoffContext.h:36-41).
// bit0 emit, bit1 phase is -1, bit2 term is in the cos set.
template <int NW, int PAULI>
__device__ __forceinline__ uint32_t classify(const uint64_t *m, const GateDesc &gd, const Trunc &tr, double v,
                                             uint64_t *t) {
    const int mono_pop = mpa::slot_pop<NW, PAULI>(m);
    const bool in_cos = tr.only_rot < 0 || mono_pop <= tr.only_rot;
    const double abs_c = fabs(v);
    bool emit = in_cos && !(tr.check_lower && gd.abs_s * abs_c <= tr.lower);
#pragma unroll
    for (int w = 0; w < NW; ++w) t[w] = m[w] ^ gd.g[w];
    int phase = 1;
    if (emit) {
        const bool keep = keep_cut<NW, PAULI>(t, tr);
        emit = keep || (tr.check_upper && gd.abs_s * abs_c >= tr.upper);
        if (emit) phase = mpa::emit_phase<NW, PAULI>(m, gd.g, t);
    }
    return (emit ? 1u : 0u) | (phase < 0 ? 2u : 0u) | (in_cos ? 4u : 0u);
}

template <int NW, int PAULI>
__global__ __launch_bounds__(kBlock) void k_leaders(const GateDesc *__restrict__ gdesc, int gi, Trunc tr,
                                                    const uint64_t *__restrict__ rows,
                                                    const double *__restrict__ coef,
                                                    const uint32_t *__restrict__ AL, uint32_t nL,
                                                    uint64_t *__restrict__ table, uint64_t tmask, uint64_t epoch,
                                                    uint8_t *__restrict__ rec, double *__restrict__ lv,
                                                    uint8_t *__restrict__ match, uint64_t *__restrict__ counters) {
    constexpr int SW = NW + 1;
    const GateDesc &gd = gdesc[gi];
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    bool emit = false;
    if (b < nL) {
        const uint32_t r = AL[b];
        uint64_t m[NW], t[NW];
        load_row<NW>(rows, r, m);
        const double v = coef[r];
        const uint32_t cls = classify<NW, PAULI>(m, gd, tr, v, t);
        emit = cls & 1u;
        rec[b] = static_cast<uint8_t>(cls);
        lv[b] = v;
        match[b] = 0;
        // insert the leader's monomial: claim a slot of this gate's epoch, then write the key
        const uint64_t tag = (epoch << 32) | b;
        uint64_t h = mpa::hash_key<NW>(m) & tmask;
        bool done = false;
        for (uint64_t probe = 0; probe <= tmask; ++probe) {
            uint64_t *slot = table + h * SW;
            const uint64_t cur = __hip_atomic_load(slot + NW, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((cur >> 32) != epoch) {
                const unsigned long long old =
                    atomicCAS(reinterpret_cast<unsigned long long *>(slot + NW), cur, tag);
                if (old == cur) {
#pragma unroll
                    for (int w = 0; w < NW; ++w) slot[w] = m[w];
                    done = true;
                    break;
                }
            }
            h = (h + 1) & tmask;
        }
        if (!done) atomicOr(reinterpret_cast<unsigned long long *>(counters + C_ERR), 1ull);
    }
    const uint64_t ball = __ballot(emit);
    if (lane_id() == 0 && ball) atomicAdd(reinterpret_cast<unsigned long long *>(counters + C_NEMIT),
                                          static_cast<unsigned long long>(__popcll(ball)));
}

// Coefficient update of a pair.  S is the emitting endpoint with phase ph; T the other.  The
// expression order is the reference's: c[S] = k v_S + (s (-ph)) v_T~, c[T] = k v_T + (s ph) v_S,
// where v_T~ is the post-cos stored value times 1/k in fused mode (Engine.h:209-211) and v_T otherwise.
__device__ __forceinline__ void rotate_pair(const GateDesc &gd, double vS, double vT, bool cosS, bool cosT,
                                            int ph, double *cS, double *cT) {
    const double wS = cosS ? vS * gd.k : vS;
    const double wT = cosT ? vT * gd.k : vT;
    const double vTu = gd.fused ? wT * gd.inv_k : vT;
    *cS = wS + gd.s * static_cast<double>(-ph) * vTu;
    *cT = wT + gd.s * static_cast<double>(ph) * vS;
}

// Graph recording (build_graph): the rotation pairs of the layer, appended in any order (each row is
// touched by at most one pair, so replay order inside a layer is irrelevant).  Cos-only rows are only
// counted (C_GCOS): the replay recomputes them from the column index.
struct GraphRec {
    uint2 *rot;  // {source row, target row | phase bit}; null when not recording
};

template <int NW, int PAULI>
__global__ __launch_bounds__(kBlock) void k_followers(const GateDesc *__restrict__ gdesc, int gi, Trunc tr,
                                                      const uint64_t *__restrict__ rows, double *__restrict__ coef,
                                                      const uint32_t *__restrict__ AL, uint32_t nL, uint32_t nF,
                                                      const uint64_t *__restrict__ table, uint64_t tmask,
                                                      uint64_t epoch, const uint8_t *__restrict__ rec,
                                                      const double *__restrict__ lv, uint8_t *__restrict__ match,
                                                      uint8_t *__restrict__ miss, double *__restrict__ ins,
                                                      uint64_t *__restrict__ counters, GraphRec grec,
                                                      const uint64_t *__restrict__ occ_dev) {
    constexpr int SW = NW + 1;
    const GateDesc &gd = gdesc[gi];
    const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
    bool emit = false, paired = false;
    // graph records of this thread: up to two cos-only rows, or one rotation
    bool rc0 = false, rc1 = false, rr = false;
    uint2 rot = {0u, 0u};
    if (f < nF) {
        const uint32_t item = nL + f;
        const uint32_t r = AL[item];
        uint64_t m[NW], t[NW];
        load_row<NW>(rows, r, m);
        const double v = coef[r];
        const uint32_t cls = classify<NW, PAULI>(m, gd, tr, v, t);
        emit = cls & 1u;
        const int ph = (cls & 2u) ? -1 : 1;
        const bool in_cos = cls & 4u;
        // partner lookup: t = m ^ g is the leader form of this pair
        int64_t b = -1;
        uint64_t h = mpa::hash_key<NW>(t) & tmask;
        for (uint64_t probe = 0; probe <= tmask; ++probe) {
            const uint64_t *slot = table + h * SW;
            const uint64_t tag = slot[NW];
            if ((tag >> 32) != epoch) break;
            bool eq = true;
#pragma unroll
            for (int w = 0; w < NW; ++w) eq = eq && (slot[w] == t[w]);
            if (eq) {
                b = static_cast<int64_t>(tag & 0xffffffffull);
                break;
            }
            h = (h + 1) & tmask;
        }
        uint8_t is_miss = 0;
        double new_coef = 0.0;
        if (b >= 0) {
            paired = true;
            match[b] = 1;
            const uint32_t rl = AL[b];
            const double vL = lv[b];
            const uint32_t lc = rec[b];
            const bool eL = lc & 1u, cosL = lc & 4u;
            const int phL = (lc & 2u) ? -1 : 1;
            double cL, cF;
            if (eL) {
                rotate_pair(gd, vL, v, cosL, in_cos, phL, &cL, &cF);
                rr = true;
                rot = make_uint2(rl, r | (phL < 0 ? kPhaseBit : 0u));
            } else if (emit) {
                rotate_pair(gd, v, vL, in_cos, cosL, ph, &cF, &cL);
                rr = true;
                rot = make_uint2(r, rl | (ph < 0 ? kPhaseBit : 0u));
            } else {
                cL = cosL ? vL * gd.k : vL;
                cF = in_cos ? v * gd.k : v;
                rc0 = rc1 = true;
            }
            coef[rl] = cL;
            coef[r] = cF;
        } else {
            const double w = in_cos ? v * gd.k : v;
            if (emit) {  // the new partner starts at vT0 (occ_dev set: Schrodinger picture)
                const double vT0 = occ_dev ? fresh_score<NW, PAULI>(t, occ_dev) : 0.0;
                coef[r] = w + gd.s * static_cast<double>(-ph) * vT0;
                new_coef = gd.k * vT0 + gd.s * static_cast<double>(ph) * v;
                is_miss = static_cast<uint8_t>(1u | (ph < 0 ? 2u : 0u));  // rotation recorded by k_insert
            } else {
                coef[r] = w;
                rc0 = true;
            }
        }
        miss[item] = is_miss;
        ins[item] = new_coef;
    }
    const uint64_t ball = __ballot(emit);
    if (lane_id() == 0 && ball) atomicAdd(reinterpret_cast<unsigned long long *>(counters + C_NEMIT),
                                          static_cast<unsigned long long>(__popcll(ball)));
    const uint64_t pball = __ballot(paired);
    if (lane_id() == 0 && pball) atomicAdd(reinterpret_cast<unsigned long long *>(counters + C_NPAIR),
                                           static_cast<unsigned long long>(__popcll(pball)));
    if (grec.rot != nullptr) {
        const uint64_t nc = __popcll(__ballot(rc0)) + __popcll(__ballot(rc1));
        if (lane_id() == 0 && nc)
            atomicAdd(reinterpret_cast<unsigned long long *>(counters + C_GCOS), static_cast<unsigned long long>(nc));
        const uint64_t pr = wave_append(rr, counters + C_GROT);
        if (rr) grec.rot[pr] = rot;
    }
}

template <int NW, int PAULI>
__global__ __launch_bounds__(kBlock) void k_leaders_unmatched(const GateDesc *__restrict__ gdesc, int gi,
                                                              const uint64_t *__restrict__ rows,
                                                              const uint64_t *__restrict__ occ_dev,
                                                              double *__restrict__ coef,
                                                              const uint32_t *__restrict__ AL, uint32_t nL,
                                                              const uint8_t *__restrict__ rec,
                                                              const double *__restrict__ lv,
                                                              const uint8_t *__restrict__ match,
                                                              uint8_t *__restrict__ miss, double *__restrict__ ins,
                                                              uint64_t *__restrict__ counters, GraphRec grec) {
    const GateDesc &gd = gdesc[gi];
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    bool rc = false;
    if (b < nL) {
        if (match[b]) {
            miss[b] = 0;
        } else {
            const uint32_t lc = rec[b];
            const bool e = lc & 1u, in_cos = lc & 4u;
            const int ph = (lc & 2u) ? -1 : 1;
            const double v = lv[b];
            const double w = in_cos ? v * gd.k : v;
            const uint32_t row = AL[b];
            if (e) {
                double vT0 = 0.0;
                if (occ_dev) {
                    uint64_t t[NW];
                    load_row<NW>(rows, row, t);
#pragma unroll
                    for (int k = 0; k < NW; ++k) t[k] ^= gd.g[k];
                    vT0 = fresh_score<NW, PAULI>(t, occ_dev);
                }
                coef[row] = w + gd.s * static_cast<double>(-ph) * vT0;
                ins[b] = gd.k * vT0 + gd.s * static_cast<double>(ph) * v;
                miss[b] = static_cast<uint8_t>(1u | (ph < 0 ? 2u : 0u));
            } else {
                coef[row] = w;
                miss[b] = 0;
                rc = true;
            }
        }
    }
    if (grec.rot != nullptr) (void)wave_append(rc, counters + C_GCOS);
}

// Per-block counts of miss flags over the combined item range [0, nA).
__global__ __launch_bounds__(kBlock) void k_count_miss(const uint8_t *__restrict__ miss, uint32_t nA,
                                                       uint32_t *__restrict__ blk) {
    __shared__ uint32_t lds[kWaves];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t v = (i < nA) ? (miss[i] & 1u) : 0;
    const uint32_t s = block_sum(v, lds);
    if (threadIdx.x == 0) blk[blockIdx.x] = s;
}

// Append the partners to create at N + rank (rank over leaders then followers).  `zero_fill`, when
// set, receives the new rows' starting value (the operator's own coefficients while a graph is being
// built): 0, or with occ_dev set (Schrodinger picture) the rows' state score.
template <int NW, int PAULI>
__global__ __launch_bounds__(kBlock) void k_insert(const GateDesc *__restrict__ gdesc, int gi,
                                                   uint64_t *__restrict__ rows, double *__restrict__ coef,
                                                   double *__restrict__ zero_fill, const uint32_t *__restrict__ AL,
                                                   uint32_t nA, const uint8_t *__restrict__ miss,
                                                   const double *__restrict__ ins, const uint32_t *__restrict__ blk,
                                                   uint64_t *__restrict__ counters, GraphRec grec,
                                                   const uint64_t *__restrict__ occ_dev) {
    __shared__ uint32_t lds[kWaves];
    const GateDesc &gd = gdesc[gi];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t flag = (i < nA) ? miss[i] : 0u;
    const uint32_t v = flag & 1u;
    uint32_t tot;
    const uint32_t rank = block_excl_scan(v, lds, &tot);
    uint64_t dest = 0;
    uint32_t src = 0;
    if (v) {
        dest = counters[C_N] + blk[blockIdx.x] + rank;
        src = AL[i];
        uint64_t t[NW];
        load_row<NW>(rows, src, t);
        uint64_t *pd = rows + dest * NW;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            t[w] ^= gd.g[w];
            pd[w] = t[w];
        }
        coef[dest] = ins[i];
        if (zero_fill != nullptr) zero_fill[dest] = occ_dev ? fresh_score<NW, PAULI>(t, occ_dev) : 0.0;
    }
    if (grec.rot != nullptr) {
        const uint64_t p = wave_append(v != 0, counters + C_GROT);
        if (v) grec.rot[p] = make_uint2(src, static_cast<uint32_t>(dest) | ((flag & 2u) ? kPhaseBit : 0u));
    }
}

// Transpose rows [start, start + count) into the column index: one wave per 64-term word, one
// ballot per column.  start/count come from counters (C_N, C_NINS) when `from_counters`.
template <int NW, int PAULI>
__global__ __launch_bounds__(kBlock) void k_cols_update(const uint64_t *__restrict__ rows,
                                                        uint64_t *__restrict__ cols, uint64_t colstride,
                                                        int n_modes, const uint64_t *__restrict__ counters,
                                                        uint64_t start_arg, uint64_t count_arg, int from_counters) {
    if (from_counters && counters[C_STOP]) return;
    const uint64_t start = from_counters ? counters[C_N] : start_arg;
    const uint64_t count = from_counters ? counters[C_NINS] : count_arg;
    if (count == 0) return;
    const uint64_t w0 = start / 64;
    const uint64_t wend = (start + count + 63) / 64;
    const bool lane0 = lane_id() == 0;
    // grid-stride over words; a whole wave shares each word
    for (uint64_t w = w0 + static_cast<uint64_t>(blockIdx.x) * kWaves + (threadIdx.x >> 6); w < wend;
         w += static_cast<uint64_t>(gridDim.x) * kWaves) {
        const uint64_t r = w * 64 + lane_id();
        uint64_t m[NW];
        const bool live = r >= start && r < start + count;
#pragma unroll
        for (int k = 0; k < NW; ++k) m[k] = live ? rows[r * NW + k] : 0;
        if (PAULI) {
            constexpr int L = NW / 2;
#pragma unroll
            for (int l = 0; l < L; ++l) {
                for (int bit = 0; bit < 64; ++bit) {
                    const int q = l * 64 + bit;
                    if (q >= n_modes) break;
                    const uint64_t bx = __ballot((m[l] >> bit) & 1u);
                    const uint64_t bz = __ballot((m[L + l] >> bit) & 1u);
                    if (lane0) {
                        if (bx) cols[static_cast<uint64_t>(q) * colstride + w] |= bx;
                        if (bz) cols[static_cast<uint64_t>(n_modes + q) * colstride + w] |= bz;
                    }
                }
            }
        } else {
            int pop = 0;
#pragma unroll
            for (int k = 0; k < NW; ++k) {
                pop += __popcll(m[k]);
                for (int bit = 0; bit < 64; ++bit) {
                    const int sl = k * 64 + bit;
                    if (sl >= 2 * n_modes) break;
                    const uint64_t bb = __ballot((m[k] >> bit) & 1u);
                    if (lane0 && bb) cols[static_cast<uint64_t>(sl) * colstride + w] |= bb;
                }
            }
            const uint64_t bp = __ballot(pop & 1);
            if (lane0 && bp) cols[static_cast<uint64_t>(2 * n_modes) * colstride + w] |= bp;
        }
    }
}

__global__ void k_advance(uint64_t *__restrict__ counters, int64_t *__restrict__ stats, int gi, int identity) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t *s = stats + static_cast<int64_t>(gi) * kStatW;
    s[0] = static_cast<int64_t>(counters[C_N]);
    s[6] = static_cast<int64_t>(counters[C_GCOS]);
    s[7] = static_cast<int64_t>(counters[C_GROT]);
    if (identity) {
        s[1] = s[2] = s[3] = s[4] = s[5] = 0;
        return;
    }
    s[1] = static_cast<int64_t>(counters[C_NL] + counters[C_NF]);
    s[2] = static_cast<int64_t>(counters[C_NEMIT]);
    s[3] = static_cast<int64_t>(counters[C_NINS]);
    s[4] = static_cast<int64_t>(counters[C_NL]);
    s[5] = static_cast<int64_t>(counters[C_NPAIR]);
    counters[C_N] += counters[C_NINS];
    counters[C_NINS] = 0;
    counters[C_NEMIT] = 0;
    counters[C_NPAIR] = 0;
}

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
            if (!GRAPH) coef[r] = v[j] * k;  // a graph build scales after its resolve (k_build_cos)
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


// Ordered sums.  The reference adds the terms of <O> one after another in ascending row order
// (EvalState::dot and inner_product, MPFunctions.cpp:86-94,135-148): phase * c over the fully paired
// rows, or state * op over every row.  That sum is reproduced to the bit: each block owns a contiguous
// segment of rows and compacts its nonzero terms in row order (k_terms, a counting pass then a writing
// pass at the scanned block offsets), then one thread adds them in that order (k_seq_sum).  A term
// that is exactly zero leaves a running sum that starts at +0.0 unchanged (it cannot turn it into
// -0.0), so leaving such terms out changes no bit.  With `member` set (a batch), every term carries its
// row's member and each member's terms are added in row order, which is the order of that member's
// own store.
enum : int { TERM_PAIRED = 0, TERM_DOT = 1 };

template <int NW, int PAULI, int MODE, int EMIT>
__global__ __launch_bounds__(kBlock) void k_terms(const uint64_t *__restrict__ rows, const double *__restrict__ a,
                                                  const double *__restrict__ b, const uint16_t *__restrict__ member,
                                                  uint64_t N, uint64_t seg, const uint64_t *__restrict__ occ_dev,
                                                  uint32_t *__restrict__ blk, double *__restrict__ out_v,
                                                  uint16_t *__restrict__ out_m) {
    __shared__ uint32_t lds[kWaves];
    uint64_t occ[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) occ[w] = MODE == TERM_PAIRED ? occ_dev[w] : 0;
    const uint64_t lo = static_cast<uint64_t>(blockIdx.x) * seg;
    const uint64_t hi = lo + seg < N ? lo + seg : N;
    uint32_t run = EMIT ? blk[blockIdx.x] : 0u;
    for (uint64_t base = lo; base < hi; base += kBlock) {
        const uint64_t i = base + threadIdx.x;
        double x = 0.0;
        if (i < hi) {
            if (MODE == TERM_PAIRED) {
                uint64_t m[NW];
                load_row<NW>(rows, static_cast<uint32_t>(i), m);
                if (mpa::is_paired<NW, PAULI>(m)) x = mpa::state_phase<NW, PAULI>(m, occ) * a[i];
            } else {
                x = a[i] * b[i];
            }
        }
        const uint32_t f = x != 0.0 ? 1u : 0u;  // NaN is kept
        uint32_t tot;
        const uint32_t r = block_excl_scan(f, lds, &tot);
        if (EMIT && f) {
            out_v[run + r] = x;
            if (member) out_m[run + r] = member[i];
        }
        run += tot;
    }
    if (!EMIT && threadIdx.x == 0) blk[blockIdx.x] = run;
}

// In-place exclusive scan of nb block counts in one block; the total goes to *total.
__global__ __launch_bounds__(kScanBlock) void k_scan_counts(uint32_t *__restrict__ a, uint32_t nb,
                                                            uint64_t *__restrict__ total) {
    __shared__ uint32_t lds[kScanBlock / 64];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += kScanBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nb ? a[i] : 0u;
        uint32_t tot;
        const uint32_t x = block_excl_scan<kScanBlock / 64>(v, lds, &tot);
        if (i < nb) a[i] = carry + x;
        __syncthreads();
        if (threadIdx.x == 0) carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// Scratch of the allocation-free <O> (mpa_expectation_value_noalloc): the ordered sum's block counts
// and total, its compacted terms and its result, in one static array per GPU.  Host code never uses the
// array's name as an address: the library looks the address up once per device at engine set-up
// (noalloc_scratch) and passes that.
constexpr uint64_t kNoAllocTerms = uint64_t{1} << 20;  // nonzero <O> terms one call can add (8 MiB)
constexpr uint32_t kNoAllocBlocks = 2048;              // ordered-sum blocks (their counts and the total)
constexpr size_t kNoAllocCountBytes = (kNoAllocBlocks + 2) * 4 + 8;
constexpr size_t kNoAllocBytes = ((kNoAllocCountBytes + 255) / 256) * 256 + kNoAllocTerms * 8 + 256;
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Adds the K compacted terms in order: out[0] (no members), or out[q] for member q = global thread q.
constexpr int kSeqStage = 4096;

__global__ __launch_bounds__(kBlock) void k_seq_sum(const double *__restrict__ v, const uint16_t *__restrict__ m,
                                                    uint64_t K, uint32_t nb, double *__restrict__ out) {
    __shared__ double sv[kSeqStage];
    __shared__ uint16_t sm[kSeqStage];
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    const bool owner = m ? q < nb : q == 0;
    double r = 0.0;
    for (uint64_t base = 0; base < K; base += kSeqStage) {
        const uint32_t n = static_cast<uint32_t>(K - base < kSeqStage ? K - base : kSeqStage);
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < n; j += kBlock) {
            sv[j] = v[base + j];
            if (m) sm[j] = m[base + j];
        }
        __syncthreads();
        if (owner) {
            if (m) {
                for (uint32_t j = 0; j < n; ++j)
                    if (sm[j] == q) r += sv[j];
            } else {
                for (uint32_t j = 0; j < n; ++j) r += sv[j];
            }
        }
    }
    if (owner) out[q] = r;
}

// Graph layers (reference evolve_step, Evolution.cpp:471-512).  A layer stores only its rotation
// records; its cos-only rows are recomputed from the column index (CosineRecompute.h:17-20): the rows
// below the layer's live count n_before that anticommute with G (XOR of G's fold columns) and belong to
// no rotation record.  Every anticommuting row of the layer is in exactly one of the two sets, so the
// rotation kernel marks its rows below n_before in `pmark` (one byte per row, plain stores: two
// records never share a row) and the cosine kernel, launched after it, skips the marked rows and
// clears their marks, leaving `pmark` zero.
//
// The cosine kernels run one wavefront per 16 words (1024 rows): lanes (column slot, word) load G's
// columns for the 16 words at once and XOR them across lanes, the set bits are listed in row order in
// LDS (10-bit offsets in the chunk), and the list is walked densely, 64 rows per round and two rounds in
// flight, each lane loading a row's mark and coefficient together.  A wave's cos-only rows are numbered
// in row order from its chunk offset `cofs[chunk]` (computed when the graph was built), which places
// the pre-layer values of the adjoint's snapshots.
constexpr int kCosChunkWords = 16;
constexpr int kCosChunkRows = kCosChunkWords * 64;

__device__ __forceinline__ void pmark_pair(uint8_t *pmark, uint32_t S, uint32_t T, uint64_t n_before) {
    pmark[S] = 1;
    if (T < n_before) pmark[T] = 1;
}

// G's fold columns of one graph layer, passed by value so that no load precedes the column loads
// (generators with more than kLayerCols columns read the rest from `fold` in memory).
constexpr int kLayerCols = 8;
struct LayerCols {
    int32_t col[kLayerCols];
    int32_t n;
};

// G's anticommuting bits among rows [0, n_before) of word w0 + (lane % 16) of the chunk starting at
// word w0 (0 past them): lane = column slot * 16 + word, four slots per pass, XOR across the slots.
__device__ __forceinline__ uint64_t layer_fold_bits(const uint64_t *__restrict__ cols, uint64_t colstride,
                                                    const LayerCols &lc, const int32_t *__restrict__ fold,
                                                    uint64_t n_before, uint64_t w0) {
    const uint32_t lane = lane_id();
    const uint64_t nwords = (n_before + 63) / 64;
    const uint64_t w = w0 + (lane & (kCosChunkWords - 1));
    const int sl = static_cast<int>(lane / kCosChunkWords);
    constexpr int kSlots = 64 / kCosChunkWords;
    uint64_t bits = 0;
    if (w < nwords) {
        uint64_t y[kLayerCols / kSlots];
#pragma unroll
        for (int pass = 0; pass < kLayerCols / kSlots; ++pass) {
            const int f = pass * kSlots + sl;
            y[pass] = f < lc.n ? cols[static_cast<uint64_t>(lc.col[f]) * colstride + w] : 0;
        }
#pragma unroll
        for (int pass = 0; pass < kLayerCols / kSlots; ++pass) bits ^= y[pass];
        for (int f = kLayerCols + sl; f < lc.n; f += kSlots) bits ^= cols[static_cast<uint64_t>(fold[f]) * colstride + w];
    }
#pragma unroll
    for (int d = kCosChunkWords; d < 64; d <<= 1) bits ^= __shfl_xor(bits, d, 64);
    if (w >= nwords) return 0;
    if (w == nwords - 1 && (n_before & 63)) bits &= (uint64_t{1} << (n_before & 63)) - 1;
    return bits;
}

// The cos-only rows of this wave's chunk in row order, rank counted from `rank0`; marked rows have their
// mark cleared.  MODE 0: replay (c *= k, pre-layer value to snap[rank] when snap is set), 1: adjoint
// (d += lam (-2 s snap[rank]), lam *= k), 2: count only.  `list` is the wave's LDS list.  Returns the
// number of cos-only rows.
template <int MODE>
__device__ __forceinline__ uint64_t layer_cos_chunk(const uint64_t *__restrict__ cols, uint64_t colstride,
                                                    const LayerCols &lc, const int32_t *__restrict__ fold,
                                                    uint64_t n_before,
                                                    uint8_t *__restrict__ pmark, uint64_t chunk, double k, double s,
                                                    double *__restrict__ c, double *__restrict__ snap,
                                                    uint64_t rank0, double *__restrict__ lam, double *d_out,
                                                    uint16_t *list) {
    const uint32_t lane = lane_id();
    const uint64_t row0 = chunk * kCosChunkRows;
    const uint64_t mine = layer_fold_bits(cols, colstride, lc, fold, n_before, chunk * kCosChunkWords);
    const uint32_t cnt = lane < kCosChunkWords ? static_cast<uint32_t>(__popcll(mine)) : 0u;  // lanes >= 16 repeat
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t total = lane_u32(incl, 63);
    if (total == 0) return 0;
    {
        uint32_t pos = incl - cnt;
        for (uint64_t b = cnt ? mine : 0; b; b &= b - 1)
            list[pos++] = static_cast<uint16_t>(lane * 64 + static_cast<uint32_t>(__builtin_ctzll(b)));
    }
    wave_lds_sync();
    const uint64_t below = (uint64_t{1} << lane) - 1;
    uint64_t rank = rank0;
    double d = 0.0;
    for (uint32_t e0 = 0; e0 < total; e0 += 128) {
        uint64_t r[2];
        uint8_t pm[2];
        double v[2];
        bool in[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint32_t e = e0 + q * 64 + lane;
            in[q] = e < total;
            r[q] = in[q] ? row0 + list[e] : 0;
            pm[q] = 0;
            v[q] = 0.0;
            if (in[q]) {
                pm[q] = pmark[r[q]];
                if (MODE == 0) v[q] = c[r[q]];
                if (MODE == 1) v[q] = lam[r[q]];
            }
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const bool cos_row = in[q] && pm[q] == 0;
            const uint64_t cm = __ballot(cos_row);
            if (in[q] && pm[q]) pmark[r[q]] = 0;
            if (cos_row) {
                const uint64_t at = rank + static_cast<uint64_t>(__popcll(cm & below));
                if (MODE == 0) {
                    if (snap) snap[at] = v[q];
                    c[r[q]] = v[q] * k;
                } else if (MODE == 1) {
                    d += v[q] * (-2.0 * s * snap[at]);
                    lam[r[q]] = v[q] * k;
                }
            }
            rank += static_cast<uint64_t>(__popcll(cm));
        }
    }
    wave_lds_sync();  // the list is reused by the wave's next caller only after every lane has read it
    if (MODE == 1) *d_out = d;
    return rank - rank0;
}

// Sum of one block's derivative terms into partial[blockIdx.x] (a fixed butterfly per wave, then the
// wave sums in wave order).
__device__ __forceinline__ void block_partial(double d, double *red, double *__restrict__ partial) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) d += __shfl_xor(d, o, 64);
    if (lane_id() == 0) red[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int w = 1; w < kWaves; ++w) t += red[w];
        partial[blockIdx.x] = t;
    }
}

// One graph layer in one launch, by block range:
//   [0, rot_blocks)                rotations: c_S = c_S k + (s (-ph)) c_T, c_T = c_T k + (s ph) c_S from the
//                                  pre-layer values (kept in snap_r from the segment's first rotation sr0
//                                  when set); ADJ: d<O>/d alpha from snap_r, then lambda <- R^T lambda
//                                  (pair: -2s c_S - 2k ph c_T and -2s c_T + 2k ph c_S)
//   [rot_blocks, + cos_blocks)     cos-only rows (marks of this layer in pm_cur, cleared): c *= k, pre-layer
//                                  values to snap_c when set; ADJ: -2 s c_i lambda_i, lambda_i *= k
//   [.., + mark_blocks)            marks of the next layer of the pass into pm_next (the other buffer)
// ADJ writes one derivative partial per rotation and cosine block (partial[blockIdx.x]).
struct LayerArgs {
    uint64_t rb, nrot, n_before, nchunks, snap_base, sr0;
    double k, s;
    LayerCols lc;
    const int32_t *fold;
    const uint32_t *cofs;
    double *snap_c, *snap_r;
    uint32_t rot_blocks, cos_blocks;
    uint64_t next_rb, next_nrot, next_n_before;  // the next layer's pair rows to mark (next_nrot 0: none)
};

template <int ADJ>
__global__ __launch_bounds__(kBlock) void k_layer(const uint2 *__restrict__ grot, const uint64_t *__restrict__ cols,
                                                  uint64_t colstride, double *__restrict__ v,
                                                  uint8_t *__restrict__ pm_cur, uint8_t *__restrict__ pm_next,
                                                  LayerArgs a, double *__restrict__ partial) {
    __shared__ uint16_t list[kWaves][kCosChunkRows];
    __shared__ double red[kWaves];
    const uint32_t b = blockIdx.x;
    if (b < a.rot_blocks) {
        const uint64_t i = static_cast<uint64_t>(b) * kBlock + threadIdx.x;
        double d = 0.0;
        if (i < a.nrot) {
            const uint64_t j = a.rb + i;
            const uint2 e = grot[j];
            const uint32_t S = e.x, T = e.y & ~kPhaseBit;
            if (!ADJ) {
                const int ph = (e.y & kPhaseBit) ? -1 : 1;
                const double vS = v[S], vT = v[T];
                if (a.snap_r) {
                    a.snap_r[2 * (j - a.sr0)] = vS;
                    a.snap_r[2 * (j - a.sr0) + 1] = vT;
                }
                v[S] = vS * a.k + a.s * static_cast<double>(-ph) * vT;
                v[T] = vT * a.k + a.s * static_cast<double>(ph) * vS;
            } else {
                const double ph = (e.y & kPhaseBit) ? -1.0 : 1.0;
                const double cs = a.snap_r[2 * (j - a.sr0)], ct = a.snap_r[2 * (j - a.sr0) + 1];
                const double ls = v[S], lt = v[T];
                d = ls * (-2.0 * a.s * cs - 2.0 * a.k * ph * ct);
                d += lt * (-2.0 * a.s * ct + 2.0 * a.k * ph * cs);
                v[S] = a.k * ls + a.s * ph * lt;
                v[T] = a.k * lt - a.s * ph * ls;
            }
        }
        if (ADJ) block_partial(d, red, partial);
        return;
    }
    if (b < a.rot_blocks + a.cos_blocks) {
        const uint64_t chunk = static_cast<uint64_t>(b - a.rot_blocks) * kWaves + (threadIdx.x >> 6);
        double d = 0.0;
        if (chunk < a.nchunks) {
            const uint64_t rank0 = (ADJ || a.snap_c) ? a.snap_base + a.cofs[chunk] : 0;
            if (!ADJ)
                (void)layer_cos_chunk<0>(cols, colstride, a.lc, a.fold, a.n_before, pm_cur, chunk, a.k, 0.0, v,
                                         a.snap_c, rank0, nullptr, nullptr, list[threadIdx.x >> 6]);
            else
                (void)layer_cos_chunk<1>(cols, colstride, a.lc, a.fold, a.n_before, pm_cur, chunk, a.k, a.s, nullptr,
                                         a.snap_c, rank0, v, &d, list[threadIdx.x >> 6]);
        }
        if (ADJ) block_partial(d, red, partial);  // the cosine blocks' partials follow the rotation blocks'
        return;
    }
    const uint64_t i = static_cast<uint64_t>(b - a.rot_blocks - a.cos_blocks) * kBlock + threadIdx.x;
    if (i < a.next_nrot) {
        const uint2 e = grot[a.next_rb + i];
        pmark_pair(pm_next, e.x, e.y & ~kPhaseBit, a.next_n_before);
    }
}

// Graph build on the fused kernels: after a gate's resolve, its cos-only rows (the anticommuting rows
// below the gate's live count that the resolve did not mark) take c *= k on the seed vector, and the
// resolve's marks are cleared.  The live count is read on the device; a stopped device skips the gate.
__global__ __launch_bounds__(kBlock) void k_build_cos(const uint64_t *__restrict__ cols, uint64_t colstride,
                                                      LayerCols lc, const int32_t *__restrict__ fold, double k,
                                                      double *__restrict__ c, uint8_t *__restrict__ pmark,
                                                      const uint64_t *__restrict__ counters, int gi) {
    __shared__ uint16_t list[kWaves][kCosChunkRows];
    if (counters[C_STOP]) return;
    const uint64_t n_before = counters[C_NB + (gi & 1)];
    const uint64_t nchunks = ((n_before + 63) / 64 + kCosChunkWords - 1) / kCosChunkWords;
    const uint64_t chunk = static_cast<uint64_t>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
    if (chunk >= nchunks) return;
    (void)layer_cos_chunk<0>(cols, colstride, lc, fold, n_before, pmark, chunk, k, 0.0, c, nullptr, 0, nullptr,
                             nullptr, list[threadIdx.x >> 6]);
}

// Graph set-up: marks of one layer's rotation rows (below n_before).
__global__ __launch_bounds__(kBlock) void k_layer_mark(const uint2 *__restrict__ grot, uint64_t rb, uint64_t nrot,
                                                       uint8_t *__restrict__ pmark, uint64_t n_before) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= nrot) return;
    const uint2 e = grot[rb + i];
    pmark_pair(pmark, e.x, e.y & ~kPhaseBit, n_before);
}

// Graph set-up: cos-only rows of one layer per wave chunk (out[chunk]).
__global__ __launch_bounds__(kBlock) void k_layer_cos_count(const uint64_t *__restrict__ cols, uint64_t colstride,
                                                            LayerCols lc, const int32_t *__restrict__ fold,
                                                            uint64_t n_before, uint8_t *__restrict__ pmark,
                                                            uint32_t *__restrict__ out, uint64_t nchunks) {
    __shared__ uint16_t list[kWaves][kCosChunkRows];
    const uint64_t chunk = static_cast<uint64_t>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
    if (chunk >= nchunks) return;
    const uint64_t n = layer_cos_chunk<2>(cols, colstride, lc, fold, n_before, pmark, chunk, 1.0, 0.0, nullptr,
                                          nullptr, 0, nullptr, nullptr, list[threadIdx.x >> 6]);
    if (lane_id() == 0) out[chunk] = static_cast<uint32_t>(n);
}

// lambda_i = <ref| M_i |ref>: the diagonal phase of fully paired rows, 0 elsewhere.
template <int NW, int PAULI>
__global__ __launch_bounds__(kBlock) void k_init_lambda(const uint64_t *__restrict__ rows, uint64_t N,
                                                        const uint64_t *__restrict__ occ_dev,
                                                        double *__restrict__ lam) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= N) return;
    uint64_t occ[NW], m[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) occ[w] = occ_dev[w];
    load_row<NW>(rows, static_cast<uint32_t>(i), m);
    lam[i] = mpa::is_paired<NW, PAULI>(m) ? mpa::state_phase<NW, PAULI>(m, occ) : 0.0;
}

// Rows that hold the observable's terms (Schrodinger picture, MPOperator.h:119-141).  `tab` is an
// open-addressing table of the terms, NW key words then the term's index + 1 (0: empty slot), at load
// <= 1/2.  A row holding term j writes its index and value c[row] at hit_row[j] / hit_val[j] and, when
// `scatter` is set, the term's coefficient obs_c[j] at scatter[row].
template <int NW>
__global__ __launch_bounds__(kBlock) void k_obs_match(const uint64_t *__restrict__ rows, const double *__restrict__ c,
                                                      uint64_t N, const uint64_t *__restrict__ tab, uint64_t tmask,
                                                      int64_t *__restrict__ hit_row, double *__restrict__ hit_val,
                                                      const double *__restrict__ obs_c, double *__restrict__ scatter) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= N) return;
    uint64_t m[NW];
    load_row<NW>(rows, static_cast<uint32_t>(i), m);
    uint64_t h = mpa::hash_key<NW>(m) & tmask;
    for (uint64_t probe = 0; probe <= tmask; ++probe) {
        const uint64_t *slot = tab + h * (NW + 1);
        const uint64_t id = slot[NW];
        if (!id) return;
        bool eq = true;
#pragma unroll
        for (int w = 0; w < NW; ++w) eq = eq && (slot[w] == m[w]);
        if (eq) {
            hit_row[id - 1] = static_cast<int64_t>(i);
            hit_val[id - 1] = c[i];
            if (scatter) scatter[i] = obs_c[id - 1];
            return;
        }
        h = (h + 1) & tmask;
    }
}

constexpr uint32_t kMaxBatch = 2048;  // members per batch (16-bit member ids; the ordered sum's threads)

// Stable partition of a batch store by member: side[member] = 1 moves the row out.  Per 256-row block,
// the number of moving rows.
__global__ __launch_bounds__(kBlock) void k_part_count(const uint16_t *__restrict__ member, uint64_t N,
                                                       const uint8_t *__restrict__ side, uint32_t *__restrict__ blk) {
    __shared__ uint32_t lds[kWaves];
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const uint32_t f = i < N ? side[member[i]] : 0u;
    const uint32_t c = block_sum(f, lds);
    if (threadIdx.x == 0) blk[blockIdx.x] = c;
}

// blk: moving rows before each block (exclusive scan of k_part_count).  Row i goes to position
// (moving rows before i) of the moving side or (staying rows before i) of the staying side, with its
// member renumbered to renum[member].
template <int NW>
__global__ __launch_bounds__(kBlock) void k_part_scatter(const uint64_t *__restrict__ rows, const double *__restrict__ coef,
                                                         const uint16_t *__restrict__ member, uint64_t N,
                                                         const uint8_t *__restrict__ side,
                                                         const uint16_t *__restrict__ renum,
                                                         const uint32_t *__restrict__ blk,
                                                         uint64_t *__restrict__ rows_mv, double *__restrict__ coef_mv,
                                                         uint16_t *__restrict__ mem_mv, uint64_t *__restrict__ rows_st,
                                                         double *__restrict__ coef_st, uint16_t *__restrict__ mem_st) {
    __shared__ uint32_t lds[kWaves];
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
    uint32_t f = 0, q = 0;
    if (i < N) {
        q = member[i];
        f = side[q];
    }
    uint32_t tot;
    const uint32_t rank = block_excl_scan(f, lds, &tot);
    if (i >= N) return;
    const uint64_t moved_before = static_cast<uint64_t>(blk[blockIdx.x]) + rank;
    const uint64_t dest = f ? moved_before : i - moved_before;
    uint64_t *rd = (f ? rows_mv : rows_st) + dest * NW;
#pragma unroll
    for (int k = 0; k < NW; ++k) rd[k] = rows[i * NW + k];
    (f ? coef_mv : coef_st)[dest] = coef[i];
    (f ? mem_mv : mem_st)[dest] = renum[q];
}

// Stored rows per batch member (counts[nb], zeroed by the caller).
__global__ __launch_bounds__(kBlock) void k_count_members(const uint16_t *__restrict__ member, uint64_t N, uint32_t nb,
                                                          unsigned long long *__restrict__ counts) {
    extern __shared__ uint32_t hist[];
    for (uint32_t i = threadIdx.x; i < nb; i += kBlock) hist[i] = 0;
    __syncthreads();
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < N;
         i += static_cast<uint64_t>(gridDim.x) * kBlock)
        atomicAdd(&hist[member[i]], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += kBlock)
        if (hist[i]) atomicAdd(&counts[i], static_cast<unsigned long long>(hist[i]));
}

// Per-layer sum of the adjoint partials, in block order: one block per layer.
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double *__restrict__ partial,
                                                            const uint64_t *__restrict__ offs,
                                                            double *__restrict__ out) {
    __shared__ double red[kBlock];
    const uint64_t a = offs[blockIdx.x], b = offs[blockIdx.x + 1];
    double acc = 0.0;
    for (uint64_t i = a + threadIdx.x; i < b; i += kBlock) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
        if (threadIdx.x < static_cast<unsigned>(st)) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

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

// ------------------------------------------------------------------------------------------------
// C ABI

namespace {

template <typename F>
int guarded(F &&f) {
    try {
        f();
        g_error.clear();
        return 0;
    } catch (const ArgError &e) {
        return set_error(-1, e.what);
    } catch (const RunError &e) {
        return set_error(-2, e.what);
    } catch (const HipFailure &e) {
        return set_error(-3, e.what);
    } catch (const std::bad_alloc &) {
        return set_error(-2, "host allocation failed");
    } catch (const std::exception &e) {
        return set_error(-2, e.what());
    }
}

bool is_none(double x) { return std::isnan(x); }

// Template functors handed to mpa_engine::dispatch (the runtime basis / row width picks the instance).
struct InitCols {
    mpa_engine *e;
    uint64_t n0;
    template <int NW, int PAULI>
    void operator()() {
        e->launch_cols_update<NW, PAULI>(0, n0, false, n0);
    }
};

struct RunPropagate {
    mpa_engine *e;
    const std::vector<GateDesc> &gds;
    const std::vector<int32_t> &folds;
    bool graph;
    Trunc trl;
    bool fused;  // every gate takes the fused cos sweep: run the fused pipeline
    template <int NW, int PAULI>
    void operator()() {
        if (fused) {
            e->propagate_fused<NW, PAULI, 0>(gds, folds, graph ? trl : e->tr, graph);
        } else {
            e->gidx_rows = static_cast<uint64_t>(-1);  // the unfused pipeline does not maintain the index
            e->propagate_impl<NW, PAULI>(gds, folds, graph, trl);
        }
    }
};

struct RunBatch {
    mpa_engine *e;
    const std::vector<GateDesc> &gds;
    const std::vector<int32_t> &folds;
    template <int NW, int PAULI>
    void operator()() {
        e->propagate_fused<NW, PAULI, 1>(gds, folds);
    }
};

struct RunExpectBatch {
    mpa_engine *e;
    double *out;
    template <int NW, int PAULI>
    void operator()() {
        e->expect_batch<NW, PAULI>(out);
    }
};

struct RunExpect {
    mpa_engine *e;
    const double *c;
    double *out;
    template <int NW, int PAULI>
    void operator()() {
        *out = e->expect_impl<NW, PAULI>(c);
    }
};

struct RunExpectFixed {
    mpa_engine *e;
    unsigned char *scratch_dev;
    double *out;
    template <int NW, int PAULI>
    void operator()() {
        double r = 0.0;
        e->ordered_sum<NW, PAULI, TERM_PAIRED>(e->coef.as<double>(), nullptr, nullptr, 1, &r, scratch_dev);
        *out = e->core + r;
    }
};

// Schrodinger <O>: without a graph the state against the observable's rows; with one, the observable
// replayed newest layer first and dotted with the state (the reference's functional).
struct RunSchrExpect {
    mpa_engine *e;
    const double *params;
    int64_t n_params;
    double *out;
    template <int NW, int PAULI>
    void operator()() {
        if (e->layers.empty()) {
            if (n_params != 0) throw ArgError("Expected 0 parameter value(s), got " + std::to_string(n_params) + ".");
            *out = e->observable_expect<NW>(e->coef.as<double>());
            return;
        }
        if (n_params != e->graph_nparams)
            throw ArgError("Expected " + std::to_string(e->graph_nparams) + " parameter value(s), got " +
                           std::to_string(n_params) + ".");
        e->rep.reserve(std::max<uint64_t>(e->N, 1) * 8);
        std::vector<int64_t> hr;
        std::vector<double> hv;
        e->match_observable<NW>(e->coef.as<double>(), hr, hv, e->rep.as<double>());
        e->replay(params, mpa_engine::kObservable);
        *out = e->dot_core<NW, PAULI>(e->coef.as<double>(), e->rep.as<double>());
    }
};

// The checkpointed forward replay of the stored graph at params and its <O> (the reference's
// ev / ev_and_grad value, MPFunctions.cpp:196-228): Heisenberg, the operator from the graph's start
// vector; Schrodinger, the observable newest layer first, dotted with the state.
struct RunForwardCheckpointed {
    mpa_engine *e;
    const double *params;
    int64_t n_params;
    double *out;
    template <int NW, int PAULI>
    void operator()() {
        if (n_params != e->graph_nparams)
            throw ArgError("Expected " + std::to_string(e->graph_nparams) + " parameter value(s), got " +
                           std::to_string(n_params) + ".");
        if (e->schr) {
            e->rep.reserve(std::max<uint64_t>(e->N, 1) * 8);
            std::vector<int64_t> hr;
            std::vector<double> hv;
            e->match_observable<NW>(e->coef.as<double>(), hr, hv, e->rep.as<double>());
            e->forward_checkpointed(mpa_engine::kObservable, params, n_params);
            *out = e->dot_core<NW, PAULI>(e->coef.as<double>(), e->rep.as<double>());
        } else {
            e->replay_init(mpa_engine::kOperator);
            e->forward_checkpointed(mpa_engine::kOperator, params, n_params);
            *out = e->expect_impl<NW, PAULI>(e->rep.as<double>());
        }
        e->ckpt_value = *out;
    }
};

// update_initial_operator: terms -> the rows holding them (the observable's matcher), the new
// coefficient vector scattered over the store into `dst`; returns the terms no row holds.
struct RunMatchTerms {
    mpa_engine *e;
    double *dst;
    std::vector<int64_t> *hit_row;
    template <int NW, int PAULI>
    void operator()() {
        std::vector<double> hv;
        e->match_observable<NW>(e->coef.as<double>(), *hit_row, hv, dst);
    }
};

struct RunGradient {
    mpa_engine *e;
    const double *params;
    int64_t n_params;
    double *grad;
    template <int NW, int PAULI>
    void operator()() {
        e->gradient<NW, PAULI>(params, n_params, grad);
    }
};

// Layer arrays (authoring order) -> per-gate descriptors in processing order (Heisenberg: back to
// front; Schrodinger: front to back, MonomialPropagator.inl:720-733).  alpha = (1.0 * theta[map]) * g (MPFunctions.cpp:57-69);
// without params (structural graph build) alpha = 0.  `fused`: the contract-immediately cos sweep.
// validate_parameters_length (Validation.cpp:79-91) for a lone propagator's propagate and seeded
// build_graph; the batch entry points resume with a suffix of the layers and whole parameter rows.
void check_params_length(const int64_t *param_map, int64_t n_layers, const double *params, int64_t n_params) {
    if (params == nullptr || n_layers <= 0) return;
    const int64_t need = *std::max_element(param_map, param_map + n_layers) + 1;
    if (n_params != need)
        throw RunError("The length of parameters (" + std::to_string(n_params) +
                       ") must be the same as max(parameter_mapping)+1 (" + std::to_string(need) + ").");
}

void make_gates(const mpa_engine *e, const int64_t *gen_offsets, const int32_t *gen_slots, int64_t n_layers,
                const int64_t *param_map, const double *gen_coeffs, const double *params, int64_t n_params,
                bool fused_allowed, std::vector<GateDesc> &gds, std::vector<int32_t> &folds) {
    gds.clear();
    folds.clear();
    gds.reserve(n_layers);
    std::vector<uint64_t> g(mpa::kMaxWords);
    for (int64_t i = 0; i < n_layers; ++i) {
        const int64_t idx = e->schr ? i : n_layers - 1 - i;
        if (param_map[idx] < 0 || (params != nullptr && param_map[idx] >= n_params))
            throw ArgError("parameter index out of range for the given parameters");
        GateDesc gd{};
        std::fill(g.begin(), g.end(), 0);
        e->slots_to_words(gen_slots + gen_offsets[idx], gen_offsets[idx + 1] - gen_offsets[idx], g.data());
        std::copy(g.begin(), g.begin() + e->nw, gd.g);
        gd.param_index = param_map[idx];
        gd.gen_coeff = gen_coeffs[idx];
        if (params != nullptr) {
            // build angle alpha (cutoffs, cos sweep); the rotation applies it negated in the
            // Schrodinger picture (MonomialPropagator.h:411-416)
            const double alpha = 1.0 * params[param_map[idx]] * gen_coeffs[idx];
            const double apply = e->schr ? -alpha : alpha;
            sin_cos(2 * apply, &gd.s, &gd.k);
            double sin_build, cos_build;
            sin_cos(2 * alpha, &sin_build, &cos_build);
            gd.abs_s = std::abs(sin_build);
            gd.fused = (fused_allowed && cos_build != 0.0) ? 1 : 0;
            gd.inv_k = gd.fused ? 1.0 / cos_build : 1.0;
        } else {
            gd.k = 1.0;
            gd.s = 0.0;
            gd.abs_s = 1.0;
            gd.fused = 0;
            gd.inv_k = 1.0;
        }
        gd.fold_off = static_cast<int32_t>(folds.size());
        e->gate_columns(gd.g, folds, gd.piv0, gd.piv1);
        gd.fold_cnt = static_cast<int32_t>(folds.size()) - gd.fold_off;
        gd.identity = gd.fold_cnt == 0 ? 1 : 0;
        gds.push_back(gd);
    }
}

// The coefficient vector of the current picture at params: the replay when a graph is stored (n_params
// must cover it; the Schrodinger picture replays the state), else the contracted coefficients
// (n_params must be 0).
const double *evaluated_coeffs(mpa_engine *e, const double *params, int64_t n_params) {
    if (e->layers.empty()) {
        if (n_params != 0) throw ArgError("Expected 0 parameter value(s), got " + std::to_string(n_params) + ".");
        return e->coef.as<double>();
    }
    if (n_params != e->graph_nparams)
        throw ArgError("Expected " + std::to_string(e->graph_nparams) + " parameter value(s), got " +
                       std::to_string(n_params) + ".");
    e->replay(params, e->schr ? mpa_engine::kState : mpa_engine::kOperator);
    return e->rep.as<double>();
}

void set_trunc(mpa_engine *e, uint32_t cutoff, int cutoff_type, double lower, double upper) {
    if (e->basis == MPA_BASIS_PAULI && cutoff_type != MPA_CUTOFF_SUPPORT)
        throw ArgError("Pauli basis requires cutoff_type == Support (Length has no Pauli-weight meaning under the "
                       "Pauli encoding).");
    if (!is_none(lower) && !is_none(upper) && upper < lower)
        throw RunError("upper_atol must be greater than or equal to lower_atol.");
    e->tr.cutoff = cutoff;
    e->tr.cutoff_type = cutoff_type;
    e->tr.check_lower = is_none(lower) ? 0 : 1;
    e->tr.check_upper = is_none(upper) ? 0 : 1;
    e->tr.lower = is_none(lower) ? 0.0 : lower;
    e->tr.upper = is_none(upper) ? 0.0 : upper;
}

// A fresh engine with its stream, truncation and initial state; no terms yet.
std::unique_ptr<mpa_engine> new_engine(int basis, int num_modes, uint32_t cutoff, int cutoff_type, double lower_atol,
                                       double upper_atol, const int32_t *initial_state, int64_t n_state, int device) {
    if (basis != MPA_BASIS_PAULI && basis != MPA_BASIS_MAJORANA) throw ArgError("unknown basis");
    if (num_modes <= 0 || num_modes > 256)
        throw RunError("logical_num_modes (" + std::to_string(num_modes) + ") must be in the range [1, 256].");
    auto e = std::make_unique<mpa_engine>();
    e->basis = basis;
    e->n_modes = num_modes;
    e->L = (num_modes + 63) / 64;
    e->nw = basis == MPA_BASIS_PAULI ? 2 * e->L : (2 * num_modes + 63) / 64;
    e->ncols = basis == MPA_BASIS_PAULI ? 2 * num_modes : 2 * num_modes + 1;
    set_trunc(e.get(), cutoff, cutoff_type, lower_atol, upper_atol);
    e->tr.only_rot = -1;
    e->tr.bchg = nullptr;
    e->device = device;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&e->h_counters), (C_COUNT + 1) * 8, hipHostMallocDefault));
    (void)noalloc_scratch(device);  // the scratch array's address on this device, looked up at set-up
    e->occ.assign(mpa::kMaxWords, 0);
    for (int64_t k = 0; k < n_state; ++k) {
        const int j = initial_state[k];
        if (j < 0 || j >= num_modes) throw ArgError("initial_state index out of range");
        if (basis == MPA_BASIS_PAULI) e->occ[j >> 6] |= uint64_t{1} << (j & 63);
        else e->occ[(2 * j) >> 6] |= uint64_t{1} << ((2 * j) & 63);
    }
    return e;
}

// An engine with the configuration (basis, width, truncation, initial state, device) of `e`, its own
// stream and counters, and no terms.
std::unique_ptr<mpa_engine> clone_config(const mpa_engine *e) {
    auto o = std::make_unique<mpa_engine>();
    o->basis = e->basis;
    o->n_modes = e->n_modes;
    o->L = e->L;
    o->nw = e->nw;
    o->ncols = e->ncols;
    o->tr = e->tr;
    o->occ = e->occ;
    o->device = e->device;
    o->fused_enabled = e->fused_enabled;
    o->res_grid_cap = e->res_grid_cap;
    o->min_group_log2 = e->min_group_log2;
    HIP_TRY(hipSetDevice(o->device));
    HIP_TRY(hipStreamCreateWithFlags(&o->stream, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&o->h_counters), (C_COUNT + 1) * 8, hipHostMallocDefault));
    o->counters.reserve(C_COUNT * 8);
    HIP_TRY(hipMemsetAsync(o->counters.p, 0, C_COUNT * 8, o->stream));
    o->occ_dev.reserve(mpa::kMaxWords * 8);
    HIP_TRY(hipMemcpyAsync(o->occ_dev.p, o->occ.data(), mpa::kMaxWords * 8, hipMemcpyHostToDevice, o->stream));
    HIP_TRY(hipStreamSynchronize(o->stream));
    o->bchg_host = e->bchg_host;  // the copied truncation still points at e's table until this
    o->upload_basis();
    return o;
}

// Capacity of a store that holds n rows exactly (rounded up to the allocation granule).
uint64_t exact_capacity(uint64_t n) {
    const uint64_t g = 64ull * kBlock;
    return std::max<uint64_t>(g, (n + g - 1) / g * g);
}

struct RunPartition {
    mpa_engine *e;
    const uint8_t *side;
    const uint16_t *renum;
    const uint32_t *blk;
    uint64_t *rows_mv, *rows_st;
    double *coef_mv, *coef_st;
    uint16_t *mem_mv, *mem_st;
    template <int NW, int PAULI>
    void operator()() {
        const uint64_t nblk = (e->N + kBlock - 1) / kBlock;
        hipLaunchKernelGGL((k_part_scatter<NW>), dim3(static_cast<uint32_t>(nblk)), dim3(kBlock), 0, e->stream,
                           e->rows.as<uint64_t>(), e->coef.as<double>(), e->bmember.as<uint16_t>(), e->N, side, renum,
                           blk, rows_mv, coef_mv, mem_mv, rows_st, coef_st, mem_st);
        HIP_TRY(hipGetLastError());
    }
};

// Terms [t0, t1) of a CSR operator, in lexicographic slot-tuple order (std::map<VecZ, complex>
// iteration), encoded and appended to rows / coefs; the empty tuple's coefficient goes to *core.
// Distinct sorted slot tuples give distinct rows, so no key repeats.
void encode_operator(const mpa_engine *e, const int64_t *term_offsets, const int32_t *term_slots,
                     const double *coeff_re, const double *coeff_im, int64_t t0, int64_t t1,
                     std::vector<uint64_t> &rows, std::vector<double> &coefs, double *core) {
    std::vector<int64_t> order;
    order.reserve(static_cast<size_t>(t1 - t0));
    for (int64_t t = t0; t < t1; ++t) order.push_back(t);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return std::lexicographical_compare(term_slots + term_offsets[a], term_slots + term_offsets[a + 1],
                                            term_slots + term_offsets[b], term_slots + term_offsets[b + 1]);
    });
    std::vector<uint64_t> m(e->nw);
    for (int64_t t : order) {
        const int64_t n = term_offsets[t + 1] - term_offsets[t];
        e->slots_to_words(term_slots + term_offsets[t], n, m.data());
        const std::complex<double> c(coeff_re[t], coeff_im[t]);
        double enc;
        if (e->basis == MPA_BASIS_PAULI) {  // PauliAlgebra.h:150-155
            if (std::abs(c.imag()) > 1e-10) throw RunError("Non-real Pauli coeffs detected");
            enc = c.real();
        } else {  // MajoranaAlgebra.h:162-171
            static const std::complex<double> kPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
            const size_t un = static_cast<size_t>(n);
            const std::complex<double> d = c / kPow[(un * (un - 1) / 2) % 4];
            if (std::abs(d.imag()) > 1e-10) throw RunError("Non-Hermitian coeffs detected");
            enc = d.real();
        }
        if (n == 0) {
            *core = enc;
            continue;
        }
        rows.insert(rows.end(), m.begin(), m.end());
        coefs.push_back(enc);
    }
}

// The Schrodinger picture's initial rows (MonomialPropagator.inl:162-181, MajoranaAlgebra.h:125-160):
// every fully paired monomial of at most `pairs` modes, by pair count; within a count the mode sets in
// the order std::prev_permutation walks a selector that starts as `pairs` ones then zeros (the k-subsets
// of selector positions in lexicographic order), selector position i being mode n-1-i.  A row starts at
// its diagonal phase <b|M|b> (MPOperator.h:235-254).
void paired_state_rows(const mpa_engine *e, int pairs, std::vector<uint64_t> &rows, std::vector<double> &coefs) {
    const int n = e->n_modes;
    double count = 0.0, binom = 1.0;
    for (int k = 0; k <= pairs; ++k) {
        count += binom;
        binom = binom * (n - k) / (k + 1);
    }
    if (count > static_cast<double>(kMaxSchrodingerRows))
        throw ArgError("schrodinger_cutoff gives " + std::to_string(static_cast<long double>(count)) +
                       " initial state terms, more than the " + std::to_string(kMaxSchrodingerRows) +
                       " this engine starts from");
    rows.reserve(static_cast<size_t>(count) * e->nw);
    coefs.reserve(static_cast<size_t>(count));
    std::vector<int> pos;
    std::vector<uint64_t> m(e->nw);
    for (int k = 0; k <= pairs; ++k) {
        pos.resize(k);
        for (int i = 0; i < k; ++i) pos[i] = i;
        for (;;) {
            std::fill(m.begin(), m.end(), 0);
            int occupied = 0;
            for (int i : pos) {
                const int j = n - 1 - i;
                if (e->basis == MPA_BASIS_PAULI) {
                    m[e->L + (j >> 6)] |= uint64_t{1} << (j & 63);  // Z_j
                    occupied += static_cast<int>((e->occ[j >> 6] >> (j & 63)) & 1u);
                } else {
                    m[(2 * j) >> 6] |= uint64_t{3} << ((2 * j) & 63);  // slots 2j, 2j+1
                    occupied += static_cast<int>((e->occ[(2 * j) >> 6] >> ((2 * j) & 63)) & 1u);
                }
            }
            // Pauli: (-1)^{|Z support & occupied|}; Majorana: (-1)^{occupied + pairs} (MajoranaAlgebra.h:52-60)
            const int odd = e->basis == MPA_BASIS_PAULI ? (occupied & 1) : ((occupied + k) & 1);
            rows.insert(rows.end(), m.begin(), m.end());
            coefs.push_back(odd ? -1.0 : 1.0);
            int i = k - 1;
            while (i >= 0 && pos[i] == n - k + i) --i;
            if (i < 0) break;
            ++pos[i];
            for (int t = i + 1; t < k; ++t) pos[t] = pos[t - 1] + 1;
        }
    }
}

// Uploads the initial rows (and, for a batch, their members) and indexes their columns.
void load_initial(mpa_engine *e, uint64_t n0, const uint64_t *hrows, const double *hcoef, const uint16_t *members) {
    e->counters.reserve(C_COUNT * 8);
    HIP_TRY(hipMemsetAsync(e->counters.p, 0, C_COUNT * 8, e->stream));
    e->occ_dev.reserve(mpa::kMaxWords * 8);
    HIP_TRY(hipMemcpyAsync(e->occ_dev.p, e->occ.data(), mpa::kMaxWords * 8, hipMemcpyHostToDevice, e->stream));
    e->ensure_capacity(std::max<uint64_t>(n0, 1));
    if (members) {
        e->bmember.reserve(e->cap * 2);
        if (n0) HIP_TRY(hipMemcpyAsync(e->bmember.p, members, n0 * 2, hipMemcpyHostToDevice, e->stream));
    }
    if (n0) {
        HIP_TRY(hipMemcpyAsync(e->rows.p, hrows, n0 * e->nw * 8, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(e->coef.p, hcoef, n0 * 8, hipMemcpyHostToDevice, e->stream));
        InitCols init{e, n0};
        e->dispatch(init);
        const int64_t c0[2] = {static_cast<int64_t>(n0), 0};
        HIP_TRY(hipMemcpyAsync(e->counters.p, c0, 8, hipMemcpyHostToDevice, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->N = n0;
    e->n_init = n0;
    if (!members) e->init_coef.assign(hcoef, hcoef + n0);  // the graph path restarts from these
}

// Entry points of a lone propagator refuse a batch and the batch entry points a lone propagator.
void require_single(const mpa_engine *e) {
    if (e->nbatch) throw ArgError("this engine holds a batch; use the mpa_batch_* entry points");
}
void require_batch(const mpa_engine *e) {
    if (!e->nbatch) throw ArgError("this engine is not a batch (create it with mpa_batch_create)");
}

// The cutoff basis of a lone Majorana propagator (update_basis_change, MonomialPropagator.h:204-211,
// validated as MonomialPropagator.inl:461-495): exactly 2n rows, every slot below 2n; the whole table
// is checked before anything changes.  n_rows < 0 restores the native basis.
void set_basis_change(mpa_engine *e, const int64_t *offsets, const int32_t *slots, int64_t n_rows) {
    require_single(e);
    if (n_rows < 0) {
        e->bchg_host.clear();
        e->upload_basis();
        return;
    }
    if (e->basis == MPA_BASIS_PAULI)
        throw ArgError("Pauli basis does not accept a basis_change (the encoding is already the Jordan-Wigner image).");
    const int64_t want = 2 * static_cast<int64_t>(e->n_modes);
    if (n_rows != want)
        throw ArgError("basis_change must have exactly 2*logical_num_modes (" + std::to_string(want) + ") rows; got " +
                       std::to_string(n_rows) + ".");
    std::vector<uint64_t> tab(static_cast<size_t>(64) * e->nw * e->nw, 0);
    for (int64_t s = 0; s < n_rows; ++s) {
        if (offsets[s + 1] < offsets[s]) throw ArgError("basis_change offsets must be non-decreasing");
        for (int64_t k = offsets[s]; k < offsets[s + 1]; ++k) {
            const int32_t j = slots[k];
            if (j < 0 || j >= want)
                throw RunError("Majorana/Pauli index " + std::to_string(j) + " is out of range; must be less than " +
                               std::to_string(want) + ".");
            tab[static_cast<size_t>(s) * e->nw + (j >> 6)] |= uint64_t{1} << (j & 63);  // a repeated index sets once
        }
    }
    e->bchg_host.swap(tab);
    e->upload_basis();
}

}  // namespace

extern "C" {

const char *mpa_last_error(void) { return g_error.c_str(); }

int mpa_create(mpa_engine **out, int basis, int num_modes, uint32_t cutoff, int cutoff_type, double lower_atol,
               double upper_atol, const int64_t *term_offsets, const int32_t *term_slots, const double *coeff_re,
               const double *coeff_im, int64_t n_terms, const int32_t *initial_state, int64_t n_state, int device) {
    return guarded([&] {
        if (!out) throw ArgError("out is null");
        *out = nullptr;
        auto e = new_engine(basis, num_modes, cutoff, cutoff_type, lower_atol, upper_atol, initial_state, n_state, device);
        std::vector<uint64_t> hrows;
        std::vector<double> hcoef;
        encode_operator(e.get(), term_offsets, term_slots, coeff_re, coeff_im, 0, n_terms, hrows, hcoef, &e->core);
        load_initial(e.get(), hcoef.size(), hrows.data(), hcoef.data(), nullptr);
        *out = e.release();
    });
}

int mpa_create_schrodinger(mpa_engine **out, int basis, int num_modes, uint32_t cutoff, int cutoff_type,
                           double lower_atol, double upper_atol, int64_t schrodinger_cutoff,
                           const int64_t *term_offsets, const int32_t *term_slots, const double *coeff_re,
                           const double *coeff_im, int64_t n_terms, const int32_t *initial_state, int64_t n_state,
                           int device) {
    return guarded([&] {
        if (!out) throw ArgError("out is null");
        *out = nullptr;
        if (schrodinger_cutoff < 0) throw ArgError("schrodinger_cutoff must be non-negative");
        auto e = new_engine(basis, num_modes, cutoff, cutoff_type, lower_atol, upper_atol, initial_state, n_state, device);
        e->schr = true;
        encode_operator(e.get(), term_offsets, term_slots, coeff_re, coeff_im, 0, n_terms, e->obs_rows, e->obs_coef,
                        &e->core);
        // ceil(sc / 2) pairs of the clamped slot cutoff, at most one per mode (MonomialPropagator.inl:162-164)
        const int64_t sc = std::min<int64_t>(schrodinger_cutoff, 2 * static_cast<int64_t>(num_modes));
        const int pairs = static_cast<int>(std::min<int64_t>(sc / 2 + sc % 2, num_modes));
        std::vector<uint64_t> hrows;
        std::vector<double> hcoef;
        paired_state_rows(e.get(), pairs, hrows, hcoef);
        load_initial(e.get(), hcoef.size(), hrows.data(), hcoef.data(), nullptr);
        *out = e.release();
    });
}

int mpa_update_initial_operator(mpa_engine *e, const int64_t *term_offsets, const int32_t *term_slots,
                                const double *coeff_re, const double *coeff_im, int64_t n_terms) {
    return guarded([&] {
        require_single(e);
        e->ckpt_valid = false;
        std::vector<uint64_t> rows;
        std::vector<double> coefs;
        // the core term changes as soon as it is read, before any term is looked up (MonomialPropagator.inl:363-376)
        encode_operator(e, term_offsets, term_slots, coeff_re, coeff_im, 0, n_terms, rows, coefs, &e->core);
        if (e->schr) {  // the observable is replaced; absent terms wait for a row to hold them (MPOperator.h:218-226)
            e->obs_rows = std::move(rows);
            e->obs_coef = std::move(coefs);
            e->obs_tmask = 0;
            return;
        }
        // Heisenberg: every term must be a row; the operator becomes exactly these terms (MPOperator.h:193-235)
        std::vector<uint64_t> keep_rows = std::move(e->obs_rows);
        std::vector<double> keep_coef = std::move(e->obs_coef);
        e->obs_rows = rows;
        e->obs_coef = coefs;
        e->obs_tmask = 0;
        DevArray fresh{&e->stream};
        fresh.reserve(std::max<uint64_t>(e->N, 1) * 8);
        std::vector<int64_t> hit;
        RunMatchTerms mt{e, fresh.as<double>(), &hit};
        try {
            e->dispatch(mt);
        } catch (...) {
            e->obs_rows = std::move(keep_rows);
            e->obs_coef = std::move(keep_coef);
            e->obs_tmask = 0;
            throw;
        }
        e->obs_rows.clear();
        e->obs_coef.clear();
        e->obs_tmask = 0;
        for (size_t j = 0; j < hit.size(); ++j) {
            if (hit[j] >= 0) continue;
            std::vector<int32_t> slots;
            e->words_to_slots(rows.data() + j * e->nw, slots);
            std::string repr = "[";
            for (size_t k = 0; k < slots.size(); ++k) repr += (k ? ", " : "") + std::to_string(slots[k]);
            throw RunError("Operator term " + repr + "] not found in the operator.");
        }
        if (e->N) {
            HIP_TRY(hipMemcpyAsync(e->coef.p, fresh.p, e->N * 8, hipMemcpyDeviceToDevice, e->stream));
            if (!e->layers.empty()) {  // the graph replays from the new operator
                e->base.reserve(e->N * 8);
                HIP_TRY(hipMemcpyAsync(e->base.p, fresh.p, e->N * 8, hipMemcpyDeviceToDevice, e->stream));
                e->base_n = e->N;
            }
        }
        HIP_TRY(hipStreamSynchronize(e->stream));
    });
}

int mpa_set_parameter_mapping(mpa_engine *e, const int64_t *param_index, int64_t n_layers) {
    return guarded([&] {
        require_single(e);
        e->ckpt_valid = false;
        if (n_layers != static_cast<int64_t>(e->layers.size()))
            throw RunError("parameter_mapping has " + std::to_string(n_layers) + " entries; expected " +
                           std::to_string(e->layers.size()) + " (per graph layer).");
        int64_t np = 0;
        for (int64_t i = 0; i < n_layers; ++i) {
            if (param_index[i] < 0) throw ArgError("parameter indices must be non-negative");
            np = std::max(np, param_index[i] + 1);
        }
        for (int64_t i = 0; i < n_layers; ++i) e->layers[static_cast<size_t>(i)].param_index = param_index[i];
        e->graph_nparams = np;
    });
}

int mpa_is_schrodinger(mpa_engine *e, int32_t *out) {
    *out = e->schr ? 1 : 0;
    return 0;
}

int mpa_batch_create(mpa_engine **out, int basis, int num_modes, uint32_t cutoff, int cutoff_type, double lower_atol,
                     double upper_atol, int32_t n_members, const int64_t *member_offsets, const int64_t *term_offsets,
                     const int32_t *term_slots, const double *coeff_re, const double *coeff_im,
                     const int32_t *initial_state, int64_t n_state, int device) {
    return guarded([&] {
        if (!out) throw ArgError("out is null");
        *out = nullptr;
        if (n_members < 1 || static_cast<uint32_t>(n_members) > kMaxBatch)
            throw ArgError("a batch holds 1 to " + std::to_string(kMaxBatch) + " members, got " +
                           std::to_string(n_members));
        auto e = new_engine(basis, num_modes, cutoff, cutoff_type, lower_atol, upper_atol, initial_state, n_state, device);
        e->nbatch = static_cast<uint32_t>(n_members);
        e->bcore.assign(static_cast<size_t>(n_members), 0.0);
        std::vector<uint64_t> hrows;
        std::vector<double> hcoef;
        std::vector<uint16_t> hmem;
        for (int32_t q = 0; q < n_members; ++q) {
            if (member_offsets[q + 1] < member_offsets[q]) throw ArgError("member_offsets must be non-decreasing");
            encode_operator(e.get(), term_offsets, term_slots, coeff_re, coeff_im, member_offsets[q],
                            member_offsets[q + 1], hrows, hcoef, &e->bcore[static_cast<size_t>(q)]);
            hmem.resize(hcoef.size(), static_cast<uint16_t>(q));
        }
        load_initial(e.get(), hcoef.size(), hrows.data(), hcoef.data(), hmem.data());
        *out = e.release();
    });
}

int mpa_destroy(mpa_engine *e) {
    return guarded([&] { delete e; });
}

int mpa_set_truncation(mpa_engine *e, uint32_t cutoff, int cutoff_type, double lower_atol, double upper_atol) {
    return guarded([&] { set_trunc(e, cutoff, cutoff_type, lower_atol, upper_atol); });
}

int mpa_set_basis_change(mpa_engine *e, const int64_t *offsets, const int32_t *slots, int64_t n_rows) {
    return guarded([&] { set_basis_change(e, offsets, slots, n_rows); });
}

int mpa_propagate(mpa_engine *e, const int64_t *gen_offsets, const int32_t *gen_slots, int64_t n_layers,
                  const int64_t *param_map, const double *gen_coeffs, const double *params, int64_t n_params,
                  int32_t only_rotate_len_k) {
    return guarded([&] {
        require_single(e);
        if (only_rotate_len_k == 0 || only_rotate_len_k > 2 * e->n_modes)
            throw ArgError("only_rotate_len_k is out of range; must be 0 < k <= 2*num_modes");
        if (!e->layers.empty())
            throw RunError("Cannot propagate() on top of a non-empty graph of " + std::to_string(e->layers.size()) +
                           " layer(s): propagate() evolves and contracts in place and assumes no stored graph.");
        if (n_layers == 0) return;
        check_params_length(param_map, n_layers, params, n_params);
        std::vector<GateDesc> gds;
        std::vector<int32_t> folds;
        make_gates(e, gen_offsets, gen_slots, n_layers, param_map, gen_coeffs, params, n_params,
                   only_rotate_len_k < 0, gds, folds);
        Trunc trl = e->tr;
        trl.only_rot = only_rotate_len_k < 0 ? -1 : only_rotate_len_k;
        bool fused = only_rotate_len_k < 0 && e->fused_enabled;
        for (const GateDesc &gd : gds) fused = fused && (gd.identity || gd.fused);
        RunPropagate run{e, gds, folds, false, trl, fused};
        e->dispatch(run);
    });
}

int mpa_batch_propagate(mpa_engine *e, const int64_t *gen_offsets, const int32_t *gen_slots, int64_t n_layers,
                        const int64_t *param_map, const double *gen_coeffs, const double *params, int64_t n_params,
                        int64_t max_capacity, int64_t *layers_done) {
    return guarded([&] {
        require_batch(e);
        if (layers_done) *layers_done = 0;
        if (n_layers == 0) return;
        if (max_capacity < 0) {  // automatic: a growth (old + new store) stays within 3/4 of what is available
            size_t free_b = 0, total_b = 0;
            HIP_TRY(hipMemGetInfo(&free_b, &total_b));
            const double avail = static_cast<double>(free_b) +
                                 static_cast<double>(BlockCache::get().cached_bytes(e->device)) +
                                 e->bytes_per_row() * static_cast<double>(e->cap);
            // the last growth holds the old store and the new one (1.5 times larger) at once
            const double rows = 0.75 * avail / (2.5 * e->bytes_per_row()) / static_cast<double>(-max_capacity);
            e->cap_limit = std::max<uint64_t>(static_cast<uint64_t>(rows) / mpa_engine::kGranule, 1) * mpa_engine::kGranule;
        } else {
            e->cap_limit = static_cast<uint64_t>(max_capacity);
        }
        std::vector<GateDesc> gds;
        std::vector<int32_t> folds;
        make_gates(e, gen_offsets, gen_slots, n_layers, param_map, gen_coeffs, params, n_params, true, gds, folds);
        // member q's constants from its own parameter row, exactly as a lone propagator derives them
        const uint32_t nb = e->nbatch;
        std::vector<Trig> trig(gds.size() * nb);
        for (size_t g = 0; g < gds.size(); ++g) {
            const GateDesc &gd = gds[g];
            for (uint32_t q = 0; q < nb; ++q) {
                Trig &t = trig[g * nb + q];
                if (gd.identity) {
                    t = Trig{1.0, 0.0, 1.0, 0.0};
                    continue;
                }
                const double alpha = 1.0 * params[static_cast<int64_t>(q) * n_params + gd.param_index] * gd.gen_coeff;
                sin_cos(2 * alpha, &t.s, &t.k);
                t.abs_s = std::abs(t.s);
                const double cos_build = t.k;
                if (cos_build == 0.0)
                    throw RunError("batch member " + std::to_string(q) + " has cos(2 alpha) == 0 at a gate; the batched "
                                   "pipeline needs the fused cos sweep, propagate that member alone");
                t.inv_k = 1.0 / cos_build;
            }
        }
        e->btrig.reserve(std::max<size_t>(trig.size(), 1) * sizeof(Trig));
        HIP_TRY(hipMemcpyAsync(e->btrig.p, trig.data(), trig.size() * sizeof(Trig), hipMemcpyHostToDevice, e->stream));
        RunBatch run{e, gds, folds};
        try {
            e->dispatch(run);
        } catch (...) {
            e->cap_limit = 0;
            throw;
        }
        e->cap_limit = 0;
        HIP_TRY(hipStreamSynchronize(e->stream));
        if (layers_done) *layers_done = static_cast<int64_t>(e->fused_done);
    });
}

int mpa_batch_create_rows(mpa_engine **out, int basis, int num_modes, uint32_t cutoff, int cutoff_type,
                          double lower_atol, double upper_atol, int32_t n_members, const double *cores, int64_t n_rows,
                          const uint64_t *words, const double *coeffs, const int32_t *members,
                          const int32_t *initial_state, int64_t n_state, int device) {
    return guarded([&] {
        if (!out) throw ArgError("out is null");
        *out = nullptr;
        if (n_members < 1 || static_cast<uint32_t>(n_members) > kMaxBatch)
            throw ArgError("a batch holds 1 to " + std::to_string(kMaxBatch) + " members, got " +
                           std::to_string(n_members));
        if (n_rows < 0 || n_rows >= (int64_t{1} << 31)) throw ArgError("row count out of range");
        auto e = new_engine(basis, num_modes, cutoff, cutoff_type, lower_atol, upper_atol, initial_state, n_state, device);
        e->nbatch = static_cast<uint32_t>(n_members);
        e->bcore.assign(cores, cores + n_members);
        const size_t nr = static_cast<size_t>(n_rows);
        std::vector<uint16_t> hmem(nr);
        for (size_t i = 0; i < nr; ++i) {
            if (members[i] < 0 || members[i] >= n_members) throw ArgError("row member out of range");
            hmem[i] = static_cast<uint16_t>(members[i]);
        }
        load_initial(e.get(), static_cast<uint64_t>(nr), words, coeffs, hmem.data());
        *out = e.release();
    });
}

int mpa_batch_split(mpa_engine *e, const uint8_t *move, mpa_engine **out) {
    return guarded([&] {
        require_batch(e);
        if (!out) throw ArgError("out is null");
        *out = nullptr;
        const uint32_t nb = e->nbatch;
        std::vector<uint8_t> side(nb);
        uint32_t n_mv = 0;
        for (uint32_t q = 0; q < nb; ++q) {
            side[q] = move[q] ? 1 : 0;
            n_mv += side[q];
        }
        if (n_mv == 0 || n_mv == nb) throw ArgError("a split moves at least one member and keeps at least one");
        std::vector<int64_t> sizes(nb);
        e->member_sizes(sizes.data());
        uint64_t rows_mv = 0;
        std::vector<uint16_t> renum(nb);
        std::vector<double> core_mv, core_st;
        for (uint32_t q = 0; q < nb; ++q) {
            if (side[q]) {
                rows_mv += static_cast<uint64_t>(sizes[q]);
                renum[q] = static_cast<uint16_t>(core_mv.size());
                core_mv.push_back(e->bcore[q]);
            } else {
                renum[q] = static_cast<uint16_t>(core_st.size());
                core_st.push_back(e->bcore[q]);
            }
        }
        const uint64_t rows_st = e->N - rows_mv;
        auto o = clone_config(e);
        o->nbatch = n_mv;
        o->bcore = core_mv;
        o->alloc_store(exact_capacity(rows_mv));
        HIP_TRY(hipStreamSynchronize(o->stream));
        // the staying members' new store, on e's stream
        const uint64_t cst = exact_capacity(rows_st);
        DevArray r2{&e->stream}, c2{&e->stream}, m2{&e->stream}, k2{&e->stream}, dside{&e->stream},
            drenum{&e->stream}, dblk{&e->stream};
        r2.reserve(cst * e->nw * 8);
        c2.reserve(cst * 8);
        m2.reserve(cst * 2);
        k2.reserve(static_cast<size_t>(e->ncols) * (cst / 64) * 8);
        HIP_TRY(hipMemsetAsync(k2.p, 0, static_cast<size_t>(e->ncols) * (cst / 64) * 8, e->stream));
        const uint64_t nblk = (e->N + kBlock - 1) / kBlock;
        dside.reserve(nb);
        drenum.reserve(static_cast<size_t>(nb) * 2);
        dblk.reserve(std::max<uint64_t>(nblk, 1) * 4);
        HIP_TRY(hipMemcpyAsync(dside.p, side.data(), nb, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(drenum.p, renum.data(), static_cast<size_t>(nb) * 2, hipMemcpyHostToDevice, e->stream));
        if (e->N) {
            hipLaunchKernelGGL(k_part_count, dim3(static_cast<uint32_t>(nblk)), dim3(kBlock), 0, e->stream,
                               e->bmember.as<uint16_t>(), e->N, dside.as<uint8_t>(), dblk.as<uint32_t>());
            HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, e->stream, dblk.as<uint32_t>(), nullptr,
                               static_cast<uint32_t>(nblk), e->counters.as<uint64_t>(), -1, -1, 0, 0);
            HIP_TRY(hipGetLastError());
            RunPartition part{e, dside.as<uint8_t>(), drenum.as<uint16_t>(), dblk.as<uint32_t>(),
                              o->rows.as<uint64_t>(), r2.as<uint64_t>(), o->coef.as<double>(), c2.as<double>(),
                              o->bmember.as<uint16_t>(), m2.as<uint16_t>()};
            e->dispatch(part);
        }
        HIP_TRY(hipStreamSynchronize(e->stream));
        std::swap(e->rows.p, r2.p);
        std::swap(e->rows.bytes, r2.bytes);
        std::swap(e->coef.p, c2.p);
        std::swap(e->coef.bytes, c2.bytes);
        std::swap(e->bmember.p, m2.p);
        std::swap(e->bmember.bytes, m2.bytes);
        std::swap(e->cols.p, k2.p);
        std::swap(e->cols.bytes, k2.bytes);
        e->cap = cst;
        e->N = rows_st;
        e->nbatch = nb - n_mv;
        e->bcore = core_st;
        e->release_work_arrays();
        o->N = rows_mv;
        for (mpa_engine *x : {e, o.get()}) {
            if (x->N) {
                InitCols init{x, x->N};
                x->dispatch(init);
            }
            const int64_t c0[1] = {static_cast<int64_t>(x->N)};
            HIP_TRY(hipMemcpyAsync(x->counters.p, c0, 8, hipMemcpyHostToDevice, x->stream));
            HIP_TRY(hipStreamSynchronize(x->stream));
        }
        *out = o.release();
    });
}

int mpa_clone(mpa_engine *e, mpa_engine **out) {
    return guarded([&] {
        require_single(e);
        if (!out) throw ArgError("out is null");
        *out = nullptr;
        HIP_TRY(hipStreamSynchronize(e->stream));
        auto o = clone_config(e);
        o->core = e->core;
        o->schr = e->schr;
        o->obs_rows = e->obs_rows;
        o->obs_coef = e->obs_coef;
        o->n_init = e->n_init;
        o->init_coef = e->init_coef;
        // the store at the same capacity: rows, coefficients and the column index copied as they are;
        // the row index is rebuilt by the next fused propagate (gidx_rows = 0)
        o->alloc_store(e->cap);
        const uint64_t N = e->N;
        if (N) {
            HIP_TRY(hipMemcpyAsync(o->rows.p, e->rows.p, N * e->nw * 8, hipMemcpyDeviceToDevice, o->stream));
            HIP_TRY(hipMemcpyAsync(o->coef.p, e->coef.p, N * 8, hipMemcpyDeviceToDevice, o->stream));
        }
        HIP_TRY(hipMemcpyAsync(o->cols.p, e->cols.p, static_cast<size_t>(e->ncols) * (e->cap / 64) * 8,
                               hipMemcpyDeviceToDevice, o->stream));
        o->N = N;
        const int64_t c0[1] = {static_cast<int64_t>(N)};
        HIP_TRY(hipMemcpyAsync(o->counters.p, c0, 8, hipMemcpyHostToDevice, o->stream));
        // the graph: layer list, cos-only and rotation records, the vector it replays from
        o->layers = e->layers;
        o->graph_nparams = e->graph_nparams;
        o->gcos_total = e->gcos_total;
        o->grot_total = e->grot_total;
        o->cofs_total = e->cofs_total;
        o->graph_folds = e->graph_folds;
        o->base_n = e->base_n;
        if (!e->graph_folds.empty()) {
            o->g_fold.reserve(e->graph_folds.size() * 4);
            HIP_TRY(hipMemcpyAsync(o->g_fold.p, e->graph_folds.data(), e->graph_folds.size() * 4, hipMemcpyHostToDevice,
                                   o->stream));
        }
        if (e->cofs_total) {
            o->g_cofs.reserve(e->cofs_total * 4);
            HIP_TRY(hipMemcpyAsync(o->g_cofs.p, e->g_cofs.p, e->cofs_total * 4, hipMemcpyDeviceToDevice, o->stream));
        }
        if (e->grot_total) {
            o->g_rot.reserve(e->grot_total * 8);
            HIP_TRY(hipMemcpyAsync(o->g_rot.p, e->g_rot.p, e->grot_total * 8, hipMemcpyDeviceToDevice, o->stream));
        }
        if (e->base_n) {
            o->base.reserve(e->base_n * 8);
            HIP_TRY(hipMemcpyAsync(o->base.p, e->base.p, e->base_n * 8, hipMemcpyDeviceToDevice, o->stream));
        }
        o->stat_n = e->stat_n;
        o->stat_a = e->stat_a;
        o->stat_e = e->stat_e;
        o->stat_i = e->stat_i;
        o->stat_l = e->stat_l;
        o->stat_p = e->stat_p;
        HIP_TRY(hipStreamSynchronize(o->stream));
        *out = o.release();
    });
}

int mpa_memory_breakdown(mpa_engine *e, int64_t *out, int32_t n) {
    return guarded([&] {
        if (n < kMemFields) throw ArgError("memory breakdown needs " + std::to_string(kMemFields) + " entries");
        const int64_t coef_b = static_cast<int64_t>(e->coef.bytes);
        const int64_t obs_b = static_cast<int64_t>(e->obs_coef.size() * 8 + e->obs_tab.bytes + e->obs_c.bytes);
        const int64_t init_b = static_cast<int64_t>(e->obs_rows.size() * 8 + e->init_coef.size() * 8);
        int64_t f[kMemFields] = {};
        f[0] = static_cast<int64_t>(e->rows.bytes + e->bmember.bytes);   // operator_terms_bytes
        f[1] = e->schr ? obs_b : coef_b;                                  // op_coeffs_bytes
        f[2] = e->schr ? coef_b : 0;                                      // state_coeffs_bytes (Heisenberg: scored on the fly)
        f[3] = static_cast<int64_t>(e->gidx.bytes + e->table.bytes);      // indexing_bytes
        f[4] = init_b;                                                    // init_operator_bytes
        f[5] = static_cast<int64_t>(e->occ_dev.bytes);                    // initial_state_bytes
        f[6] = static_cast<int64_t>(e->cols.bytes);                       // inverted_index_bytes (bit columns)
        f[7] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6];            // total_bytes
        f[8] = static_cast<int64_t>((e->cap - std::min(e->cap, e->N)) * e->nw * 8);  // d_terms_slack_bytes
        f[9] = static_cast<int64_t>(e->g_rot.bytes + e->g_fold.bytes + e->g_cofs.bytes + e->base.bytes);  // graph bytes
        size_t work = 0;  // propagate / evaluation scratch the engine keeps between calls
        for (const DevArray *a : {&e->f_elead, &e->t_val, &e->t_item, &e->t_bsum, &e->t_gsum, &e->desc_m,
                                  &e->AL, &e->lead, &e->foll, &e->blk_a, &e->blk_b, &e->rec, &e->lv, &e->match,
                                  &e->miss, &e->ins, &e->seedc, &e->rep, &e->snap_c, &e->snap_r, &e->lam, &e->pmark, &e->pmark1,
                                  &e->gpartial, &e->sum_blk, &e->sum_v, &e->sum_m, &e->sum_out})
            work += a->bytes;
        f[10] = static_cast<int64_t>(work);
        std::copy(f, f + kMemFields, out);
    });
}

int mpa_batch_members(mpa_engine *e, int32_t *out) {
    return guarded([&] {
        require_batch(e);
        *out = static_cast<int32_t>(e->nbatch);
    });
}

int mpa_batch_expectation_values(mpa_engine *e, double *out) {
    return guarded([&] {
        require_batch(e);
        RunExpectBatch ev{e, out};
        e->dispatch(ev);
    });
}

int mpa_batch_sizes(mpa_engine *e, int64_t *out) {
    return guarded([&] {
        require_batch(e);
        e->member_sizes(out);
    });
}

int mpa_batch_row_members(mpa_engine *e, int32_t *out) {
    return guarded([&] {
        require_batch(e);
        std::vector<uint16_t> h(e->N);
        if (e->N) HIP_TRY(hipMemcpyAsync(h.data(), e->bmember.p, e->N * 2, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
        for (uint64_t i = 0; i < e->N; ++i) out[i] = h[i];
    });
}

int mpa_build_graph(mpa_engine *e, const int64_t *gen_offsets, const int32_t *gen_slots, int64_t n_layers,
                    const int64_t *param_map, const double *gen_coeffs, const double *params, int64_t n_params,
                    int32_t only_rotate_len_k) {
    return guarded([&] {
        require_single(e);
        if (only_rotate_len_k == 0 || only_rotate_len_k > 2 * e->n_modes)
            throw ArgError("only_rotate_len_k is out of range; must be 0 < k <= 2*num_modes");
        if (n_layers == 0) return;
        e->ckpt_valid = false;
        check_params_length(param_map, n_layers, params, n_params);
        std::vector<GateDesc> gds;
        std::vector<int32_t> folds;
        make_gates(e, gen_offsets, gen_slots, n_layers, param_map, gen_coeffs, params, n_params, false, gds, folds);
        // the seed the coefficient-informed build evolves (MonomialPropagator.inl:653-689)
        e->seedc.reserve(std::max<uint64_t>(e->cap, 1) * 8);
        if (e->layers.empty()) {
            e->base.reserve(std::max<uint64_t>(e->N, 1) * 8);
            if (e->N) HIP_TRY(hipMemcpyAsync(e->base.p, e->coef.p, e->N * 8, hipMemcpyDeviceToDevice, e->stream));
            e->base_n = e->N;
            if (e->N) HIP_TRY(hipMemcpyAsync(e->seedc.p, e->coef.p, e->N * 8, hipMemcpyDeviceToDevice, e->stream));
        } else if (params != nullptr) {
            if (n_params < e->graph_nparams)
                throw ArgError("Coefficient-informed build_graph() needs at least " +
                               std::to_string(e->graph_nparams) + " parameter value(s) to replay the existing graph.");
            e->replay(params, e->schr ? mpa_engine::kState : mpa_engine::kOperator);
            if (e->N) HIP_TRY(hipMemcpyAsync(e->seedc.p, e->rep.p, e->N * 8, hipMemcpyDeviceToDevice, e->stream));
        } else if (e->N) {
            HIP_TRY(hipMemcpyAsync(e->seedc.p, e->coef.p, e->N * 8, hipMemcpyDeviceToDevice, e->stream));
        }
        HIP_TRY(hipStreamSynchronize(e->stream));
        // The orbital cap gates emission and the seed's cosine (Scan.h:419-433); the records keep every
        // anticommuting row in the cosine set, since a stored layer's replay recomputes that set over all
        // rows before its scaled_count (CosineRecompute.h:17-20, MonomialPropagator.inl:784-790).
        Trunc trl = e->tr;
        trl.only_rot = only_rotate_len_k < 0 ? -1 : only_rotate_len_k;
        if (params == nullptr) {  // structural build: no coefficient checks (Scan.h:81-95)
            trl.check_lower = 0;
            trl.check_upper = 0;
        }
        const size_t first_layer = e->layers.size();
        // without an orbital cap the build runs on the fused kernels (GraphArgs); the cap's two-pass
        // cosine arm, and mpa_set_pipeline(e, 0), take the per-gate pipeline
        const bool fused = only_rotate_len_k < 0 && e->fused_enabled;
        e->building = true;
        try {
            RunPropagate run{e, gds, folds, true, trl, fused};
            e->dispatch(run);
        } catch (...) {
            e->building = false;
            throw;
        }
        e->building = false;
        for (int64_t i = 0; i < n_layers; ++i) e->graph_nparams = std::max(e->graph_nparams, param_map[i] + 1);
        e->shrink_keep(e->g_rot, e->grot_total * 8);  // the records' array at its exact size
        e->graph_finalize(first_layer);
    });
}

int mpa_expectation_value(mpa_engine *e, const double *params, int64_t n_params, double *out) {
    return guarded([&] {
        require_single(e);
        if (e->keep_ckpt && !e->layers.empty()) {  // the forward a following gradient starts from
            RunForwardCheckpointed fw{e, params, n_params, out};
            e->dispatch(fw);
            return;
        }
        if (e->schr) {
            RunSchrExpect ev{e, params, n_params, out};
            e->dispatch(ev);
            return;
        }
        const double *c = evaluated_coeffs(e, params, n_params);
        RunExpect ev{e, c, out};
        e->dispatch(ev);
    });
}

int mpa_expectation_value_noalloc(mpa_engine *e, double *out) {
    return guarded([&] {
        require_single(e);
        if (e->schr) throw ArgError("the allocation-free <O> serves the Heisenberg picture only");
        if (!e->layers.empty())
            throw ArgError("the allocation-free <O> needs a contracted operator (no stored graph); use "
                           "mpa_expectation_value");
        HIP_TRY(hipSetDevice(e->device));
        RunExpectFixed ev{e, static_cast<unsigned char *>(noalloc_scratch(e->device)), out};
        e->dispatch(ev);
    });
}

int mpa_expectation_value_and_gradient(mpa_engine *e, const double *params, int64_t n_params, double *value,
                                       double *grad) {
    return guarded([&] {
        require_single(e);
        if (e->layers.empty()) {  // no graph: the contracted value, no parameters to differentiate
            if (e->schr) {
                RunSchrExpect ev{e, params, n_params, value};
                e->dispatch(ev);
            } else {
                const double *c = evaluated_coeffs(e, params, n_params);
                RunExpect ev{e, c, value};
                e->dispatch(ev);
            }
            return;
        }
        const bool cached = e->ckpt_valid && n_params == static_cast<int64_t>(e->ckpt_params.size()) &&
                            std::memcmp(params, e->ckpt_params.data(), static_cast<size_t>(n_params) * 8) == 0;
        if (cached) {
            *value = e->ckpt_value;
        } else {
            RunForwardCheckpointed fw{e, params, n_params, value};
            e->dispatch(fw);
        }
        RunGradient rg{e, params, n_params, grad};
        e->dispatch(rg);
    });
}

int mpa_set_gradient_checkpointing(mpa_engine *e, int keep, int64_t budget_bytes) {
    return guarded([&] {
        if (budget_bytes < 0) throw ArgError("budget_bytes must be non-negative");
        e->keep_ckpt = keep != 0;
        e->grad_budget = budget_bytes;
        e->ckpt_valid = false;
    });
}

int mpa_gradient_segments(mpa_engine *e, int64_t *segments, int64_t *checkpoint_bytes, int64_t *snapshot_bytes) {
    return guarded([&] {
        // what the last plan needs (its arrays may be larger: they only grow, and cached device blocks
        // can be larger than asked for)
        const size_t ns = e->seg_start.empty() ? 0 : e->seg_start.size() - 1;
        *segments = static_cast<int64_t>(ns);
        *checkpoint_bytes = ns > 1 ? static_cast<int64_t>((ns - 1) * std::max<uint64_t>(e->N, 1) * 8) : 0;
        *snapshot_bytes = ns ? static_cast<int64_t>(e->seg_snap_c * 8 + e->seg_snap_r * 16) : 0;
    });
}

int mpa_contract(mpa_engine *e, const double *params, int64_t n_params, int inplace, double *out) {
    return guarded([&] {
        require_single(e);
        const double *c = evaluated_coeffs(e, params, n_params);
        if (out && e->N) {
            HIP_TRY(hipMemcpyAsync(out, c, e->N * 8, hipMemcpyDeviceToHost, e->stream));
        }
        if (inplace && !e->layers.empty()) {
            if (e->N) HIP_TRY(hipMemcpyAsync(e->coef.p, c, e->N * 8, hipMemcpyDeviceToDevice, e->stream));
            e->layers.clear();
            e->graph_folds.clear();
            e->graph_nparams = 0;
            e->gcos_total = e->grot_total = e->cofs_total = 0;
            e->base_n = 0;
        }
        HIP_TRY(hipStreamSynchronize(e->stream));
    });
}

int mpa_size(mpa_engine *e, int64_t *out) {
    return guarded([&] { *out = static_cast<int64_t>(e->N); });
}

int mpa_graph_layers(mpa_engine *e, int64_t *layers, int64_t *n_params) {
    *layers = static_cast<int64_t>(e->layers.size());
    *n_params = e->graph_nparams;
    return 0;
}

int mpa_graph_records(mpa_engine *e, int64_t *cos_only, int64_t *rotations) {
    *cos_only = static_cast<int64_t>(e->gcos_total);
    *rotations = static_cast<int64_t>(e->grot_total);
    return 0;
}

int mpa_words_per_row(mpa_engine *e, int32_t *out) {
    *out = e->nw;
    return 0;
}

int mpa_read_rows(mpa_engine *e, uint64_t *words, double *coeffs) {
    return guarded([&] {
        if (e->N == 0) return;
        HIP_TRY(hipMemcpyAsync(words, e->rows.p, e->N * e->nw * 8, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipMemcpyAsync(coeffs, e->coef.p, e->N * 8, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    });
}

int mpa_export_terms(mpa_engine *e, const double *params, int64_t n_params, double atol, int64_t *n_terms,
                     int64_t *n_slots, int64_t *offsets, int32_t *slots, double *re, double *im) {
    return guarded([&] {
        require_single(e);
        const double *cdev = evaluated_coeffs(e, params, n_params);
        std::vector<uint64_t> w(e->N * e->nw);
        std::vector<double> c(e->N);
        if (e->N) {
            HIP_TRY(hipMemcpyAsync(w.data(), e->rows.p, e->N * e->nw * 8, hipMemcpyDeviceToHost, e->stream));
            HIP_TRY(hipMemcpyAsync(c.data(), cdev, e->N * 8, hipMemcpyDeviceToHost, e->stream));
            HIP_TRY(hipStreamSynchronize(e->stream));
        }
        std::vector<int32_t> tmp;
        int64_t nt = 0, ns = 0;
        static const std::complex<double> kPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
        for (uint64_t i = 0; i < e->N; ++i) {
            if (std::abs(c[i]) < atol) continue;
            e->words_to_slots(w.data() + i * e->nw, tmp);
            if (slots) {
                offsets[nt] = ns;
                std::copy(tmp.begin(), tmp.end(), slots + ns);
                std::complex<double> dec(c[i], 0.0);
                if (e->basis == MPA_BASIS_MAJORANA) dec = dec * kPow[(tmp.size() * (tmp.size() - 1) / 2) % 4];
                re[nt] = std::round(dec.real() * 1e12) / 1e12;
                im[nt] = std::round(dec.imag() * 1e12) / 1e12;
            }
            ++nt;
            ns += static_cast<int64_t>(tmp.size());
        }
        if (slots) offsets[nt] = ns;
        // core term, Heisenberg picture only (binder.h:212-215)
        if (!e->schr && std::abs(e->core) >= atol) {
            if (slots) {
                offsets[nt + 1] = ns;
                re[nt] = e->core;
                im[nt] = 0.0;
            }
            ++nt;
        }
        *n_terms = nt;
        *n_slots = ns;
    });
}

int mpa_gate_stats(mpa_engine *e, int64_t *len, int64_t *n, int64_t *a, int64_t *em, int64_t *ins) {
    return guarded([&] {
        const int64_t have = static_cast<int64_t>(e->stat_n.size());
        if (!n) {
            *len = have;
            return;
        }
        const int64_t k = std::min(*len, have);
        for (int64_t i = 0; i < k; ++i) {
            n[i] = e->stat_n[i];
            a[i] = e->stat_a[i];
            em[i] = e->stat_e[i];
            ins[i] = e->stat_i[i];
        }
        *len = k;
    });
}

int mpa_set_launch(mpa_engine *e, int32_t resolve_grid, int32_t min_group_log2) {
    return guarded([&] {
        if (resolve_grid < 1 || resolve_grid > 65536) throw ArgError("resolve_grid must be in 1..65536");
        if (min_group_log2 < 2 || min_group_log2 > kMaxGroupLog2)
            throw ArgError("min_group_log2 must be in 2.." + std::to_string(kMaxGroupLog2));
        e->res_grid_cap = static_cast<uint32_t>(resolve_grid);
        if (e->min_group_log2 != min_group_log2) {
            e->min_group_log2 = min_group_log2;
            e->fused_cap = 0;  // the count groups are laid out again before the next fused gate
        }
    });
}

int mpa_set_store_limit(mpa_engine *e, int64_t rows) {
    return guarded([&] {
        if (rows < 0) throw ArgError("rows must be non-negative");
        e->store_limit = static_cast<uint64_t>(rows);
    });
}

int mpa_set_pipeline(mpa_engine *e, int fused) {
    e->fused_enabled = fused != 0;
    return 0;
}

int mpa_set_profiling(mpa_engine *e, int on) {
    return guarded([&] {
        e->prof_on = on != 0;
        for (int k = 0; k < K_KINDS; ++k) {
            e->prof_ms[k] = 0.0;
            e->prof_bytes[k] = 0.0;
            e->prof_launches[k] = 0;
        }
    });
}

int mpa_kernel_profile(mpa_engine *e, int kind, const char **name, double *ms, double *bytes, int64_t *launches) {
    if (kind < 0 || kind >= K_KINDS) return set_error(-1, "kernel kind out of range");
    *name = kKindNames[kind];
    *ms = e->prof_ms[kind];
    *bytes = e->prof_bytes[kind];
    *launches = e->prof_launches[kind];
    return 0;
}

int mpa_last_propagate_ms(mpa_engine *e, double *ms) {
    *ms = e->last_propagate_ms;
    return 0;
}

}  // extern "C"
