// Launch constants, error types, wave-level device helpers and the per-gate constants.
// Part of the monoprop_amd engine; included by engine.hip (one translation unit).

#pragma once

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

}  // namespace
