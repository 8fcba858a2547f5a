// Kernels of the per-gate pipeline (orbital cap, mpa_set_pipeline(e, 0)).
// Part of the monoprop_amd engine; included by engine.hip (one translation unit).

#pragma once

namespace {

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

}  // namespace
