// Ordered <O> sums, graph replay / adjoint layer kernels, batch partition and reductions.
// Part of the monoprop_amd engine; included by engine.hip (one translation unit).

#pragma once

namespace {

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

}  // namespace
