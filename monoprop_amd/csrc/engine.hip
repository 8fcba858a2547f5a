// monoprop_amd engine: Heisenberg operator propagation on one MI355X (gfx950).
//
// One translation unit: this file (the C ABI) includes, in order, device_common.h (constants, wave
// helpers, per-gate constants), kernels_pergate.h, kernels_fused.h (row index, k_sweep, k_resolve),
// kernels_reduce_graph.h (ordered <O> sums, graph layers, batch partition) and engine_host.h (device
// memory cache, mpa_engine).
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

#include "device_common.h"
#include "kernels_pergate.h"
#include "kernels_fused.h"
#include "kernels_reduce_graph.h"
#include "engine_host.h"

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
