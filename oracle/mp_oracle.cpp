// Oracle for monoprop_amd -- TEST INFRASTRUCTURE ONLY.
//
// A serial CPU restatement of the reference's single-partition (monoprop_PARTITIONS=off) operator
// propagation, used as the checker for the HIP engine.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library; the product path never calls it.
//
// Storage is this project's own canonical row layout (the one the GPU engine uses), not the
// reference's MSb0 Majorana bitset, so every layout-dependent rule (pivot bit, rotation sign, cutoff
// sums, state phase) is restated for it with the reference line it follows:
//   Pauli, L = ceil(n/64) limbs:  words[0..L)  = x-plane, words[L..2L) = z-plane, qubit q -> bit q.
//   Majorana, W = ceil(2n/64):    words[0..W)  = slot mask, Majorana slot s -> bit s.
// The reference's Pauli slot encoding X_q -> {2q}, Y_q -> {2q+1}, Z_q -> {2q,2q+1}
// (src/monoprop/conversion_utils.py:66-88) is converted at the boundary.
//
// Every floating-point step follows the reference's operation order, and this file is built with
// -ffp-contract=off (the reference compiles in ISO C++ mode without -march flags for its wheels,
// cmake/compiler_flags/CXXFlags.cmake:47-72, pyproject.toml:387, so no FMA contraction happens).
//
// Parity pins (checked in tests/): dense U^dag O U brute force for Pauli gates and circuits
// (pauli_build_layer_tests.cpp:126-259 method), exhaustive Pauli algebra on all 4^n strings,
// the reference's closed-form truncation table (tests/test_coeff_trunc.py:59-99, 142-227),
// the reference's msgpack fixtures with exact energies (tests/test_infinite_cutoff.py:42-70),
// the native-Pauli vs Jordan-Wigner-Majorana equivalence (pauli_build_layer_tests.cpp:325-390;
// checked through the Jordan-Wigner cutoff basis, gate by gate),
// and central finite differences / parameter shift for the gradient.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr int kMaxWords = 8;
constexpr uint64_t kEven = 0x5555555555555555ULL;  // even Majorana slots (2j) in each limb

thread_local std::string g_last_error;

// Threads of the per-gate scan, resolve and apply (orc_set_threads); 1 = the serial restatement.  The
// parallel ranges are joined in their serial order, so every result is the same for any count.
int g_threads = 1;
constexpr size_t kParMin = 1 << 14;  // rows or queries below which a step stays serial

inline size_t ranges_for(size_t n) {
    if (g_threads <= 1 || n < kParMin) return 1;
    return std::min<size_t>(static_cast<size_t>(g_threads) * 8, (n + kParMin / 4 - 1) / (kParMin / 4));
}

inline int popc(uint64_t x) { return __builtin_popcountll(x); }

// sin and cos of one angle from one glibc sincos call.  The reference evaluates std::sin and std::cos
// of the same angle side by side (FusedApply.h:47-48, Evolution.cpp:475-476, Scan.h:44 next to
// Engine.h:552), which GCC, the compiler of its Linux builds, turns into a single sincos call; glibc's
// sincos can differ from its sin or cos in the last bit (2a = -2.831787767853196: sin gives
// -0x1.38309465f345bp-2, sincos -0x1.38309465f345ap-2), so every angle here goes through sincos,
// whatever the compiler of this file decides.
inline void sin_cos(double x, double *s, double *c) { ::sincos(x, s, c); }

// n*(n-1)/2 in size_t arithmetic, as reference cpp/monoprop/Utilities.h:64-66.
inline size_t n_choose_2(size_t n) { return n * (n - 1) / 2; }

inline uint64_t splitmix(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}

struct Rotation {
    uint32_t src;
    uint32_t tgt;
    int32_t phase;
};

// One graph layer (reference LayerCore, graph_encoding/MPGraphEncodingTypes.h:115-167): the cos set
// is recomputed from the generator over [0, scaled_count) exactly as CosineRecompute.h:15-22 does.
struct Layer {
    std::vector<uint64_t> gen;
    uint64_t scaled_count;
    std::vector<Rotation> rots;  // hits in resolve order, then inserts (Engine.h:85-98)
    int64_t param_index;
    double gen_coeff;
};

struct Engine {
    int basis;  // 0 Majorana, 1 Pauli
    int n_modes;
    int L;      // Pauli limbs per plane
    int nw;     // words per row
    unsigned cutoff;
    int cutoff_type;  // 0 length, 1 support
    bool has_lower, has_upper;
    double lower_atol, upper_atol;
    std::vector<uint64_t> bchg;  // cutoff basis: 2n rows of nw words (row s = image of slot s); empty: native
    double core = 0.0;
    std::vector<uint64_t> rows;  // nw words per row, append-only
    std::vector<double> coeffs;  // current-picture coefficients (MPOperator::op_coeffs)
    std::vector<double> init_coeffs;  // initial operator coefficients at their rows
    std::vector<uint64_t> occ;   // occupation mask of the initial product state
    // Schrodinger picture (MonomialPropagator.inl:86-89): the rows hold the evolved state and the
    // observable's terms (core excluded) are looked up among them at evaluation time.
    bool schr = false;
    std::vector<uint64_t> obs_rows;
    std::vector<double> obs_coef;
    std::vector<uint32_t> table; // keyless open-addressing index over rows (OperatorIndex.h:278-367)
    uint64_t table_mask = 0;
    std::vector<Layer> graph;
    // per-gate counters of the last propagate: anticommuting, emitted, inserted, live-before
    std::vector<int64_t> stat_n, stat_a, stat_e, stat_i;

    size_t size() const { return coeffs.size(); }
    const uint64_t *row(size_t i) const { return rows.data() + i * nw; }

    uint64_t hash_words(const uint64_t *m) const {
        uint64_t h = 0;
        for (int w = 0; w < nw; ++w) h ^= splitmix(m[w] + static_cast<uint64_t>(w));
        return h;
    }
    bool equal(const uint64_t *a, const uint64_t *b) const {
        for (int w = 0; w < nw; ++w)
            if (a[w] != b[w]) return false;
        return true;
    }
    void rehash(size_t want) {
        size_t cap = 16;
        while (cap < 2 * want) cap <<= 1;
        table.assign(cap, UINT32_MAX);
        table_mask = cap - 1;
        for (size_t i = 0; i < size(); ++i) index_insert(static_cast<uint32_t>(i));
    }
    void index_insert(uint32_t idx) {
        uint64_t h = hash_words(row(idx)) & table_mask;
        while (table[h] != UINT32_MAX) h = (h + 1) & table_mask;
        table[h] = idx;
    }
    int64_t find(const uint64_t *key) const {
        if (table.empty()) return -1;
        uint64_t h = hash_words(key) & table_mask;
        while (table[h] != UINT32_MAX) {
            if (equal(row(table[h]), key)) return table[h];
            h = (h + 1) & table_mask;
        }
        return -1;
    }
    void append_row(const uint64_t *m, double c) {
        rows.insert(rows.end(), m, m + nw);
        coeffs.push_back(c);
        if (2 * size() > table.size()) rehash(size());
        else index_insert(static_cast<uint32_t>(size() - 1));
    }

    // --- algebra -------------------------------------------------------------------------------------

    // Anticommutation. Pauli: parity(x_M.z_G + z_M.x_G) (PauliAlgebra.h:78-83, the symplectic form of
    // parity(M & pair_swap(G))). Majorana: parity(|M&G|) xor (|G| odd and |M| odd) (Scan.h:125-185,
    // Algebra.h:76-78 odd-|G| row-parity correction).
    bool anticommutes(const uint64_t *m, const uint64_t *g, bool g_odd) const {
        uint64_t acc = 0;
        if (basis == 1) {
            for (int l = 0; l < L; ++l) acc ^= (m[l] & g[L + l]) ^ (m[L + l] & g[l]);
            return popc(acc) & 1;
        }
        uint64_t par = 0;
        for (int w = 0; w < nw; ++w) {
            acc ^= m[w] & g[w];
            par ^= m[w];
        }
        bool a = popc(acc) & 1;
        if (g_odd) a ^= (popc(par) & 1);
        return a;
    }

    // Slot popcount of a row: the reference's Monomial::count() (Pauli: X,Y one slot, Z two).
    int slot_pop(const uint64_t *m) const {
        int p = 0;
        if (basis == 1) {
            for (int l = 0; l < L; ++l) p += popc(m[l] ^ m[L + l]) + popc(m[L + l]);
            return p;
        }
        for (int w = 0; w < nw; ++w) p += popc(m[w]);
        return p;
    }

    // Structural keep test on the partner (AlgebraCommon.h:285-342): keep iff fully paired
    // (xor_sum == 0) or the measure <= cutoff.  Pauli: xor_sum = |x-plane|, support = |x|z|
    // (PauliAlgebra.h:17-21).  Majorana: per-mode pair sums over even/odd slots.
    // With a cutoff basis (MonomialPropagator.inl:484-495, MonomialPropagatorCommon.h:44-63) the test
    // runs on change_basis(t) (MajoranaAlgebra.h:178-189): the XOR of the images of t's slots.
    bool keep(const uint64_t *t) const {
        if (!bchg.empty()) {
            uint64_t mapped[kMaxWords] = {};
            for (int w = 0; w < nw; ++w)
                for (uint64_t b = t[w]; b; b &= b - 1) {
                    const size_t s = static_cast<size_t>(w) * 64 + static_cast<size_t>(__builtin_ctzll(b));
                    for (int k = 0; k < nw; ++k) mapped[k] ^= bchg[s * nw + k];
                }
            return keep_native(mapped);
        }
        return keep_native(t);
    }

    bool keep_native(const uint64_t *t) const {
        size_t xs = 0, measure = 0;
        if (basis == 1) {
            for (int l = 0; l < L; ++l) {
                xs += popc(t[l]);
                measure += popc(t[l] | t[L + l]);
            }
        } else {
            for (int w = 0; w < nw; ++w) {
                const uint64_t a = t[w] & kEven, b = (t[w] >> 1) & kEven;
                xs += popc(a ^ b);
                measure += cutoff_type == 0 ? popc(t[w]) : popc(a | b);
            }
        }
        return xs == 0 || measure <= cutoff;
    }

    // Pivot = the reference's lowest physical set bit of G (Scan.h:265-269, gen.find_first()); a
    // term with the pivot set is a follower.  Slot s sits at physical bit 2N-1-s
    // (AlgebraCommon.h:204-213), so the lowest physical bit is G's highest slot.  For Pauli that is
    // the highest qubit q* of supp(G); its even physical bit (slot 2q*+1, the z-plane) ranks below the
    // odd one (slot 2q*, the u = x^z plane).
    struct Pivot {
        int word;
        uint64_t bit;
        bool pauli_u;  // Pauli X-letter pivot: test x^z
    };
    Pivot pivot_of(const uint64_t *g) const {
        Pivot p{-1, 0, false};
        if (basis == 1) {
            for (int l = L - 1; l >= 0; --l) {
                const uint64_t s = g[l] | g[L + l];
                if (!s) continue;
                const int q = 63 - __builtin_clzll(s);
                p.bit = uint64_t{1} << q;
                if (g[L + l] & p.bit) {
                    p.word = L + l;  // v bit = z-plane
                } else {
                    p.word = l;
                    p.pauli_u = true;
                }
                return p;
            }
            return p;
        }
        for (int w = nw - 1; w >= 0; --w) {
            if (!g[w]) continue;
            p.word = w;
            p.bit = uint64_t{1} << (63 - __builtin_clzll(g[w]));
            return p;
        }
        return p;
    }
    bool is_follower(const uint64_t *m, const Pivot &p) const {
        if (p.word < 0) return false;
        if (basis == 1 && p.pauli_u) return ((m[p.word] ^ m[L + p.word]) & p.bit) != 0;
        return (m[p.word] & p.bit) != 0;
    }

    // Emitted sine phase of the partner M^G.
    // Pauli (PauliAlgebra.h:116-140): e = y_G + y_M - y_T + 2|z_M & x_G|, phase = -1 iff e = 1 mod 4,
    // with y = |x & z| (a Y letter has v=1,u=0, i.e. x=z=1).
    // Majorana (Algebra.h:51-59, MajoranaAlgebra.h:82-122): (-1)^{#(m in M, g in G): g < m} times
    // REAL_PARTS[(C(|M|,2) + C(|G|,2) - C(|T|,2) + 3) mod 4].
    int emit_phase(const uint64_t *m, const uint64_t *g, const uint64_t *t) const {
        if (basis == 1) {
            long e = 0;
            for (int l = 0; l < L; ++l) {
                e += popc(g[l] & g[L + l]) + popc(m[l] & m[L + l]) - popc(t[l] & t[L + l]);
                e += 2 * popc(m[L + l] & g[l]);
            }
            const int r = static_cast<int>(((e % 4) + 4) % 4);
            return r == 1 ? -1 : 1;
        }
        // interleave parity: for each slot of M count G slots strictly below it
        size_t par = 0;
        uint64_t carry = 0;  // parity of G slots in lower words
        for (int w = 0; w < nw; ++w) {
            // prefix parity of g strictly below each bit: prefix_xor(g) << 1
            uint64_t x = g[w];
            x ^= x << 1;
            x ^= x << 2;
            x ^= x << 4;
            x ^= x << 8;
            x ^= x << 16;
            x ^= x << 32;
            const uint64_t below = (x << 1) ^ (uint64_t{0} - carry);
            par ^= static_cast<size_t>(popc(below & m[w]));
            carry ^= static_cast<uint64_t>(popc(g[w]) & 1);
        }
        const int sign = (par & 1) ? -1 : 1;
        size_t pm = 0, pg = 0, pt = 0;
        for (int w = 0; w < nw; ++w) {
            pm += popc(m[w]);
            pg += popc(g[w]);
            pt += popc(t[w]);
        }
        static const int kRealParts[4] = {1, 0, -1, 0};
        const size_t power = (n_choose_2(pm) + n_choose_2(pg) - n_choose_2(pt) + 3) % 4;
        return sign * kRealParts[power];
    }

    // Fully paired + diagonal phase against the initial product state (MPOperator.h:235-254,
    // PauliAlgebra.h:142-147, MajoranaAlgebra.h:52-60).
    bool paired(const uint64_t *m) const {
        if (basis == 1) {
            for (int l = 0; l < L; ++l)
                if (m[l]) return false;
            return true;
        }
        for (int w = 0; w < nw; ++w)
            if ((m[w] & kEven) != ((m[w] >> 1) & kEven)) return false;
        return true;
    }
    double state_phase(const uint64_t *m) const {
        if (basis == 1) {
            int c = 0;
            for (int l = 0; l < L; ++l) c += popc(m[L + l] & occ[l]);
            return (c & 1) ? -1.0 : 1.0;
        }
        int c = 0, pop = 0;
        for (int w = 0; w < nw; ++w) {
            c += popc(m[w] & occ[w]);
            pop += popc(m[w]);
        }
        return ((c + pop / 2) % 2) ? -1.0 : 1.0;
    }

    // Value a row takes when it enters the Schrodinger state: its diagonal phase if fully paired, else
    // 0 (MPOperator.h:158-181 dense_state() scoring of every row it has not seen yet).
    double score(const uint64_t *m) const { return paired(m) ? state_phase(m) : 0.0; }

    // The observable's coefficients at the rows that hold its terms (MPOperator.h:119-141 get_operator():
    // a term joins the vector once some row holds it).
    std::vector<double> observable_vector() const {
        std::vector<double> op(size(), 0.0);
        for (size_t j = 0; j < obs_coef.size(); ++j) {
            const int64_t r = find(obs_rows.data() + j * nw);
            if (r >= 0) op[static_cast<size_t>(r)] = obs_coef[j];
        }
        return op;
    }

    // Schrodinger <O> = core + state . op in row order (MPFunctions.cpp:86-94, 196-205).
    double expectation_schr(const std::vector<double> &state, const std::vector<double> &op) const {
        double r = 0.0;
        for (size_t i = 0; i < state.size(); ++i) r += state[i] * op[i];
        return core + r;
    }

    // --- per-gate build ------------------------------------------------------------------------------

    // Emitted partner queries, flat: key words, phase, source row, pre-cos source value.
    struct Queries {
        std::vector<uint64_t> key;
        std::vector<int32_t> phase;
        std::vector<uint32_t> src;
        std::vector<double> v_src;
        size_t size() const { return src.size(); }
        void clear() {
            key.clear();
            phase.clear();
            src.clear();
            v_src.clear();
        }
    };
    Queries leaders_, followers_;
    std::vector<uint32_t> epoch_;  // MatchedEpochSet (layer_build/Common.h:29-48)
    uint32_t cur_epoch_ = 0;

    // One layer of the contract-immediately path (Engine.h:522-637 + FusedApply.h:32-83) when
    // `contract` is true, or the graph-build path (GraphSink, Engine.h:69-180) when false.
    // `coeff_view` is the coefficient vector the cutoffs read (nullptr = structural build).
    // `alpha` is the build angle (the cutoffs and the fused cos sweep); `apply_alpha` the angle of the
    // rotation, the build angle negated in the Schrodinger picture (MonomialPropagator.h:411-416).
    void build_layer(const uint64_t *g, bool have_param, double alpha, double apply_alpha, int only_rot,
                     bool contract, std::vector<double> *coeff_view, Layer *out_layer,
                     std::vector<uint32_t> *out_cos) {
        std::vector<double> *cp = contract ? &coeffs : coeff_view;
        const bool local_coeffs = cp != nullptr;
        const bool check_atol = has_lower && local_coeffs && have_param;
        const bool check_upper = has_upper && local_coeffs;
        double sin2 = 1.0, cos2 = 1.0;
        if (have_param) sin_cos(2 * alpha, &sin2, &cos2);
        const double sin_cut = sin2;  // Scan.h:81-95
        const double abs_sin = std::abs(sin_cut);
        const bool use_coeff_checks = check_atol || check_upper;
        const double cos_build = contract ? cos2 : 1.0;  // Engine.h:552
        const bool fused = contract && only_rot < 0 && have_param && cos_build != 0.0;
        const double inv_cos = fused ? 1.0 / cos_build : 1.0;  // Engine.h:614

        int gpop = 0;
        for (int w = 0; w < nw; ++w) gpop += popc(g[w]);
        const bool g_odd = basis == 0 && (gpop % 2 != 0);
        const size_t n0 = size();
        stat_n.push_back(static_cast<int64_t>(n0));
        int64_t n_anti = 0, n_emit = 0;

        Queries &leaders = leaders_, &followers = followers_;
        leaders.clear();
        followers.clear();
        std::vector<uint32_t> cos_set;
        const Pivot pv = pivot_of(g);

        auto is_below = [&](double abs_c) { return check_atol && (abs_sin * abs_c <= lower_atol); };
        // One range of rows: anticommutation, the fused cos sweep, the emission gate; queries and the cos
        // set in ascending row order.  Rows are independent, so ranges run in parallel and are joined in
        // range order (the serial order).
        struct ScanOut {
            Queries lead, foll;
            std::vector<uint32_t> cos;
            int64_t n_anti = 0, n_emit = 0;
        };
        auto scan_range = [&](size_t lo, size_t hi, ScanOut &out) {
            uint64_t tbuf[kMaxWords];
            auto emit = [&](int mono_pop, uint32_t i, double abs_c, double v_src, bool foll) {
                if (only_rot >= 0 && mono_pop > only_rot) return;
                if (is_below(abs_c)) return;
                const uint64_t *m = row(i);
                for (int w = 0; w < nw; ++w) tbuf[w] = m[w] ^ g[w];
                const bool struct_pass = keep(tbuf);
                if (!struct_pass && !(check_upper && (abs_sin * abs_c >= upper_atol))) return;
                const int phase = emit_phase(m, g, tbuf);
                ++out.n_emit;
                Queries &q = foll ? out.foll : out.lead;
                q.key.insert(q.key.end(), tbuf, tbuf + nw);
                q.phase.push_back(phase);
                q.src.push_back(i);
                q.v_src.push_back(v_src);
            };
            for (size_t i = lo; i < hi; ++i) {
                const uint64_t *m = row(i);
                if (!anticommutes(m, g, g_odd)) continue;
                ++out.n_anti;
                const bool foll = is_follower(m, pv);
                if (only_rot < 0 && fused) {  // Scan.h:387-401 fused cos sweep
                    std::vector<double> &c = *cp;
                    const double v_src = c[i];
                    c[i] = v_src * cos_build;
                    const double abs_c = std::abs(v_src);
                    if (is_below(abs_c)) continue;
                    emit(slot_pop(m), static_cast<uint32_t>(i), abs_c, v_src, foll);
                    continue;
                }
                int mono_pop = 0;
                if (only_rot >= 0) {  // Scan.h:419-433 orbital gate precedes the cos push
                    mono_pop = slot_pop(m);
                    if (mono_pop > only_rot) continue;
                }
                out.cos.push_back(static_cast<uint32_t>(i));
                double v_src = 0.0, abs_c = 0.0;
                if (contract) {
                    v_src = (*cp)[i];
                    abs_c = use_coeff_checks ? std::abs(v_src) : 0.0;
                } else if (use_coeff_checks) {
                    abs_c = std::abs((*cp)[i]);
                }
                if (only_rot < 0) {  // Scan.h:403-417
                    if (is_below(abs_c)) continue;
                    mono_pop = slot_pop(m);
                }
                emit(mono_pop, static_cast<uint32_t>(i), abs_c, v_src, foll);
            }
        };
        auto append_queries = [nw = nw](Queries &dst, Queries &src) {
            dst.key.insert(dst.key.end(), src.key.begin(), src.key.end());
            dst.phase.insert(dst.phase.end(), src.phase.begin(), src.phase.end());
            dst.src.insert(dst.src.end(), src.src.begin(), src.src.end());
            dst.v_src.insert(dst.v_src.end(), src.v_src.begin(), src.v_src.end());
            (void)nw;
        };
        if (gpop != 0) {  // Scan.h:272-274: an empty fold generator builds nothing
            const size_t nr = ranges_for(n0);
            std::vector<ScanOut> parts(nr);
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads) if (nr > 1)
            for (size_t r = 0; r < nr; ++r) scan_range(n0 * r / nr, n0 * (r + 1) / nr, parts[r]);
            for (ScanOut &o : parts) {
                append_queries(leaders, o.lead);
                append_queries(followers, o.foll);
                cos_set.insert(cos_set.end(), o.cos.begin(), o.cos.end());
                n_anti += o.n_anti;
                n_emit += o.n_emit;
            }
        }
        stat_a.push_back(n_anti);
        stat_e.push_back(n_emit);

        // Two resolve passes, leaders then followers (Engine.h:364-387, 465-513).
        struct Rec {
            uint32_t src, tgt;
            double v_src, v_tgt;
            int phase;
        };
        std::vector<Rec> hits;
        std::vector<std::pair<const Queries *, size_t>> deferred;
        if (epoch_.size() < n0) epoch_.resize(n0, 0);
        if (cur_epoch_ == UINT32_MAX) {
            std::fill(epoch_.begin(), epoch_.end(), 0);
            cur_epoch_ = 0;
        }
        ++cur_epoch_;
        // A pass over one side's queries; lookups are read-only and the leader pass marks distinct rows
        // (f = src ^ G is one-to-one), so ranges run in parallel and are joined in query order.
        auto resolve = [&](const Queries &qs, bool leader_pass) {
            const size_t nr = ranges_for(qs.size());
            std::vector<std::vector<Rec>> rh(nr);
            std::vector<std::vector<std::pair<const Queries *, size_t>>> rd(nr);
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads) if (nr > 1)
            for (size_t r = 0; r < nr; ++r) {
                for (size_t q = qs.size() * r / nr; q < qs.size() * (r + 1) / nr; ++q) {
                    if (!leader_pass && epoch_[qs.src[q]] == cur_epoch_) continue;
                    const int64_t f = find(qs.key.data() + q * nw);
                    if (f >= 0 && static_cast<size_t>(f) < n0) {
                        if (leader_pass) epoch_[f] = cur_epoch_;
                        double vt = 0.0;
                        if (contract) vt = fused ? (*cp)[f] * inv_cos : (*cp)[f];
                        rh[r].push_back({qs.src[q], static_cast<uint32_t>(f), qs.v_src[q], vt, qs.phase[q]});
                    } else {
                        rd[r].emplace_back(&qs, q);
                    }
                }
            }
            for (size_t r = 0; r < nr; ++r) {
                hits.insert(hits.end(), rh[r].begin(), rh[r].end());
                deferred.insert(deferred.end(), rd[r].begin(), rd[r].end());
            }
        };
        resolve(leaders, true);
        resolve(followers, false);
        // Deferred misses land at n0 + k in leader-then-follower order (Engine.h:428-444).
        std::vector<Rec> inserts;
        inserts.reserve(deferred.size());
        for (size_t k = 0; k < deferred.size(); ++k) {
            const Queries &qs = *deferred[k].first;
            const size_t q = deferred[k].second;
            // A fresh row starts at 0 in the Heisenberg picture and at its state score in the
            // Schrodinger picture, where that value is also the insert's v_tgt (FusedApply.h:38-44).
            const uint64_t *key = qs.key.data() + q * nw;
            const double fresh = schr ? score(key) : 0.0;
            append_row(key, fresh);
            inserts.push_back({qs.src[q], static_cast<uint32_t>(n0 + k), qs.v_src[q], fresh, qs.phase[q]});
        }
        init_coeffs.resize(size(), 0.0);
        stat_i.push_back(static_cast<int64_t>(deferred.size()));

        if (contract) {
            // FusedApply.h:46-82
            std::vector<double> &c = *cp;
            double sin_val, cos_val;
            sin_cos(2 * apply_alpha, &sin_val, &cos_val);
            if (!fused) {
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1 && cos_set.size() > kParMin)
                for (size_t j = 0; j < cos_set.size(); ++j) c[cos_set[j]] *= cos_val;
                for (size_t k = n0; k < size(); ++k) c[k] *= cos_val;
            }
            // every row is an endpoint of at most one record, so the records apply in parallel
            const size_t n_hit = hits.size();
#pragma omp parallel for num_threads(g_threads) if (g_threads > 1 && n_hit + inserts.size() > kParMin)
            for (size_t k = 0; k < n_hit + inserts.size(); ++k) {
                const bool is_insert = k >= n_hit;
                const Rec &r = is_insert ? inserts[k - n_hit] : hits[k];
                c[r.src] += sin_val * static_cast<double>(-r.phase) * r.v_tgt;
                if (fused && is_insert)
                    c[r.tgt] = cos_val * c[r.tgt] + sin_val * static_cast<double>(r.phase) * r.v_src;
                else
                    c[r.tgt] += sin_val * static_cast<double>(r.phase) * r.v_src;
            }
        } else {
            // the seed grows by the current picture's values of the new rows
            // (MonomialPropagator.inl:523-538 extend_coeffs_from_current_picture_if_needed_)
            if (coeff_view)
                for (size_t k = coeff_view->size(); k < size(); ++k) coeff_view->push_back(coeffs[k]);
            Layer &ly = *out_layer;
            ly.gen.assign(g, g + nw);
            ly.scaled_count = size();
            for (const Rec &r : hits) ly.rots.push_back({r.src, r.tgt, r.phase});
            for (const Rec &r : inserts) ly.rots.push_back({r.src, r.tgt, r.phase});
            if (out_cos) {
                *out_cos = cos_set;
                for (size_t k = n0; k < size(); ++k) out_cos->push_back(static_cast<uint32_t>(k));
            }
        }
    }

    // Cos set of a graph layer, recomputed as CosineRecompute.h:15-22 does.
    void layer_cos(const Layer &ly, std::vector<uint32_t> &out) const {
        out.clear();
        int gpop = 0;
        for (int w = 0; w < nw; ++w) gpop += popc(ly.gen[w]);
        if (gpop == 0) return;
        const bool g_odd = basis == 0 && (gpop % 2 != 0);
        for (size_t i = 0; i < ly.scaled_count; ++i)
            if (anticommutes(row(i), ly.gen.data(), g_odd)) out.push_back(static_cast<uint32_t>(i));
    }

    // evolve_step (Evolution.cpp:471-512): snapshot the pre-cos partner values, cos-scale, add sines.
    void evolve_step(std::vector<double> &op, const Layer &ly, const std::vector<uint32_t> &cos, double alpha) const {
        double sin_val, cos_val;
        sin_cos(2 * alpha, &sin_val, &cos_val);
        std::vector<double> snap_t(ly.rots.size()), snap_s(ly.rots.size());
        for (size_t k = 0; k < ly.rots.size(); ++k) {
            snap_t[k] = op[ly.rots[k].tgt];
            snap_s[k] = op[ly.rots[k].src];
        }
        for (uint32_t i : cos) op[i] *= cos_val;
        for (size_t k = 0; k < ly.rots.size(); ++k)
            op[ly.rots[k].src] += sin_val * static_cast<double>(-ly.rots[k].phase) * snap_t[k];
        for (size_t k = 0; k < ly.rots.size(); ++k)
            op[ly.rots[k].tgt] += sin_val * static_cast<double>(ly.rots[k].phase) * snap_s[k];
    }

    // <O> = core + sum over fully paired rows, ascending, of phase * c (MPFunctions.cpp:135-148,196-205).
    double expectation_of(const std::vector<double> &c) const {
        double r = 0.0;
        for (size_t i = 0; i < c.size(); ++i) {
            const uint64_t *m = row(i);
            if (!paired(m)) continue;
            r += state_phase(m) * c[i];
        }
        return core + r;
    }
};

// Boundary conversion of one reference slot tuple into this layout.
bool slots_to_words(const Engine &e, const int32_t *slots, int64_t n, uint64_t *out) {
    std::fill(out, out + e.nw, 0);
    for (int64_t k = 0; k < n; ++k) {
        const int32_t s = slots[k];
        if (s < 0 || s >= 2 * e.n_modes) return false;
        if (e.basis == 1) {
            const int q = s / 2;
            const int l = q >> 6;
            const uint64_t b = uint64_t{1} << (q & 63);
            // X-slot 2q flips x^z (u), Y-slot 2q+1 is z (v); x = u ^ v.
            if (s % 2 == 0) out[l] ^= b;  // toggles x
            else { out[l] ^= b; out[e.L + l] ^= b; }  // toggles x and z
        } else {
            out[s >> 6] |= uint64_t{1} << (s & 63);
        }
    }
    return true;
}

void words_to_slots(const Engine &e, const uint64_t *m, std::vector<int32_t> &out) {
    out.clear();
    if (e.basis == 1) {
        for (int q = 0; q < e.n_modes; ++q) {
            const int l = q >> 6;
            const uint64_t b = uint64_t{1} << (q & 63);
            const bool x = m[l] & b, z = m[e.L + l] & b;
            const bool u = x ^ z, v = z;
            if (u) out.push_back(2 * q);
            if (v) out.push_back(2 * q + 1);
        }
        return;
    }
    for (int s = 0; s < 2 * e.n_modes; ++s)
        if (m[s >> 6] & (uint64_t{1} << (s & 63))) out.push_back(s);
}

int fail(const std::string &msg, int code = -1) {
    g_last_error = msg;
    return code;
}

}  // namespace

extern "C" {

const char *orc_last_error() { return g_last_error.c_str(); }

// Create: initial operator in reference slot-tuple CSR form, complex coefficients.  Terms are taken in
// lexicographic slot-tuple order, the iteration order of the reference's std::map<VecZ, complex>
// (binder.h:60-100, MonomialPropagator.inl:147-182).
// The Schrodinger picture's initial rows (MonomialPropagator.inl:162-181, MajoranaAlgebra.h:125-160):
// every fully paired monomial of at most max_pairs modes, by pair count, each count's mode sets in the
// order std::prev_permutation walks a selector that starts as max_pairs ones then zeros; selector
// position i is mode n-1-i.  Each row starts at its diagonal phase.
void add_paired_rows(Engine &e, int max_pairs) {
    const int n = e.n_modes;
    std::vector<int> pos;
    std::vector<uint64_t> m(e.nw);
    for (int k = 0; k <= max_pairs; ++k) {
        // the selector's set positions run through the k-subsets of [0, n) in lexicographic order
        pos.resize(k);
        for (int i = 0; i < k; ++i) pos[i] = i;
        for (;;) {
            std::fill(m.begin(), m.end(), 0);
            for (int i : pos) {
                const int j = n - 1 - i;
                if (e.basis == 1) m[e.L + (j >> 6)] |= uint64_t{1} << (j & 63);  // Z_j
                else m[(2 * j) >> 6] |= uint64_t{3} << ((2 * j) & 63);           // slots 2j, 2j+1
            }
            e.append_row(m.data(), e.score(m.data()));
            int i = k - 1;
            while (i >= 0 && pos[i] == n - k + i) --i;
            if (i < 0) break;
            ++pos[i];
            for (int t = i + 1; t < k; ++t) pos[t] = pos[t - 1] + 1;
        }
    }
}

// schr_cutoff < 0: Heisenberg picture; otherwise the Schrodinger picture with that cutoff in slots.
void *create_impl(int basis, int n_modes, unsigned cutoff, int cutoff_type, double lower_atol, double upper_atol,
                  int64_t schr_cutoff, const int64_t *offsets, const int32_t *slots, const double *re,
                  const double *im, int64_t n_terms, const int32_t *state, int64_t n_state) {
    try {
        if (n_modes <= 0 || n_modes > 256) {
            fail("num_modes out of range");
            return nullptr;
        }
        auto *e = new Engine();
        e->basis = basis;
        e->n_modes = n_modes;
        e->L = (n_modes + 63) / 64;
        e->nw = basis == 1 ? 2 * e->L : (2 * n_modes + 63) / 64;
        if (e->nw > kMaxWords) {
            delete e;
            fail("too many words per row");
            return nullptr;
        }
        e->cutoff = cutoff;
        e->cutoff_type = cutoff_type;
        e->has_lower = !std::isnan(lower_atol);
        e->has_upper = !std::isnan(upper_atol);
        e->lower_atol = lower_atol;
        e->upper_atol = upper_atol;
        // lexicographic order of the slot tuples
        std::vector<int64_t> order(n_terms);
        for (int64_t t = 0; t < n_terms; ++t) order[t] = t;
        std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
            return std::lexicographical_compare(slots + offsets[a], slots + offsets[a + 1], slots + offsets[b],
                                                slots + offsets[b + 1]);
        });
        e->occ.assign(e->nw, 0);
        for (int64_t k = 0; k < n_state; ++k) {
            const int j = state[k];
            if (basis == 1) e->occ[j >> 6] |= uint64_t{1} << (j & 63);
            else e->occ[(2 * j) >> 6] |= uint64_t{1} << ((2 * j) & 63);
        }
        e->rehash(std::max<int64_t>(16, n_terms));
        std::vector<uint64_t> m(e->nw);
        for (int64_t t : order) {
            const int64_t n = offsets[t + 1] - offsets[t];
            if (!slots_to_words(*e, slots + offsets[t], n, m.data())) {
                delete e;
                fail("index out of range");
                return nullptr;
            }
            const std::complex<double> c(re[t], im[t]);
            double enc;
            if (basis == 1) {  // PauliAlgebra.h:150-155
                if (std::abs(c.imag()) > 1e-10) {
                    delete e;
                    fail("Non-real Pauli coeffs detected");
                    return nullptr;
                }
                enc = c.real();
            } else {  // MajoranaAlgebra.h:162-171: divide by i^{C(|M|,2)}
                static const std::complex<double> kPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
                const std::complex<double> d = c / kPow[n_choose_2(static_cast<size_t>(n)) % 4];
                if (std::abs(d.imag()) > 1e-10) {
                    delete e;
                    fail("Non-Hermitian coeffs detected");
                    return nullptr;
                }
                enc = d.real();
            }
            if (n == 0) {
                e->core = enc;
                continue;
            }
            if (schr_cutoff >= 0) {
                e->obs_rows.insert(e->obs_rows.end(), m.begin(), m.end());
                e->obs_coef.push_back(enc);
                continue;
            }
            e->append_row(m.data(), enc);
        }
        if (schr_cutoff >= 0) {
            e->schr = true;
            const int64_t sc = std::min<int64_t>(schr_cutoff, 2 * static_cast<int64_t>(n_modes));
            add_paired_rows(*e, static_cast<int>(std::min<int64_t>(sc / 2 + sc % 2, n_modes)));
        }
        e->init_coeffs = e->coeffs;
        return e;
    } catch (const std::exception &ex) {
        fail(ex.what());
        return nullptr;
    }
}

void *orc_create(int basis, int n_modes, unsigned cutoff, int cutoff_type, double lower_atol, double upper_atol,
                 const int64_t *offsets, const int32_t *slots, const double *re, const double *im, int64_t n_terms,
                 const int32_t *state, int64_t n_state) {
    return create_impl(basis, n_modes, cutoff, cutoff_type, lower_atol, upper_atol, -1, offsets, slots, re, im,
                       n_terms, state, n_state);
}

// Schrodinger picture (MonomialPropagator.inl:73-190 with schrodinger_cutoff set): the initial operator
// is the observable, the rows start as the paired state terms of at most ceil(schr_cutoff / 2) modes.
void *orc_create_schrodinger(int basis, int n_modes, unsigned cutoff, int cutoff_type, double lower_atol,
                             double upper_atol, int64_t schr_cutoff, const int64_t *offsets, const int32_t *slots,
                             const double *re, const double *im, int64_t n_terms, const int32_t *state,
                             int64_t n_state) {
    if (schr_cutoff < 0) {
        fail("schrodinger_cutoff must be non-negative");
        return nullptr;
    }
    return create_impl(basis, n_modes, cutoff, cutoff_type, lower_atol, upper_atol, schr_cutoff, offsets, slots, re,
                       im, n_terms, state, n_state);
}

int orc_is_schrodinger(void *h) { return static_cast<Engine *>(h)->schr ? 1 : 0; }

// update_initial_operator (MonomialPropagator.inl:353-377, MPOperator.h:193-235): the core term is
// replaced first; Schrodinger: the observable becomes the given terms; Heisenberg: every term must be
// a row, else "Operator term [..] not found in the operator."; the operator (and the vector a stored
// graph replays from) becomes exactly the given terms at their rows.
int orc_update_initial_operator(void *h, const int64_t *offsets, const int32_t *slots, const double *re,
                                const double *im, int64_t n_terms) {
    auto *e = static_cast<Engine *>(h);
    std::vector<uint64_t> rows;
    std::vector<double> coefs;
    std::vector<uint64_t> m(e->nw);
    for (int64_t t = 0; t < n_terms; ++t) {
        const int64_t n = offsets[t + 1] - offsets[t];
        if (!slots_to_words(*e, slots + offsets[t], n, m.data())) return fail("index out of range");
        std::complex<double> c(re[t], im[t]);
        if (e->basis == 0) {
            static const std::complex<double> kPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
            c = c / kPow[n_choose_2(static_cast<size_t>(n)) % 4];
        }
        if (std::abs(c.imag()) > 1e-10) return fail(e->basis == 1 ? "Non-real Pauli coeffs detected" : "Non-Hermitian coeffs detected");
        if (n == 0) {
            e->core = c.real();
            continue;
        }
        rows.insert(rows.end(), m.begin(), m.end());
        coefs.push_back(c.real());
    }
    if (e->schr) {
        e->obs_rows = rows;
        e->obs_coef = coefs;
        return 0;
    }
    std::vector<double> op(e->size(), 0.0);
    for (size_t j = 0; j < coefs.size(); ++j) {
        const int64_t r = e->find(rows.data() + j * e->nw);
        if (r < 0) {
            std::vector<int32_t> tmp;
            words_to_slots(*e, rows.data() + j * e->nw, tmp);
            std::string repr = "[";
            for (size_t k = 0; k < tmp.size(); ++k) repr += (k ? ", " : "") + std::to_string(tmp[k]);
            return fail("Operator term " + repr + "] not found in the operator.");
        }
        op[static_cast<size_t>(r)] = coefs[j];
    }
    e->coeffs = op;
    e->init_coeffs = op;
    return 0;
}

// update_basis_change (MonomialPropagator.h:204-211): Majorana only, exactly 2n rows, slots below 2n
// (MonomialPropagator.inl:461-495, AlgebraCommon.h:56-65); n_rows < 0 restores the native basis.
int orc_set_basis_change(void *h, const int64_t *offsets, const int32_t *slots, int64_t n_rows) {
    auto *e = static_cast<Engine *>(h);
    if (n_rows < 0) {
        e->bchg.clear();
        return 0;
    }
    if (e->basis == 1) return fail("Pauli basis does not accept a basis_change");
    const int64_t want = 2 * static_cast<int64_t>(e->n_modes);
    if (n_rows != want) return fail("basis_change must have exactly 2*logical_num_modes rows");
    std::vector<uint64_t> tab(static_cast<size_t>(want) * e->nw, 0);
    for (int64_t r = 0; r < n_rows; ++r)
        for (int64_t k = offsets[r]; k < offsets[r + 1]; ++k) {
            const int32_t j = slots[k];
            if (j < 0 || j >= want) return fail("index out of range");
            tab[static_cast<size_t>(r) * e->nw + (j >> 6)] |= uint64_t{1} << (j & 63);
        }
    e->bchg.swap(tab);
    return 0;
}

void orc_destroy(void *h) { delete static_cast<Engine *>(h); }

// Threads of the parallel restatement (1: serial); results are identical for every count.
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

int64_t orc_size(void *h) { return static_cast<int64_t>(static_cast<Engine *>(h)->size()); }
int orc_words(void *h) { return static_cast<Engine *>(h)->nw; }

// Layers in authoring order; processed back to front in the Heisenberg picture, front to back in the
// Schrodinger picture (MonomialPropagator.inl:720-733), alpha = (1.0 * theta[map[i]]) * g[i]
// (MPFunctions.cpp:57-69), applied negated in the Schrodinger picture (MonomialPropagator.h:411-416).
int orc_propagate(void *h, const int64_t *gen_offsets, const int32_t *gen_slots, int64_t n_layers,
                  const int64_t *param_map, const double *gen_coeffs, const double *params, int64_t n_params,
                  int only_rot) {
    auto *e = static_cast<Engine *>(h);
    if (!e->graph.empty()) return fail("Cannot propagate() on top of a non-empty graph");
    e->stat_n.clear();
    e->stat_a.clear();
    e->stat_e.clear();
    e->stat_i.clear();
    std::vector<uint64_t> g(e->nw);
    for (int64_t i = 0; i < n_layers; ++i) {
        const int64_t idx = e->schr ? i : n_layers - 1 - i;
        if (param_map[idx] < 0 || param_map[idx] >= n_params) return fail("parameter index out of range");
        if (!slots_to_words(*e, gen_slots + gen_offsets[idx], gen_offsets[idx + 1] - gen_offsets[idx], g.data()))
            return fail("generator index out of range");
        const double alpha = 1.0 * params[param_map[idx]] * gen_coeffs[idx];
        e->build_layer(g.data(), true, alpha, e->schr ? -alpha : alpha, only_rot, true, nullptr, nullptr, nullptr);
    }
    return 0;
}

// Graph build (MonomialPropagator.inl:537-592); seeded when params != nullptr.  The orbital cap
// (only_rot >= 0) gates emission and the seed's cosine set (Scan.h:419-433); the stored layer's cosine
// set stays "every anticommuting row before scaled_count" (layer_cos, CosineRecompute.h:17-20).
int orc_build_graph(void *h, const int64_t *gen_offsets, const int32_t *gen_slots, int64_t n_layers,
                    const int64_t *param_map, const double *gen_coeffs, const double *params, int64_t n_params,
                    int only_rot) {
    auto *e = static_cast<Engine *>(h);
    std::vector<double> seed;
    if (params) seed = e->coeffs;  // current picture (no stored graph yet in our use)
    std::vector<uint64_t> g(e->nw);
    std::vector<uint32_t> cos;
    // Layers are kept in processing order (the reference keeps Schrodinger layers newest first,
    // MPGraph.h:41-43; the evaluations below walk them accordingly).
    for (int64_t i = 0; i < n_layers; ++i) {
        const int64_t idx = e->schr ? i : n_layers - 1 - i;
        if (!slots_to_words(*e, gen_slots + gen_offsets[idx], gen_offsets[idx + 1] - gen_offsets[idx], g.data()))
            return fail("generator index out of range");
        Layer ly;
        ly.param_index = param_map[idx];
        ly.gen_coeff = gen_coeffs[idx];
        if (params) {
            if (param_map[idx] < 0 || param_map[idx] >= n_params) return fail("parameter index out of range");
            const double alpha = 1.0 * params[param_map[idx]] * gen_coeffs[idx];
            const double apply = e->schr ? -alpha : alpha;
            e->build_layer(g.data(), true, alpha, apply, only_rot, false, &seed, &ly, &cos);
            e->evolve_step(seed, ly, cos, apply);
        } else {
            e->build_layer(g.data(), false, 0.0, 0.0, only_rot, false, nullptr, &ly, &cos);
        }
        e->graph.push_back(std::move(ly));
    }
    e->coeffs.resize(e->size(), 0.0);
    return 0;
}

int64_t orc_graph_layers(void *h) { return static_cast<int64_t>(static_cast<Engine *>(h)->graph.size()); }

// Per-gate counters of the last propagate: live terms before the gate, anticommuting, emitted, inserted.
void orc_gate_stats(void *h, int64_t *n, int64_t *a, int64_t *em, int64_t *ins) {
    auto *e = static_cast<Engine *>(h);
    for (size_t k = 0; k < e->stat_n.size(); ++k) {
        n[k] = e->stat_n[k];
        a[k] = e->stat_a[k];
        em[k] = e->stat_e[k];
        ins[k] = e->stat_i[k];
    }
}
int64_t orc_gate_stats_len(void *h) { return static_cast<int64_t>(static_cast<Engine *>(h)->stat_n.size()); }

// Expectation with no graph: core + dot (inl:1034-1041 -> MPFunctions.cpp:196-205).
double orc_expectation(void *h) {
    auto *e = static_cast<Engine *>(h);
    if (e->schr) return e->expectation_schr(e->coeffs, e->observable_vector());
    return e->expectation_of(e->coeffs);
}

// Functional replay of the stored graph at params (evolve_operator, Evolution.cpp:529-540).
int orc_graph_ev(void *h, const double *params, int64_t n_params, double *out) {
    auto *e = static_cast<Engine *>(h);
    if (e->schr) {
        // the observable evolves through the layers newest first at +alpha and meets the state
        // (MonomialPropagator.inl:918-1000 make_functional_, MPFunctions.cpp:71-82, 196-205)
        std::vector<double> op = e->observable_vector();
        std::vector<uint32_t> cos;
        for (size_t l = e->graph.size(); l-- > 0;) {
            const Layer &ly = e->graph[l];
            if (ly.param_index >= n_params) return fail("parameters too short");
            e->layer_cos(ly, cos);
            e->evolve_step(op, ly, cos, 1.0 * params[ly.param_index] * ly.gen_coeff);
        }
        *out = e->expectation_schr(e->coeffs, op);
        return 0;
    }
    std::vector<double> op = e->init_coeffs;
    op.resize(e->size(), 0.0);
    std::vector<uint32_t> cos;
    for (const Layer &ly : e->graph) {
        if (ly.param_index >= n_params) return fail("parameters too short");
        const double alpha = 1.0 * params[ly.param_index] * ly.gen_coeff;
        e->layer_cos(ly, cos);
        e->evolve_step(op, ly, cos, alpha);
    }
    *out = e->expectation_of(op);
    return 0;
}

// Replayed coefficient vector at params (contract_partially(params, inplace=False)).
// Schrodinger: the state evolves oldest layer first at -alpha (MonomialPropagator.inl:1069-1083).
int orc_graph_contract(void *h, const double *params, int64_t n_params, double *out) {
    auto *e = static_cast<Engine *>(h);
    std::vector<double> op = e->schr ? e->coeffs : e->init_coeffs;
    op.resize(e->size(), 0.0);
    std::vector<uint32_t> cos;
    for (const Layer &ly : e->graph) {
        if (ly.param_index >= n_params) return fail("parameters too short");
        e->layer_cos(ly, cos);
        e->evolve_step(op, ly, cos, (e->schr ? -1.0 : 1.0) * params[ly.param_index] * ly.gen_coeff);
    }
    std::memcpy(out, op.data(), op.size() * sizeof(double));
    return 0;
}

// Reverse-mode gradient of the replay (exact, no 1/cos un-evolution): the forward replay keeps the
// pre-layer values of the entries each layer touches (its cos set and rotation endpoints); the reverse
// sweep applies the transposed layer to the adjoint vector lambda.
// d/dalpha of a layer: cos-only rows -2s c_i; pair (S,T,phi): -2s c_S - 2k phi c_T, -2s c_T + 2k phi c_S.
int orc_graph_grad(void *h, const double *params, int64_t n_params, double *value, double *grad) {
    auto *e = static_cast<Engine *>(h);
    const size_t nl = e->graph.size();
    std::vector<std::vector<uint32_t>> cos_only(nl);
    std::vector<std::vector<double>> snap_cos(nl), snap_rot(nl);
    // replay order: the stored layers, reversed in the Schrodinger picture (the observable evolves)
    auto at = [&](size_t l) -> const Layer & { return e->graph[e->schr ? nl - 1 - l : l]; };
    std::vector<double> op = e->schr ? e->observable_vector() : e->init_coeffs;
    op.resize(e->size(), 0.0);
    std::vector<char> endpoint(e->size(), 0);
    std::vector<uint32_t> cos;
    for (size_t l = 0; l < nl; ++l) {
        const Layer &ly = at(l);
        if (ly.param_index >= n_params) return fail("parameters too short");
        e->layer_cos(ly, cos);
        for (const Rotation &r : ly.rots) endpoint[r.src] = endpoint[r.tgt] = 1;
        for (uint32_t i : cos)
            if (!endpoint[i]) {
                cos_only[l].push_back(i);
                snap_cos[l].push_back(op[i]);
            }
        for (const Rotation &r : ly.rots) {
            snap_rot[l].push_back(op[r.src]);
            snap_rot[l].push_back(op[r.tgt]);
            endpoint[r.src] = endpoint[r.tgt] = 0;
        }
        e->evolve_step(op, ly, cos, 1.0 * params[ly.param_index] * ly.gen_coeff);
    }
    std::vector<double> lam(e->size(), 0.0);
    if (e->schr) {
        *value = e->expectation_schr(e->coeffs, op);
        lam = e->coeffs;
    } else {
        *value = e->expectation_of(op);
        for (size_t i = 0; i < e->size(); ++i)
            if (e->paired(e->row(i))) lam[i] = e->state_phase(e->row(i));
    }
    for (int64_t p = 0; p < n_params; ++p) grad[p] = 0.0;
    for (size_t l = nl; l-- > 0;) {
        const Layer &ly = at(l);
        const double alpha = 1.0 * params[ly.param_index] * ly.gen_coeff;
        double s, k;
        sin_cos(2 * alpha, &s, &k);
        double d = 0.0;
        for (size_t j = 0; j < cos_only[l].size(); ++j) {
            const uint32_t i = cos_only[l][j];
            d += lam[i] * (-2.0 * s * snap_cos[l][j]);
            lam[i] *= k;
        }
        for (size_t j = 0; j < ly.rots.size(); ++j) {
            const Rotation &r = ly.rots[j];
            const double ph = r.phase, cs = snap_rot[l][2 * j], ct = snap_rot[l][2 * j + 1];
            d += lam[r.src] * (-2.0 * s * cs - 2.0 * k * ph * ct);
            d += lam[r.tgt] * (-2.0 * s * ct + 2.0 * k * ph * cs);
            const double ls = lam[r.src], lt = lam[r.tgt];
            lam[r.src] = k * ls + s * ph * lt;
            lam[r.tgt] = k * lt - s * ph * ls;
        }
        std::vector<uint32_t>().swap(cos_only[l]);
        std::vector<double>().swap(snap_cos[l]);
        std::vector<double>().swap(snap_rot[l]);
        grad[ly.param_index] += ly.gen_coeff * d;
    }
    return 0;
}

// Raw rows in store order: nw words per row and the current coefficients.
void orc_rows(void *h, uint64_t *words, double *coeffs) {
    auto *e = static_cast<Engine *>(h);
    std::memcpy(words, e->rows.data(), e->rows.size() * sizeof(uint64_t));
    std::memcpy(coeffs, e->coeffs.data(), e->coeffs.size() * sizeof(double));
}

double orc_core(void *h) { return static_cast<Engine *>(h)->core; }

// evolved_operator (inl:1099-1127, binder.h:200-220): rows with |c| >= atol, decoded and rounded to
// 1e-12, as slot-tuple CSR.  Two-call protocol: with slots == nullptr only the counts are returned.
int orc_export(void *h, double atol, int64_t *n_terms, int64_t *n_slots, int64_t *offsets, int32_t *slots,
               double *re, double *im) {
    auto *e = static_cast<Engine *>(h);
    std::vector<int32_t> tmp;
    int64_t nt = 0, ns = 0;
    for (size_t i = 0; i < e->size(); ++i) {
        const double c = e->coeffs[i];
        if (std::abs(c) < atol) continue;
        words_to_slots(*e, e->row(i), tmp);
        if (slots) {
            offsets[nt] = ns;
            std::copy(tmp.begin(), tmp.end(), slots + ns);
            std::complex<double> dec(c, 0.0);
            if (e->basis == 0) {
                static const std::complex<double> kPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
                dec = dec * kPow[n_choose_2(tmp.size()) % 4];
            }
            re[nt] = std::round(dec.real() * 1e12) / 1e12;
            im[nt] = std::round(dec.imag() * 1e12) / 1e12;
        }
        ++nt;
        ns += static_cast<int64_t>(tmp.size());
    }
    if (slots) offsets[nt] = ns;
    *n_terms = nt;
    *n_slots = ns;
    return 0;
}

}  // extern "C"
