/*
 * monoprop_amd -- C ABI of the MI355X operator-propagation engine.
 *
 * One opaque handle per propagator; plain pointers and sizes only.  Every function returns 0 on
 * success and a negative code on failure; mpa_last_error() then names the failure (thread-local).
 * Codes: -1 invalid argument (reference ValueError), -2 runtime failure (reference RuntimeError),
 * -3 HIP/device failure.
 *
 * Encodings at the boundary are the reference's: operators and generators are CSR lists of slot
 * tuples (Pauli: X_q -> 2q, Y_q -> 2q+1, Z_q -> {2q, 2q+1}; Majorana: the index itself), with
 * complex coefficients for operators (re, im arrays) and real structural coefficients for
 * generators.  `lower_atol` / `upper_atol` take NaN for "None".  Layer arrays are in authoring
 * order; the engine applies them back to front (Heisenberg picture) or front to back (Schrodinger
 * picture, an engine made by mpa_create_schrodinger).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef MONOPROP_AMD_H
#define MONOPROP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpa_engine mpa_engine;

enum { MPA_BASIS_MAJORANA = 0, MPA_BASIS_PAULI = 1 };
enum { MPA_CUTOFF_LENGTH = 0, MPA_CUTOFF_SUPPORT = 1 };

/* Last error message of the calling thread ("" if none). */
const char *mpa_last_error(void);

/* Replaces MonomialPropagator<N>::MonomialPropagator (reference
 * cpp/monoprop/detail/monomial_propagator/MonomialPropagator.inl:73-188, bound at
 * src/monoprop/bindings/binder.h:60-100).  Terms are taken in lexicographic slot-tuple order; the
 * empty tuple is the core (identity) term. */
int mpa_create(mpa_engine **out, int basis, int num_modes, uint32_t cutoff, int cutoff_type,
               double lower_atol, double upper_atol, const int64_t *term_offsets,
               const int32_t *term_slots, const double *coeff_re, const double *coeff_im,
               int64_t n_terms, const int32_t *initial_state, int64_t n_state, int device);

/* Replaces the same constructor with schrodinger_cutoff set (MonomialPropagator.inl:86-89, 162-181;
 * binder.h:60-100): the Schrodinger picture.  schrodinger_cutoff is in slots (the reference's
 * PauliPropagator passes 2 x its qubit value, pauli_propagator.py:71-74).  The engine starts from the
 * fully paired state terms of at most ceil(schrodinger_cutoff / 2) modes at their <b|M|b>; the given
 * operator is the observable, matched against the evolved state's terms when <O> is evaluated.  Every
 * other entry point works on such an engine except the mpa_batch_* ones. */
int mpa_create_schrodinger(mpa_engine **out, int basis, int num_modes, uint32_t cutoff,
                           int cutoff_type, double lower_atol, double upper_atol,
                           int64_t schrodinger_cutoff, const int64_t *term_offsets,
                           const int32_t *term_slots, const double *coeff_re,
                           const double *coeff_im, int64_t n_terms, const int32_t *initial_state,
                           int64_t n_state, int device);

/* Replaces MonomialPropagator::update_initial_operator (MonomialPropagator.inl:353-377,
 * MPOperator.h:193-235, binder.h:153-156): the core term (empty tuple) is replaced first; in the
 * Heisenberg picture every other term must already be a row (-2, "Operator term [..] not found in the
 * operator.", nothing else changed) and the operator -- also the vector a stored graph replays from --
 * becomes exactly the given terms; in the Schrodinger picture the observable becomes the given terms. */
int mpa_update_initial_operator(mpa_engine *e, const int64_t *term_offsets, const int32_t *term_slots,
                                const double *coeff_re, const double *coeff_im, int64_t n_terms);

/* Replaces MonomialPropagator::set_parameter_mapping (MonomialPropagator.inl:809-852, binder.h:241-244):
 * the parameter index of every stored graph layer, in the engine's processing order (the Python
 * adapter maps the reference's per-layer or per-gate mapping onto it); the graph then takes
 * max + 1 parameters. */
int mpa_set_parameter_mapping(mpa_engine *e, const int64_t *param_index, int64_t n_layers);

/* Replaces MonomialPropagator::schrodinger (binder.h:191-193): 1 for the Schrodinger picture. */
int mpa_is_schrodinger(mpa_engine *e, int32_t *out);

/* Replaces MonomialPropagator's copy constructor behind __deepcopy__ (MonomialPropagator.inl:190-212,
 * binder.h:112-116): an independent engine with the same store, coefficients, observable, truncation,
 * initial state and stored graph, on the same device. */
int mpa_clone(mpa_engine *e, mpa_engine **out);

/* Replaces the memory probes operator_memory_bytes / graph_memory_bytes / operator_memory_breakdown
 * (binder.h:247-273): device bytes held, into out[0..10]: operator_terms (rows), op_coeffs,
 * state_coeffs (Schrodinger: the evolved state), indexing (row index), init_operator, initial_state,
 * inverted_index (bit columns), total (sum of the seven), terms slack (capacity beyond the live rows),
 * graph (records and the vector it replays from), work arrays kept between calls.  n >= 11. */
int mpa_memory_breakdown(mpa_engine *e, int64_t *out, int32_t n);

/* Releases all device memory of the engine. */
int mpa_destroy(mpa_engine *e);

/* Replaces MonomialPropagator::propagate (MonomialPropagator.inl:692-733, binder.h:118-125):
 * contract-immediately Heisenberg evolution, alpha_i = params[param_map[i]] * gen_coeffs[i].
 * only_rotate_len_k < 0 means None. */
int mpa_propagate(mpa_engine *e, const int64_t *gen_offsets, const int32_t *gen_slots,
                  int64_t n_layers, const int64_t *param_map, const double *gen_coeffs,
                  const double *params, int64_t n_params, int32_t only_rotate_len_k);

/* Replaces MonomialPropagator::build_graph (MonomialPropagator.inl:620-690, binder.h:102-110).
 * params == NULL builds structurally; otherwise the build is coefficient-seeded at params.
 * only_rotate_len_k < 0: no orbital cap; else only terms of at most k slots rotate (their layers'
 * cos sets and emitters), as in mpa_propagate. */
int mpa_build_graph(mpa_engine *e, const int64_t *gen_offsets, const int32_t *gen_slots,
                    int64_t n_layers, const int64_t *param_map, const double *gen_coeffs,
                    const double *params, int64_t n_params, int32_t only_rotate_len_k);

/* Replaces MonomialPropagator::expectation_value (MonomialPropagator.inl:1034-1041,
 * binder.h:127-130): replays the stored graph at params (n_params must equal the graph's
 * parameter count; 0 with no graph). */
int mpa_expectation_value(mpa_engine *e, const double *params, int64_t n_params, double *out);

/* No counterpart in the reference (it keeps no snapshots: its adjoint un-evolves with 1/cos,
 * Evolution.cpp:355-387).  Memory of the exact adjoint: budget_bytes bounds the coefficient-vector
 * checkpoints plus one segment's pre-layer values (default 24 GiB; the gradient replays each earlier
 * segment once more, and its result does not depend on the budget; 0 asks for one segment per
 * layer, up to 256).  keep != 0 makes
 * mpa_expectation_value run that checkpointed forward, so that mpa_expectation_value_and_gradient
 * at the same parameters (bit for bit) starts from it instead of replaying again. */
int mpa_set_gradient_checkpointing(mpa_engine *e, int keep, int64_t budget_bytes);

/* No counterpart in the reference: the last gradient's segment count and the device bytes its plan
 * needs for the checkpoints and for one segment's snapshots. */
int mpa_gradient_segments(mpa_engine *e, int64_t *segments, int64_t *checkpoint_bytes,
                          int64_t *snapshot_bytes);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation): <O> of a contracted
 * Heisenberg lone propagator (no stored graph), the value mpa_expectation_value(e, NULL, 0, out)
 * gives, to the bit.  It allocates nothing: the ordered sum's counts, terms and result live in one
 * static device array per GPU (about 8 MiB), which holds at most 2^20 nonzero <O> terms (more is an
 * error, -1).  Every call on a GPU shares that GPU's array: calls that overlap on one GPU race. */
int mpa_expectation_value_noalloc(mpa_engine *e, double *out);

/* Replaces MonomialPropagator::expectation_value_and_gradient (MonomialPropagator.inl:1043-1050,
 * MPFunctions.cpp:207-246): value and d<O>/d params[p] for every p. */
int mpa_expectation_value_and_gradient(mpa_engine *e, const double *params, int64_t n_params,
                                       double *value, double *grad);

/* Replaces MonomialPropagator::contract_partially (MonomialPropagator.inl:1052-1097, binder.h:147-151):
 * the evolved coefficients of every stored term at params (the graph replayed, or the contracted
 * operator when no graph is stored), written to out[size] when out != NULL; inplace != 0 folds the
 * graph into the operator and drops it. */
int mpa_contract(mpa_engine *e, const double *params, int64_t n_params, int inplace, double *out);

/* Replaces MonomialPropagator::size (binder.h:233): number of stored terms. */
int mpa_size(mpa_engine *e, int64_t *out);

/* Graph metadata (binder.h:235-240): stored layers and the parameter count they reference; the
 * layers' cos-only rows (recomputed from the column index at replay, not stored) and their stored
 * rotation records (reference graph_size, binder.h:236). */
int mpa_graph_layers(mpa_engine *e, int64_t *layers, int64_t *n_params);
int mpa_graph_records(mpa_engine *e, int64_t *cos_only, int64_t *rotations);

/* Replaces MonomialPropagator::evolved_operator_terms (MonomialPropagator.inl:1099-1127) plus the
 * core-term rule of binder.h:200-220: terms with |c| >= atol, decoded to complex and rounded to
 * 1e-12.  Two-call protocol: call with slots == NULL to get n_terms / n_slots, then again with
 * buffers of those sizes (offsets has n_terms + 1 entries).  params as in mpa_expectation_value. */
int mpa_export_terms(mpa_engine *e, const double *params, int64_t n_params, double atol,
                     int64_t *n_terms, int64_t *n_slots, int64_t *offsets, int32_t *slots,
                     double *re, double *im);

/* Raw store in insertion order: words_per_row u64 words per row (Pauli: x-plane limbs then z-plane
 * limbs, qubit q at bit q; Majorana: slot s at bit s) and the current coefficients.  No counterpart
 * in the reference (its store is private); used by parity tests. */
int mpa_words_per_row(mpa_engine *e, int32_t *out);
int mpa_read_rows(mpa_engine *e, uint64_t *words, double *coeffs);

/* Per-gate counters of the last propagate, in processing order: live terms before the gate,
 * anticommuting terms, emitted partners, inserted terms.  `len` in/out. */
int mpa_gate_stats(mpa_engine *e, int64_t *len, int64_t *n, int64_t *a, int64_t *em, int64_t *ins);

/* Profiler (no counterpart in the reference): with profiling on, every kernel launch of propagate
 * is bracketed by HIP events on the engine's stream.  mpa_kernel_profile reports, per kernel class
 * `kind` (0 <= kind < 10; `name` is the kernel's name), the summed device time in ms, the summed
 * algorithmic bytes and the launch count, accumulated since profiling was switched on.  Algorithmic
 * bytes of the fused kernels per gate: k_sweep 16 w N + 16 A + 16 (L - w) A + (16 L + 8) E, k_resolve
 * (16 L + 8)(E + I) (N live, A anticommuting, E emitted, I inserted terms, L limbs per plane, w the
 * limbs meeting the generator; Majorana: 8 bytes per slot word for 16 per limb pair); of the per-gate
 * pipeline's kernels, the columns, rows, coefficients and hash slots they touch.  Returns -1 past the
 * last kind. */
int mpa_set_profiling(mpa_engine *e, int on);
int mpa_kernel_profile(mpa_engine *e, int kind, const char **name, double *ms, double *bytes,
                       int64_t *launches);

/* Pipeline of mpa_propagate (no counterpart in the reference): fused != 0 (default) runs the
 * single-sweep pipeline whenever every gate takes the reference's fused cos sweep (no orbital cap,
 * cos 2a != 0); 0 forces the per-gate pipeline that also serves build_graph.  Both produce the same
 * store and coefficients. */
int mpa_set_pipeline(mpa_engine *e, int fused);

/* Launch parameters of the fused pipeline (no counterpart in the reference; set by the autotuner,
 * monoprop_amd/tune.py): the number of k_resolve blocks that take emitter chunks (default 128) and the
 * smallest count group, 2^min_group_log2 sweep tiles (default 2, one sweep block; 2..10).  Results do
 * not depend on them. */
int mpa_set_launch(mpa_engine *e, int32_t resolve_grid, int32_t min_group_log2);

/* No counterpart in the reference (its store grows until host memory runs out): the largest term store
 * a lone propagator may hold, in rows (0: as much as device memory takes).  When the store cannot grow
 * to the worst case of the next gate (twice the live terms) within the limit or device memory, each
 * gate is bounded by its own anticommuting terms instead, counted on the device; a gate that does not
 * fit even so raises -2.  Results do not depend on it. */
int mpa_set_store_limit(mpa_engine *e, int64_t rows);

/* Device time of the last propagate call from first to last launch (HIP events), in ms. */
int mpa_last_propagate_ms(mpa_engine *e, double *ms);

/* Replaces the cutoff / cutoff_type / lower_atol / upper_atol setters (binder.h:158-189); applies to
 * gates propagated afterwards. */
int mpa_set_truncation(mpa_engine *e, uint32_t cutoff, int cutoff_type, double lower_atol,
                       double upper_atol);

/* Replaces the basis_change constructor argument and setter (binder.h:71,185-189,
 * MonomialPropagator.h:204-211): a CSR list of n_rows = 2 * num_modes slot lists, row s the image of
 * Majorana slot s; the structural cutoff then tests the XOR of the images of a partner's slots
 * (MonomialPropagatorCommon.h:44-63, MajoranaAlgebra.h:178-189).  n_rows < 0 restores the native
 * basis.  -1 for a Pauli or batch engine or a row count other than 2 * num_modes, -2 for a slot
 * >= 2 * num_modes; nothing changes on error.  Applies to gates propagated afterwards. */
int mpa_set_basis_change(mpa_engine *e, const int64_t *offsets, const int32_t *slots, int64_t n_rows);

/* Batched propagation.  One engine holds `n_members` independent propagations of one layer sequence:
 * member q starts from terms [member_offsets[q], member_offsets[q + 1]) of the CSR term list and is
 * propagated at its own parameter row.  It replaces n_members reference MonomialPropagator objects
 * built with the same truncation and initial state (MonomialPropagator.inl:73-188, binder.h:60-100),
 * each propagated through the same layers (MonomialPropagator.inl:692-733, binder.h:118-125) and
 * read with expectation_value (MPFunctions.cpp:135-148,196-205).  Member q's rows, in store order,
 * are the rows its own propagator would hold, with identical coefficients.  1 <= n_members <= 2048.
 * The entry points above that take a lone propagator (propagate, build_graph, expectation_value*,
 * contract, export_terms) refuse a batch engine with -1; mpa_size, mpa_read_rows, mpa_gate_stats
 * and mpa_last_propagate_ms report the whole store. */
int mpa_batch_create(mpa_engine **out, int basis, int num_modes, uint32_t cutoff, int cutoff_type,
                     double lower_atol, double upper_atol, int32_t n_members,
                     const int64_t *member_offsets, const int64_t *term_offsets,
                     const int32_t *term_slots, const double *coeff_re, const double *coeff_im,
                     const int32_t *initial_state, int64_t n_state, int device);

/* params: n_members rows of n_params values (row-major); alpha of member q at layer i is
 * params[q * n_params + param_map[i]] * gen_coeffs[i].  Runs the fused pipeline for every member
 * (-2 if some member has cos 2a == 0 at a gate: propagate that member alone).
 * max_capacity bounds the store: propagation stops before the first layer that would grow it past
 * max_capacity rows (0: no bound; < 0: automatic, the largest multiple of 16384 rows whose growth,
 * the old and the new store at once, fits in 3/4 of the device memory available now, divided among
 * -max_capacity stores growing at once).  *layers_done receives the layers applied: the last
 * *layers_done layers in authoring order; the first n_layers - *layers_done are still to apply. */
int mpa_batch_propagate(mpa_engine *e, const int64_t *gen_offsets, const int32_t *gen_slots,
                        int64_t n_layers, const int64_t *param_map, const double *gen_coeffs,
                        const double *params, int64_t n_params, int64_t max_capacity,
                        int64_t *layers_done);

/* A batch engine holding given rows: words (words_per_row per row, mpa_read_rows layout),
 * coefficients and each row's member, in store order, plus each member's identity coefficient.
 * With mpa_read_rows / mpa_batch_row_members this moves members of a stopped batch into new
 * engines that continue where it stopped. */
int mpa_batch_create_rows(mpa_engine **out, int basis, int num_modes, uint32_t cutoff,
                          int cutoff_type, double lower_atol, double upper_atol, int32_t n_members,
                          const double *cores, int64_t n_rows, const uint64_t *words,
                          const double *coeffs, const int32_t *members,
                          const int32_t *initial_state, int64_t n_state, int device);

/* Moves the members with move[q] != 0 (one flag per member) into a new batch engine *out; e keeps
 * the others.  Both sides keep their rows in store order and number their members in ascending order
 * of the old numbers; both continue from where e stopped (mpa_batch_propagate with the layers still
 * to apply).  The partition runs on the device; each new store is sized to its rows.  At least one
 * member must move and one stay. */
int mpa_batch_split(mpa_engine *e, const uint8_t *move, mpa_engine **out);

/* Number of members. */
int mpa_batch_members(mpa_engine *e, int32_t *out);

/* <ref| O_q |ref> of every member into out[n_members] (each summed in a fixed order). */
int mpa_batch_expectation_values(mpa_engine *e, double *out);

/* Stored terms of every member into out[n_members]. */
int mpa_batch_sizes(mpa_engine *e, int64_t *out);

/* Member of every stored row (mpa_read_rows order) into out[size]. */
int mpa_batch_row_members(mpa_engine *e, int32_t *out);

#ifdef __cplusplus
}
#endif

#endif /* MONOPROP_AMD_H */
