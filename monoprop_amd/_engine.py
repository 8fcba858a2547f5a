"""ctypes binding of ``libmonoprop_amd.so`` and the compiled-core adapter class.

``DeviceSimulator`` mirrors the reference's nanobind class ``MonomialPropagatorNNN``
(src/monoprop/bindings/binder.h:53-274): same constructor keywords, methods, properties and error
types (``ValueError`` for invalid arguments, ``RuntimeError`` for engine failures).  It is the only
place the C ABI (include/monoprop_amd.h) is called from Python.

The library must be built (``python -m monoprop_amd.build``); a missing library raises instead of
falling back to any CPU path.
"""

from __future__ import annotations

import atexit
import ctypes as C
import itertools
import json
import math
import os
import threading
import weakref
from collections.abc import Mapping, Sequence
from pathlib import Path

import numpy as np

from .exceptions import NumberOfModesInvalidError

_LIB_PATH = Path(__file__).resolve().parent / "libmonoprop_amd.so"
_LIB = None
TUNING_FILE = Path(__file__).resolve().parent / "tuning.json"
_LAUNCH: tuple[int, int] | None = None  # (resolve_grid, min_group_log2) every new engine takes
_LAUNCH_READ = False

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_P = C.c_void_p

# name -> (restype, argtypes); every function returns int status unless stated
_SIGNATURES = {
    "mpa_last_error": (C.c_char_p, []),
    "mpa_create": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_double, C.c_double,
                             _i64p, _i32p, _f64p, _f64p, C.c_int64, _i32p, C.c_int64, C.c_int]),
    "mpa_create_schrodinger": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_double,
                                         C.c_double, C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_int64, _i32p,
                                         C.c_int64, C.c_int]),
    "mpa_is_schrodinger": (C.c_int, [_P, _i32p]),
    "mpa_update_initial_operator": (C.c_int, [_P, _i64p, _i32p, _f64p, _f64p, C.c_int64]),
    "mpa_set_parameter_mapping": (C.c_int, [_P, _i64p, C.c_int64]),
    "mpa_clone": (C.c_int, [_P, C.POINTER(_P)]),
    "mpa_memory_breakdown": (C.c_int, [_P, _i64p, C.c_int32]),
    "mpa_destroy": (C.c_int, [_P]),
    "mpa_propagate": (C.c_int, [_P, _i64p, _i32p, C.c_int64, _i64p, _f64p, _f64p, C.c_int64, C.c_int32]),
    "mpa_build_graph": (C.c_int, [_P, _i64p, _i32p, C.c_int64, _i64p, _f64p, _f64p, C.c_int64, C.c_int32]),
    "mpa_expectation_value": (C.c_int, [_P, _f64p, C.c_int64, _f64p]),
    "mpa_expectation_value_noalloc": (C.c_int, [_P, _f64p]),
    "mpa_set_gradient_checkpointing": (C.c_int, [_P, C.c_int, C.c_int64]),
    "mpa_gradient_segments": (C.c_int, [_P, _i64p, _i64p, _i64p]),
    "mpa_expectation_value_and_gradient": (C.c_int, [_P, _f64p, C.c_int64, _f64p, _f64p]),
    "mpa_contract": (C.c_int, [_P, _f64p, C.c_int64, C.c_int, _f64p]),
    "mpa_size": (C.c_int, [_P, _i64p]),
    "mpa_graph_records": (C.c_int, [_P, _i64p, _i64p]),
    "mpa_graph_layers": (C.c_int, [_P, _i64p, _i64p]),
    "mpa_export_terms": (C.c_int, [_P, _f64p, C.c_int64, C.c_double, _i64p, _i64p, _i64p, _i32p, _f64p, _f64p]),
    "mpa_words_per_row": (C.c_int, [_P, _i32p]),
    "mpa_read_rows": (C.c_int, [_P, _u64p, _f64p]),
    "mpa_gate_stats": (C.c_int, [_P, _i64p, _i64p, _i64p, _i64p, _i64p]),
    "mpa_set_pipeline": (C.c_int, [_P, C.c_int]),
    "mpa_set_launch": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "mpa_set_store_limit": (C.c_int, [_P, C.c_int64]),
    "mpa_set_profiling": (C.c_int, [_P, C.c_int]),
    "mpa_kernel_profile": (C.c_int, [_P, C.c_int, C.POINTER(C.c_char_p), _f64p, _f64p, _i64p]),
    "mpa_last_propagate_ms": (C.c_int, [_P, _f64p]),
    "mpa_set_truncation": (C.c_int, [_P, C.c_uint32, C.c_int, C.c_double, C.c_double]),
    "mpa_set_basis_change": (C.c_int, [_P, _i64p, _i32p, C.c_int64]),
    "mpa_batch_create": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_double, C.c_double,
                                   C.c_int32, _i64p, _i64p, _i32p, _f64p, _f64p, _i32p, C.c_int64, C.c_int]),
    "mpa_batch_propagate": (C.c_int, [_P, _i64p, _i32p, C.c_int64, _i64p, _f64p, _f64p, C.c_int64, C.c_int64,
                                      _i64p]),
    "mpa_batch_create_rows": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_double,
                                        C.c_double, C.c_int32, _f64p, C.c_int64, _u64p, _f64p, _i32p, _i32p,
                                        C.c_int64, C.c_int]),
    "mpa_batch_split": (C.c_int, [_P, C.POINTER(C.c_uint8), C.POINTER(_P)]),
    "mpa_batch_members": (C.c_int, [_P, _i32p]),
    "mpa_batch_expectation_values": (C.c_int, [_P, _f64p]),
    "mpa_batch_sizes": (C.c_int, [_P, _i64p]),
    "mpa_batch_row_members": (C.c_int, [_P, _i32p]),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)

# Engines alive in this process.  At exit they are destroyed by an atexit hook, which Python runs
# before the C runtime's exit handlers (where the HIP runtime and any profiler attached to it shut
# down); a destructor that runs after the hook, during interpreter teardown, makes no HIP call.
_LIVE: weakref.WeakSet = weakref.WeakSet()
_LIVE_LOCK = threading.Lock()  # engines may be created on several host threads at once
_SHUTDOWN = False


def _destroy_live_engines() -> None:
    global _SHUTDOWN
    with _LIVE_LOCK:
        live = list(_LIVE)
    for obj in live:
        obj._destroy()
    _SHUTDOWN = True


class _EngineHandle:
    """Owner of one engine handle (``self._h``), destroyed once: by ``__del__`` or at process exit."""

    def _own(self, h) -> None:
        self._h = h
        with _LIVE_LOCK:
            _LIVE.add(self)

    def _destroy(self) -> None:
        h = getattr(self, "_h", None)
        self._h = None
        if h and _LIB is not None and not _SHUTDOWN:
            _LIB.mpa_destroy(h)

    def __del__(self):
        self._destroy()


def library_path() -> Path:
    return _LIB_PATH


def load_library():
    """Load the HIP engine; raise ``ImportError`` if it has not been built."""
    with _LIVE_LOCK:
        return _load_library_locked()


def _load_library_locked():
    global _LIB
    if _LIB is None:
        if not _LIB_PATH.exists():
            raise ImportError(
                f"monoprop_amd engine library not found at {_LIB_PATH}; build it with `python -m monoprop_amd.build`."
            )
        lib = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
        atexit.register(_destroy_live_engines)
    return _LIB


def default_launch() -> tuple[int, int] | None:
    """Launch parameters new engines take: those of set_default_launch(), else the autotuner's choice in
    tuning.json (monoprop_amd/tune.py), else None (the engine's built-in defaults)."""
    global _LAUNCH, _LAUNCH_READ
    if _LAUNCH is None and not _LAUNCH_READ:
        _LAUNCH_READ = True
        try:
            best = json.loads(TUNING_FILE.read_text())["best"]
            _LAUNCH = (int(best["resolve_grid"]), int(best["min_group_log2"]))
        except (OSError, KeyError, TypeError, ValueError):
            _LAUNCH = None
    return _LAUNCH


def set_default_launch(resolve_grid: int | None, min_group_log2: int | None = None) -> None:
    """Launch parameters for engines created from now on (None: the engine's built-in defaults)."""
    global _LAUNCH, _LAUNCH_READ
    _LAUNCH_READ = True
    _LAUNCH = None if resolve_grid is None else (int(resolve_grid), int(min_group_log2))


def _apply_launch(h) -> None:
    launch = default_launch()
    if launch is not None:
        _check(load_library().mpa_set_launch(h, *launch))


def _check(rc: int) -> None:
    if rc == 0:
        return
    msg = load_library().mpa_last_error().decode()
    if rc == -1:
        raise ValueError(msg)
    raise RuntimeError(msg)


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _csr(tuples: Sequence[Sequence[int]]):
    """Offsets (len + 1, int64) and concatenated indices (int32) of a list of index tuples.  Built with
    numpy iterators: a propagate call converts its whole gate list (thousands of monomials) while
    holding the interpreter lock, which propagations running side by side on other threads wait for."""
    n = len(tuples)
    offsets = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.fromiter(map(len, tuples), dtype=np.int64, count=n), out=offsets[1:])
    total = int(offsets[-1])
    if total == 0:
        return offsets, np.zeros(1, dtype=np.int32)
    flat = np.fromiter(itertools.chain.from_iterable(tuples), dtype=np.int64, count=total)
    if flat.min() < 0:
        raise ValueError(f"Majorana/Pauli index {int(flat[flat < 0][0])} is out of range; must be non-negative.")
    if flat.max() > np.iinfo(np.int32).max:
        raise ValueError(f"Majorana/Pauli index {int(flat.max())} is out of range.")
    return offsets, flat.astype(np.int32)


def _f64(xs) -> np.ndarray:
    a = np.asarray(list(xs), dtype=np.float64)
    return a if a.size else np.zeros(1, dtype=np.float64)


def _nan_if_none(x) -> float:
    return math.nan if x is None else float(x)


def default_device() -> int:
    return int(os.environ.get("MONOPROP_AMD_DEVICE", "0"))


class DeviceSimulator(_EngineHandle):
    """HIP-engine counterpart of the reference's compiled ``MonomialPropagatorNNN``."""

    storage_num_modes = 256

    def __init__(
        self,
        initial_operator: Mapping[Sequence[int], complex],
        cutoff: int,
        initial_state: Sequence[int],
        comm=None,
        schrodinger_cutoff: int | None = None,
        lower_atol: float | None = None,
        upper_atol: float | None = None,
        cutoff_type: str = "length",
        basis_change=None,
        logical_num_modes: int | None = None,
        basis: str = "majorana",
        partitions: int = 0,
    ) -> None:
        lib = load_library()
        if comm is not None:
            raise ValueError("MPI communicators are not supported: the engine holds the whole operator on one GPU.")
        if schrodinger_cutoff is not None and (isinstance(schrodinger_cutoff, bool) or int(schrodinger_cutoff) < 0):
            raise ValueError(f"schrodinger_cutoff must be a non-negative integer; got {schrodinger_cutoff!r}.")
        if basis not in ("majorana", "pauli"):
            raise ValueError(f"Unknown basis {basis!r}; expected 'majorana' or 'pauli'.")
        if cutoff_type not in ("length", "support"):
            raise ValueError(f"Unknown cutoff type {cutoff_type!r}; expected 'length' or 'support'.")
        if isinstance(cutoff, bool) or int(cutoff) < 0:
            raise ValueError(f"cutoff must be a non-negative integer; got {cutoff!r}.")
        n = int(logical_num_modes) if logical_num_modes is not None else self.storage_num_modes
        keys = [tuple(int(i) for i in k) for k in initial_operator]
        vals = [complex(initial_operator[k]) for k in initial_operator]
        offs, slots = _csr(keys)
        re = _f64(v.real for v in vals)
        im = _f64(v.imag for v in vals)
        st = np.asarray([int(i) for i in initial_state] or [0], dtype=np.int32)
        self._basis = basis
        self._num_modes = n
        self._cutoff = int(cutoff)
        self._cutoff_type = cutoff_type
        self._lower = lower_atol
        self._upper = upper_atol
        h = _P()
        head = (C.byref(h), 1 if basis == "pauli" else 0, n, int(cutoff), 0 if cutoff_type == "length" else 1,
                _nan_if_none(lower_atol), _nan_if_none(upper_atol))
        tail = (_ptr(offs, C.c_int64), _ptr(slots, C.c_int32), _ptr(re, C.c_double), _ptr(im, C.c_double), len(keys),
                _ptr(st, C.c_int32), len(list(initial_state)), default_device())
        if schrodinger_cutoff is None:
            _check(lib.mpa_create(*head, *tail))
        else:
            _check(lib.mpa_create_schrodinger(*head, int(schrodinger_cutoff), *tail))
        self._schrodinger = schrodinger_cutoff is not None
        self._own(h)
        self._basis_change = None
        if basis_change is not None:
            self.basis_change = basis_change  # validated by the engine; __del__ frees it on error
        _apply_launch(h)
        self._parameter_mapping: list[int] = []  # per graph layer, optimizer order
        self._layer_gates: list[int] = []  # the gate of each graph layer, optimizer order
        self._n_gates = 0
        self._operator_epoch = 0  # update_initial_operator() calls; functionals created before go stale

    def __deepcopy__(self, memo):
        """An independent engine with the same state (the reference's __deepcopy__, binder.h:112-116)."""
        import copy

        h = _P()
        _check(load_library().mpa_clone(self._h, C.byref(h)))
        other = DeviceSimulator.__new__(DeviceSimulator)
        for k, v in self.__dict__.items():
            if k != "_h":
                setattr(other, k, copy.deepcopy(v, memo))
        other._own(h)
        return other

    def __copy__(self):
        return self.__deepcopy__({})

    # --- memory probes (binder.h:247-273) -----------------------------------------------------------
    _MEM_KEYS = ("operator_terms_bytes", "op_coeffs_bytes", "state_coeffs_bytes", "indexing_bytes",
                 "init_operator_bytes", "initial_state_bytes", "inverted_index_bytes", "total_bytes",
                 "d_terms_slack_bytes", "graph_bytes", "d_work_bytes")

    def operator_memory_breakdown(self) -> dict[str, int]:
        """Device bytes held by the operator, by part (total_bytes sums the first seven)."""
        out = np.zeros(len(self._MEM_KEYS), dtype=np.int64)
        _check(load_library().mpa_memory_breakdown(self._h, _ptr(out, C.c_int64), len(out)))
        return {k: int(v) for k, v in zip(self._MEM_KEYS, out) if k != "graph_bytes"}

    def operator_memory_bytes(self) -> int:
        return self.operator_memory_breakdown()["total_bytes"]

    def graph_memory_bytes(self) -> int:
        out = np.zeros(len(self._MEM_KEYS), dtype=np.int64)
        _check(load_library().mpa_memory_breakdown(self._h, _ptr(out, C.c_int64), len(out)))
        return int(out[self._MEM_KEYS.index("graph_bytes")])

    # --- evolution -------------------------------------------------------------------------------
    @staticmethod
    def _layers(majoranas, parameter_mapping, gen_coeffs):
        if len(parameter_mapping) != len(majoranas) or len(gen_coeffs) != len(majoranas):
            raise ValueError(
                f"parameter_mapping ({len(parameter_mapping)}) and gen_coeffs ({len(gen_coeffs)}) must have one "
                f"entry per monomial ({len(majoranas)})."
            )
        offs, slots = _csr(majoranas)
        pm = (np.fromiter(parameter_mapping, dtype=np.int64, count=len(parameter_mapping))
              if len(parameter_mapping) else np.zeros(1, dtype=np.int64))
        gc = _f64(gen_coeffs)
        return offs, slots, pm, gc

    def propagate(self, majoranas, parameter_mapping, gen_coeffs, parameters, only_rotate_len_k=None) -> None:
        if not len(majoranas):
            return
        if self.graph_layers() > 0:
            raise RuntimeError(
                f"Cannot propagate() on top of a non-empty graph of {self.graph_layers()} layer(s): propagate() "
                "evolves and contracts in place and assumes no stored graph."
            )
        offs, slots, pm, gc = self._layers(majoranas, parameter_mapping, gen_coeffs)
        params = list(parameters)  # a length other than max(mapping) + 1 is the engine's RuntimeError
        th = _f64(params)
        k = -1 if only_rotate_len_k is None else int(only_rotate_len_k)
        _check(load_library().mpa_propagate(self._h, _ptr(offs, C.c_int64), _ptr(slots, C.c_int32), len(majoranas),
                                            _ptr(pm, C.c_int64), _ptr(gc, C.c_double), _ptr(th, C.c_double),
                                            len(params), k))

    def build_graph(self, majoranas, parameter_mapping, gen_coeffs, gate_indices=None, parameters=None,
                    only_rotate_len_k=None) -> None:
        if not len(majoranas):
            return
        offs, slots, pm, gc = self._layers(majoranas, parameter_mapping, gen_coeffs)
        if parameters is None:
            th_ptr, n = None, 0
        else:
            th = _f64(parameters)
            th_ptr, n = _ptr(th, C.c_double), len(list(parameters))
        k = -1 if only_rotate_len_k is None else int(only_rotate_len_k)
        _check(load_library().mpa_build_graph(self._h, _ptr(offs, C.c_int64), _ptr(slots, C.c_int32), len(majoranas),
                                              _ptr(pm, C.c_int64), _ptr(gc, C.c_double), th_ptr, n, k))
        gates = list(gate_indices) if gate_indices is not None else list(range(len(majoranas)))
        offset = self._n_gates
        self._n_gates = offset + (max(gates) + 1 if gates else 0)
        # parameter_mapping in optimizer order, the reverse of the reference's layer order
        # (MonomialPropagator.inl:855-870): a Heisenberg call's layers go after the stored ones, a
        # Schrodinger call's before them (MPGraph.h:41-43)
        new = [int(p) for p in parameter_mapping]
        new_gates = [offset + int(g) for g in gates]
        if self._schrodinger:
            self._parameter_mapping = list(self._parameter_mapping) + new
            self._layer_gates = list(self._layer_gates) + new_gates
        else:
            self._parameter_mapping = new + list(self._parameter_mapping)
            self._layer_gates = new_gates + list(self._layer_gates)

    # --- evaluation ------------------------------------------------------------------------------
    def expectation_value(self, parameters) -> float:
        th = _f64(parameters)
        out = C.c_double()
        _check(load_library().mpa_expectation_value(self._h, _ptr(th, C.c_double), len(list(parameters)),
                                                    C.byref(out)))
        return out.value

    def set_gradient_checkpointing(self, keep: bool, budget_bytes: int = 24 << 30) -> None:
        """Memory of the exact adjoint (checkpoints + one segment's pre-layer values, at most
        ``budget_bytes`` when possible); with ``keep`` every expectation_value runs the checkpointed
        forward so that a gradient at the same parameters starts from it."""
        _check(load_library().mpa_set_gradient_checkpointing(self._h, 1 if keep else 0, int(budget_bytes)))

    def gradient_segments(self) -> dict[str, int]:
        """Segments of the last gradient and the device bytes of its checkpoints and snapshots."""
        seg, ck, sn = C.c_int64(), C.c_int64(), C.c_int64()
        _check(load_library().mpa_gradient_segments(self._h, C.byref(seg), C.byref(ck), C.byref(sn)))
        return {"segments": seg.value, "checkpoint_bytes": ck.value, "snapshot_bytes": sn.value}

    def expectation_value_and_gradient(self, parameters):
        params = list(parameters)
        th = _f64(params)
        g = np.zeros(max(len(params), 1), dtype=np.float64)
        out = C.c_double()
        _check(load_library().mpa_expectation_value_and_gradient(self._h, _ptr(th, C.c_double), len(params),
                                                                 C.byref(out), _ptr(g, C.c_double)))
        return out.value, list(g[: len(params)])

    def _functional(self, fn, pare_threshold):
        """A callable bound to the operator and graph of now: it raises once either has changed
        (Validation.cpp:102-117), as the reference's functionals do."""
        if pare_threshold is not None:
            raise ValueError("graph paring is not supported by the HIP engine.")
        epoch, layers, mapping = self._operator_epoch, self.graph_layers(), list(self._parameter_mapping)

        def call(params):
            if self._operator_epoch != epoch:
                raise RuntimeError("MP object has been modified since the functional was created. The initial "
                                   "operator was re-weighted, so the coefficients the functional snapshotted are "
                                   "stale; create a new functional.")
            if self.graph_layers() != layers:
                raise RuntimeError("MP object has been modified since the functional was created. Previous number "
                                   f"of graph layers was {layers} and now is {self.graph_layers()}.")
            if self._parameter_mapping == mapping:
                return fn(params)
            # re-wired since: this functional keeps the mapping it was built with
            now = list(self._parameter_mapping)
            self._set_engine_mapping(mapping)
            try:
                return fn(params)
            finally:
                self._set_engine_mapping(now)

        return call

    def expectation_value_functional(self, pare_threshold=None):
        return self._functional(self.expectation_value, pare_threshold)

    def expectation_value_and_gradient_functional(self, pare_threshold=None):
        return self._functional(self.expectation_value_and_gradient, pare_threshold)

    def update_initial_operator(self, op_dict) -> None:
        """Re-weight the initial operator (Heisenberg: existing terms only; Schrodinger: the observable)."""
        keys = [tuple(int(i) for i in k) for k in op_dict]
        vals = [complex(op_dict[k]) for k in op_dict]
        offs, slots = _csr(keys)
        re = _f64(v.real for v in vals)
        im = _f64(v.imag for v in vals)
        _check(load_library().mpa_update_initial_operator(self._h, _ptr(offs, C.c_int64), _ptr(slots, C.c_int32),
                                                          _ptr(re, C.c_double), _ptr(im, C.c_double), len(keys)))
        self._operator_epoch += 1

    def evolved_operator(self, parameters, atol: float) -> dict:
        lib = load_library()
        params = list(parameters)
        th = _f64(params)
        nt, ns = C.c_int64(), C.c_int64()
        _check(lib.mpa_export_terms(self._h, _ptr(th, C.c_double), len(params), float(atol), C.byref(nt),
                                    C.byref(ns), None, None, None, None))
        offs = np.zeros(nt.value + 1, dtype=np.int64)
        slots = np.zeros(max(ns.value, 1), dtype=np.int32)
        re = np.zeros(max(nt.value, 1), dtype=np.float64)
        im = np.zeros(max(nt.value, 1), dtype=np.float64)
        _check(lib.mpa_export_terms(self._h, _ptr(th, C.c_double), len(params), float(atol), C.byref(nt),
                                    C.byref(ns), _ptr(offs, C.c_int64), _ptr(slots, C.c_int32),
                                    _ptr(re, C.c_double), _ptr(im, C.c_double)))
        out = {}
        for k in range(nt.value):
            out[tuple(int(s) for s in slots[offs[k]:offs[k + 1]])] = complex(re[k], im[k])
        return out

    # --- introspection ---------------------------------------------------------------------------
    def size(self) -> int:
        v = C.c_int64()
        _check(load_library().mpa_size(self._h, C.byref(v)))
        return v.value

    def graph_layers(self) -> int:
        layers, n = C.c_int64(), C.c_int64()
        _check(load_library().mpa_graph_layers(self._h, C.byref(layers), C.byref(n)))
        return layers.value

    def n_gates(self) -> int:
        return self._n_gates if self.graph_layers() > 0 else 0

    def graph_size(self) -> tuple[int, int]:
        """``(cos-only rows, rotations)`` of the graph's layers (the rotations are stored, the cos-only rows
        recomputed at replay)."""
        c, r = C.c_int64(), C.c_int64()
        _check(load_library().mpa_graph_records(self._h, C.byref(c), C.byref(r)))
        return c.value, r.value

    def contract_partially(self, parameters, inplace: bool) -> list[float]:
        params = list(parameters)
        th = _f64(params)
        out = np.zeros(max(self.size(), 1), dtype=np.float64)
        _check(load_library().mpa_contract(self._h, _ptr(th, C.c_double), len(params), 1 if inplace else 0,
                                           _ptr(out, C.c_double)))
        if inplace:
            self._parameter_mapping = []
            self._layer_gates = []
            self._n_gates = 0
        return list(out[: self.size()])

    def rows(self):
        """Raw store: ``(words[N, nw] uint64, coeffs[N] float64)`` in insertion order."""
        lib = load_library()
        nw = C.c_int32()
        _check(lib.mpa_words_per_row(self._h, C.byref(nw)))
        n = self.size()
        w = np.zeros((max(n, 1), nw.value), dtype=np.uint64)
        c = np.zeros(max(n, 1), dtype=np.float64)
        _check(lib.mpa_read_rows(self._h, _ptr(w, C.c_uint64), _ptr(c, C.c_double)))
        return w[:n], c[:n]

    def gate_stats(self):
        """Per processed gate: (live terms before, anticommuting, emitted, inserted)."""
        lib = load_library()
        ln = C.c_int64(0)
        _check(lib.mpa_gate_stats(self._h, C.byref(ln), None, None, None, None))
        arrs = [np.zeros(max(ln.value, 1), dtype=np.int64) for _ in range(4)]
        _check(lib.mpa_gate_stats(self._h, C.byref(ln), *[_ptr(a, C.c_int64) for a in arrs]))
        return tuple(a[: ln.value] for a in arrs)

    def set_pipeline(self, fused: bool) -> None:
        """Select the fused single-sweep pipeline (default) or the per-gate pipeline for propagate()."""
        _check(load_library().mpa_set_pipeline(self._h, 1 if fused else 0))

    def set_launch(self, resolve_grid: int, min_group_log2: int) -> None:
        """Launch parameters of the fused pipeline (results do not depend on them)."""
        _check(load_library().mpa_set_launch(self._h, int(resolve_grid), int(min_group_log2)))

    def set_store_limit(self, rows: int) -> None:
        """Largest term store in rows (0: device memory decides).  Past the 2N worst case of a gate the
        store is bounded gate by gate by the gate's anticommuting terms (results do not depend on it); a
        gate that does not fit raises RuntimeError."""
        _check(load_library().mpa_set_store_limit(self._h, int(rows)))

    def set_profiling(self, on: bool) -> None:
        """Bracket every kernel launch with HIP events (resets the accumulators)."""
        _check(load_library().mpa_set_profiling(self._h, 1 if on else 0))

    def kernel_profile(self) -> dict[str, dict[str, float]]:
        """Per kernel: summed device ms, algorithmic bytes and launches since profiling was enabled."""
        lib = load_library()
        out = {}
        kind = 0
        while True:
            name = C.c_char_p()
            ms, by, n = C.c_double(), C.c_double(), C.c_int64()
            if lib.mpa_kernel_profile(self._h, kind, C.byref(name), C.byref(ms), C.byref(by), C.byref(n)) != 0:
                break
            out[name.value.decode()] = {"ms": ms.value, "bytes": by.value, "launches": n.value}
            kind += 1
        return out

    def last_propagate_ms(self) -> float:
        v = C.c_double()
        _check(load_library().mpa_last_propagate_ms(self._h, C.byref(v)))
        return v.value

    # --- truncation properties (binder.h:158-189) ------------------------------------------------
    def _push_truncation(self):
        _check(load_library().mpa_set_truncation(self._h, self._cutoff, 0 if self._cutoff_type == "length" else 1,
                                                 _nan_if_none(self._lower), _nan_if_none(self._upper)))

    @property
    def cutoff(self) -> int:
        return self._cutoff

    def _set_truncation_field(self, name: str, value) -> None:
        """Assign one truncation field and push it; on any error the old value stays in place."""
        old = getattr(self, name)
        setattr(self, name, value)
        try:
            self._push_truncation()
        except Exception:
            setattr(self, name, old)
            raise

    @staticmethod
    def _optional_float(v, name: str):
        """None or a real number, as the binding's std::optional<double> conversion accepts."""
        if v is None:
            return None
        if isinstance(v, bool) or not isinstance(v, (int, float, np.integer, np.floating)):
            raise TypeError(f"{name} must be a float or None (not {type(v).__name__}).")
        return float(v)

    @cutoff.setter
    def cutoff(self, v: int) -> None:
        # an unsigned int in the binding: bools, floats and strings are type errors, negatives too
        if isinstance(v, bool) or not isinstance(v, (int, np.integer)):
            raise TypeError(f"cutoff must be a non-negative integer (not {type(v).__name__}).")
        if v < 0 or v >= 2**32:
            raise TypeError(f"cutoff must be a non-negative 32-bit integer; got {v}.")
        self._set_truncation_field("_cutoff", int(v))

    @property
    def lower_atol(self):
        return self._lower

    @lower_atol.setter
    def lower_atol(self, v) -> None:
        # MonomialPropagator.h:168-176
        v = self._optional_float(v, "lower_atol")
        if v is not None and self._upper is not None and v > self._upper:
            raise RuntimeError(f"New lower_atol ({v!r}) must be less than or equal to current upper_atol "
                               f"({self._upper!r}).")
        self._set_truncation_field("_lower", v)

    @property
    def upper_atol(self):
        return self._upper

    @upper_atol.setter
    def upper_atol(self, v) -> None:
        # MonomialPropagator.h:178-186
        v = self._optional_float(v, "upper_atol")
        if v is not None and self._lower is not None and v < self._lower:
            raise RuntimeError(f"New upper_atol ({v!r}) must be greater than or equal to current lower_atol "
                               f"({self._lower!r}).")
        self._set_truncation_field("_upper", v)

    @property
    def cutoff_type(self) -> str:
        return self._cutoff_type

    @cutoff_type.setter
    def cutoff_type(self, v: str) -> None:
        if not isinstance(v, str):
            raise TypeError(f"cutoff_type must be 'length' or 'support' (not {type(v).__name__}).")
        if v not in ("length", "support"):
            raise ValueError(f"Unknown cutoff type {v!r}; expected 'length' or 'support'.")
        old = self._cutoff_type
        self._cutoff_type = v
        try:
            self._push_truncation()
        except Exception:
            self._cutoff_type = old
            raise

    @property
    def basis_change(self):
        """The cutoff basis: 2 * num_modes slot lists (row s the image of Majorana slot s), or None for
        the native basis (binder.h:185-189)."""
        return None if self._basis_change is None else [list(r) for r in self._basis_change]

    @basis_change.setter
    def basis_change(self, rows) -> None:
        lib = load_library()
        if rows is None:
            _check(lib.mpa_set_basis_change(self._h, None, None, -1))
            self._basis_change = None
            return
        table = [tuple(int(i) for i in r) for r in rows]
        offs, slots = _csr(table)
        _check(lib.mpa_set_basis_change(self._h, _ptr(offs, C.c_int64), _ptr(slots, C.c_int32), len(table)))
        self._basis_change = table

    @property
    def schrodinger(self) -> bool:
        return self._schrodinger

    @property
    def basis(self) -> str:
        return self._basis

    @property
    def num_modes(self) -> int:
        return self._num_modes

    @property
    def parameter_mapping(self) -> list[int]:
        return list(self._parameter_mapping)

    @parameter_mapping.setter
    def parameter_mapping(self, mapping) -> None:
        """Re-wire the stored graph (MonomialPropagator.inl:809-852): one entry per layer in optimizer
        order, or one per gate (every layer of a gate takes the gate's entry); per layer wins when the
        two lengths coincide."""
        m = [int(x) for x in mapping]
        if len(m) == len(self._parameter_mapping):
            per_layer = m
        elif len(m) == self.n_gates():
            per_layer = [m[g] for g in self._layer_gates]
        else:
            raise RuntimeError(f"parameter_mapping has {len(m)} entries; expected {len(self._parameter_mapping)} "
                               f"(per graph layer) or {self.n_gates()} (per gate).")
        self._set_engine_mapping(per_layer)
        self._parameter_mapping = per_layer

    def _set_engine_mapping(self, per_layer) -> None:
        # the engine keeps layers in processing order: the optimizer order reversed (Heisenberg) or as is
        stored = list(per_layer) if self._schrodinger else list(per_layer)[::-1]
        a = np.asarray(stored or [0], dtype=np.int64)
        _check(load_library().mpa_set_parameter_mapping(self._h, _ptr(a, C.c_int64), len(stored)))


# Largest logical mode count the front-end accepts: the reference's monoprop_MAX_NUM_MODES
# (CMakeLists.txt:45-50).  The engine itself stores up to storage_num_modes (256) modes per row.
MAX_LOGICAL_NUM_MODES = 250


def dispatch(num_modes: int):
    """Adapter factory for a mode count, with the reference's checks and messages
    (tools/generate-dispatch.py:145-160)."""
    if not isinstance(num_modes, int) or num_modes <= 0:
        raise NumberOfModesInvalidError(f"Number of Fermionic modes {num_modes} invalid. num_modes must be > 0.")
    if num_modes > MAX_LOGICAL_NUM_MODES:
        raise NumberOfModesInvalidError(
            f"Number of Fermionic modes {num_modes} invalid."
            f" num_modes must be <= {MAX_LOGICAL_NUM_MODES}."
            f" Contact monoprop developers if more than {MAX_LOGICAL_NUM_MODES}"
            " Fermionic modes are required."
        )

    def make(**kwargs):
        kwargs.setdefault("logical_num_modes", num_modes)
        return DeviceSimulator(**kwargs)

    return make


class BatchSimulator(_EngineHandle):
    """Many independent propagations of one gate sequence in one device store.

    Member ``q`` is the propagation of ``operators[q]`` at parameter row ``q``; all members share the
    truncation settings, the initial state and the layer structure.  It stands for ``len(operators)``
    reference ``MonomialPropagatorNNN`` objects (binder.h:53-274) constructed with the same keywords and
    each given the same ``propagate`` call at its own parameters.
    """

    max_members = 2048

    def __init__(self, operators: Sequence[Mapping[Sequence[int], complex]], cutoff: int, initial_state: Sequence[int],
                 *, lower_atol: float | None = None, upper_atol: float | None = None, cutoff_type: str = "length",
                 logical_num_modes: int, basis: str = "majorana", _rows=None) -> None:
        lib = load_library()
        if basis not in ("majorana", "pauli"):
            raise ValueError(f"Unknown basis {basis!r}; expected 'majorana' or 'pauli'.")
        if cutoff_type not in ("length", "support"):
            raise ValueError(f"Unknown cutoff type {cutoff_type!r}; expected 'length' or 'support'.")
        if isinstance(cutoff, bool) or int(cutoff) < 0:
            raise ValueError(f"cutoff must be a non-negative integer; got {cutoff!r}.")
        n_members = len(operators) if _rows is None else len(_rows[0])
        if not 1 <= n_members <= self.max_members:
            raise ValueError(f"a batch holds 1 to {self.max_members} operators, got {n_members}.")
        st = np.asarray([int(i) for i in initial_state] or [0], dtype=np.int32)
        self._basis = basis
        self._num_modes = int(logical_num_modes)
        self._members = n_members
        h = _P()
        if _rows is not None:  # (cores[members], words[N, nw], coeffs[N], members[N]) of a stopped batch
            cores, words, coeffs, members = _rows
            cores = np.ascontiguousarray(cores, dtype=np.float64)
            nr = len(coeffs)
            words = np.ascontiguousarray(words, dtype=np.uint64).reshape(-1) if nr else np.zeros(1, np.uint64)
            coeffs = np.ascontiguousarray(coeffs, dtype=np.float64) if nr else np.zeros(1, np.float64)
            members = np.ascontiguousarray(members, dtype=np.int32) if nr else np.zeros(1, np.int32)
            _check(lib.mpa_batch_create_rows(C.byref(h), 1 if basis == "pauli" else 0, self._num_modes, int(cutoff),
                                             0 if cutoff_type == "length" else 1, _nan_if_none(lower_atol),
                                             _nan_if_none(upper_atol), n_members, _ptr(cores, C.c_double), nr,
                                             _ptr(words, C.c_uint64), _ptr(coeffs, C.c_double),
                                             _ptr(members, C.c_int32), _ptr(st, C.c_int32), len(list(initial_state)),
                                             default_device()))
            self._own(h)
            _apply_launch(h)
            return
        keys: list[tuple[int, ...]] = []
        vals: list[complex] = []
        moffs = np.zeros(len(operators) + 1, dtype=np.int64)
        for q, op in enumerate(operators):
            for k in op:
                keys.append(tuple(int(i) for i in k))
                vals.append(complex(op[k]))
            moffs[q + 1] = len(keys)
        offs, slots = _csr(keys)
        re = _f64(v.real for v in vals)
        im = _f64(v.imag for v in vals)
        _check(lib.mpa_batch_create(C.byref(h), 1 if basis == "pauli" else 0, self._num_modes, int(cutoff),
                                    0 if cutoff_type == "length" else 1, _nan_if_none(lower_atol),
                                    _nan_if_none(upper_atol), len(operators), _ptr(moffs, C.c_int64),
                                    _ptr(offs, C.c_int64), _ptr(slots, C.c_int32), _ptr(re, C.c_double),
                                    _ptr(im, C.c_double), _ptr(st, C.c_int32), len(list(initial_state)),
                                    default_device()))
        self._own(h)
        _apply_launch(h)

    @property
    def members(self) -> int:
        return self._members

    def propagate(self, majoranas, parameter_mapping, gen_coeffs, parameters, max_capacity: int = 0) -> int:
        """``parameters``: one row per member (shape ``(members, n_params)``).  Returns the layers applied
        (all of them unless ``max_capacity`` stopped the store from growing: then the last ones)."""
        if not len(majoranas):
            return 0
        offs, slots, pm, gc = DeviceSimulator._layers(majoranas, parameter_mapping, gen_coeffs)
        th = np.ascontiguousarray(np.asarray(parameters, dtype=np.float64))
        if th.ndim != 2 or th.shape[0] != self._members:
            raise ValueError(f"parameters must have one row per member ({self._members}), got shape {th.shape}.")
        need = (max(int(p) for p in parameter_mapping) + 1) if len(parameter_mapping) else 0
        if th.shape[1] < need:
            raise ValueError(f"Expected at least {need} parameter value(s) per member, got {th.shape[1]}.")
        flat = th.reshape(-1) if th.size else np.zeros(1, dtype=np.float64)
        done = C.c_int64(0)
        _check(load_library().mpa_batch_propagate(self._h, _ptr(offs, C.c_int64), _ptr(slots, C.c_int32),
                                                  len(majoranas), _ptr(pm, C.c_int64), _ptr(gc, C.c_double),
                                                  _ptr(flat, C.c_double), th.shape[1], int(max_capacity),
                                                  C.byref(done)))
        return done.value

    def split(self, move) -> BatchSimulator:
        """Move the members with ``move[q]`` true into a new batch (both renumbered in ascending order,
        rows in store order); both continue from the gate this one stopped at."""
        mv = np.ascontiguousarray(np.asarray(move, dtype=bool), dtype=np.uint8)
        if mv.shape != (self._members,):
            raise ValueError(f"move must have one flag per member ({self._members}), got shape {mv.shape}")
        h = _P()
        _check(load_library().mpa_batch_split(self._h, _ptr(mv, C.c_uint8), C.byref(h)))
        other = BatchSimulator.__new__(BatchSimulator)
        other._basis, other._num_modes = self._basis, self._num_modes
        other._members = int(mv.sum())
        other._own(h)
        self._members -= other._members
        return other

    def expectation_values(self) -> np.ndarray:
        out = np.zeros(self._members, dtype=np.float64)
        _check(load_library().mpa_batch_expectation_values(self._h, _ptr(out, C.c_double)))
        return out

    def sizes(self) -> np.ndarray:
        out = np.zeros(self._members, dtype=np.int64)
        _check(load_library().mpa_batch_sizes(self._h, _ptr(out, C.c_int64)))
        return out

    def size(self) -> int:
        v = C.c_int64()
        _check(load_library().mpa_size(self._h, C.byref(v)))
        return v.value

    def rows(self):
        """Raw store ``(words[N, nw], coeffs[N], member[N])`` in insertion order."""
        lib = load_library()
        nw = C.c_int32()
        _check(lib.mpa_words_per_row(self._h, C.byref(nw)))
        n = self.size()
        w = np.zeros((max(n, 1), nw.value), dtype=np.uint64)
        c = np.zeros(max(n, 1), dtype=np.float64)
        m = np.zeros(max(n, 1), dtype=np.int32)
        _check(lib.mpa_read_rows(self._h, _ptr(w, C.c_uint64), _ptr(c, C.c_double)))
        _check(lib.mpa_batch_row_members(self._h, _ptr(m, C.c_int32)))
        return w[:n], c[:n], m[:n]

    def gate_stats(self):
        """Per processed gate, summed over members: (live terms before, anticommuting, emitted, inserted)."""
        lib = load_library()
        ln = C.c_int64(0)
        _check(lib.mpa_gate_stats(self._h, C.byref(ln), None, None, None, None))
        arrs = [np.zeros(max(ln.value, 1), dtype=np.int64) for _ in range(4)]
        _check(lib.mpa_gate_stats(self._h, C.byref(ln), *[_ptr(a, C.c_int64) for a in arrs]))
        return tuple(a[: ln.value] for a in arrs)

    def last_propagate_ms(self) -> float:
        v = C.c_double()
        _check(load_library().mpa_last_propagate_ms(self._h, C.byref(v)))
        return v.value

    def set_profiling(self, on: bool) -> None:
        """Bracket every kernel launch with HIP events (resets the accumulators)."""
        _check(load_library().mpa_set_profiling(self._h, 1 if on else 0))

    def set_launch(self, resolve_grid: int, min_group_log2: int) -> None:
        """Launch parameters of the fused pipeline (results do not depend on them)."""
        _check(load_library().mpa_set_launch(self._h, int(resolve_grid), int(min_group_log2)))

    def kernel_profile(self) -> dict[str, dict[str, float]]:
        """Per kernel: summed device ms, algorithmic bytes and launches since profiling was enabled."""
        return DeviceSimulator.kernel_profile(self)

