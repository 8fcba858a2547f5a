"""Oracle: CPU restatement of the reference propagate path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
this package, and only as the checker (or the timed CPU baseline).  The product path
(``monoprop_amd``) never imports it.

The native code is ``oracle/mp_oracle.cpp`` (see its header for the reference files it follows and
the checks that pin it); ``oracle/build.py`` compiles it.
"""

from __future__ import annotations

import ctypes as C
import math
from collections.abc import Mapping, Sequence
from pathlib import Path

import numpy as np

from . import build as _build

_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = Path(_build.LIB)
        if not path.exists():
            _build.build()
        lib = C.CDLL(str(path))
        P = C.c_void_p
        i64p = C.POINTER(C.c_int64)
        i32p = C.POINTER(C.c_int32)
        f64p = C.POINTER(C.c_double)
        lib.orc_last_error.restype = C.c_char_p
        lib.orc_create.restype = P
        lib.orc_create.argtypes = [C.c_int, C.c_int, C.c_uint, C.c_int, C.c_double, C.c_double,
                                   i64p, i32p, f64p, f64p, C.c_int64, i32p, C.c_int64]
        lib.orc_create_schrodinger.restype = P
        lib.orc_create_schrodinger.argtypes = [C.c_int, C.c_int, C.c_uint, C.c_int, C.c_double, C.c_double,
                                               C.c_int64, i64p, i32p, f64p, f64p, C.c_int64, i32p, C.c_int64]
        lib.orc_is_schrodinger.restype = C.c_int
        lib.orc_is_schrodinger.argtypes = [P]
        lib.orc_set_threads.restype = None
        lib.orc_set_threads.argtypes = [C.c_int]
        lib.orc_update_initial_operator.restype = C.c_int
        lib.orc_update_initial_operator.argtypes = [P, i64p, i32p, f64p, f64p, C.c_int64]
        lib.orc_set_basis_change.argtypes = [P, i64p, i32p, C.c_int64]
        lib.orc_destroy.argtypes = [P]
        lib.orc_size.restype = C.c_int64
        lib.orc_size.argtypes = [P]
        lib.orc_words.restype = C.c_int
        lib.orc_words.argtypes = [P]
        lib.orc_propagate.restype = C.c_int
        lib.orc_propagate.argtypes = [P, i64p, i32p, C.c_int64, i64p, f64p, f64p, C.c_int64, C.c_int]
        lib.orc_build_graph.restype = C.c_int
        lib.orc_build_graph.argtypes = [P, i64p, i32p, C.c_int64, i64p, f64p, f64p, C.c_int64, C.c_int]
        lib.orc_graph_layers.restype = C.c_int64
        lib.orc_graph_layers.argtypes = [P]
        lib.orc_expectation.restype = C.c_double
        lib.orc_expectation.argtypes = [P]
        lib.orc_graph_ev.restype = C.c_int
        lib.orc_graph_ev.argtypes = [P, f64p, C.c_int64, f64p]
        lib.orc_graph_contract.restype = C.c_int
        lib.orc_graph_contract.argtypes = [P, f64p, C.c_int64, f64p]
        lib.orc_graph_grad.restype = C.c_int
        lib.orc_graph_grad.argtypes = [P, f64p, C.c_int64, f64p, f64p]
        lib.orc_rows.argtypes = [P, C.POINTER(C.c_uint64), f64p]
        lib.orc_core.restype = C.c_double
        lib.orc_core.argtypes = [P]
        lib.orc_export.restype = C.c_int
        lib.orc_export.argtypes = [P, C.c_double, i64p, i64p, i64p, i32p, f64p, f64p]
        lib.orc_gate_stats.argtypes = [P, i64p, i64p, i64p, i64p]
        lib.orc_gate_stats_len.restype = C.c_int64
        lib.orc_gate_stats_len.argtypes = [P]
        _LIB = lib
    return _LIB


def set_threads(n: int) -> None:
    """Threads of the per-gate scan, resolve and apply (1: the serial restatement); the store, the
    coefficients and every counter are the same for any count."""
    _lib().orc_set_threads(int(n))


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _csr(tuples: Sequence[Sequence[int]]):
    offsets = np.zeros(len(tuples) + 1, dtype=np.int64)
    flat: list[int] = []
    for k, t in enumerate(tuples):
        flat.extend(int(i) for i in t)
        offsets[k + 1] = len(flat)
    return offsets, np.asarray(flat if flat else [0], dtype=np.int32)


def _none_nan(x):
    return math.nan if x is None else float(x)


class Oracle:
    """Serial CPU engine with the reference's propagate / graph / expectation semantics.

    ``basis`` is ``"pauli"`` or ``"majorana"``; ``initial_operator`` maps engine slot tuples to complex
    coefficients exactly as the reference's compiled core receives them.  ``schrodinger_cutoff`` (in
    slots, as the compiled core takes it) selects the Schrodinger picture.
    """

    def __init__(
        self,
        initial_operator: Mapping[tuple[int, ...], complex],
        num_modes: int,
        *,
        basis: str,
        cutoff: int,
        cutoff_type: str = "support",
        lower_atol: float | None = None,
        upper_atol: float | None = None,
        initial_state: Sequence[int] = (),
        schrodinger_cutoff: int | None = None,
        basis_change=None,
    ) -> None:
        lib = _lib()
        keys = list(initial_operator.keys())
        offs, slots = _csr(keys)
        vals = [complex(initial_operator[k]) for k in keys]
        re = np.asarray([v.real for v in vals] or [0.0], dtype=np.float64)
        im = np.asarray([v.imag for v in vals] or [0.0], dtype=np.float64)
        st = np.asarray(list(initial_state) or [0], dtype=np.int32)
        self.num_modes = num_modes
        self.basis = basis
        head = (1 if basis == "pauli" else 0, num_modes, cutoff, 0 if cutoff_type == "length" else 1,
                _none_nan(lower_atol), _none_nan(upper_atol))
        tail = (_ptr(offs, C.c_int64), _ptr(slots, C.c_int32), _ptr(re, C.c_double), _ptr(im, C.c_double),
                len(keys), _ptr(st, C.c_int32), len(list(initial_state)))
        if schrodinger_cutoff is None:
            self._h = lib.orc_create(*head, *tail)
        else:
            self._h = lib.orc_create_schrodinger(*head, int(schrodinger_cutoff), *tail)
        if not self._h:
            raise RuntimeError(lib.orc_last_error().decode())
        if basis_change is not None:
            self.set_basis_change(basis_change)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _LIB is not None:
            _LIB.orc_destroy(h)
            self._h = None

    @staticmethod
    def _layers(majoranas, parameter_mapping, gen_coeffs):
        offs, slots = _csr(majoranas)
        pm = np.asarray(list(parameter_mapping) or [0], dtype=np.int64)
        gc = np.asarray(list(gen_coeffs) or [0.0], dtype=np.float64)
        return offs, slots, pm, gc

    def propagate(self, majoranas, parameter_mapping, gen_coeffs, parameters, only_rotate_len_k=None):
        offs, slots, pm, gc = self._layers(majoranas, parameter_mapping, gen_coeffs)
        th = np.asarray(list(parameters) or [0.0], dtype=np.float64)
        rc = _lib().orc_propagate(
            self._h, _ptr(offs, C.c_int64), _ptr(slots, C.c_int32), len(majoranas), _ptr(pm, C.c_int64),
            _ptr(gc, C.c_double), _ptr(th, C.c_double), len(list(parameters)),
            -1 if only_rotate_len_k is None else int(only_rotate_len_k),
        )
        if rc != 0:
            raise RuntimeError(_lib().orc_last_error().decode())

    def build_graph(self, majoranas, parameter_mapping, gen_coeffs, parameters=None, only_rotate_len_k=None):
        offs, slots, pm, gc = self._layers(majoranas, parameter_mapping, gen_coeffs)
        if parameters is None:
            th_ptr, n = None, 0
        else:
            th = np.asarray(list(parameters) or [0.0], dtype=np.float64)
            th_ptr, n = _ptr(th, C.c_double), len(list(parameters))
            self._keep = th
        rc = _lib().orc_build_graph(self._h, _ptr(offs, C.c_int64), _ptr(slots, C.c_int32), len(majoranas),
                                    _ptr(pm, C.c_int64), _ptr(gc, C.c_double), th_ptr, n,
                                    -1 if only_rotate_len_k is None else int(only_rotate_len_k))
        if rc != 0:
            raise RuntimeError(_lib().orc_last_error().decode())

    def expectation(self) -> float:
        return float(_lib().orc_expectation(self._h))

    def graph_ev(self, parameters) -> float:
        th = np.asarray(list(parameters) or [0.0], dtype=np.float64)
        out = C.c_double()
        rc = _lib().orc_graph_ev(self._h, _ptr(th, C.c_double), len(list(parameters)), C.byref(out))
        if rc != 0:
            raise RuntimeError(_lib().orc_last_error().decode())
        return out.value

    def graph_contract(self, parameters) -> np.ndarray:
        th = np.asarray(list(parameters) or [0.0], dtype=np.float64)
        out = np.zeros(max(self.size(), 1), dtype=np.float64)
        rc = _lib().orc_graph_contract(self._h, _ptr(th, C.c_double), len(list(parameters)), _ptr(out, C.c_double))
        if rc != 0:
            raise RuntimeError(_lib().orc_last_error().decode())
        return out[: self.size()]

    def graph_grad(self, parameters):
        th = np.asarray(list(parameters), dtype=np.float64)
        g = np.zeros(max(1, len(th)), dtype=np.float64)
        out = C.c_double()
        rc = _lib().orc_graph_grad(self._h, _ptr(th, C.c_double), len(th), C.byref(out), _ptr(g, C.c_double))
        if rc != 0:
            raise RuntimeError(_lib().orc_last_error().decode())
        return out.value, g[: len(th)]

    def size(self) -> int:
        return int(_lib().orc_size(self._h))

    def graph_layers(self) -> int:
        return int(_lib().orc_graph_layers(self._h))

    def rows(self):
        """``(words[N, nw] uint64, coeffs[N] float64)`` in store (insertion) order."""
        n, nw = self.size(), int(_lib().orc_words(self._h))
        w = np.zeros((max(n, 1), nw), dtype=np.uint64)
        c = np.zeros(max(n, 1), dtype=np.float64)
        _lib().orc_rows(self._h, _ptr(w, C.c_uint64), _ptr(c, C.c_double))
        return w[:n], c[:n]

    def core(self) -> float:
        return float(_lib().orc_core(self._h))

    def gate_stats(self):
        """Per processed gate: (live terms before, anticommuting, emitted, inserted)."""
        n = int(_lib().orc_gate_stats_len(self._h))
        arrs = [np.zeros(max(n, 1), dtype=np.int64) for _ in range(4)]
        _lib().orc_gate_stats(self._h, *[_ptr(a, C.c_int64) for a in arrs])
        return tuple(a[:n] for a in arrs)

    def update_initial_operator(self, terms: Mapping[tuple[int, ...], complex]) -> None:
        keys = list(terms.keys())
        offs, slots = _csr(keys)
        vals = [complex(terms[k]) for k in keys]
        re = np.asarray([v.real for v in vals] or [0.0], dtype=np.float64)
        im = np.asarray([v.imag for v in vals] or [0.0], dtype=np.float64)
        rc = _lib().orc_update_initial_operator(self._h, _ptr(offs, C.c_int64), _ptr(slots, C.c_int32),
                                                _ptr(re, C.c_double), _ptr(im, C.c_double), len(keys))
        if rc != 0:
            raise RuntimeError(_lib().orc_last_error().decode())

    def set_basis_change(self, rows) -> None:
        """Cutoff basis: 2 * num_modes slot lists (row s the image of slot s), or None for native."""
        if rows is None:
            rc = _lib().orc_set_basis_change(self._h, None, None, -1)
        else:
            offs, slots = _csr([tuple(r) for r in rows])
            rc = _lib().orc_set_basis_change(self._h, _ptr(offs, C.c_int64), _ptr(slots, C.c_int32), len(rows))
        if rc != 0:
            raise RuntimeError(_lib().orc_last_error().decode())

    @property
    def schrodinger(self) -> bool:
        return bool(_lib().orc_is_schrodinger(self._h))

    def evolved_operator(self, atol: float = 1e-12) -> dict[tuple[int, ...], complex]:
        """Reference ``evolved_operator`` dict (slot tuples -> rounded complex); the core term is
        included in the Heisenberg picture only (binder.h:212-214)."""
        lib = _lib()
        nt, ns = C.c_int64(), C.c_int64()
        lib.orc_export(self._h, atol, C.byref(nt), C.byref(ns), None, None, None, None)
        offs = np.zeros(nt.value + 1, dtype=np.int64)
        slots = np.zeros(max(ns.value, 1), dtype=np.int32)
        re = np.zeros(max(nt.value, 1), dtype=np.float64)
        im = np.zeros(max(nt.value, 1), dtype=np.float64)
        lib.orc_export(self._h, atol, C.byref(nt), C.byref(ns), _ptr(offs, C.c_int64), _ptr(slots, C.c_int32),
                       _ptr(re, C.c_double), _ptr(im, C.c_double))
        out = {}
        for k in range(nt.value):
            key = tuple(int(s) for s in slots[offs[k]:offs[k + 1]])
            out[key] = complex(re[k], im[k])
        core = self.core()
        if not self.schrodinger and abs(core) >= atol:
            out[()] = complex(core, 0.0)
        return out
