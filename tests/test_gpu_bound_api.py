"""The compiled-core object as the reference's own smoke tests drive it (tests/test_monoprop_smoke.py:
24-170): every constructor keyword, the mutators, build_graph / contract_partially, the functionals,
evolved_operator, the size probes and update_initial_operator, by their declared argument names, in
both pictures, on the reference's fixtures."""

from __future__ import annotations

import numpy as np
import pytest

import fixtures
from monoprop_amd import MajoranaPropagator, jordan_wigner_basis_change

pytestmark = pytest.mark.gpu


def _core(p, schrodinger: bool):
    bound_type = type(MajoranaPropagator(p.operator, p.initial_state, cutoff=2 * p.n_modes)._simulator)
    return bound_type(initial_operator=p.operator.terms, cutoff=2 * p.n_modes, initial_state=p.initial_state,
                      comm=None, schrodinger_cutoff=2 * p.n_modes if schrodinger else None, lower_atol=None,
                      upper_atol=None, cutoff_type="length", basis_change=None,
                      logical_num_modes=p.n_modes)


def _evolve(core, p):
    params = [float(x) for x in p.parameters]
    core.build_graph(p.majoranas, [int(i) for i in p.param_inds], [float(g) for g in p.gen_coeffs])
    return params, core.contract_partially(params, inplace=False)


@pytest.mark.parametrize("name", fixtures.EXACT_CASES)
@pytest.mark.parametrize("schrodinger", [False, True], ids=["heisenberg", "schrodinger"])
def test_bound_object_accepts_declared_arguments(name, schrodinger):
    print(f"case: bound api {name} schrodinger={schrodinger}", flush=True)
    p = fixtures.load(name)
    core = _core(p, schrodinger)
    assert core.num_modes == p.n_modes
    assert core.schrodinger is schrodinger

    core.lower_atol = 1e-12
    core.lower_atol = None
    core.upper_atol = 1e-12
    core.upper_atol = None
    core.cutoff = 2 * p.n_modes
    core.cutoff_type = "length"
    core.basis_change = jordan_wigner_basis_change(p.n_modes)
    core.basis_change = None
    assert core.cutoff == 2 * p.n_modes and core.cutoff_type == "length" and core.basis_change is None

    params, coeffs = _evolve(core, p)
    assert len(coeffs) > 0
    ev = core.expectation_value_functional(pare_threshold=None)(params)
    assert isinstance(ev, float)
    ev2, grad = core.expectation_value_and_gradient_functional(pare_threshold=None)(params)
    assert isinstance(ev2, float) and len(grad) == len(params)
    assert abs(ev - p.exact_energy) <= 1e-12 and abs(ev2 - ev) <= 1e-12

    evolved = core.evolved_operator(params, 1e-12)
    assert isinstance(evolved, dict)
    assert core.size() > 0
    assert len(core.graph_size()) == 2
    assert core.graph_layers() is not None
    core.update_initial_operator(op_dict=p.operator.terms)
    assert abs(core.expectation_value_functional(pare_threshold=None)(params) - ev) <= 1e-12
