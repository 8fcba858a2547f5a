"""PyTorch autograd bridge for the graph-mode expectation value.

``expectation(prop, theta)`` returns ``<O>(theta)`` of a propagator whose graph was built with
``prop.build_graph(circuit)`` as a differentiable scalar tensor: the forward pass is the engine's
checkpointed replay of the recorded layers (it keeps the coefficient vector at segment starts and the
last segment's pre-layer values), the backward pass its exact adjoint sweep over the same layers,
which starts from that forward (the parameters are compared bit for bit) and replays each earlier
segment once more for its pre-layer values (no ``1/cos`` recovery), so ``theta.grad`` equals
``prop.expectation_value_and_gradient(theta)[1]`` scaled by the incoming gradient.  Without a
gradient to compute (``torch.no_grad()`` or ``theta`` not requiring one) only the replay runs.  This is the
engine-side counterpart of the reference's ``expval_and_grad`` (``MPFunctions.cpp:207-246``,
``monomial_propagator.py:352-403``).

The engine runs on the GPU whatever device ``theta`` lives on; the result comes back on ``theta``'s
device and dtype (float64 expected).
"""

from __future__ import annotations

import numpy as np
import torch


class _GraphExpectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, prop) -> torch.Tensor:
        params = theta.detach().to("cpu", torch.float64).numpy().copy()
        prop._simulator.set_gradient_checkpointing(True)
        value = prop.expectation_value(params)
        ctx.prop, ctx.params, ctx.device = prop, params, theta.device
        return torch.tensor(value, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        _, grad = ctx.prop.expectation_value_and_gradient(ctx.params)
        g = torch.as_tensor(np.asarray(grad, dtype=np.float64)).to(ctx.device)
        return grad_out.to(g.dtype) * g, None


def expectation(prop, theta: torch.Tensor) -> torch.Tensor:
    """Differentiable ``<O>(theta)`` of ``prop`` (a propagator with a built graph)."""
    if prop.graph_layers == 0:
        raise RuntimeError("build_graph must be called before a differentiable expectation value")
    if theta.dim() != 1 or theta.numel() != prop.n_parameters:
        raise ValueError(f"theta must be a vector of {prop.n_parameters} parameters, got shape {tuple(theta.shape)}")
    if not (torch.is_grad_enabled() and theta.requires_grad):  # no backward can follow: the replay alone
        params = theta.detach().to("cpu", torch.float64).numpy()
        return torch.tensor(prop.expectation_value(params), dtype=theta.dtype, device=theta.device)
    return _GraphExpectation.apply(theta, prop)
