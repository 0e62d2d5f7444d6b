"""Adjoint-method gradients behind torch.autograd.Function.

`Expectation.apply(params, engine)` returns E_b = <psi(params_b)|O|psi(params_b)>
for a batch of parameter sets (rows of `params`), computed on the GPU by
qse_adjoint_expect; its backward returns dL/dparams = dL/dE_b * dE_b/dparams from
the same adjoint sweep (one forward + one reverse sweep per parameter set,
instead of 2 circuit evaluations per parameterised gate for parameter shift).

The circuit is a qse_op list (quantumsimevo_amd.circuits.GateList.ops()) whose
parameterised gates (RX, RY, RZ, PHASE, CPHASE) carry `param` indices into the
parameter vector; the observable is a weighted sum of Z-strings (masks, weights).
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi


class AdjointEngine:
    """A circuit + observable bound to a HIP context."""

    def __init__(self, ctx: capi.Context, nqubits: int, ops, obs_masks, obs_weights,
                 precision: int = capi.C128, ckpt_every: int = 0):
        self.ctx = ctx
        self.n = nqubits
        self.ops = capi.as_ops(ops)
        self.masks = np.ascontiguousarray(obs_masks, dtype=np.uint64)
        self.weights = np.ascontiguousarray(obs_weights, dtype=np.float64)
        self.precision = precision
        self.ckpt_every = ckpt_every
        self.nparams = int(self.ops["param"].max() + 1) if len(self.ops) else 0

    def value_and_grad(self, params: np.ndarray):
        return self.ctx.adjoint_expect(self.n, self.precision, self.ops, params, self.masks, self.weights,
                                       self.ckpt_every)

    def value(self, params: np.ndarray):
        return self.ctx.batch_expect(self.n, self.precision, self.ops, params, self.masks, self.weights)

    def parameter_shift(self, params: np.ndarray) -> np.ndarray:
        """Plain parameter-shift gradient on the same GPU: for every parameterised
        gate occurrence, two circuit evaluations at angle +-pi/2."""
        P = np.atleast_2d(np.asarray(params, dtype=np.float64))
        G = np.zeros_like(P)
        for k in np.nonzero(self.ops["param"] >= 0)[0]:
            o = self.ops[k]
            for sign in (+1.0, -1.0):
                sh = self.ops.copy()
                sh["theta"][k] += sign * np.pi / 2
                E = self.ctx.batch_expect(self.n, self.precision, sh, P, self.masks, self.weights)
                G[:, int(o["param"])] += sign * 0.5 * float(o["scale"]) * E
        return G


class Expectation(torch.autograd.Function):
    """E(params) with adjoint-method backward. params: (batch, nparams) or (nparams,)."""

    @staticmethod
    def forward(ctx, params: torch.Tensor, engine: AdjointEngine):
        squeeze = params.dim() == 1
        P = params.detach().reshape(-1, params.shape[-1]).to("cpu", torch.float64).numpy()
        E, G = engine.value_and_grad(P)
        ctx.save_for_backward(torch.from_numpy(G).to(params.device, params.dtype))
        ctx.squeeze = squeeze
        out = torch.from_numpy(E).to(params.device, params.dtype)
        return out[0] if squeeze else out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (G,) = ctx.saved_tensors
        g = grad_out.reshape(-1, 1) * G
        if ctx.squeeze:
            g = g[0]
        return g, None


def expectation(params: torch.Tensor, engine: AdjointEngine) -> torch.Tensor:
    return Expectation.apply(params, engine)
