"""torch.autograd wrapper of the engine's unrolled mean field, for gradient
checks only.

``MeanField.apply(U, p_0, ..., p_{K-1}, crf=..., iterations=T)`` runs the
engine's forward (checkpointing every iterate on the device) and, in
backward, its hand-written adjoint.  U is an (N, L) float32 CPU tensor, p_k
the parameters of potential k in the layout of
``PairwisePotential::parameters()``; the tensors are staged through host
memory, so the wrapper needs no GPU support in torch itself.

The engine keeps the checkpoints of a CRF's last forward only: a backward
after another forward on the same CRF raises, and so does one after the
unary or a parameter changed (the C ABI refuses it).
"""

from __future__ import annotations

import numpy as np


def _function():
    import torch

    class MeanField(torch.autograd.Function):
        @staticmethod
        def forward(ctx, U, *params_and_cfg):
            *params, crf, iterations = params_and_cfg
            crf.setUnaryEnergy(U.detach().cpu().numpy().astype(np.float32))
            for k, p in enumerate(params):
                crf.set_pairwise_parameters(k, p.detach().cpu().numpy().astype(np.float32))
            Q = crf.meanfield(int(iterations))
            ctx.crf = crf
            ctx.mf_id = crf.meanfield_id()
            ctx.nparams = len(params)
            return torch.from_numpy(Q)

        @staticmethod
        def backward(ctx, gQ):
            import torch
            if ctx.crf.meanfield_id() != ctx.mf_id:
                # the engine keeps the checkpoints of its last forward only
                raise RuntimeError("MeanField.backward: another meanfield ran on this CRF after this forward; "
                                   "run backward before the next forward or use one CRF per forward")
            dU, dps = ctx.crf.meanfield_backward(np.ascontiguousarray(gQ.detach().cpu().numpy(), np.float32))
            grads = [torch.from_numpy(dU)] + [torch.from_numpy(g.astype(np.float32)) for g in dps[: ctx.nparams]]
            return tuple(grads) + (None, None)

    return MeanField


def meanfield(U, params, crf, iterations):
    """Q_T = MeanField(U, *params) with gradients to U and every parameter tensor."""
    return _function().apply(U, *params, crf, iterations)
