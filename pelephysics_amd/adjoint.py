"""torch.autograd view of the discrete adjoint, for gradient checks.

``BDFSensitivity.apply(dlnA, y0, rX, reactor, nmacro, dt)`` returns the
per-cell outputs (T_final, Y_a, Y_b) of the fixed-step BDF integration of the
1-D cells y0 (the ``ReactorBDF`` map with exact linear solves), as a function
of the offsets ``dlnA`` of the logarithms of the pre-exponential factors.  The
forward pass runs at the mechanism's own constants, so it accepts only
``dlnA == 0``; the backward pass contracts the upstream gradient with the
adjoint sensitivities d out / d ln A_j computed by ``pp_chem_adjoint_1d``.
"""
from __future__ import annotations

import torch


class BDFSensitivity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dlnA, y0, rX, reactor, nmacro: int, dt: float):
        if torch.count_nonzero(dlnA).item() != 0:
            raise ValueError("the device integrator runs at the tabulated rate constants (dlnA = 0)")
        y = y0.detach().clone().contiguous()
        out, grad = reactor.adjoint_1d(y, rX.contiguous(), nmacro, dt)
        ctx.save_for_backward(grad)
        return out

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        # d L / d lnA_j = sum_cells sum_o gout[c, o] * grad[c, j, o]
        g = (grad * gout.to(grad.dtype)[:, None, :]).sum(dim=(0, 2))
        return g, None, None, None, None, None
