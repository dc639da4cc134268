"""torch.autograd binding over the engine's C ABI (gradient checks only).

QMCPrice.apply(S0, sigma, mu) returns the RQMC price of a one-factor GBM
contract as a 0-d tensor; its backward multiplies the incoming gradient by the
pathwise-adjoint Greeks (delta, vega, rho) that the same fused kernel pass
computed.  The compute runs in libfdmc.so; torch only carries the values.
"""
import torch

import fdmc


class QMCPrice(torch.autograd.Function):
    @staticmethod
    def forward(ctx, S0, sigma, mu, engine, payoff, n_steps, R, n_points, mode, dev_bb):
        model = fdmc.GBMParams(float(S0), float(mu), float(sigma), 1.0)
        rows = engine.greeks_accumulate(model, payoff, n_steps, R, 42, 0, n_points, mode, dev_bb=dev_bb)
        n = R * n_points
        price = sum(r[0] for r in rows) / n
        g = [sum(r[2 + 2 * k] for r in rows) / n for k in range(3)]
        ctx.save_for_backward(torch.tensor(g, dtype=torch.float64))
        return torch.tensor(price, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        (g,) = ctx.saved_tensors
        return grad * g[0], grad * g[1], grad * g[2], None, None, None, None, None, None, None
