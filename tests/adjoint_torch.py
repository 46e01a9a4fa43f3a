"""torch.autograd.Function over the C-ABI adjoint, for the gradient tests only:
forward runs flb_adjoint_forward (recorded), backward returns
flb_adjoint_backward's dJ/ds scaled by the incoming gradient."""
import numpy as np
import torch


class AdjointObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, s, adj, steps, checkpoint_every):
        adj.set_fraction(s.detach().cpu().numpy().astype(np.float64))
        J = adj.forward(steps, checkpoint_every)
        ctx.adj = adj
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        g = torch.from_numpy(ctx.adj.backward())
        return gJ * g, None, None, None
