"""torch.autograd.Function over the C-ABI adjoint (test harness only)."""
import torch

from bubbledynamics_amd.adjoint import adjoint_gradient


class WesterveltObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, amp, phase, p, n, n_steps, obj, checkpoint_every):
        J, gA, gP, _ = adjoint_gradient(p, n, n_steps, amp.detach().cpu().numpy(),
                                        phase.detach().cpu().numpy(), obj, checkpoint_every, torch.float64)
        ctx.save_for_backward(torch.as_tensor(gA), torch.as_tensor(gP))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gA, gP = ctx.saved_tensors
        return g * gA, g * gP, None, None, None, None, None
