"""Check harness: the GPU differentiable-flocking forward/backward pair
wrapped as a torch.autograd.Function over CPU tensors, so losses written in
torch differentiate through the hand-written adjoint kernels."""
import torch


class DflockFunction(torch.autograd.Function):
    """(p0, v0, weights) -> (p_final, v_final) through DiffFlock."""

    @staticmethod
    def forward(ctx, p0, v0, weights, flock, steps):
        p, v = flock.forward(p0.detach().numpy(), v0.detach().numpy(), weights.detach().numpy(), steps)
        ctx.flock = flock
        return torch.from_numpy(p), torch.from_numpy(v)

    @staticmethod
    def backward(ctx, grad_p, grad_v):
        gp0, gv0, gw = ctx.flock.backward(grad_p.contiguous().numpy(),
                                          None if grad_v is None else grad_v.contiguous().numpy())
        return torch.from_numpy(gp0), torch.from_numpy(gv0), torch.from_numpy(gw), None, None


def simulate(flock, p0, v0, weights, steps):
    return DflockFunction.apply(p0, v0, weights, flock, steps)
