"""torch.autograd.Function over the hand-written adjoint of the WCSPH step.

    x_T = rollout(vel0, alpha, sim, dt, n_steps, checkpoint_every)

runs n_steps steps of the scene held by `sim` (a wcsph.FluidSimulator with
the benchmark physics and a fixed dt) from its creation positions with
initial velocities vel0 ((n, 4) float32 CUDA tensor) and artificial
viscosity alpha (0-d tensor), and returns the final positions (n, 4) in
creation order.  The backward pass recomputes each segment of
`checkpoint_every` steps from its checkpoint and runs the adjoint kernels;
it returns gradients for vel0 and alpha.  The library runs on its own HIP
stream, so inputs are synchronised before the call.

A simulator holds the data of its last forward run only: backward() of an
earlier rollout on the same simulator raises instead of returning the
gradients of another run (its generation number no longer matches).  The
backward pass uses that run's own alpha, whatever was set since, and may be
called again (retain_graph).
"""
import torch


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vel0, alpha, sim, dt, n_steps, checkpoint_every):
        v = vel0.detach().to(torch.float32).contiguous()
        sim.set_viscosity(float(alpha))
        torch.cuda.synchronize(v.device)
        out = torch.empty_like(v)
        sim.grad_forward(v, dt, n_steps, checkpoint_every, out=out)
        ctx.sim = sim
        ctx.generation = sim.grad_generation()
        ctx.alpha_meta = (alpha.dtype, alpha.device)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        if ctx.sim.grad_generation() != ctx.generation:
            raise RuntimeError("rollout backward: the simulator has run another forward rollout since this one; "
                               "only its last rollout can be differentiated")
        g = grad_out.detach().to(torch.float32).contiguous()
        torch.cuda.synchronize(g.device)
        gv = torch.empty_like(g)
        _, ga = ctx.sim.grad_backward(g, out=gv)
        dtype, device = ctx.alpha_meta
        return gv, torch.tensor(ga, dtype=dtype, device=device), None, None, None, None


def rollout(vel0, alpha, sim, dt, n_steps, checkpoint_every=20):
    return Rollout.apply(vel0, alpha, sim, dt, n_steps, checkpoint_every)
