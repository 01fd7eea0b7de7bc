"""torch.autograd binding of the engine's adjoint, for fitting scripts.

`final_positions(net, k, l0, q, nsteps)` runs nsteps steps of `net` from its
current state with spring stiffnesses k, rest lengths l0 and charges q (torch
tensors, one per kept spring / per particle; any may be None to keep the
network's values) and returns the final positions as an (n, 3) float64
tensor.  Its backward is bs_adjoint_backward: gradients of any loss of the
final positions with respect to the given k, l0, q.  The engine runs on its
own HIP stream; tensors cross as host arrays.  Nothing else in the package
imports torch.
"""
import numpy as np
import torch


def _refuse_reload(net):
    """Charges change through setParticles, which also clears the handle's
    systems, rigid bodies, trajectory writers and step count: refuse rather
    than lose them."""
    if net.numSystems() > 1 or net.numRigidBodies() > 0:
        raise ValueError("final_positions: q cannot be set on a network with independent systems or rigid bodies "
                         "(setParticles would drop them); pass q=None")


class _FinalPositions(torch.autograd.Function):
    @staticmethod
    def forward(ctx, k, l0, q, net, system, nsteps, checkpoint_every):
        if k is not None or l0 is not None:
            a, b, l0_0, k_0 = system.spring_arrays()
            net.setSprings(a, b, l0_0 if l0 is None else l0.detach().cpu().numpy().astype(np.float32),
                           k_0 if k is None else k.detach().cpu().numpy().astype(np.float32))
        if q is not None:
            _refuse_reload(net)
            arrs = list(system.particle_arrays())
            x, v = net.getState()
            arrs[0], arrs[1], arrs[3] = x, v, q.detach().cpu().numpy().astype(np.float32)
            net.setParticles(*arrs)
            net.setSprings(*system.spring_arrays() if k is None and l0 is None else
                           (system.id1, system.id2,
                            system.l0 if l0 is None else l0.detach().cpu().numpy().astype(np.float32),
                            system.k if k is None else k.detach().cpu().numpy().astype(np.float32)))
        ctx.tape = net.adjointForward(nsteps, checkpoint_every)
        x, _ = net.getState()
        ctx.net = net
        ctx.has = (k is not None, l0 is not None, q is not None)
        return torch.from_numpy(x.astype(np.float64))

    @staticmethod
    def backward(ctx, grad_x):
        if ctx.net.tape != ctx.tape:
            # the handle holds one record: a later forward replaced this one
            raise RuntimeError("final_positions: the network recorded another forward pass since this one; "
                               "call backward before the next forward, or use one network per pass")
        dk, dl0, dq = ctx.net.adjointBackward(grad_x.detach().cpu().numpy().astype(np.float64))
        out = [torch.from_numpy(g) if h else None for g, h in zip((dk, dl0, dq), ctx.has)]
        return out[0], out[1], out[2], None, None, None, None


def final_positions(net, system, k=None, l0=None, q=None, nsteps=1, checkpoint_every=0):
    """Final positions after nsteps steps as a differentiable function of
    k, l0, q (see the module docstring); `system` supplies the values not
    given and the spring endpoints."""
    return _FinalPositions.apply(k, l0, q, net, system, int(nsteps), int(checkpoint_every))
