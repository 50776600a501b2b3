"""Reverse-mode differentiation through the GPU energy path.

``energies(batch, coords, weights)`` returns per-structure total energies
and per-atom energies as torch tensors that carry gradients to the atomic
coordinates and to the network weights.  Forward and backward both run in
libaenet_amd.so: the forward pass is the batch evaluation, the backward
pass is ``aenet_amd_batch_vjp``, which forms sum_i g_i dE_i/dR with the
force kernels (the reference's analytic forces, src/aenet.f90:471-477,
weighted per atom) and sum_i g_i dE_i/dW with the network-gradient kernel
(ff_wderiv, src/ext/feedforward.f90:763-860).  No torch autograd graph is
built inside; torch only routes the upstream gradients.
"""
import torch


class _PotentialEnergy(torch.autograd.Function):
    """The weights an evaluation used are kept with it: backward
    differentiates at those weights, whatever the model holds by then, and
    neither pass leaves the model's weights changed."""

    @staticmethod
    def forward(ctx, coords, weights, batch, model):
        W_used = None
        if model is not None:
            W_prev = model.weights()
            W_used = W_prev if weights is None else weights.detach().to(torch.float32).contiguous().clone()
        x = coords.detach().to(torch.float64).contiguous().view(-1)
        if weights is not None and model is not None:
            model.set_weights(W_used)
        try:
            E_tot, _, E_at, _ = batch.eval(coords=x, forces=False, check=True, model=model)
            E_tot, E_at = E_tot.clone(), E_at.clone()
        finally:
            if weights is not None and model is not None:
                model.set_weights(W_prev)
        ctx.batch = batch
        ctx.model = model
        ctx.W_used = W_used
        ctx.has_w = weights is not None
        ctx.save_for_backward(x)
        return E_tot, E_at

    @staticmethod
    def backward(ctx, g_tot, g_at):
        (x,) = ctx.saved_tensors
        batch, model = ctx.batch, ctx.model
        counts = torch.as_tensor(batch.natoms_per, device=x.device, dtype=torch.long)
        g = torch.zeros(batch.natoms, dtype=torch.float64, device=x.device)
        if g_tot is not None:
            g = g + torch.repeat_interleave(g_tot.to(torch.float64), counts)
        if g_at is not None:
            g = g + g_at.to(torch.float64)
        gW = None
        if ctx.has_w and ctx.needs_input_grad[1]:
            gW = torch.zeros(model.wsize, dtype=torch.float32, device=x.device)
        W_now = None
        if model is not None:
            W_now = model.weights()
            model.set_weights(ctx.W_used)
        try:
            gR = batch.vjp(g.contiguous(), grad_W=gW, coords=x, model=model)
        finally:
            if W_now is not None:
                model.set_weights(W_now)
        return gR.view(-1, 3), gW, None, None


def energies(batch, coords, weights=None, model=None):
    """(E_tot per structure, E_atoms) of the batch at coords ((natoms, 3)
    float64 device tensor), differentiable with respect to coords and, when
    given, the weights (a float32 tensor of the model's weight layout; the
    model evaluates with it and gets its own weights back afterwards)."""
    return _PotentialEnergy.apply(coords, weights, batch, model if model is not None else batch.model)


def forces(batch, coords, model=None):
    """-dE_tot/dR through the autograd path (for checks against the direct
    force kernels)."""
    x = coords.detach().clone().requires_grad_(True)
    E_tot, _ = energies(batch, x, model=model)
    (g,) = torch.autograd.grad(E_tot.sum(), x)
    return -g
