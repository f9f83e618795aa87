"""torch.autograd.Function over the C-ABI's forward and adjoint entry points:
the L2 loss between a rendered Stokes image and a target as a
differentiable function of M_unit, R_high, R_low and a per-cell density
(n_e) scale of an iharm dump.

    loss_fn = ImageLoss(ctx, prims, units, image, target)
    M = torch.tensor(3e26, dtype=torch.float64, requires_grad=True)
    ...
    L = loss_fn(M, R_high, R_low, scale)      # scale: (N1, N2, N3) tensor
    L.backward()

forward re-derives the dump on the device at the given M_unit, R_high,
R_low (ipgpu_upload_prims), applies the density scale
(ipgpu_scale_density) and runs ipgpu_adjoint, which returns the loss and
its gradient in one pass; backward scales the stored gradient by the
incoming one.  Tensors are CPU float64 (the library takes host buffers).
"""
import numpy as np
import torch

from . import ipgpu as G


class _ImageLossFn(torch.autograd.Function):
    @staticmethod
    def forward(actx, M_unit, R_high, R_low, scale, owner):
        u = owner.units
        units = G.iharm_units(float(M_unit), MBH=u.MBH_solar, trat_small=float(R_low), trat_large=float(R_high),
                              beta_crit=u.beta_crit, gam=u.gam)
        owner.ctx.upload_prims(owner.prims, units)
        s = scale.detach().cpu().numpy().astype(np.float64)
        owner.ctx.scale_density(s)
        r = owner.ctx.adjoint(owner.image, owner.target, owner.weights, density=True)
        owner.last = r
        # dL/ds_c at s: the adjoint gives dL/d(ln s_c) (a multiplicative
        # perturbation of the scaled n_e)
        gs = torch.from_numpy(r["density"] / np.where(s != 0., s, 1.))
        actx.save_for_backward(torch.tensor([r["dM_unit"], r["dR_high"], r["dR_low"]], dtype=torch.float64), gs)
        return torch.tensor(r["loss"], dtype=torch.float64)

    @staticmethod
    def backward(actx, grad):
        g3, gs = actx.saved_tensors
        return grad * g3[0], grad * g3[1], grad * g3[2], grad * gs, None


class ImageLoss:
    """The L2 Stokes-image loss of one camera against a target image.

    ctx: ipgpu.Context with geometry and model set; prims: the dump's
    primitives (N1, N2, N3, nprim); units: ipgpu.Units carrying MBH, beta_crit
    and gamma (M_unit, R_high, R_low come from the call); image: ipgpu.Image;
    target: (nx, ny, 5) imageS layout; weights: I, Q, U, V weights."""

    def __init__(self, ctx, prims, units, image, target, weights=(1., 1., 1., 1.)):
        self.ctx, self.prims, self.units, self.image = ctx, prims, units, image
        self.target = np.ascontiguousarray(target, dtype=np.float64)
        self.weights = np.asarray(weights, dtype=np.float64)
        self.last = None

    def __call__(self, M_unit, R_high, R_low, scale):
        return _ImageLossFn.apply(M_unit, R_high, R_low, scale, self)
