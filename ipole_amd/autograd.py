"""torch.autograd.Functions over the C-ABI's forward and adjoint entry points.

Render: the polarized Stokes image (nx, ny, 4: I, Q, U, V in the output
units of ipgpu_render) as a differentiable function of M_unit, R_high,
R_low and a per-cell density (n_e) scale of an iharm dump; its backward
takes grad_output and runs ipgpu_backward (the vector-Jacobian product), so
any loss written in torch on the image differentiates through it.

    render = Render(ctx, prims, units, image)
    S = render(M, R_high, R_low, scale)       # scale: (N1, N2, N3) tensor
    loss = ((S - target) ** 2).sum()
    loss.backward()

ImageLoss: the L2 loss between a rendered Stokes image and a target as a
differentiable function of the same parameters, in one adjoint pass
(ipgpu_adjoint returns the loss and its gradient together).

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
        s = _prepare(owner, M_unit, R_high, R_low, scale)
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


def _prepare(owner, M_unit, R_high, R_low, scale):
    """Re-derive the dump at the given parameters and apply the density
    scale; returns the scale as a numpy array."""
    u = owner.units
    units = G.iharm_units(float(M_unit), MBH=u.MBH_solar, trat_small=float(R_low), trat_large=float(R_high),
                          beta_crit=u.beta_crit, gam=u.gam)
    owner.ctx.upload_prims(owner.prims, units)
    s = scale.detach().cpu().numpy().astype(np.float64)
    owner.ctx.scale_density(s)
    return s


class _RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(actx, M_unit, R_high, R_low, scale, owner):
        s = _prepare(owner, M_unit, R_high, R_low, scale)
        S = owner.ctx.render(owner.image)["imageS"][..., :4]
        actx.owner = owner
        actx.params = (float(M_unit), float(R_high), float(R_low))
        actx.save_for_backward(scale)
        return torch.from_numpy(np.ascontiguousarray(S))

    @staticmethod
    def backward(actx, grad_output):
        owner = actx.owner
        (scale,) = actx.saved_tensors
        M, Rh, Rl = actx.params
        # the dump of this forward call (another call may have re-derived it)
        s = _prepare(owner, M, Rh, Rl, scale)
        r = owner.ctx.backward(owner.image, grad_output.detach().cpu().numpy())
        owner.last = r
        gs = torch.from_numpy(r["density"] / np.where(s != 0., s, 1.))
        one = lambda v: torch.tensor(v, dtype=torch.float64)  # noqa: E731
        return one(r["dM_unit"]), one(r["dR_high"]), one(r["dR_low"]), gs, None


class Render:
    """The polarized image of one camera as a differentiable function of
    M_unit, R_high, R_low and a per-cell n_e scale.

    ctx: ipgpu.Context with geometry and model set; prims: the dump's
    primitives (N1, N2, N3, nprim); units: ipgpu.Units carrying MBH, beta_crit
    and gamma (M_unit, R_high, R_low come from the call); image: ipgpu.Image.
    Returns an (nx, ny, 4) float64 tensor (I, Q, U, V)."""

    def __init__(self, ctx, prims, units, image):
        self.ctx, self.prims, self.units, self.image = ctx, prims, units, image
        self.last = None

    def __call__(self, M_unit, R_high, R_low, scale):
        return _RenderFn.apply(M_unit, R_high, R_low, scale, self)


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
