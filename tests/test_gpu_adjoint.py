"""The adjoint entry point (ipgpu_adjoint) on the GPU: against the same code
built for the CPU (tests/kernel_emu, itself checked against finite
differences), against central finite differences of GPU renders for
M_unit, R_high, R_low and single-cell density scales (bar 1e-3 relative),
and through the torch.autograd.Function wrapper."""
import sys

import numpy as np
import pytest

import emu
import oracle as O
from helpers import small_dump
from ipole_amd import ipgpu as G

pytestmark = pytest.mark.gpu

M, RH, RL = 3e26, 40., 1.
W = np.array([1., 3., 3., 5.])


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fold(g3, N1, N2, N3):
    """The ghost fold of ipgpu_adjoint, in numpy (even N3)."""
    g = g3.reshape(N1 + 2, N2 + 2, N3 + 2).copy()
    g[:, :, N3] += g[:, :, 0]
    g[:, :, 1] += g[:, :, N3 + 1]
    kk = np.arange(1, N3 + 1)
    kflip = ((kk - 1) + N3 // 2) % N3 + 1
    for k, kf in zip(kk, kflip):
        g[:, 1, kf] += g[:, 0, k]
        g[:, N2, kf] += g[:, N2 + 1, k]
    g[1, 1:N2 + 1, 1:N3 + 1] += g[0, 1:N2 + 1, 1:N3 + 1]
    g[N1, 1:N2 + 1, 1:N3 + 1] += g[N1 + 1, 1:N2 + 1, 1:N3 + 1]
    return g[1:N1 + 1, 1:N2 + 1, 1:N3 + 1]


@pytest.fixture(scope="module")
def setup():
    h, prims = small_dump(N=(32, 24, 24), seed=5)
    c = G.Context(0)
    g, m = G.iharm_geom(h), G.iharm_model()
    c.set_geometry(g)
    c.set_model(m)
    im, _, _ = G.make_image(g, m.L_unit, nx=12, ny=12, thetacam=60., freqcgs=230e9)
    c.upload_prims(prims, G.iharm_units(1.3 * M, trat_large=0.8 * RH, trat_small=RL))
    target = c.render(im)["imageS"]
    yield h, prims, c, g, m, im, target
    c.close()


def gpu_loss(c, prims, im, target, Mu=M, rh=RH, rl=RL, scale=None):
    c.upload_prims(prims, G.iharm_units(Mu, trat_large=rh, trat_small=rl))
    if scale is not None:
        c.scale_density(scale)
    S = c.render(im)["imageS"]
    return float(np.sum(W * (S[..., :4] - target[..., :4]) ** 2))


def test_adjoint_vs_cpu_build(setup):
    h, prims, c, g, m, im, target = setup
    say("case adjoint GPU vs CPU build, 12x12")
    c.upload_prims(prims, G.iharm_units(M, trat_large=RH, trat_small=RL))
    r = c.adjoint(im, target, W)
    _, f11 = O.iharm_oracle(h, prims, M_unit=M, trat_large=RH, trat_small=RL)
    d = emu.pack_dump_adjoint(f11, G.iharm_units(M, trat_large=RH, trat_small=RL))
    out, gcell, S = emu.adjoint(g, m, im, d, target.reshape(-1), W)
    say("  loss %r %r; dlnM %r %r" % (r["loss"], out[0], r["dM_unit"] * M, out[1]))
    assert abs(r["loss"] / out[0] - 1) < 1e-5
    assert abs(r["dM_unit"] * M / out[1] - 1) < 1e-5
    assert abs(r["dR_high"] / out[2] - 1) < 1e-5
    assert abs(r["dR_low"] / out[3] - 1) < 1e-5
    ref = fold(gcell, g.N1, g.N2, g.N3)
    scale = np.abs(ref).max()
    assert np.abs(r["density"] - ref).max() < 1e-5 * scale


def test_adjoint_vs_finite_differences(setup):
    h, prims, c, g, m, im, target = setup
    say("case adjoint GPU vs GPU finite differences, 12x12")
    c.upload_prims(prims, G.iharm_units(M, trat_large=RH, trat_small=RL))
    r = c.adjoint(im, target, W)
    L0 = gpu_loss(c, prims, im, target)
    assert abs(L0 / r["loss"] - 1) < 1e-10
    hh = 1e-5
    fd_M = (gpu_loss(c, prims, im, target, Mu=M * (1 + hh)) - gpu_loss(c, prims, im, target, Mu=M * (1 - hh))) / (2 * hh * M)
    fd_Rh = (gpu_loss(c, prims, im, target, rh=RH * (1 + hh)) - gpu_loss(c, prims, im, target, rh=RH * (1 - hh))) / (2 * hh * RH)
    fd_Rl = (gpu_loss(c, prims, im, target, rl=RL * (1 + hh)) - gpu_loss(c, prims, im, target, rl=RL * (1 - hh))) / (2 * hh * RL)
    say("  dM %r %r dRh %r %r dRl %r %r" % (r["dM_unit"], fd_M, r["dR_high"], fd_Rh, r["dR_low"], fd_Rl))
    assert abs(r["dM_unit"] / fd_M - 1) < 1e-3
    assert abs(r["dR_high"] / fd_Rh - 1) < 1e-3
    assert abs(r["dR_low"] / fd_Rl - 1) < 1e-3
    gd = r["density"]
    for idx in np.argsort(-np.abs(gd).ravel())[:3]:
        ijk = np.unravel_index(idx, gd.shape)
        sp, sm = np.ones(gd.shape), np.ones(gd.shape)
        sp[ijk] += hh
        sm[ijk] -= hh
        fd = (gpu_loss(c, prims, im, target, scale=sp) - gpu_loss(c, prims, im, target, scale=sm)) / (2 * hh)
        say("  cell %s %r %r" % (ijk, gd[ijk], fd))
        assert abs(gd[ijk] / fd - 1) < 1e-3


def test_autograd_function(setup):
    import torch
    from ipole_amd.autograd import ImageLoss
    h, prims, c, g, m, im, target = setup
    say("case autograd wrapper")
    fn = ImageLoss(c, prims, G.iharm_units(M, trat_large=RH, trat_small=RL), im, target, W)
    Mt = torch.tensor(M, dtype=torch.float64, requires_grad=True)
    Rh = torch.tensor(RH, dtype=torch.float64, requires_grad=True)
    Rl = torch.tensor(RL, dtype=torch.float64, requires_grad=True)
    s = torch.full((g.N1, g.N2, g.N3), 1.1, dtype=torch.float64, requires_grad=True)
    L = fn(Mt, Rh, Rl, s)
    L.backward()
    assert abs(L.item() / fn.last["loss"] - 1) < 1e-14
    assert Mt.grad.item() == fn.last["dM_unit"]
    # dL/ds at s = 1.1 by a finite difference in one cell
    gd = s.grad.numpy()
    idx = np.unravel_index(np.argmax(np.abs(gd)), gd.shape)
    hh = 1e-5
    sp, sm = np.full(gd.shape, 1.1), np.full(gd.shape, 1.1)
    sp[idx] += hh
    sm[idx] -= hh
    fd = (gpu_loss(c, prims, im, target, scale=sp) - gpu_loss(c, prims, im, target, scale=sm)) / (2 * hh)
    say("  ds %r %r" % (gd[idx], fd))
    assert abs(gd[idx] / fd - 1) < 1e-3
