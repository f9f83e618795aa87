"""The adjoint entry points (ipgpu_adjoint, ipgpu_backward) on the GPU: against the same code
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


def oracle_loss(h, prims, im_params, target, Mu=M, rh=RH, rl=RL, cell=None, sc=1.):
    """The loss of the oracle's image (the reference's algorithm on the
    CPU) against the target: n_e of one interior cell scaled by sc."""
    orc, _ = O.iharm_oracle(h, prims, M_unit=Mu, trat_large=rh, trat_small=rl)
    if cell is not None:
        g = orc.geom
        nc = (g.N1 + 2) * (g.N2 + 2) * (g.N3 + 2)
        i, j, k = cell
        idx = ((i + 1) * (g.N2 + 2) + (j + 1)) * (g.N3 + 2) + (k + 1)
        orc.fields[6 * nc + idx] *= sc
        orc.activate()
    S = orc.render(im_params)["imageS"]
    return float(np.sum(W * (S[..., :4] - target[..., :4]) ** 2))


def interior_cells(gd, n=2):
    """The cells of largest |gradient| away from the grid's boundary faces
    (no ghost copies, so scaling n_e there is one entry of the oracle's
    fields)."""
    N1, N2, N3 = gd.shape
    out = []
    for idx in np.argsort(-np.abs(gd).ravel()):
        i, j, k = np.unravel_index(idx, gd.shape)
        if 0 < i < N1 - 1 and 0 < j < N2 - 1 and 0 < k < N3 - 1:
            out.append((int(i), int(j), int(k)))
        if len(out) == n:
            return out
    return out


def test_adjoint_vs_oracle_finite_differences(setup):
    """The GPU's gradient against central finite differences of the oracle's
    loss (the reference algorithm on the CPU, rendered at M_unit, R_high,
    R_low and one cell's n_e +- h): bar 1e-3."""
    h, prims, c, g, m, im, target = setup
    say("case adjoint GPU vs oracle finite differences, 12x12")
    c.upload_prims(prims, G.iharm_units(M, trat_large=RH, trat_small=RL))
    r = c.adjoint(im, target, W)
    p = O.default_params(nx=12, ny=12, thetacam=60., freqcgs=230e9)
    L0 = oracle_loss(h, prims, p, target)
    assert abs(L0 / r["loss"] - 1) < 1e-4, (L0, r["loss"])
    hh = 1e-5
    fd = {}
    fd["dM_unit"] = (oracle_loss(h, prims, p, target, Mu=M * (1 + hh)) -
                     oracle_loss(h, prims, p, target, Mu=M * (1 - hh))) / (2 * hh * M)
    fd["dR_high"] = (oracle_loss(h, prims, p, target, rh=RH * (1 + hh)) -
                     oracle_loss(h, prims, p, target, rh=RH * (1 - hh))) / (2 * hh * RH)
    fd["dR_low"] = (oracle_loss(h, prims, p, target, rl=RL * (1 + hh)) -
                    oracle_loss(h, prims, p, target, rl=RL * (1 - hh))) / (2 * hh * RL)
    for k, v in fd.items():
        say("  %s gpu %r oracle fd %r" % (k, r[k], v))
        assert abs(r[k] / v - 1) < 1e-3, (k, r[k], v)
    gd = r["density"]
    for cell in interior_cells(gd):
        v = (oracle_loss(h, prims, p, target, cell=cell, sc=1 + hh) -
             oracle_loss(h, prims, p, target, cell=cell, sc=1 - hh)) / (2 * hh)
        say("  cell %s gpu %r oracle fd %r" % (cell, gd[cell], v))
        assert abs(gd[cell] / v - 1) < 1e-3


def test_backward_vjp_vs_finite_differences(setup):
    """ipgpu_backward with an arbitrary cotangent C of the image: its
    M_unit, R_high, R_low and cell gradients against central differences of
    sum(C * S) over GPU renders (bar 1e-3); with C = 2 w (S - target) it
    equals ipgpu_adjoint's gradient."""
    h, prims, c, g, m, im, target = setup
    say("case backward (VJP) vs finite differences, 12x12")
    rng = np.random.default_rng(11)
    c.upload_prims(prims, G.iharm_units(M, trat_large=RH, trat_small=RL))
    S0 = c.render(im)["imageS"]
    Cot = rng.normal(size=(12, 12, 4)) * np.abs(S0[..., :1]).mean() ** -1

    def phi(Mu=M, rh=RH, rl=RL, scale=None):
        c.upload_prims(prims, G.iharm_units(Mu, trat_large=rh, trat_small=rl))
        if scale is not None:
            c.scale_density(scale)
        return float(np.sum(Cot * c.render(im)["imageS"][..., :4]))

    c.upload_prims(prims, G.iharm_units(M, trat_large=RH, trat_small=RL))
    r = c.backward(im, Cot)
    hh = 1e-5
    fd = dict(dM_unit=(phi(Mu=M * (1 + hh)) - phi(Mu=M * (1 - hh))) / (2 * hh * M),
              dR_high=(phi(rh=RH * (1 + hh)) - phi(rh=RH * (1 - hh))) / (2 * hh * RH),
              dR_low=(phi(rl=RL * (1 + hh)) - phi(rl=RL * (1 - hh))) / (2 * hh * RL))
    for k, v in fd.items():
        say("  %s vjp %r fd %r" % (k, r[k], v))
        assert abs(r[k] / v - 1) < 1e-3, (k, r[k], v)
    gd = r["density"]
    for cell in interior_cells(gd):
        sp, sm = np.ones(gd.shape), np.ones(gd.shape)
        sp[cell] += hh
        sm[cell] -= hh
        v = (phi(scale=sp) - phi(scale=sm)) / (2 * hh)
        say("  cell %s vjp %r fd %r" % (cell, gd[cell], v))
        assert abs(gd[cell] / v - 1) < 1e-3
    # the L2 loss's cotangent reproduces ipgpu_adjoint
    c.upload_prims(prims, G.iharm_units(M, trat_large=RH, trat_small=RL))
    a = c.adjoint(im, target, W)
    b = c.backward(im, 2 * W * (a["imageS"][..., :4] - target[..., :4]))
    for k in ("dM_unit", "dR_high", "dR_low"):
        assert abs(b[k] / a[k] - 1) < 1e-12, (k, b[k], a[k])
    assert np.allclose(b["density"], a["density"], rtol=1e-12, atol=1e-12 * np.abs(a["density"]).max())


def test_render_autograd(setup):
    """The Render function: a torch loss on the rendered image backward
    through ipgpu_backward gives ipgpu_adjoint's gradient."""
    import torch
    from ipole_amd.autograd import Render
    h, prims, c, g, m, im, target = setup
    say("case Render autograd")
    render = Render(c, prims, G.iharm_units(M, trat_large=RH, trat_small=RL), im)
    Mt = torch.tensor(M, dtype=torch.float64, requires_grad=True)
    Rh = torch.tensor(RH, dtype=torch.float64, requires_grad=True)
    Rl = torch.tensor(RL, dtype=torch.float64, requires_grad=True)
    s = torch.ones((g.N1, g.N2, g.N3), dtype=torch.float64, requires_grad=True)
    S = render(Mt, Rh, Rl, s)
    w = torch.tensor(W, dtype=torch.float64)
    T = torch.from_numpy(np.ascontiguousarray(target[..., :4]))
    loss = (w * (S - T) ** 2).sum()
    loss.backward()
    c.upload_prims(prims, G.iharm_units(M, trat_large=RH, trat_small=RL))
    a = c.adjoint(im, target, W)
    assert abs(loss.item() / a["loss"] - 1) < 1e-12
    assert abs(Mt.grad.item() / a["dM_unit"] - 1) < 1e-10
    assert abs(Rh.grad.item() / a["dR_high"] - 1) < 1e-10
    assert abs(Rl.grad.item() / a["dR_low"] - 1) < 1e-10
    assert np.allclose(s.grad.numpy(), a["density"], rtol=1e-10, atol=1e-10 * np.abs(a["density"]).max())


def test_c3_adjoint_vs_finite_differences():
    """The differentiation workload: the 96x64x64 dump (the 288x128x128
    workload's generator), 128x128 px, thetacam 163, 230 GHz, loss against a
    target rendered at M_unit x1.2, R_high x0.9 and a seeded 5% lognormal n_e
    scale.  GPU adjoint against central differences of GPU renders for
    M_unit, R_high, R_low and three cells (bar 1e-3), and, on a 32x32 image of
    the same dump, against central differences of the oracle's loss."""
    from ipole_amd import synth
    say("case C3 adjoint 96x64x64 at 128x128")
    h, prims = synth.make_dump(96, 64, 64, a=0.9375, seed=1)
    c = G.Context(0)
    try:
        g, m = G.iharm_geom(h), G.iharm_model()
        c.set_geometry(g)
        c.set_model(m)
        rng = np.random.default_rng(3)
        dscale = np.exp(0.05 * rng.normal(size=(96, 64, 64)))
        hh = 1e-5
        for nx in (128, 32):
            im, _, _ = G.make_image(g, m.L_unit, nx=nx, ny=nx, thetacam=163., freqcgs=230e9)
            c.upload_prims(prims, G.iharm_units(1.2 * M, trat_large=0.9 * RH))
            c.scale_density(dscale)
            target = c.render(im)["imageS"]
            c.upload_prims(prims, G.iharm_units(M, trat_large=RH, trat_small=RL))
            r = c.adjoint(im, target, W)
            if nx == 128:
                fd = dict(dM_unit=(gpu_loss(c, prims, im, target, Mu=M * (1 + hh))
                                   - gpu_loss(c, prims, im, target, Mu=M * (1 - hh))) / (2 * hh * M),
                          dR_high=(gpu_loss(c, prims, im, target, rh=RH * (1 + hh))
                                   - gpu_loss(c, prims, im, target, rh=RH * (1 - hh))) / (2 * hh * RH),
                          dR_low=(gpu_loss(c, prims, im, target, rl=RL * (1 + hh))
                                  - gpu_loss(c, prims, im, target, rl=RL * (1 - hh))) / (2 * hh * RL))
                gd = r["density"]
                for cell in interior_cells(gd, 3):
                    sp, sm = np.ones(gd.shape), np.ones(gd.shape)
                    sp[cell] += hh
                    sm[cell] -= hh
                    v = (gpu_loss(c, prims, im, target, scale=sp) - gpu_loss(c, prims, im, target, scale=sm)) / (2 * hh)
                    say("  128: cell %s adjoint %r fd %r" % (cell, gd[cell], v))
                    assert abs(gd[cell] / v - 1) < 1e-3
                tag = "GPU fd"
            else:
                p = O.default_params(nx=32, ny=32, thetacam=163., freqcgs=230e9)
                fd = dict(dM_unit=(oracle_loss(h, prims, p, target, Mu=M * (1 + hh))
                                   - oracle_loss(h, prims, p, target, Mu=M * (1 - hh))) / (2 * hh * M),
                          dR_high=(oracle_loss(h, prims, p, target, rh=RH * (1 + hh))
                                   - oracle_loss(h, prims, p, target, rh=RH * (1 - hh))) / (2 * hh * RH),
                          dR_low=(oracle_loss(h, prims, p, target, rl=RL * (1 + hh))
                                  - oracle_loss(h, prims, p, target, rl=RL * (1 - hh))) / (2 * hh * RL))
                tag = "oracle fd"
            for k, v in fd.items():
                say("  %d: %s adjoint %r %s %r" % (nx, k, r[k], tag, v))
                assert abs(r[k] / v - 1) < 1e-3, (nx, k, r[k], v)
    finally:
        c.close()
