"""The mixed-precision transfer (ipgpu_image.transfer_fp32: the coefficients'
shape functions and the plasma-frame Stokes step in FP32; geodesics,
connection, tetrads and N in FP64) on the GPU, single image and batch,
against the FP64 oracle: image Ftot (polarized and unpolarized) within 1e-5
(the mode's bar is |dFtot|/Ftot <= 1e-4; measured ~4e-7), LP, CP and EVPA
within the mode's 1e-4 (measured up to ~1e-5 rad in EVPA, where the FP32
Faraday rotation's rounding accumulates over the ray)."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import gpu_image_kwargs, small_dump
from ipole_amd import ipgpu as G

pytestmark = pytest.mark.gpu


def summary(imageS):
    S = imageS.reshape(-1, 5)
    sI, sQ, sU, sV = S[:, 0].sum(), S[:, 1].sum(), S[:, 2].sum(), S[:, 3].sum()
    return sI, np.hypot(sQ, sU) / sI, sV / sI, 0.5 * np.arctan2(sU, sQ)


def errors(ref, got):
    rI, rLP, rCP, rEV = summary(ref["imageS"])
    gI, gLP, gCP, gEV = summary(got["imageS"])
    return dict(Ftot=abs(gI / rI - 1), LP=abs(gLP - rLP), CP=abs(gCP - rCP), EVPA=abs(gEV - rEV),
                Funpol=abs(got["image"].sum() / ref["image"].sum() - 1))


@pytest.fixture(scope="module")
def setup():
    h, prims = small_dump()
    c = G.Context(0)
    g, m = G.iharm_geom(h), G.iharm_model()
    c.set_geometry(g)
    c.set_model(m)
    c.upload_prims(prims, G.iharm_units(3e26))
    yield h, prims, c, g, m
    c.close()


@pytest.mark.parametrize("thetacam,etype", [(163., 4), (60., 1)])
def test_fp32_single_and_batch(setup, thetacam, etype):
    h, prims, c, g, m = setup
    print("case fp32 th=%g emission_type=%d 32x32" % (thetacam, etype), flush=True)
    orc, _ = O.iharm_oracle(h, prims, M_unit=3e26)
    p = O.default_params(nx=32, ny=32, thetacam=thetacam, freqcgs=230e9, emission_type=etype)
    ref = orc.render(p)
    kw = gpu_image_kwargs(p)
    kw["transfer_fp32"] = 1
    im, _, _ = G.make_image(g, m.L_unit, **kw)
    single = c.render(im)
    batch = c.render_batch([(im, 1.)])[0]
    for name, got in (("single", single), ("batch", batch)):
        assert np.array_equal(got["nstep"], ref["nstep"])
        e = errors(ref, got)
        print("  %s %s" % (name, e), flush=True)
        for k, v in e.items():
            assert v < (1e-5 if k in ("Ftot", "Funpol") else 1e-4), (name, k, v)
