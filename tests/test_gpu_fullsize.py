"""The benchmarked workload at its full size: the 160x160 image of the
288x128x128 FMKS dump (a = 0.9375, seed 1, thetacam 163, 230 GHz, M_unit
3e26, R_high 40) on the GPU.

* against the oracle on a strided sample of its pixels (every 8th row and
  column, 400 rays, the photon-ring rays of ~10^4 steps among them), at the
  small cases' parity bar: nstep identical, per pixel |dS| / max(|I|, 1e-6
  max I) below 1e-4 everywhere and above 1e-6 on at most 1% of the pixels,
  the sample's summed I within 1e-6;
* the streamed render (bounded phases, finished rays on streams of their
  own, the longest rays' first steps composed during the last phase)
  against the one-pass render of the same image, every output bit for bit.
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import gpu_image_kwargs, stokes_errors
from ipole_amd import ipgpu as G
from ipole_amd import synth

pytestmark = pytest.mark.gpu

N, A, SEED, M_UNIT, TRAT = (288, 128, 128), 0.9375, 1, 3e26, 40.
STRIDE = 8


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def full():
    say("full-size dump %s" % (N,))
    h, prims = synth.make_dump(*N, a=A, seed=SEED)
    orc, _ = O.iharm_oracle(h, prims, M_unit=M_UNIT, trat_large=TRAT)
    p = O.default_params(nx=160, ny=160, thetacam=163., freqcgs=230e9)
    c = G.Context(0)
    g, m = G.iharm_geom(h), G.iharm_model()
    c.set_geometry(g)
    c.set_model(m)
    c.upload_prims(prims, G.iharm_units(M_UNIT, trat_large=TRAT))
    im, _, _ = G.make_image(g, m.L_unit, **gpu_image_kwargs(p))
    yield orc, p, c, im
    c.close()


def test_fullsize_sample_vs_oracle(full):
    orc, p, c, im = full
    say("full-size render vs oracle on every %d-th row and column" % STRIDE)
    got = c.render(im)
    pix = np.array([(i, j) for i in range(0, 160, STRIDE) for j in range(0, 160, STRIDE)], np.int32)
    rc, out7, ns, fl = orc.render_pixels(p, orc.camera(p), pix, nthreads=16)
    assert rc == 0
    n = 160 // STRIDE
    f3 = p.freqcgs ** 3
    ref_I = (out7[:, 0] * f3).reshape(n, n)
    ref_S = np.stack([out7[:, 1] * f3, -out7[:, 2] * f3, -out7[:, 3] * f3, out7[:, 4] * f3, out7[:, 6]],
                     axis=1).reshape(n, n, 5)
    got_I = got["image"][::STRIDE, ::STRIDE]
    got_S = got["imageS"][::STRIDE, ::STRIDE]
    assert np.array_equal(np.asarray(ns).reshape(n, n), got["nstep"][::STRIDE, ::STRIDE])
    say("  longest sampled ray %d steps" % int(np.max(ns)))
    assert abs(got_I.sum() / ref_I.sum() - 1) < 1e-6
    assert abs(got_S[..., 0].sum() / ref_S[..., 0].sum() - 1) < 1e-6
    errs = stokes_errors(ref_S, ref_I, got_S, got_I)
    say("  errors %s" % errs)
    for k, (mx, frac) in errs.items():
        assert mx < 1e-4 and frac <= 1e-2, (k, mx, frac)


def test_fullsize_streamed_equals_one_pass(full):
    orc, p, c, im = full
    say("full-size streamed vs one-pass")
    streamed = c.render(im)
    assert c.stage_times_mode()[1] is True
    c.set_tuning(trace_budget=0)
    try:
        one = c.render(im)
        assert c.stage_times_mode()[1] is False
    finally:
        c.set_tuning(trace_budget=3584)
    for k in ("image", "imageS", "taus", "nstep", "flag"):
        assert np.array_equal(streamed[k], one[k]), k
    assert int(streamed["nstep"].max()) > 10000
