"""The benchmarked workload at its full size: the 160x160 image of the
288x128x128 FMKS dump (a = 0.9375, seed 1, thetacam 163, 230 GHz, M_unit
3e26, R_high 40) on the GPU.

* against the oracle on a strided sample of its pixels (every 8th row and
  column, 400 rays, the photon-ring rays of ~10^4 steps among them), at the
  small cases' parity bar: nstep identical, per pixel |dS| / max(|I|, 1e-6
  max I) below 1e-5 everywhere and above 1e-6 on at most 1% of the pixels,
  the sample's summed I within 1e-6;
* the streamed render (bounded phases, finished rays on streams of their
  own, the longest rays' first steps composed during the last phase, rays
  past the record capacity continued in regions of their own) against the
  one-pass render of the same image, every output bit for bit;
* the image rendered in chunks of rays under a small render memory against
  the whole render, bit for bit;
* a batch of three jobs on the same dump (its first trace keeps the records
  in per-ray slots and continues the rays longer than a slot), FP64 and
  mixed precision, against each job's plain render.
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
        assert mx < 1e-5 and frac <= 1e-2, (k, mx, frac)


def test_fullsize_streamed_equals_one_pass(full, monkeypatch):
    """On a context of its own (so independent of the other tests): the
    first render is streamed, its per-ray slots holding the bounded phases'
    3,585 records and the 808 rays still live after them (up to 11,267
    steps) taking pooled full-length regions, so nothing is retraced; it and
    the next render equal the one-pass render bit for bit.  Without the pool
    (IPGPU_RECORD_POOL_MAX=0) the longest rays exceed the default record
    capacity of 8,192 on the first render and continue from their last
    record in regions of their own, with the same result."""
    orc, p, c, im = full
    say("full-size streamed (fresh context) vs one-pass")
    h, prims = synth.make_dump(*N, a=A, seed=SEED)
    renders = []
    for pool_max in ("-1", "0"):
        monkeypatch.setenv("IPGPU_RECORD_POOL_MAX", pool_max)
        c2 = G.Context(0)
        try:
            g, m = G.iharm_geom(h), G.iharm_model()
            c2.set_geometry(g)
            c2.set_model(m)
            c2.upload_prims(prims, G.iharm_units(M_UNIT, trat_large=TRAT))
            first = c2.render(im)
            assert c2.stage_times_mode()[1] is True
            chunks, cont, nbytes = c2.render_info()
            say("  pool max %s, first render: %d chunk(s), %d continued rays, %.1f GB held" %
                (pool_max, chunks, cont, nbytes / 1e9))
            assert chunks == 1 and (cont == 0 if pool_max == "-1" else cont > 0)
            second = c2.render(im)
            assert c2.stage_times_mode()[1] is True and c2.render_info()[1] == 0
            renders += [first, second]
        finally:
            c2.close()
    c.set_tuning(trace_budget=0)
    try:
        one = c.render(im)
        assert c.stage_times_mode()[1] is False
    finally:
        c.set_tuning(trace_budget=3584)
    for k in ("image", "imageS", "taus", "nstep", "flag"):
        for r in renders:
            assert np.array_equal(r[k], one[k]), k
    assert int(one["nstep"].max()) > 10000


@pytest.mark.parametrize("fp32", [0, 1])
def test_fullsize_chunked_equals_whole(full, fp32):
    """The 160x160 image rendered in chunks of rays under a small render
    memory (8 GiB) equals the render of all rays at once, bit for bit; in
    mixed precision too, where every chunk's render rewrites the FP32 copy of
    the cell records its packet kernels gather."""
    orc, p, c, im = full
    if fp32:
        im = type(im).from_buffer_copy(im)
        im.transfer_fp32 = 1
    say("full-size chunked (8 GiB) vs whole, fp32=%d" % fp32)
    whole = c.render(im)
    assert c.render_info()[0] == 1
    c.set_render_memory(8 << 30)
    try:
        parts = c.render(im)
        chunks, cont, nbytes = c.render_info()
    finally:
        c.set_render_memory(64 << 30)
    say("  %d chunks, %.1f GB held" % (chunks, nbytes / 1e9))
    assert chunks >= 3
    for k in ("image", "imageS", "taus", "nstep", "flag"):
        assert np.array_equal(whole[k], parts[k]), k


def test_fullsize_batch_vs_plain(full):
    """Three jobs on the full dump in one batch (two frequencies and M_unit
    scales of the 160x160 camera, one 96x96 camera at 50 degrees), FP64 and
    mixed precision, against the plain render of each job with the dump
    derived at its M_unit: FP64 nstep identical, |dFtot|/Ftot < 1e-9, pixels
    within 1e-6; mixed |dFtot|/Ftot within the mode's bar of 1e-4."""
    orc, p, c, im = full
    h, prims = synth.make_dump(*N, a=A, seed=SEED)
    g, m = G.iharm_geom(h), G.iharm_model()
    specs = [(160, 163., 230e9, 1.0), (160, 163., 86e9, 0.4), (96, 50., 345e9, 2.5)]
    say("full-size batch of %d jobs vs plain" % len(specs))
    plain = []
    for n, th, f, s in specs:
        c.upload_prims(prims, G.iharm_units(M_UNIT * s, trat_large=TRAT))
        ji, _, _ = G.make_image(g, m.L_unit, nx=n, ny=n, thetacam=th, freqcgs=f)
        plain.append(c.render(ji))
    c.upload_prims(prims, G.iharm_units(M_UNIT, trat_large=TRAT))
    for f32 in (0, 1):
        jobs = []
        for n, th, f, s in specs:
            ji, _, _ = G.make_image(g, m.L_unit, nx=n, ny=n, thetacam=th, freqcgs=f, transfer_fp32=f32)
            jobs.append((ji, s))
        got = c.render_batch(jobs)
        for k, (ref, out) in enumerate(zip(plain, got)):
            assert np.array_equal(ref["nstep"], out["nstep"])
            dF = abs(out["imageS"][..., 0].sum() / ref["imageS"][..., 0].sum() - 1)
            say("  %s job %d: dFtot %.2e, longest ray %d" % ("mixed" if f32 else "fp64", k, dF,
                                                              int(out["nstep"].max())))
            if f32:
                assert dF < 1e-4
            else:
                assert dF < 1e-9
                for name, (mx, frac) in stokes_errors(ref["imageS"], ref["image"], out["imageS"],
                                                      out["image"]).items():
                    assert mx < 1e-6, (k, name, mx)


def test_fullsize_512_within_render_memory(full):
    """A 512x512 image of the same dump under a render memory of 32 GiB
    (all its rays at once would need ~400 GB of records and packets): it
    runs in chunks, the context holds at most the budget plus the dump and
    the image, and a strided sample of pixels (every 32nd row and column)
    matches the oracle at the parity bar."""
    orc, p, c, im = full
    budget = 32 << 30
    say("512x512 under %d GiB" % (budget >> 30))
    h, prims = synth.make_dump(*N, a=A, seed=SEED)
    p512 = O.default_params(nx=512, ny=512, thetacam=163., freqcgs=230e9)
    c3 = G.Context(0)
    try:
        g, m = G.iharm_geom(h), G.iharm_model()
        c3.set_geometry(g)
        c3.set_model(m)
        c3.upload_prims(prims, G.iharm_units(M_UNIT, trat_large=TRAT))
        c3.set_render_memory(budget)
        im512, _, _ = G.make_image(g, m.L_unit, **gpu_image_kwargs(p512))
        got = c3.render(im512)
        chunks, cont, nbytes = c3.render_info()
    finally:
        c3.close()
    dump_bytes = 290 * 130 * 130 * 128
    image_bytes = 512 * 512 * 9 * 8
    say("  %d chunks, %d continued rays, %.2f GB held (budget %.2f GB + dump %.2f GB)"
        % (chunks, cont, nbytes / 1e9, budget / 1e9, dump_bytes / 1e9))
    assert chunks >= 8
    assert nbytes <= budget + dump_bytes + image_bytes + (64 << 20), nbytes
    S = 32
    pix = np.array([(i, j) for i in range(0, 512, S) for j in range(0, 512, S)], np.int32)
    rc, out7, ns, fl = orc.render_pixels(p512, orc.camera(p512), pix, nthreads=16)
    assert rc == 0
    n = 512 // S
    f3 = p512.freqcgs ** 3
    ref_I = (out7[:, 0] * f3).reshape(n, n)
    ref_S = np.stack([out7[:, 1] * f3, -out7[:, 2] * f3, -out7[:, 3] * f3, out7[:, 4] * f3, out7[:, 6]],
                     axis=1).reshape(n, n, 5)
    assert np.array_equal(np.asarray(ns).reshape(n, n), got["nstep"][::S, ::S])
    errs = stokes_errors(ref_S, ref_I, got["imageS"][::S, ::S], got["image"][::S, ::S])
    say("  errors %s" % errs)
    for k, (mx, frac) in errs.items():
        assert mx < 1e-5 and frac <= 1e-2, (k, mx, frac)
