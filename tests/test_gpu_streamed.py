"""The streamed render (geodesic phases with growing step budgets; the rays
each bounded phase finishes are transferred on a stream of their own while
the others resume; the longest rays last) against the one-pass render of
the same image on the GPU: the per-ray arithmetic is
the same, so every output array must be identical, whatever the budget or
the record capacity (a resumed ray that outgrows it sends the render back
to one pass with a larger capacity), given the same segment threshold (a
composed chain agrees with the sequential one to roundoff only).  One case is
also checked against the oracle at the parity bar."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import gpu_image_kwargs, small_dump
from ipole_amd import ipgpu as G
from test_gpu_parity import check

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def case():
    h, prims = small_dump()
    orc, fields = O.iharm_oracle(h, prims, M_unit=3e26)
    return h, prims, orc


def fresh(h, prims, **tune):
    c = G.Context(0)
    g, m = G.iharm_geom(h), G.iharm_model()
    c.set_geometry(g)
    c.set_model(m)
    c.upload_prims(prims, G.iharm_units(3e26))
    if tune:
        c.set_tuning(**tune)
    return c, g, m


def same(a, b):
    for k in ("image", "imageS", "taus", "nstep", "flag"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("unpol", [0, 1])
def test_streamed_equals_one_pass(case, unpol):
    h, prims, orc = case
    say("case streamed vs one-pass unpol=%d 32x32 th=163" % unpol)
    c1, g, m = fresh(h, prims, trace_budget=0)
    im, scale, _ = G.make_image(g, m.L_unit, nx=32, ny=32, thetacam=163., only_unpolarized=unpol)
    refs = {}
    for segmin in (2048, 100, 200):
        c1.set_tuning(seg_min_steps=segmin)
        refs[segmin] = c1.render(im)
        assert c1.stage_times_mode()[1] is False
    c1.close()
    nmax = int(refs[2048]["nstep"].max())
    assert nmax > 300
    for budget, segmin in ((64, 2048), (300, 100), (nmax - 1, 200), (4096, 2048)):
        say("  budget %d seg_min %d" % (budget, segmin))
        c2, _, _ = fresh(h, prims, trace_budget=budget, seg_min_steps=segmin)
        got = c2.render(im)
        assert c2.stage_times_mode()[1] is True
        same(refs[segmin], got)
        c2.close()


@pytest.mark.parametrize("seg_len", [32, 37])
def test_streamed_tail_prefix(case, seg_len, monkeypatch):
    """The first steps of the rays still traced after the last bounded phase
    (here 288 or 296 of them, whole segments) are prepared and composed on a
    stream of their own while the unbounded phase runs; the rest after it.
    Segment boundaries are counted from the camera, so the image is that of
    the one-pass render and of the streamed render without the split, bit
    for bit."""
    h, prims, orc = case
    say("case streamed tail prefix seg_len %d 32x32 th=163" % seg_len)
    tune = dict(seg_min_steps=100, seg_len=seg_len)
    c1, g, m = fresh(h, prims, trace_budget=0, **tune)
    im, _, _ = G.make_image(g, m.L_unit, nx=32, ny=32, thetacam=163.)
    ref = c1.render(im)
    c1.close()
    assert int((ref["nstep"] > 300).sum()) > 10
    c2, _, _ = fresh(h, prims, trace_budget=300, **tune)
    got = c2.render(im)
    assert c2.stage_times_mode()[1] is True
    same(ref, got)
    same(ref, c2.render(im))
    c2.close()
    monkeypatch.setenv("IPGPU_TAIL_PREFIX", "0")
    c3, _, _ = fresh(h, prims, trace_budget=300, **tune)
    same(ref, c3.render(im))
    c3.close()


@pytest.mark.parametrize("unpol,fp32", [(0, 0), (1, 0), (0, 1)], ids=["pol", "unpol", "pol-mixed"])
def test_gather_staging(case, unpol, fp32):
    """Packet kernels reading the interpolation's dump cells staged in LDS
    per 8-step group: the streamed render (prefix, composed tails, several
    phases) equals the one-pass render bit for bit under the same setting,
    matches the oracle at the parity bar, and agrees with the direct-gather
    render: unpolarized outputs bit for bit, polarized Stokes to roundoff
    (the two kernel instances fuse the step's multiply-adds differently); in
    mixed precision to the mode's per-pixel bar (the direct gather reads the
    FP32 copy of the records, the staged one the FP64 records)."""
    h, prims, orc = case
    say("case gather staging unpol=%d fp32=%d 32x32 th=163" % (unpol, fp32))
    tune = dict(seg_min_steps=100, seg_len=32)
    p = O.default_params(nx=32, ny=32, thetacam=163., freqcgs=230e9, only_unpolarized=unpol)
    c1, g, m = fresh(h, prims, trace_budget=0, **tune)
    im, scale, _ = G.make_image(g, m.L_unit, transfer_fp32=fp32, **gpu_image_kwargs(p))
    direct = c1.render(im)
    c1.set_gather_staging(True)
    ref = c1.render(im)
    assert c1.stage_times_mode()[1] is False
    c1.close()
    for budget in (64, 300):
        say("  streamed budget %d" % budget)
        c2, _, _ = fresh(h, prims, trace_budget=budget, **tune)
        c2.set_gather_staging(True)
        got = c2.render(im)
        assert c2.stage_times_mode()[1] is True
        same(ref, got)
        c2.close()
    for k in ("nstep", "flag"):
        assert np.array_equal(ref[k], direct[k]), k
    if fp32:
        # mixed precision: the direct gather reads the FP32 copy of the cell
        # records, the staged one stages the FP64 records, so the two differ
        # by the fields' FP32 rounding; compared at the mode's per-pixel bar
        for k in ("image", "taus"):
            e = np.abs(ref[k] - direct[k]) / np.maximum(np.abs(direct[k]), 1e-300)
            say("  staged vs direct %s max rel %.3e" % (k, float(e.max())))
            assert float(e.max()) < 1e-4, (k, float(e.max()))
    else:
        for k in ("image", "taus"):
            assert np.array_equal(ref[k], direct[k]), k
    if unpol:
        assert np.array_equal(ref["imageS"], direct["imageS"])
    else:
        a, b = ref["imageS"][..., :4], direct["imageS"][..., :4]
        den = np.maximum(np.abs(b[..., :1]), 1e-300)
        err = float((np.abs(a - b) / den).max())
        say("  staged vs direct polarized max rel %.3e" % err)
        assert err < (1e-4 if fp32 else 1e-9)
        if not fp32:
            orc.activate()
            check(orc.render(p), ref, scale)


def test_streamed_vs_oracle(case):
    """Small budget (most rays resume) and a low segment threshold (most
    chains composed), at the parity bar against the oracle."""
    h, prims, orc = case
    say("case streamed budget 64 seg_min 64 vs oracle 32x32 th=60")
    orc.activate()
    p = O.default_params(nx=32, ny=32, thetacam=60., freqcgs=230e9)
    ref = orc.render(p)
    c, g, m = fresh(h, prims, trace_budget=64, seg_min_steps=64)
    im, scale, _ = G.make_image(g, m.L_unit, **gpu_image_kwargs(p))
    got = c.render(im)
    assert c.stage_times_mode()[1] is True
    check(ref, got, scale)
    c.close()


@pytest.mark.parametrize("tune", [dict(trace_budget=50, record_steps=200),
                                  dict(trace_budget=128, trace_phases=2, seg_min_steps=100, seg_len=32,
                                       record_steps=200)],
                         ids=["one-phase", "prefix-composed"])
def test_streamed_record_overflow(case, tune, monkeypatch):
    """Without the record pool (IPGPU_RECORD_POOL_MAX=0), rays longer than
    the record capacity continue from their last stored record in
    full-length regions of their own: the render stays streamed and returns
    the one-pass result bit for bit, on a fresh context and on the next
    render (whose capacity follows the longest ray, so nothing continues).
    The second tuning also forms the longest rays' first segments during the
    unbounded phase."""
    h, prims, orc = case
    say("case streamed record overflow 24x24 %s" % tune)
    monkeypatch.setenv("IPGPU_RECORD_POOL_MAX", "0")
    c1, g, m = fresh(h, prims, trace_budget=0, seg_min_steps=tune.get("seg_min_steps", 3584),
                     seg_len=tune.get("seg_len", 384))
    im, _, _ = G.make_image(g, m.L_unit, nx=24, ny=24, thetacam=163.)
    ref = c1.render(im)
    c1.close()
    assert ref["nstep"].max() > 200
    c2, _, _ = fresh(h, prims, **tune)
    got = c2.render(im)
    assert c2.stage_times_mode()[1] is True
    chunks, cont, _ = c2.render_info()
    assert chunks == 1 and cont == int(np.sum(ref["nstep"] >= 200)) > 0, (chunks, cont)
    same(ref, got)
    got2 = c2.render(im)
    assert c2.stage_times_mode()[1] is True
    assert c2.render_info()[1] == 0
    same(ref, got2)
    c2.close()


@pytest.mark.parametrize("pool_max", [-1, 2])
@pytest.mark.parametrize("tune", [dict(trace_budget=50),
                                  dict(trace_budget=128, trace_phases=2, seg_min_steps=100, seg_len=32)],
                         ids=["one-phase", "prefix-composed"])
def test_streamed_record_pool(case, tune, pool_max, monkeypatch):
    """The record pool: the per-ray slots hold the bounded phases' records,
    and each ray still live when the last bounded phase ends gets a
    full-length region assigned on the device, its records copied there
    before the unbounded phase continues it.  The render equals the one-pass
    render bit for bit.  The pool is sized from the previous render's share
    of long rays (a first guess on a fresh context), so here, where most
    rays outlast the bounded phase, the first render continues the rays
    beyond its guess as outgrown ones and the second has a region for every
    ray; with the pool held to 2 regions the other long rays continue as
    outgrown ones (their count reported) on every render, still bit for
    bit."""
    h, prims, orc = case
    say("case streamed record pool 24x24 %s pool_max %d" % (tune, pool_max))
    c1, g, m = fresh(h, prims, trace_budget=0, seg_min_steps=tune.get("seg_min_steps", 3584),
                     seg_len=tune.get("seg_len", 384))
    im, _, _ = G.make_image(g, m.L_unit, nx=24, ny=24, thetacam=163.)
    ref = c1.render(im)
    c1.close()
    n_long = int(np.sum(ref["nstep"] > tune["trace_budget"]))
    assert n_long > 2
    monkeypatch.setenv("IPGPU_RECORD_POOL_MAX", str(pool_max))
    c2, _, _ = fresh(h, prims, **tune)
    try:
        for rep in range(2):
            got = c2.render(im)
            assert c2.stage_times_mode()[1] is True
            chunks, cont, _ = c2.render_info()
            want = n_long - 2 if pool_max > 0 else (0 if rep > 0 else cont)
            assert chunks == 1 and cont == want and (rep > 0 or cont < n_long), (rep, chunks, cont, n_long)
            same(ref, got)
    finally:
        c2.close()


@pytest.mark.parametrize("seg_min", [3584, 20])
def test_streamed_maxnstep(case, seg_min):
    """A resumed ray past maxnstep is reported with the one-pass pixel, also
    when every ray of the last phase's list is composed from segments (that
    list runs the composing chain kernel alone)."""
    h, prims, orc = case
    say("case streamed maxnstep 8x8 seg_min %d" % seg_min)
    errs = []
    for budget in (0, 20):
        c, g, m = fresh(h, prims, trace_budget=budget, seg_min_steps=seg_min, seg_len=16)
        im, _, _ = G.make_image(g, m.L_unit, nx=8, ny=8, thetacam=163., maxnstep=50)
        with pytest.raises(G.IpgpuError) as e:
            c.render(im)
        assert e.value.code == G.ERR_MAXNSTEP
        errs.append(str(e.value))
        c.close()
    assert errs[0] == errs[1]


def test_streamed_context_reuse(case):
    """One context rendering different images in turn (sizes, cameras,
    budgets): each result equals a one-pass render of the same image.  The
    per-phase host rows and device arenas of a render must not leak into the
    next."""
    h, prims, orc = case
    say("case streamed context reuse")
    c, g, m = fresh(h, prims, trace_budget=256, seg_min_steps=400)
    ref_ctx, _, _ = fresh(h, prims, trace_budget=0, seg_min_steps=400)
    for nx, th, budget in ((32, 163., 256), (24, 60., 256), (40, 130., 512), (16, 17., 300), (32, 163., 128)):
        say("  %dx%d th=%g budget %d" % (nx, nx, th, budget))
        c.set_tuning(trace_budget=budget)
        im, _, _ = G.make_image(g, m.L_unit, nx=nx, ny=nx, thetacam=th)
        got = c.render(im)
        assert c.stage_times_mode()[1] is True
        same(ref_ctx.render(im), got)
    c.close()
    ref_ctx.close()


def test_autotune_render(case):
    """ipgpu_autotune_render keeps a measured setting from its candidates; the
    image after tuning agrees with the default setting's to the roundoff of
    the segment regrouping, step counts and flags identical, and with the
    oracle at the parity bar."""
    h, prims, orc = case
    say("case autotune render 48x48 th=163")
    c, g, m = fresh(h, prims, trace_budget=512, trace_phases=2, seg_min_steps=512, seg_len=128)
    try:
        im, scale, _ = G.make_image(g, m.L_unit, nx=48, ny=48, thetacam=163.)
        ref = c.render(im)
        ch = c.autotune_render(im)
        say("  chosen %s" % ch)
        assert ch["seg_len"] in (128, 256, 512) and ch["trace_phases"] == 2
        assert ch["trace_budget"] in (384, 512, 640)
        assert ch["seg_min_steps"] in (256, 512, ch["trace_budget"])
        got = c.render(im)
        assert np.array_equal(ref["nstep"], got["nstep"]) and np.array_equal(ref["flag"], got["flag"])
        e = np.abs(got["imageS"][..., :4] - ref["imageS"][..., :4]).max() / np.abs(ref["imageS"][..., 0]).max()
        assert e < 1e-12, e
    finally:
        c.close()
    orc.activate()
    p = O.default_params(nx=48, ny=48, thetacam=163., freqcgs=230e9)
    check(orc.render(p), got, scale)
