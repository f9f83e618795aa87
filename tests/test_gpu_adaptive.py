"""GPU adaptive refinement (ipgpu_render_adaptive) against the oracle's
restatement of src/main.c:505-792 and against the GPU's own plain render.

* every estimate above tolerance: every pixel traced, bit for bit the plain
  GPU render (the pixel-list passes trace and transfer each ray as the
  tiled render does);
* nothing refined: the coarse pass, then the hierarchical interpolation of
  it, bit for bit (linear, nearest neighbour, -unpol);
* working tolerances: the same pixels refined as the oracle (the decision
  reads intensities that agree to ~1e-9, far inside the thresholds' margins
  on this case), traced pixels within the plain-render parity bar, Ftot
  within 1e-6, Iavg within 1e-9;
* the host's nx_min / refine keys end to end.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from helpers import gpu_image_kwargs, small_dump
from ipole_amd import h5
from ipole_amd import ipgpu as G

pytestmark = pytest.mark.gpu

M_UNIT = 3e26
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ipole_amd", "bin", "ipole")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    h, prims = small_dump()
    orc, fields = O.iharm_oracle(h, prims, M_unit=M_UNIT)
    ctx = G.Context(0)
    ctx.set_geometry(G.iharm_geom(h))
    ctx.set_model(G.iharm_model())
    ctx.upload_fields(fields[:10])
    yield h, orc, ctx
    ctx.close()


def image(h, orc, **kw):
    p = O.default_params(nx=32, ny=32, thetacam=163., freqcgs=230e9, **kw)
    im, scale, _ = G.make_image(G.iharm_geom(h), orc.model.L_unit, **gpu_image_kwargs(p))
    return p, im, scale


def test_all_traced_equals_plain(setup):
    h, orc, ctx = setup
    p, im, _ = image(h, orc)
    say("adaptive, everything refined")
    plain = ctx.render(im)
    a = ctx.render_adaptive(im, G.make_refine(8, refine_abs=0., refine_rel=0., refine_cut=-1.))
    assert a["levels"] == 3 and a["traced"] == [81, 208, 800] and a["interpolated"] == [0, 0, 0]
    for k in ("image", "imageS", "taus", "nstep", "flag"):
        assert np.array_equal(a[k], plain[k]), k
    assert not a["interp"].any()


@pytest.mark.parametrize("nearest,unpol", [(0, 0), (1, 0), (1, 1), (0, 1)])
def test_nothing_refined(setup, nearest, unpol):
    h, orc, ctx = setup
    p, im, _ = image(h, orc, only_unpolarized=unpol)
    say("adaptive, nothing refined, nearest=%d unpol=%d" % (nearest, unpol))
    a = ctx.render_adaptive(im, G.make_refine(8, refine_abs=1e300, refine_rel=1e300, nearest_neighbor=nearest))
    assert a["traced"] == [81, 0, 0] and a["interpolated"] == [0, 208, 1008]
    o = orc.render_adaptive(p, 8, refine_abs=1e300, refine_rel=1e300, nearest=nearest)
    assert np.array_equal(a["interp"], o["interp"])
    sc = np.abs(o["image"]).max()
    assert np.abs(a["image"] - o["image"]).max() < 1e-6 * sc
    assert np.abs(a["taus"] - o["taus"]).max() <= 1e-6 * np.abs(o["taus"]).max()
    # the GPU's own coarse values, interpolated: identical arithmetic on both sides
    plain = ctx.render(im)
    base = plain["image"][::4, ::4]
    assert np.array_equal(a["image"][::4, ::4], base)
    if not unpol:
        assert np.array_equal(a["imageS"][::4, ::4], plain["imageS"][::4, ::4])
        mid = a["imageS"][2:30:4, 0:32:4]
        if nearest:
            assert np.array_equal(mid, plain["imageS"][0:28:4, 0:32:4])
        else:
            assert np.array_equal(mid, .5 * (plain["imageS"][0:28:4, 0:32:4] + plain["imageS"][4:32:4, 0:32:4]))
    assert np.all(a["nstep"][a["interp"] == 1] == 0)


@pytest.mark.parametrize("cam", [dict(), dict(only_unpolarized=1)])
def test_working_tolerances_vs_oracle(setup, cam):
    h, orc, ctx = setup
    p, im, scale = image(h, orc, **cam)
    say("adaptive vs oracle %s" % cam)
    a = ctx.render_adaptive(im, G.make_refine(8, refine_abs=1e-6, refine_rel=1e-2))
    o = orc.render_adaptive(p, 8, refine_abs=1e-6, refine_rel=1e-2)
    assert a["Iavg"] == pytest.approx(o["Iavg"], rel=1e-9)
    assert np.array_equal(a["interp"], o["interp"])
    assert a["traced"] == o["traced"] and a["interpolated"] == o["interpolated"]
    say("  traced per level %s" % a["traced"])
    I = np.abs(o["image"])
    e = np.abs(a["image"] - o["image"]) / np.maximum(I, 1e-6 * I.max())
    assert e.max() < 1e-5 and np.mean(e > 1e-6) < 0.01
    F_o, F_a = o["image"].sum() * scale, a["image"].sum() * scale
    assert abs(F_a / F_o - 1) < 1e-6
    if not cam:
        Io = np.abs(o["imageS"][..., 0])
        for k in range(4):
            e = np.abs(a["imageS"][..., k] - o["imageS"][..., k]) / np.maximum(Io, 1e-6 * Io.max())
            assert e.max() < 1e-5 and np.mean(e > 1e-6) < 0.01, k


@pytest.mark.parametrize("nearest,unpol", [(0, 0), (1, 1)])
def test_speculative_equals_levels(setup, nearest, unpol, monkeypatch):
    """The whole grid traced at once and then refined in place gives the
    image, flags and counts of tracing each level's list."""
    h, orc, ctx = setup
    p, im, _ = image(h, orc, only_unpolarized=unpol)
    rf = G.make_refine(8, refine_abs=1e-6, refine_rel=1e-2, nearest_neighbor=nearest)
    say("speculative vs levels, nearest=%d unpol=%d" % (nearest, unpol))
    a = ctx.render_adaptive(im, rf)
    monkeypatch.setenv("IPGPU_REFINE_SPECULATE_MAX", "0")
    c2 = G.Context(0)
    try:
        c2.set_geometry(G.iharm_geom(h))
        c2.set_model(G.iharm_model())
        c2.upload_fields(O.iharm_oracle(h, small_dump()[1], M_unit=M_UNIT)[1][:10])
        b = c2.render_adaptive(im, rf)
    finally:
        c2.close()
    for k in ("image", "imageS", "taus", "nstep", "flag", "interp"):
        assert np.array_equal(a[k], b[k]), k
    assert a["traced"] == b["traced"] and a["interpolated"] == b["interpolated"] and a["Iavg"] == b["Iavg"]


def test_bad_sizes(setup):
    h, orc, ctx = setup
    p = O.default_params(nx=30, ny=30, thetacam=163., freqcgs=230e9)
    im, _, _ = G.make_image(G.iharm_geom(h), orc.model.L_unit, **gpu_image_kwargs(p))
    with pytest.raises(G.IpgpuError) as e:
        ctx.render_adaptive(im, G.make_refine(8))
    assert e.value.code == G.ERR_ARG
    # nx_min = nx: the plain render
    p, im, _ = image(h, orc)
    a = ctx.render_adaptive(im, G.make_refine(32))
    assert a["levels"] == 1 and np.array_equal(a["image"], ctx.render(im)["image"])


def test_host_adaptive(tmp_path):
    h, prims = small_dump()
    dump = str(tmp_path / "dump.h5")
    h5.write_dump(dump, h, prims)
    say("host --nx_min 8")
    r = subprocess.run([EXE, "--dump=%s" % dump, "--M_unit=%g" % M_UNIT, "--freqcgs=230e9", "--nx=32", "--ny=32",
                        "--thetacam=163", "--nx_min=8", "--ny_min=8", "--refine_abs=1e-6", "--refine_rel=1e-2",
                        "--outfile=%s" % (tmp_path / "img.h5")], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "refined up to 32x32 (3 levels)" in r.stderr
    assert "Refining level 3 of 3, spacing 1,1" in r.stderr
    m = re.findall(r"(\d+) of (\d+) \(\S+%\) of computed pixels at (\d+)x(\d+) were interpolated", r.stderr)
    assert [x[2:] for x in m] == [("16", "16"), ("32", "32")]
    orc, _ = O.iharm_oracle(h, prims, M_unit=M_UNIT)
    p = O.default_params(nx=32, ny=32, thetacam=163., freqcgs=230e9)
    o = orc.render_adaptive(p, 8, refine_abs=1e-6, refine_rel=1e-2)
    assert [int(x[0]) for x in m] == o["interpolated"][1:]
    img = h5.read_image(str(tmp_path / "img.h5"), 32, 32)
    F_o = o["imageS"][..., 0].sum() * o["cam"].scale
    assert abs(img["Ftot"] / F_o - 1) < 1e-6
