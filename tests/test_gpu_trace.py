"""GPU trace output (ipgpu_trace_pixel) against the oracle's restatement of
dump_var_along (src/io.c:279-543), and the host's trace file
(--trace, src/main.c:865-880).

Bar: the same trajectory length, and each field within a tolerance of the
oracle, measured as |d| / max |ref| over the pixel's steps: 1e-7 for the
geodesic (X, Kcon, dl, r, th, phi), 1e-5 for the plasma tetrad and the
tensors projected on it, 1e-6 for the rest.  The GPU forms the connection
from exact metric derivatives and the oracle from finite differences
(DEL = 1e-7), so the positions differ at the ~1e-9 level and the fields by
what that moves through the fluid interpolation; the CPU build of the same
code meets these bars (tests/test_kernel_emu.py) with margins of 2-10x.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from helpers import gpu_image_kwargs, small_dump, trace_compare as compare
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
def iharm():
    h, prims = small_dump()
    orc, _ = O.iharm_oracle(h, prims, M_unit=M_UNIT)
    ctx = G.Context(0)
    ctx.set_geometry(G.iharm_geom(h))
    ctx.set_model(G.iharm_model())
    ctx.upload_prims(prims, G.iharm_units(M_UNIT, gam=h["gam"]))
    yield h, prims, orc, ctx
    ctx.close()


@pytest.mark.parametrize("cam", [dict(thetacam=60.), dict(thetacam=163., target_nturns=0),
                                 dict(thetacam=90., only_unpolarized=1)])
def test_trace_pixel_vs_oracle(iharm, cam):
    h, prims, orc, ctx = iharm
    p = O.default_params(nx=16, ny=16, freqcgs=230e9, **cam)
    oc = orc.camera(p)
    im, _, _ = G.make_image(G.iharm_geom(h), orc.model.L_unit, **gpu_image_kwargs(p))
    for ij in [(8, 7), (3, 12), (15, 0)]:
        say("trace pixel %s %s" % (ij, cam))
        rc, ns, ref = orc.trace_pixel(p, oc, *ij)
        assert rc == 0
        nsg, got = ctx.trace_pixel(im, *ij)
        assert nsg == ns, (ij, nsg, ns)
        worst = compare(ref, got, (ij, tuple(cam.items())))
        say("  nstep %d, worst %s" % (ns, sorted(worst.items(), key=lambda kv: -kv[1])[:3]))
        if cam.get("only_unpolarized"):
            assert not np.any(G.trace_fields(got)["ncoord"])


def test_trace_pixel_sphere():
    kw = dict(R0=50., tau0=2.0143e-5, Thetae0=10., beta0=30.)
    g, m = O.sphere_model(**kw)
    orc = O.Oracle(g, m)
    p = O.default_params(nx=16, ny=16, thetacam=60., dx=100., dy=100., freqcgs=230e9)
    oc = orc.camera(p)
    gg, mm = G.sphere_setup(**kw)
    ctx = G.Context(0)
    try:
        ctx.set_geometry(gg)
        ctx.set_model(mm)
        im, _, _ = G.make_image(gg, mm.L_unit, **gpu_image_kwargs(p))
        for ij in [(8, 8), (2, 9)]:
            say("sphere trace pixel %s" % (ij,))
            rc, ns, ref = orc.trace_pixel(p, oc, *ij)
            nsg, got = ctx.trace_pixel(im, *ij)
            assert rc == 0 and nsg == ns
            compare(ref, got, ij)
            assert np.any(G.trace_fields(got)["beta"] > 0)
    finally:
        ctx.close()


def test_trace_pixel_errors(iharm):
    h, prims, orc, ctx = iharm
    p = O.default_params(nx=8, ny=8, thetacam=60., freqcgs=230e9, maxnstep=40)
    im, _, _ = G.make_image(G.iharm_geom(h), orc.model.L_unit, **gpu_image_kwargs(p))
    with pytest.raises(G.IpgpuError) as e:
        ctx.trace_pixel(im, 4, 4)
    assert e.value.code == G.ERR_MAXNSTEP
    im.maxnstep = 50000
    with pytest.raises(G.IpgpuError):
        ctx.trace_pixel(im, 8, 0)              # outside the image


def test_host_trace_file(tmp_path):
    h, prims = small_dump()
    dump = str(tmp_path / "dump.h5")
    h5.write_dump(dump, h, prims)
    common = ["--dump=%s" % dump, "--M_unit=%g" % M_UNIT, "--freqcgs=230e9", "--nx=16", "--ny=16",
              "--thetacam=60", "-quench"]
    say("host --trace_i 5 --trace_j 6")
    r = subprocess.run([EXE] + common + ["--trace=1", "--trace_i=5", "--trace_j=6", "--trace_outf=one.h5"],
                       cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    path = str(tmp_path / "one.h5")
    nst = h5.read_dataset(path, "nstep")
    assert nst.shape == (1, 1)
    orc, _ = O.iharm_oracle(h, prims, M_unit=M_UNIT)
    p = O.default_params(nx=16, ny=16, thetacam=60., freqcgs=230e9)
    rc, ns, ref = orc.trace_pixel(p, orc.camera(p), 5, 6)
    assert int(nst[0, 0]) == ns
    got = np.zeros((ns + 1, G.TRACE_NF))
    for name, off, shape in G.TRACE_SETS:
        d = h5.read_dataset(path, name)
        assert d.shape[:3] == (1, 1, 50000)
        w = int(np.prod(shape)) if shape else 1
        got[:, off:off + w] = d[0, 0, :ns + 1].reshape(ns + 1, w)
        assert not np.any(d[0, 0, ns + 1:])
    compare(ref, got, "host")
    assert h5.dataset_shape(path, "/header/camera/x") == (4,)
    assert h5.dataset_shape(path, "/fluid_header/geom/startx1") == ()
    say("host --trace_stride 8")
    r = subprocess.run([EXE] + common + ["--trace=1", "--trace_stride=8", "--trace_outf=grid.h5"],
                       cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    nst = h5.read_dataset(str(tmp_path / "grid.h5"), "nstep")
    assert nst.shape == (2, 2)
    for a in range(2):
        for b in range(2):
            rc, ns, _ = orc.trace_pixel(p, orc.camera(p), 8 * a, 8 * b)
            assert int(nst[a, b]) == ns
