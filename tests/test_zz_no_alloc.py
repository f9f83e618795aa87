"""The allocation-free render (ipgpu_render_noalloc) through the C-ABI, into
device arrays the test allocates, against ipgpu_render of the same image:
nstep and flags identical, per pixel |dS| / max(|I|, 1e-6 max I) within 1e-9
(the two transfer the rays with the same operations, grouped differently for
the longest rays: agreement to roundoff), Ftot within 1e-12.  Two calls in a
row, the second on a smaller image, so the second runs on the scratch the
first left behind; then the full 160x160 image of the 288x128x128 dump, whose
records need several chunks of the scratch."""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import small_dump, stokes_errors
from ipole_amd import ipgpu as G
from ipole_amd import synth

pytestmark = pytest.mark.gpu

PIX_TOL = 1e-9


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _hip():
    """The HIP runtime the library itself links (the test allocates its
    output arrays there; the entry point under test allocates nothing)."""
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    L.hipDeviceSynchronize.argtypes = []
    return L


HtoD, DtoH = 1, 2


def render_noalloc(ctx, im):
    hip = _hip()
    n = im.nx * im.ny
    sizes = dict(image=8 * n, imageS=40 * n, taus=8 * n, nstep=4 * n, flag=4 * n)
    dev = {}
    try:
        for k, b in sizes.items():
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), b) == 0
            dev[k] = p
            assert hip.hipMemset(p, 0xff, b) == 0     # NaN / -1 where nothing is written
        assert hip.hipDeviceSynchronize() == 0
        rc = ctx.L.ipgpu_render_noalloc(ctx.h, C.byref(im), dev["image"], dev["imageS"], dev["taus"], dev["nstep"],
                                        dev["flag"])
        ctx._chk(rc)
        out = dict(image=np.empty(n), imageS=np.empty(5 * n), taus=np.empty(n), nstep=np.empty(n, np.int32),
                   flag=np.empty(n, np.int32))
        for k, a in out.items():
            assert hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), dev[k], sizes[k], DtoH) == 0
    finally:
        for p in dev.values():
            hip.hipFree(p)
    nx, ny = im.nx, im.ny
    return dict(image=out["image"].reshape(nx, ny), imageS=out["imageS"].reshape(nx, ny, 5),
                taus=out["taus"].reshape(nx, ny), nstep=out["nstep"].reshape(nx, ny),
                flag=out["flag"].reshape(nx, ny))


def compare(ref, got, tag):
    assert np.array_equal(ref["nstep"], got["nstep"]), tag
    assert np.array_equal(ref["flag"], got["flag"]), tag
    for k in ("image", "imageS", "taus"):
        assert np.all(np.isfinite(got[k])), (tag, k)
    F_ref, F_got = ref["imageS"][..., 0].sum(), got["imageS"][..., 0].sum()
    assert abs(F_got / F_ref - 1) < 1e-12, (tag, F_got, F_ref)
    errs = stokes_errors(ref["imageS"], ref["image"], got["imageS"], got["image"])
    say("  %s errors %s" % (tag, errs))
    for k, (mx, frac) in errs.items():
        assert mx < PIX_TOL, (tag, k, mx)
    tf = ref["imageS"][..., 4]
    assert np.all(np.abs(got["imageS"][..., 4] - tf) <= PIX_TOL * np.maximum(np.abs(tf), 1e-300)), tag
    t = ref["taus"]
    assert np.all(np.abs(got["taus"] - t) <= PIX_TOL * np.maximum(np.abs(t), 1e-300)), tag


def test_noalloc_two_calls_vs_render():
    h, prims = small_dump()
    c = G.Context(0)
    try:
        g, m = G.iharm_geom(h), G.iharm_model()
        c.set_geometry(g)
        c.set_model(m)
        c.upload_prims(prims, G.iharm_units(3e26))
        for nx, th, f in ((40, 163., 230e9), (24, 60., 345e9)):
            say("case noalloc %dx%d th=%g" % (nx, nx, th))
            im, _, _ = G.make_image(g, m.L_unit, nx=nx, ny=nx, thetacam=th, freqcgs=f)
            ref = c.render(im)
            got = render_noalloc(c, im)
            compare(ref, got, "%dx%d" % (nx, nx))
    finally:
        c.close()


def test_noalloc_fullsize_chunks():
    say("case noalloc 160x160 of the 288x128x128 dump")
    h, prims = synth.make_dump(288, 128, 128, a=0.9375, seed=1)
    c = G.Context(0)
    try:
        g, m = G.iharm_geom(h), G.iharm_model()
        c.set_geometry(g)
        c.set_model(m)
        c.upload_prims(prims, G.iharm_units(3e26, trat_large=40.))
        im, _, _ = G.make_image(g, m.L_unit, nx=160, ny=160, thetacam=163., freqcgs=230e9)
        ref = c.render(im)
        got = render_noalloc(c, im)
        chunks = c.render_info()[0]
        say("  %d chunks of the scratch" % chunks)
        assert chunks >= 2
        compare(ref, got, "160x160")
    finally:
        c.close()
