"""ctypes binding of the CPU build of the GPU library's per-ray code
(tests/kernel_emu/libkernel_emu.so).  Test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

from ipole_amd import ipgpu as G

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libkernel_emu.so")
dp, ip = G.dp, G.ip
_lib = None


def lib():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(LIB)
        L.emu_metric.argtypes = [C.POINTER(G.Geom), C.c_double, C.c_double, dp]
        L.emu_accel.argtypes = [C.POINTER(G.Geom), dp, dp, dp, dp]
        L.emu_bessel.argtypes = [C.c_double, dp, dp]
        L.emu_point.argtypes = [C.POINTER(G.Geom), C.POINTER(G.Model), dp, C.c_int, dp, dp, dp]
        L.emu_math.argtypes = [C.c_int, dp, C.c_int, dp]
        L.emu_stokes_affine.argtypes = [dp, C.c_double, dp, dp]
        L.emu_stokes_affine_f32.argtypes = [dp, C.c_double, dp, dp]
        L.emu_render_multi.argtypes = [C.POINTER(G.Geom), C.POINTER(G.Model), C.POINTER(G.Image), C.c_int,
                                       C.POINTER(G.Image), dp, dp, ip, C.c_int, dp, ip, ip]
        L.emu_plasma_tetrad.argtypes = [C.POINTER(G.Geom), dp, dp, dp, dp, dp]
        L.emu_render.argtypes = [C.POINTER(G.Geom), C.POINTER(G.Model), C.POINTER(G.Image), dp, ip,
                                 C.c_int, dp, ip, ip]
        L.emu_render_scaled.argtypes = [C.POINTER(G.Geom), C.POINTER(G.Model), C.POINTER(G.Image),
                                        C.POINTER(G.Image), dp, ip, C.c_int, C.c_double, dp, ip, ip]
        L.emu_adjoint.argtypes = [C.POINTER(G.Geom), C.POINTER(G.Model), C.POINTER(G.Image), dp, dp, dp, dp,
                                  dp, dp]
        L.emu_render_segmented.argtypes = [C.POINTER(G.Geom), C.POINTER(G.Model), C.POINTER(G.Image), dp, ip,
                                           C.c_int, C.c_int, dp, ip, ip]
        L.emu_trace.argtypes = [C.POINTER(G.Geom), C.POINTER(G.Image), C.c_int, C.c_int, dp]
        L.emu_trace_pixel.argtypes = [C.POINTER(G.Geom), C.POINTER(G.Model), C.POINTER(G.Image), dp, C.c_int,
                                      C.c_int, C.c_int, dp]
        _lib = L
    return _lib


def f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def pack_dump(fields10):
    """Ghosted SoA fields (10, N1+2, N2+2, N3+2) -> 16-double cell records."""
    f = np.asarray(fields10)[:10]
    nc = f[0].size
    d = np.zeros((nc, 16))
    d[:, :10] = f.reshape(10, nc).T
    return d.ravel()


def metric(geom, X1, X2):
    o = np.zeros(43)
    lib().emu_metric(C.byref(geom), X1, X2, G.P(o))
    return o


def accel(geom, X, K):
    a, M = np.zeros(4), np.zeros(16)
    lib().emu_accel(C.byref(geom), G.P(f64(X)), G.P(f64(K)), G.P(a), G.P(M))
    return a, M.reshape(4, 4)


def bessel(x):
    k, e = np.zeros(3), C.c_double()
    lib().emu_bessel(x, G.P(k), C.byref(e))
    return k, e.value


def point(geom, model, dump, X, K, emission_type=4):
    o = np.zeros(32)
    d = dump if dump is not None else np.zeros(16)
    lib().emu_point(C.byref(geom), C.byref(model), G.P(d), emission_type, G.P(f64(X)), G.P(f64(K)), G.P(o))
    return o


def math(kind, x):
    x = f64(np.atleast_1d(x))
    out = np.zeros((len(x), 2))
    lib().emu_math(kind, G.P(x), len(x), G.P(out))
    return out


def stokes_affine(c11, x, S):
    out = np.zeros(4)
    lib().emu_stokes_affine(G.P(f64(c11)), x, G.P(f64(S)), G.P(out))
    return out


def stokes_affine_f32(c11, x, S):
    out = np.zeros(4)
    lib().emu_stokes_affine_f32(G.P(f64(c11)), x, G.P(f64(S)), G.P(out))
    return out


def render(geom, model, image, dump, pix=None):
    if pix is None:
        pix = np.array([(i, j) for i in range(image.nx) for j in range(image.ny)], np.int32)
    pix = np.ascontiguousarray(pix, dtype=np.int32)
    n = len(pix)
    out7 = np.zeros((n, 7))
    ns, fl = np.zeros(n, np.int32), np.zeros(n, np.int32)
    d = dump if dump is not None else np.zeros(16)
    rc = lib().emu_render(C.byref(geom), C.byref(model), C.byref(image), G.P(d), G.IP(pix), n,
                          G.P(out7), G.IP(ns), G.IP(fl))
    return rc, out7, ns, fl


def render_segmented(geom, model, image, dump, seg_len=256):
    """Every ray through the segmented (composed-map) chain."""
    pix = np.array([(i, j) for i in range(image.nx) for j in range(image.ny)], np.int32)
    n = len(pix)
    out7, ns, fl = np.zeros((n, 7)), np.zeros(n, np.int32), np.zeros(n, np.int32)
    rc = lib().emu_render_segmented(C.byref(geom), C.byref(model), C.byref(image), G.P(f64(dump)), G.IP(pix),
                                    n, seg_len, G.P(out7), G.IP(ns), G.IP(fl))
    return rc, out7, ns, fl


def render_multi(geom, model, image_trace, image_jobs, ne_scales, dump):
    """Multi-job transfer on the CPU (see emu_render_multi): out7 (npix, nj,
    7), nstep (npix), flag (npix, nj)."""
    nj = len(image_jobs)
    nx, ny = image_trace.nx, image_trace.ny
    pix = np.array([(i, j) for i in range(nx) for j in range(ny)], np.int32)
    n = len(pix)
    arr = (G.Image * nj)(*image_jobs)
    out7, ns, fl = np.zeros((n, nj, 7)), np.zeros(n, np.int32), np.zeros((n, nj), np.int32)
    d = f64(dump) if dump is not None else None
    rc = lib().emu_render_multi(C.byref(geom), C.byref(model), C.byref(image_trace), nj, arr,
                                G.P(f64(ne_scales)), G.P(d) if d is not None else None, G.IP(pix), n,
                                G.P(out7), G.IP(ns), G.IP(fl))
    return rc, out7, ns, fl


def render_scaled(geom, model, image_trace, image_job, dump, ne_scale=1.):
    """Batch-path arithmetic on the CPU (see emu_render_scaled)."""
    nx, ny = image_job.nx, image_job.ny
    pix = np.array([(i, j) for i in range(nx) for j in range(ny)], np.int32)
    n = len(pix)
    out7, ns, fl = np.zeros((n, 7)), np.zeros(n, np.int32), np.zeros(n, np.int32)
    d = f64(dump) if dump is not None else None
    rc = lib().emu_render_scaled(C.byref(geom), C.byref(model), C.byref(image_trace), C.byref(image_job),
                                 G.P(d) if d is not None else None, G.IP(pix), n, ne_scale, G.P(out7),
                                 G.IP(ns), G.IP(fl))
    return rc, out7, ns, fl


def pack_dump_adjoint(fields11, units):
    """16-slot cell records with slots 13, 14 = dTheta_e/dR_high, dR_low of
    the R_high/R_low electron model (model/iharm/model.c:610-717), from the
    derived Theta_e and beta (fields11[7], fields11[10])."""
    d = pack_dump(fields11[:10]).reshape(-1, 16)
    th = np.asarray(fields11[7]).ravel()
    beta = np.asarray(fields11[10]).ravel()
    betasq = beta ** 2 / units.beta_crit ** 2
    trat = units.trat_large * betasq / (1. + betasq) + units.trat_small / (1. + betasq)
    dth = np.where(th > 1e-3, -th * (units.game - 1.) / ((units.gamp - 1.) + (units.game - 1.) * trat), 0.)
    d[:, 13] = dth * betasq / (1. + betasq)
    d[:, 14] = dth / (1. + betasq)
    return d.ravel()


def pack_dump_trace(fields11, rho_u):
    """16-slot cell records with rho, u, beta in slots 10, 11, 12, as the
    device derivation leaves them (rho_u: the ghosted rho and u arrays)."""
    d = pack_dump(fields11[:10]).reshape(-1, 16)
    ru = np.asarray(rho_u).reshape(2, -1)
    d[:, 10], d[:, 11] = ru[0], ru[1]
    d[:, 12] = np.asarray(fields11[10]).ravel()
    return d.ravel()


def trace_pixel(geom, model, image, dump, i, j, prims_ok=True):
    out = np.zeros((image.maxnstep, G.TRACE_NF))
    ns = lib().emu_trace_pixel(C.byref(geom), C.byref(model), C.byref(image),
                               None if dump is None else G.P(f64(dump)), int(prims_ok), i, j, G.P(out))
    return ns, out[:ns + 1].copy()


def adjoint(geom, model, image, dump, target, w=(1., 1., 1., 1.), seg_len=0):
    """(loss, dL/dlnM, dL/dR_high, dL/dR_low), per-cell dL/d(ne scale) over
    the ghosted cells, and the image's I Q U V; seg_len > 0: the reverse
    pass in segments of that many steps, as on the GPU."""
    n = image.nx * image.ny
    nc = (geom.N1 + 2) * (geom.N2 + 2) * (geom.N3 + 2)
    out, gcell, S = np.zeros(4), np.zeros(nc), np.zeros((n, 4))
    rc = lib().emu_adjoint(C.byref(geom), C.byref(model), C.byref(image), G.P(f64(dump)),
                           G.P(f64(target)), G.P(f64(w)), G.P(out), G.P(gcell), G.P(S), int(seg_len))
    assert rc == 0
    return out, gcell, S


def trace(geom, image, i, j):
    rec = np.zeros(image.maxnstep * 8)
    ns = lib().emu_trace(C.byref(geom), C.byref(image), i, j, G.P(rec))
    return ns, rec[: 8 * (ns + 1)].reshape(ns + 1, 8)


def trace_two_pass(geom, image, i, j, budget):
    rec = np.zeros(image.maxnstep * 8)
    paused = C.c_int(0)
    ns = lib().emu_trace_two_pass(C.byref(geom), C.byref(image), i, j, budget, G.P(rec), C.byref(paused))
    return ns, rec[: 8 * (ns + 1)].reshape(ns + 1, 8), bool(paused.value)
