"""The oracle against the reference's own code (oracle/_ref: the reference's
GSL-free translation units compiled unmodified, with FMA contraction off in
both builds).  Functions the two share must agree bit for bit.

Skipped where the reference build is absent (it is built from
/root/reference, which only exists where the sources are mounted).
"""
import os
import tempfile

import numpy as np
import pytest

import oracle as O
from helpers import small_dump
from ipole_amd import h5

pytestmark = pytest.mark.skipif(not O.ref_available("iharm"), reason="oracle/_ref not built")

SPHERE = dict(R0=50., tau0=2.0143e-5, Thetae0=10., beta0=30.)
M_UNIT = 3e26


@pytest.fixture(scope="module")
def iharm_case():
    h, prims = small_dump()
    td = tempfile.mkdtemp()
    path = os.path.join(td, "dump.h5")
    h5.write_dump(path, h, prims)
    keys = dict(dump=path, M_unit=M_UNIT, MBH=6.2e9, trat_small=1, trat_large=40)
    orc, fields = O.iharm_oracle(h, prims, M_unit=M_UNIT)
    return h, prims, keys, orc, fields


def traces_equal(model, orc, keys, p, pix):
    orc.activate()
    cam = orc.camera(p)
    kk = dict(keys, nx=p.nx, ny=p.ny, thetacam=p.thetacam, freqcgs=p.freqcgs, dsource=p.dsource)
    if p.dx:
        kk.update(dx=p.dx, dy=p.dy)
    ref = O.ref_run(model, "initxk", np.array(pix, float), kk).reshape(-1, 12)
    for q, (i, j) in enumerate(pix):
        X, K = orc.init_XK(i, j, p, cam)
        assert np.array_equal(X, ref[q, 4:8]) and np.array_equal(K, ref[q, 8:12])
        ns, tr = orc.trace(X, K, p.eps, p.maxnstep, np.array(cam.Xcam))
        rr = O.ref_run(model, "trace", np.concatenate([X, K, np.array(cam.Xcam)])[None, :],
                       dict(keys, eps=p.eps, maxnstep=p.maxnstep))
        nsr = int(rr[0])
        assert ns == nsr
        assert np.array_equal(tr, rr[1:].reshape(nsr + 1, 18))


def test_sphere_traces():
    g, m = O.sphere_model(**SPHERE)
    orc = O.Oracle(g, m)
    p = O.default_params(nx=32, ny=32, thetacam=60., dx=100., dy=100., freqcgs=230e9)
    traces_equal("sphere", orc, SPHERE, p, [(0, 0), (5, 7), (16, 16), (31, 2)])


def test_analytic_traces_and_fourv():
    """Kerr a = 0.9 in MKS (hslope 1): the Gold et al. model 3 set-up."""
    keys = dict(model=3, MBH=4.063e6, freqcgs=230e9)
    g, m = O.analytic_model(3, MBH=4.063e6, freqcgs=230e9)
    orc = O.Oracle(g, m)
    p = O.default_params(nx=16, ny=16, thetacam=60., dx=30., dy=30., dsource=7778., freqcgs=230e9)
    traces_equal("analytic", orc, keys, p, [(3, 3), (8, 8), (15, 0)])
    rng = np.random.default_rng(1)
    X = np.zeros((200, 4))
    X[:, 1] = rng.uniform(0.5, 4., 200)
    X[:, 2] = rng.uniform(0.05, 0.95, 200)
    X[:, 3] = rng.uniform(0, 6, 200)
    K = rng.normal(size=(200, 4))
    ref = O.ref_run("analytic", "fluid", np.hstack([X, K]), keys).reshape(200, 21)
    for q in range(200):
        mine = np.concatenate(orc.fourv(X[q], K[q]))
        assert np.array_equal(mine, ref[q, :16])


def test_iharm_geometry(iharm_case):
    h, prims, keys, orc, fields = iharm_case
    gr, L_unit = O.ref_geom("iharm", keys)
    g = orc.geom
    for name in ("startx", "stopx", "dx", "cstartx", "cstopx"):
        assert np.array_equal(np.array(getattr(gr, name)), np.array(getattr(g, name))), name
    for name in ("a", "hslope", "poly_xt", "poly_alpha", "mks_smooth", "poly_norm", "Rin", "Rout", "Rh",
                 "rmin_geo", "rmax_geo", "N1", "N2", "N3", "metric"):
        assert getattr(gr, name) == getattr(g, name), name
    assert L_unit == orc.model.L_unit


def test_iharm_traces(iharm_case):
    h, prims, keys, orc, fields = iharm_case
    for th in (163., 60.):
        p = O.default_params(nx=24, ny=24, thetacam=th, freqcgs=230e9)
        traces_equal("iharm", orc, keys, p, [(3, 4), (12, 12), (20, 7)])


def test_iharm_fluid_and_derivation(iharm_case):
    """Four-vectors, n_e (with the sigma cut), Theta_e, B, sigma and the
    radiating-region test at random points inside and outside the grid;
    they read the fields each side derived from the same dump file, so this
    also pins the dump derivation (b, ghost zones, R_high/R_low electrons)."""
    h, prims, keys, orc, fields = iharm_case
    orc.activate()
    g = orc.geom
    rng = np.random.default_rng(0)
    n = 3000
    X = np.zeros((n, 4))
    X[:, 1] = rng.uniform(g.startx[1] - 0.05, g.stopx[1] + 0.05, n)
    X[:, 2] = rng.uniform(-0.02, 1.02, n)
    X[:, 3] = rng.uniform(-7, 13, n)
    K = rng.normal(size=(n, 4)) * 1e-9
    ref = O.ref_run("iharm", "fluid", np.hstack([X, K]), keys).reshape(n, 21)
    for q in range(n):
        mine = np.concatenate(orc.fourv(X[q], K[q]) + [np.array(orc.scalars(X[q]))])
        assert np.array_equal(mine, ref[q]), q


def test_metric_connection(iharm_case):
    h, prims, keys, orc, fields = iharm_case
    orc.activate()
    rng = np.random.default_rng(2)
    X = np.zeros((100, 4))
    X[:, 1] = rng.uniform(0.2, 6.5, 100)
    X[:, 2] = rng.uniform(0.01, 0.99, 100)
    ref = O.ref_run("iharm", "metric", X, keys).reshape(100, -1)
    for q in range(100):
        g = orc.gcov(X[q])
        assert np.array_equal(g.ravel(), ref[q, :16])
        assert np.array_equal(orc.gcon(g).ravel(), ref[q, 16:32])
        assert np.array_equal(orc.connection(X[q]).ravel(), ref[q, 32:96])
        assert orc.bl_coord(X[q]) == tuple(ref[q, 96:98])


def test_tetrads(iharm_case):
    h, prims, keys, orc, fields = iharm_case
    orc.activate()
    rng = np.random.default_rng(3)
    recs = []
    for q in range(200):
        X = np.array([0., rng.uniform(0.3, 5.), rng.uniform(0.05, 0.95), rng.uniform(0, 6)])
        U, K, B = rng.normal(size=4), rng.normal(size=4), rng.normal(size=4)
        U[0] = abs(U[0]) + 3.
        recs.append(np.concatenate([U, K, B, X]))
    recs = np.array(recs)
    ref = O.ref_run("iharm", "tetrad", recs, keys).reshape(200, 33)
    for q in range(200):
        U, K, B, X = recs[q, :4], recs[q, 4:8], recs[q, 8:12], recs[q, 12:]
        Ec, El, fl = orc.plasma_tetrad(U, K, B, orc.gcov(X))
        assert np.array_equal(Ec.ravel(), ref[q, :16]) and np.array_equal(El.ravel(), ref[q, 16:32])
        assert fl == ref[q, 32]
    Xc = recs[:20, 12:]
    ref = O.ref_run("iharm", "camtetrad", Xc, keys).reshape(20, 66)
    for q in range(20):
        Ec, El, fl = orc.camera_tetrad(Xc[q])
        assert np.array_equal(Ec.ravel(), ref[q, :16]) and np.array_equal(El.ravel(), ref[q, 16:32])
        Ec, El, fl = orc.camera_tetrad(Xc[q], old=True)
        assert np.array_equal(Ec.ravel(), ref[q, 33:49]) and np.array_equal(El.ravel(), ref[q, 49:65])


def test_push_polar_approximate_solve_project(iharm_case):
    h, prims, keys, orc, fields = iharm_case
    orc.activate()
    rng = np.random.default_rng(4)
    recs, cases = [], []
    for q in range(100):
        Xm = np.array([0., rng.uniform(0.3, 5.), rng.uniform(0.05, 0.95), rng.uniform(0, 6)])
        Km = rng.normal(size=4) * 1e-9
        Ni = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
        Nm = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
        dlam = rng.uniform(1., 100.)
        recs.append(np.concatenate([Xm, Km, Ni.view(np.float64).ravel(), Nm.view(np.float64).ravel(), [dlam]]))
        cases.append((Xm, Km, Ni, Nm, dlam))
    ref = O.ref_run("iharm", "push_polar", np.array(recs), keys).reshape(100, 32)
    for q, (Xm, Km, Ni, Nm, dlam) in enumerate(cases):
        Nf = orc.push_polar(Xm, Km, Ni, Nm, dlam)
        assert np.array_equal(Nf.view(np.float64).ravel(), ref[q])
    a = np.column_stack([rng.uniform(0, 1, 300), rng.uniform(0, 1, 300), 10 ** rng.uniform(-6, 1, 300),
                         rng.uniform(0, 1, 300), 10 ** rng.uniform(-6, 1, 300), 10 ** rng.uniform(-3, 2, 300)])
    ref = O.ref_run("iharm", "approx", a, keys).reshape(300, 2)
    for q in range(300):
        assert O.approximate_solve(*a[q]) == tuple(ref[q])
    recs, cases = [], []
    for q in range(50):
        X = np.array([0., np.log(1000.), rng.uniform(0.05, 0.95), 0.])
        N = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
        rot = rng.uniform(0, 1)
        recs.append(np.concatenate([X, N.view(np.float64).ravel(), [rot]]))
        cases.append((X, N, rot))
    ref = O.ref_run("iharm", "project", np.array(recs), keys).reshape(50, 4)
    for q, (X, N, rot) in enumerate(cases):
        assert np.array_equal(orc.project_N(X, N, rot), ref[q])
