"""The oracle's trace output (or_trace_pixel: src/io.c:279-543, the per-step
record get_pixel writes with params.trace, src/main.c:865-880) pinned by
what fixes each field independently of the restatement:

* against the reference's own functions (oracle/_ref, where built): the
  primitives, sigma, beta, fluid frequency and field angle at every traced
  point (get_model_primitives, get_model_sigma, get_model_beta,
  get_fluid_nu, get_bk_angle), and e2Transported through two calls of
  parallel_transport_vector on the trajectory's own half-step points;
* against other oracle paths that are pinned elsewhere: X, Kcon and dl equal
  the geodesic's records (or_trace), the coefficients equal or_jar at the
  recorded point, ncoord equals N of or_integrate_emission over the steps
  before the point, r and th equal bl_coord of X;
* by identities: Ecov Econ^T = 1 (the plasma tetrad with its time row
  negated), ntetrad = Ecov ncoord Ecov^T, the Stokes I, Q, U of ntetrad,
  e2TransportedCov = g e2Transported, |e2Transported| = 1 to second order in
  the step, and I_unpol = the one-step solution from zero with the
  coefficients and step length the record holds next to it.

The host writer of the trace file (host/h5io.c) is checked by a round trip.
"""
import numpy as np
import pytest

import oracle as O
from helpers import small_dump
from ipole_amd import h5
from ipole_amd import ipgpu as G

M_UNIT = 3e26
F = {name: (off, shape) for name, off, shape in G.TRACE_SETS}


def radiating(orc, X):
    return orc.L.or_radiating_region(O.P(O.f64(X))) != 0


def field(rec, name):
    off, shape = F[name]
    w = int(np.prod(shape)) if shape else 1
    a = rec[:, off:off + w]
    return a.reshape((rec.shape[0],) + shape) if shape else a[:, 0]


@pytest.fixture(scope="module")
def case():
    h, prims = small_dump()
    orc, _ = O.iharm_oracle(h, prims, M_unit=M_UNIT)
    p = O.default_params(nx=16, ny=16, thetacam=60., freqcgs=230e9)
    cam = orc.camera(p)
    out = {}
    for ij in [(8, 7), (3, 12), (0, 0)]:
        rc, ns, rec = orc.trace_pixel(p, cam, *ij)
        assert rc == 0
        X, K = orc.init_XK(ij[0], ij[1], p, cam)
        ns2, tr = orc.trace(X, K, p.eps, p.maxnstep, np.array(cam.Xcam))
        assert ns == ns2
        out[ij] = (ns, rec, tr)
    return h, prims, orc, p, cam, out


def test_layout_matches_header():
    """Field offsets of the Python table equal the C header's macros."""
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(O.HERE), "include", "ipgpu.h")).read()
    macros = dict((m.group(1), int(m.group(2))) for m in re.finditer(r"#define IPGPU_TR_(\w+) (\d+)", hdr))
    names = {"prims": "PRIMS", "b": "B", "ne": "NE", "thetae": "THETAE", "nu": "NU", "b_k_angle": "BK_ANGLE",
             "sigma": "SIGMA", "beta": "BETA", "dl": "DL", "r": "R", "th": "TH", "phi": "PHI",
             "j_unpol": "J_UNPOL", "k_unpol": "K_UNPOL", "I_unpol": "I_UNPOL", "X": "X", "Kcon": "KCON",
             "j_inv": "J_INV", "alpha_inv": "ALPHA_INV", "rho_inv": "RHO_INV", "stokes": "STOKES",
             "e2Constructed": "E2C", "e2ConstructedCov": "E2C_COV", "e2Transported": "E2T",
             "e2TransportedCov": "E2T_COV", "ntetrad": "NTETRAD", "ncoord": "NCOORD", "Econ": "ECON",
             "Ecov": "ECOV"}
    for name, (off, shape) in F.items():
        assert macros[names[name]] == off, name
    ends = sorted(off + (int(np.prod(s)) if s else 1) for off, s in F.values())
    assert ends[-1] == G.TRACE_NF == int(re.search(r"#define IPGPU_TRACE_NF (\d+)", hdr).group(1))
    offs = sorted(F.values())
    for (o1, s1), (o2, _) in zip(offs, offs[1:]):
        assert o1 + (int(np.prod(s1)) if s1 else 1) == o2          # no gaps, no overlaps


def test_geodesic_fields(case):
    h, prims, orc, p, cam, out = case
    for ij, (ns, rec, tr) in out.items():
        s = np.arange(1, ns)
        assert np.array_equal(field(rec, "X")[s], tr[s, 2:6]), ij
        assert np.array_equal(field(rec, "Kcon")[s], tr[s, 6:10]), ij
        assert np.array_equal(field(rec, "dl")[s], tr[s, 1]), ij
        for z in (0, ns):                       # the reference writes 1 .. nstep - 1 only
            assert not np.any(field(rec, "X")[z]) and not np.any(field(rec, "Kcon")[z])
        assert np.array_equal(field(rec, "phi")[s], tr[s, 5])
        for q in s[:: max(1, len(s) // 50)]:
            r, th = orc.bl_coord(tr[q, 2:6])
            assert field(rec, "r")[q] == r and field(rec, "th")[q] == th


def test_coefficients_and_unpolarized(case):
    h, prims, orc, p, cam, out = case
    for ij, (ns, rec, tr) in out.items():
        ju, ku, Iu, dl = (field(rec, k) for k in ("j_unpol", "k_unpol", "I_unpol", "dl"))
        nrad = 0
        for q in range(1, ns):
            X, K = tr[q, 2:6], tr[q, 6:10]
            c = orc.jar(p.emission_type, 1, X, K)
            assert np.array_equal(field(rec, "j_inv")[q], c[0:4])
            assert np.array_equal(field(rec, "alpha_inv")[q], c[4:8])
            assert field(rec, "rho_inv")[q, 0] == 0. and np.array_equal(field(rec, "rho_inv")[q, 1:], c[8:11])
            cu = orc.jar(p.emission_type, 0, X, K)
            assert ju[q] == cu[0] and ku[q] == cu[4]
            # I_unpol: one step from the point after, from I = 0 (integrate_emission
            # restarts I for every call, src/ipolarray.c:64-69)
            if q + 1 <= ns - 1 and radiating(orc, X):
                want = O.approximate_solve(0., ju[q + 1], ku[q + 1], ju[q], ku[q], dl[q + 1])[0]
                assert Iu[q] == pytest.approx(want, rel=1e-14, abs=0), (ij, q)
                nrad += Iu[q] > 0
            elif not radiating(orc, X):
                assert Iu[q] == 0.
        if ij != (0, 0):
            assert nrad > 10, ij                 # the pin is exercised


def test_tetrad_identities(case):
    h, prims, orc, p, cam, out = case
    for ij, (ns, rec, tr) in out.items():
        Econ, Ecov = field(rec, "Econ"), field(rec, "Ecov")
        Nc, Nt, S = field(rec, "ncoord"), field(rec, "ntetrad"), field(rec, "stokes")
        for q in range(1, ns):
            assert np.allclose(Ecov[q] @ Econ[q].T, np.eye(4), atol=1e-9), (ij, q)
            assert np.array_equal(field(rec, "e2Constructed")[q], Econ[q, 2])
            assert np.array_equal(field(rec, "e2ConstructedCov")[q], Ecov[q, 2])
            T = Ecov[q] @ Nc[q] @ Ecov[q].T
            sc = max(np.abs(T).max(), 1e-300)
            assert np.allclose(Nt[q], T, rtol=1e-10, atol=1e-12 * sc), (ij, q)
            assert S[q, 0] == pytest.approx((Nt[q, 1, 1] + Nt[q, 2, 2]) / 2, rel=1e-12, abs=1e-14 * sc)
            assert S[q, 1] == pytest.approx((Nt[q, 1, 1] - Nt[q, 2, 2]) / 2, rel=1e-12, abs=1e-14 * sc)
            assert S[q, 2] == pytest.approx((Nt[q, 1, 2] + Nt[q, 2, 1]) / 2, rel=1e-12, abs=1e-14 * sc)


def test_transported_e2(case):
    h, prims, orc, p, cam, out = case
    conv = orc.model.L_unit * O.HPL / (O.ME * O.CL * O.CL)
    for ij, (ns, rec, tr) in out.items():
        e2t, e2tc = field(rec, "e2Transported"), field(rec, "e2TransportedCov")
        assert not np.any(e2t[ns - 1])          # nothing is carried into the last point
        for q in range(1, ns):
            g = orc.gcov(tr[q, 2:6])
            assert np.allclose(e2tc[q], g @ e2t[q], rtol=1e-12, atol=1e-14 * np.abs(e2t[q]).max(initial=1.))
        for q in range(0, ns - 1):
            # e2 of point q + 1 carried back over the step into it: norm 1 to O(dl^2)
            g = orc.gcov(tr[q, 2:6])
            n = e2t[q] @ g @ e2t[q]
            K = tr[q + 1, 6:10]
            h2 = (tr[q + 1, 1] / conv) ** 2 * (K @ K + 1.)
            assert abs(n - 1.) < 50. * h2 + 1e-9, (ij, q, n)


def test_ncoord_is_transfer_state(case):
    """ncoord at point s is N after integrate_emission over steps nstep .. s + 2
    -> s + 1 (the render's own loop, stopped early)."""
    h, prims, orc, p, cam, out = case
    for ij, (ns, rec, tr) in out.items():
        Nc = field(rec, "ncoord")
        for s in sorted({1, 2, ns // 3, ns // 2, ns - 2}):
            if s < 1 or s > ns - 2:
                continue
            _, _, _, N, _ = orc.integrate_emission(tr[s + 1:ns + 1], ns - s - 1, p)
            assert np.array_equal(Nc[s], N.real), (ij, s)


@pytest.mark.skipif(not O.ref_available("iharm"), reason="oracle/_ref not built")
def test_against_reference_functions(case, tmp_path):
    h, prims, orc, p, cam, out = case
    dump = str(tmp_path / "dump.h5")
    h5.write_dump(dump, h, prims)
    keys = dict(dump=dump, M_unit=M_UNIT, MBH=6.2e9, trat_small=1, trat_large=40)
    for ij, (ns, rec, tr) in out.items():
        s = np.arange(1, ns)
        ref = O.ref_run("iharm", "trace_fields", np.hstack([tr[s, 2:6], tr[s, 6:10]]), keys).reshape(-1, 12)
        assert np.array_equal(field(rec, "prims")[s], ref[:, 0:8]), ij
        assert np.array_equal(field(rec, "sigma")[s], ref[:, 8]), ij
        assert np.array_equal(field(rec, "beta")[s], ref[:, 9]), ij
        assert np.array_equal(field(rec, "nu")[s], ref[:, 10]), ij
        assert np.array_equal(field(rec, "b_k_angle")[s], ref[:, 11]), ij
        # e2Transported[q - 1] = ptvec at (Xhalf_q, Kh_q) of ptvec at (X_q, K_q)
        E2 = field(rec, "e2Constructed")
        q = s
        r1 = np.hstack([tr[q, 2:6], tr[q, 6:10], E2[q], E2[q], tr[q, 1:2]])
        mid = O.ref_run("iharm", "ptvec", r1, keys).reshape(-1, 4)
        r2 = np.hstack([tr[q, 10:14], tr[q, 14:18], E2[q], mid, tr[q, 1:2]])
        fin = O.ref_run("iharm", "ptvec", r2, keys).reshape(-1, 4)
        assert np.array_equal(field(rec, "e2Transported")[q - 1], fin), ij


@pytest.mark.skipif(not O.ref_available("sphere"), reason="oracle/_ref not built")
def test_sphere_against_reference():
    kw = dict(R0=50., tau0=2.0143e-5, Thetae0=10., beta0=30.)
    g, m = O.sphere_model(**kw)
    orc = O.Oracle(g, m)
    p = O.default_params(nx=16, ny=16, thetacam=60., dx=100., dy=100., freqcgs=230e9)
    cam = orc.camera(p)
    rc, ns, rec = orc.trace_pixel(p, cam, 8, 8)
    assert rc == 0
    X, K = field(rec, "X")[1:ns], field(rec, "Kcon")[1:ns]
    ref = O.ref_run("sphere", "trace_fields", np.hstack([X, K]), kw).reshape(-1, 12)
    assert np.array_equal(field(rec, "prims")[1:ns], ref[:, 0:8])
    assert np.array_equal(field(rec, "sigma")[1:ns], ref[:, 8])
    assert np.array_equal(field(rec, "beta")[1:ns], ref[:, 9])
    assert np.any(ref[:, 9] > 0) and np.any(ref[:, 9] == 0)   # inside and outside R0
    assert np.array_equal(field(rec, "nu")[1:ns], ref[:, 10])
    assert np.array_equal(field(rec, "b_k_angle")[1:ns], ref[:, 11])


def test_trace_file_round_trip(tmp_path):
    """host/h5io.c: datasets of nx x ny x maxnstep (x 4, 8, 4 x 4), pixel
    hyperslabs, zeros elsewhere, append to an existing file."""
    path = str(tmp_path / "trace.h5")
    nx, ny, maxn = 2, 3, 250
    rng = np.random.default_rng(1)
    recs = {(0, 1): (40, rng.normal(size=(41, G.TRACE_NF))), (1, 2): (230, rng.normal(size=(231, G.TRACE_NF)))}
    meta = h5.ImageMeta()
    meta.nx, meta.ny, meta.freqcgs, meta.model_kind = 32, 48, 230e9, 2
    h5.write_trace(path, meta, nx, ny, maxn, [(0, 1) + (recs[(0, 1)][0], recs[(0, 1)][1])])
    h5.write_trace(path, meta, nx, ny, maxn, [(1, 2) + (recs[(1, 2)][0], recs[(1, 2)][1])])   # appends
    nst = h5.read_dataset(path, "nstep")
    assert nst.shape == (nx, ny) and nst[0, 1] == 40 and nst[1, 2] == 230 and nst.sum() == 270
    assert h5.read_dataset(path, "/header/camera/nx") == 32
    assert h5.read_dataset(path, "/header/freqcgs") == 230e9
    for name, off, shape in G.TRACE_SETS:
        d = h5.read_dataset(path, name)
        assert d.shape == (nx, ny, maxn) + shape, name
        w = int(np.prod(shape)) if shape else 1
        for (i, j), (ns, r) in recs.items():
            assert np.array_equal(d[i, j, :ns + 1].reshape(ns + 1, w), r[:, off:off + w]), name
            assert not np.any(d[i, j, ns + 1:])
        assert not np.any(d[0, 0]) and not np.any(d[1, 0])
