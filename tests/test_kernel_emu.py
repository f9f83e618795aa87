"""The GPU library's per-ray code compiled for the CPU (tests/kernel_emu)
against the oracle: checks the kernel arithmetic (exact-derivative
connection, Hermitian coherency tensor, block Stokes update, Chebyshev K_n)
in a container without a GPU.  The same bars as the GPU parity tests."""
import numpy as np
import pytest
from scipy import special

import emu
import oracle as O
from helpers import small_dump, trace_compare
from ipole_amd import ipgpu as G


def test_metric_derivatives_vs_oracle_fd():
    """dg/dX1, dg/dX2 against the oracle's central differences (DEL = 1e-7,
    as src/geometry.c:91), g and g^-1 against the oracle's metric and LU."""
    h, _ = small_dump(N=(8, 8, 8))
    g = G.iharm_geom(h)
    go = O.iharm_geom(h)
    L = O.lib()
    L.or_set_geom(O.C.byref(go))
    orc = O.Oracle.__new__(O.Oracle)
    orc.L = L
    rng = np.random.default_rng(0)
    idx = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]
    for _ in range(300):
        X1, X2 = rng.uniform(0.1, 6.8), rng.uniform(0.02, 0.98)
        e = emu.metric(g, X1, X2)
        gc_ = orc.gcov([0, X1, X2, 0])
        gcon = orc.gcon(gc_)
        for n, (i, j) in enumerate(idx):
            scale = np.sqrt(abs(gc_[i, i] * gc_[j, j])) + 1e-300
            assert abs(e[2 + n] - gc_[i, j]) < 1e-14 * scale
            sc2 = np.sqrt(abs(gcon[i, i] * gcon[j, j]))
            assert abs(e[12 + n] - gcon[i, j]) < 1e-12 * sc2
        for k, off in ((1, 22), (2, 32)):
            hh = 1e-4
            G4 = []
            for s in (2, 1, -1, -2):
                Xs = np.array([0, X1, X2, 0.])
                Xs[k] += s * hh
                G4.append(orc.gcov(Xs))
            fd = (-G4[0] + 8 * G4[1] - 8 * G4[2] + G4[3]) / (12 * hh)   # O(h^4) stencil
            for n, (i, j) in enumerate(idx):
                scale = max(np.sqrt(abs(gc_[i, i] * gc_[j, j])), abs(fd[i, j]))
                assert abs(e[off + n] - fd[i, j]) < 1e-9 * scale, (k, i, j)


def test_connection_contractions_vs_oracle():
    """Gamma^i_jk K^j K^k and Gamma^i_kl K^l from the exact derivatives
    against the oracle's finite-difference connection."""
    h, _ = small_dump(N=(8, 8, 8))
    g = G.iharm_geom(h)
    orc = O.Oracle(O.iharm_geom(h), O.Model())
    rng = np.random.default_rng(1)
    for _ in range(200):
        X = np.array([0., rng.uniform(0.1, 6.8), rng.uniform(0.02, 0.98), rng.uniform(0, 6)])
        K = rng.normal(size=4)
        a, M = emu.accel(g, X, K)
        conn = orc.connection(X)
        a_ref = np.einsum("ijk,j,k->i", conn, K, K)
        M_ref = np.einsum("ikl,l->ik", conn, K)
        sa = np.max(np.abs(np.einsum("ijk,j,k->ijk", np.abs(conn), np.abs(K), np.abs(K)).sum(axis=(1, 2))))
        assert np.max(np.abs(a - a_ref)) < 1e-7 * sa
        sm = np.max(np.abs(conn).sum(axis=2) * np.abs(K).max())
        assert np.max(np.abs(M - M_ref)) < 1e-7 * sm


def test_device_bessel_vs_scipy():
    for x in np.concatenate([10 ** np.linspace(-3, 2.85, 300), [1.9999999, 2.0, 2.0000001]]):
        ks, emx = emu.bessel(x)
        for n in (0, 1, 2):
            ref = special.kve(n, x)
            assert abs(ks[n] / ref - 1) < 3e-14, (n, x)
        assert abs(emx / np.exp(-x) - 1) < 1e-15


def test_stokes_affine_vs_oracle_step():
    """The kernel applies evolve_N's plasma-frame step as one affine map
    S -> T S + cv formed from the coefficients; against the oracle's
    sequential Faraday / absorption-emission / Faraday step over every
    branch (rho^2 and aP^2 above and below 1e-80, optical depths either side
    of the 1e-5 series switch and of approximate_solve's 1e-3 switch)."""
    rng = np.random.default_rng(11)
    n = 0
    for rot in (0., 1e-45, 1.):
        for apol in (0., 1e-45, 0.5):
            for x in (1e-9, 3e-6, 2e-4, 0.05, 1.5, 30.):
                for _ in range(6):
                    c = np.zeros(11)
                    c[4] = rng.uniform(0.3, 2.)
                    a = rng.normal(size=3)
                    c[5:8] = a / np.linalg.norm(a) * apol * c[4] * rng.uniform(0.1, 0.95)
                    c[0] = rng.uniform(0.1, 1.)
                    jp = rng.normal(size=3)
                    c[1:4] = jp / np.linalg.norm(jp) * 0.6 * c[0]
                    c[8:11] = rng.normal(size=3) * rot
                    S0 = rng.normal(size=4)
                    S0[0] = abs(S0[0]) + 2.
                    ref = O.stokes_step(c, x, S0)
                    got = emu.stokes_affine(c, x, S0)
                    scale = max(np.abs(ref).max(), 1e-300)
                    assert np.abs(got - ref).max() <= 1e-13 * scale, (rot, apol, x, got, ref)
                    n += 1
    assert n == 3 * 3 * 6 * 6


def test_stokes_affine_f32_vs_oracle_step():
    """The mixed-precision step (stokes_affine_f32: Faraday rotation,
    absorption/emission solution and their products in FP32 from optical
    depths, angles and scaled emission) against the oracle's FP64 sequential
    step, over the branches of test_stokes_affine_vs_oracle_step and at the
    coefficient scales of a dump in cgs units (j ~ 1e-30..1e-40, alpha and
    rho ~ 1e-16..1e-12 per cm, steps of 1e9..1e14 cm, including coefficients
    just above the 1e-80 thresholds): within a few FP32 ulps of the state's
    scale."""
    rng = np.random.default_rng(12)
    n = 0
    worst = 0.
    cases = []
    for rot in (0., 1e-45, 1.):
        for apol in (0., 1e-45, 0.5):
            for x in (1e-9, 3e-6, 2e-4, 0.05, 1.5, 30.):
                cases.append((rot, apol, x, 1., 1.))
    for js in (1e-30, 1e-40):
        for ascale in (1e-16, 1e-12, 1.5e-40):
            for x in (1e9, 1e12, 1e14):
                cases.append((1., 0.5, x, js, ascale))
    for rot, apol, x, js, ascale in cases:
        for _ in range(6):
            c = np.zeros(11)
            c[4] = rng.uniform(0.3, 2.) * ascale
            a = rng.normal(size=3)
            c[5:8] = a / np.linalg.norm(a) * apol * c[4] * rng.uniform(0.1, 0.95)
            c[0] = rng.uniform(0.1, 1.) * js
            jp = rng.normal(size=3)
            c[1:4] = jp / np.linalg.norm(jp) * 0.6 * c[0]
            c[8:11] = rng.normal(size=3) * rot * ascale
            S0 = rng.normal(size=4) * js * x
            S0[0] = abs(S0[0]) + 2. * js * x
            ref = O.stokes_step(c, x, S0)
            got = emu.stokes_affine_f32(c, x, S0)
            scale = max(np.abs(ref).max(), np.abs(S0).max(), 1e-300)
            err = np.abs(got - ref).max() / scale
            assert np.all(np.isfinite(got)) and err <= 2e-6, (rot, apol, x, js, ascale, got, ref)
            worst = max(worst, err)
            n += 1
    assert n == 6 * len(cases) and worst > 0.
    print("max error / scale %.2e over %d steps" % (worst, n))


def image_errors(ref, out7, f3, scale, unpol=False):
    I_ref = ref["image"].ravel()
    I_got = out7[:, 0] * f3
    e_un = np.abs(I_got - I_ref) / np.maximum(I_ref, 1e-6 * I_ref.max())
    F_rel = abs(I_got.sum() / I_ref.sum() - 1)
    if unpol:
        return F_rel, {"unpol": e_un}
    S = ref["imageS"].reshape(-1, 5)
    den = np.maximum(np.abs(S[:, 0]), 1e-6 * np.abs(S[:, 0]).max())
    errs = {"unpol": e_un}
    for k, n in enumerate("IQUV"):
        sign = -1. if n in "QU" else 1.   # qu_conv 0: save_pixel negates Q, U
        errs[n] = np.abs(out7[:, k + 1] * f3 * sign - S[:, k]) / den
    Ft = abs((out7[:, 1] * f3).sum() / S[:, 0].sum() - 1)
    return max(F_rel, Ft), errs


@pytest.mark.parametrize("unpol", [0, 1])
def test_emu_sphere(unpol):
    go, mo = O.sphere_model(R0=50., tau0=2.0143e-5, Thetae0=10., beta0=30.)
    orc = O.Oracle(go, mo)
    p = O.default_params(nx=24, ny=24, thetacam=60., dx=100., dy=100., freqcgs=230e9, only_unpolarized=unpol)
    ref = orc.render(p)
    gs, ms = G.sphere_setup(R0=50., tau0=2.0143e-5, Thetae0=10., beta0=30.)
    im, scale, _ = G.make_image(gs, ms.L_unit, nx=24, ny=24, thetacam=60., dx=100., dy=100.,
                                freqcgs=230e9, only_unpolarized=unpol)
    rc, out7, ns, fl = emu.render(gs, ms, im, None)
    assert rc == 0 and np.array_equal(ns, ref["nstep"].ravel())
    F, errs = image_errors(ref, out7, 230e9 ** 3, scale, unpol=bool(unpol))
    assert F < 1e-6
    for k, e in errs.items():
        assert e.max() < 1e-6, (k, e.max())


@pytest.mark.parametrize("thetacam,etype,unpol", [(163., 4, 0), (60., 1, 0), (90., 4, 1)])
def test_emu_iharm(thetacam, etype, unpol):
    h, prims = small_dump()
    orc, fields = O.iharm_oracle(h, prims, M_unit=3e26)
    p = O.default_params(nx=20, ny=20, thetacam=thetacam, freqcgs=230e9, emission_type=etype,
                         only_unpolarized=unpol)
    ref = orc.render(p)
    g, m = G.iharm_geom(h), G.iharm_model()
    im, scale, _ = G.make_image(g, m.L_unit, nx=20, ny=20, thetacam=thetacam, freqcgs=230e9,
                                emission_type=etype, only_unpolarized=unpol)
    rc, out7, ns, fl = emu.render(g, m, im, emu.pack_dump(fields))
    assert rc == 0 and np.array_equal(ns, ref["nstep"].ravel())
    F, errs = image_errors(ref, out7, 230e9 ** 3, scale, unpol=bool(unpol))
    assert F < 1e-6
    for k, e in errs.items():
        assert e.max() < 1e-5, (k, e.max())
        assert np.mean(e > 1e-6) < 1e-3, (k, np.mean(e > 1e-6))


@pytest.mark.parametrize("opts", [dict(counterjet=1), dict(counterjet=2), dict(stokes_floors=1),
                                  dict(old_centering=1), dict(xoff=3.5, yoff=-2.25), dict(rotcam=25.)],
                         ids=lambda o: "-".join("%s=%g" % kv for kv in o.items()))
def test_emu_iharm_options(opts):
    """Image options (counterjet, Stokes floors, old centering, pixel offsets,
    camera rotation) through the device code built for the CPU, against the
    oracle; qu_conv acts in the output layout only (k_finalize), which the GPU
    tests cover."""
    h, prims = small_dump()
    orc, fields = O.iharm_oracle(h, prims, M_unit=3e26)
    p = O.default_params(nx=16, ny=16, thetacam=60., freqcgs=230e9, **opts)
    ref = orc.render(p)
    g, m = G.iharm_geom(h), G.iharm_model()
    kw = dict(opts)
    im, scale, _ = G.make_image(g, m.L_unit, nx=16, ny=16, thetacam=60., freqcgs=230e9, **kw)
    rc, out7, ns, fl = emu.render(g, m, im, emu.pack_dump(fields))
    assert rc == 0 and np.array_equal(ns, ref["nstep"].ravel())
    F, errs = image_errors(ref, out7, 230e9 ** 3, scale)
    assert F < 1e-6
    for k, e in errs.items():
        assert e.max() < 1e-5, (k, e.max())
        assert np.mean(e > 1e-6) < 1e-2, (k, np.mean(e > 1e-6))


def test_pipeline_log_pow_cbrt_vs_mpmath():
    """log_, cbrt_ and pow_ of ipg_math.h (the thermal fits' powers and
    logarithms) against mpmath: log within 4 ulp of max(|log x|, ln 2 / 2^10)
    (absolute near x = 1, where any log has only absolute accuracy in its
    argument's rounding), cbrt within 2 ulp, x^(-8/3) = e^(b log x) within
    (4 + 2 |b log x|) ulp: the rounding of b log x, |b log x| ulp absolute,
    carries into e^(b log x) as that much relative error (up to ~600 here;
    the fits use |b log x| of a few tens)."""
    import mpmath as mp
    mp.mp.dps = 50
    rng = np.random.default_rng(9)
    xs = np.concatenate([10 ** rng.uniform(-299, 299, 2000), rng.uniform(0.5, 2., 2000),
                         [1., 1. + 2.2e-16, 1. - 1.1e-16, 2., 0.5, np.sqrt(0.5), 1e-300, 1e300]])
    got = emu.math(4, xs)[:, 0]
    for x, g in zip(xs, got):
        ref = float(mp.log(mp.mpf(x)))
        assert abs(g - ref) <= 4 * 2.2e-16 * max(abs(ref), np.log(2) / 1024), (x, g, ref)
    xs2 = 10 ** rng.uniform(-200, 200, 2000)
    got = emu.math(5, xs2)[:, 0]
    for x, g in zip(xs2, got):
        ref = float(mp.cbrt(mp.mpf(x)))
        assert abs(g / ref - 1) <= 2 * 2.2e-16, (x, g, ref)
    xs3 = 10 ** rng.uniform(-100, 100, 2000)
    got = emu.math(6, xs3)[:, 0]
    for x, g in zip(xs3, got):
        ref = float(mp.power(mp.mpf(x), mp.mpf(-8) / 3))
        if ref == 0. or np.isinf(ref):
            continue
        assert abs(g / ref - 1) <= (4 + 2 * abs(8. / 3. * np.log(x))) * 2.2e-16, (x, g, ref)
    assert emu.math(4, [0.])[0, 0] == -np.inf and np.isnan(emu.math(4, [-1.])[0, 0])


def test_pipeline_math_vs_mpmath():
    """exp_, sincos_, sincospi_ of ipg_math.h over and beyond the argument
    ranges the loops use (X1 up to ~7, smoothing exponents, theta, 2 pi X2,
    optical depths), against mpmath at 60 digits: within 4 ulp relative
    (sin/cos: within 4 ulp of max(|f|, 2^-30) near their zeros)."""
    import mpmath as mp
    mp.mp.dps = 60
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-30, 30, 3000), rng.uniform(-700, 700, 500),
                         [0., 1e-300, -1e-20, 1e-8, 0.5 * np.log(2), -0.5 * np.log(2), 709.7, -745.1, 800.]])
    got = emu.math(0, xs)[:, 0]
    for x, g in zip(xs, got):
        ref = float(mp.exp(mp.mpf(x)))
        if ref == 0. or np.isinf(ref):
            assert g == ref
        else:
            assert abs(g / ref - 1) < 4 * 2.2e-16, (x, g, ref)
    xs = np.concatenate([rng.uniform(-8, 8, 3000), rng.uniform(-2e5, 2e5, 300),
                         np.arange(-16, 17) * np.pi / 4, [0., 1e-300, 1e-10]])
    out = emu.math(1, xs)
    for x, (s, c) in zip(xs, out):
        rs, rc = float(mp.sin(mp.mpf(x))), float(mp.cos(mp.mpf(x)))
        assert abs(s - rs) <= 4 * 2.2e-16 * max(abs(rs), 2. ** -30), (x, s, rs)
        assert abs(c - rc) <= 4 * 2.2e-16 * max(abs(rc), 2. ** -30), (x, c, rc)
    ys = np.concatenate([rng.uniform(-3, 3, 3000), np.arange(-12, 13) / 4., [1e-300, 0.999999999]])
    out = emu.math(2, ys)
    for y, (s, c) in zip(ys, out):
        rs, rc = float(mp.sinpi(mp.mpf(y))), float(mp.cospi(mp.mpf(y)))
        assert abs(s - rs) <= 4 * 2.2e-16 * max(abs(rs), 2. ** -30), (y, s, rs)
        assert abs(c - rc) <= 4 * 2.2e-16 * max(abs(rc), 2. ** -30), (y, c, rc)


@pytest.mark.parametrize("unpol", [0, 1])
def test_emu_batch_rescaling(unpol):
    """The batch path's arithmetic: geodesics traced at 86 GHz reused for a
    230 GHz job (K and dl rescaled), fused transfer, n_e and B rescaled for
    half the M_unit of the uploaded dump; against the oracle rendering that
    job directly (traced at 230 GHz, dump derived at M_unit / 2).  The
    rescaled geodesic differs from a direct one by roundoff only: nstep
    matches per pixel and the image meets the plain path's bar."""
    h, prims = small_dump()
    M = 3e26
    _, fields = O.iharm_oracle(h, prims, M_unit=M)
    orc2, _ = O.iharm_oracle(h, prims, M_unit=M * 0.5)
    p = O.default_params(nx=20, ny=20, thetacam=163., freqcgs=230e9, only_unpolarized=unpol)
    ref = orc2.render(p)
    g, m = G.iharm_geom(h), G.iharm_model()
    im_t, _, _ = G.make_image(g, m.L_unit, nx=20, ny=20, thetacam=163., freqcgs=86e9, only_unpolarized=unpol)
    im_j, scale, _ = G.make_image(g, m.L_unit, nx=20, ny=20, thetacam=163., freqcgs=230e9,
                                  only_unpolarized=unpol)
    rc, out7, ns, fl = emu.render_scaled(g, m, im_t, im_j, emu.pack_dump(fields), ne_scale=0.5)
    assert rc == 0 and np.array_equal(ns, ref["nstep"].ravel())
    F, errs = image_errors(ref, out7, 230e9 ** 3, scale, unpol=bool(unpol))
    assert F < 1e-6
    for k, e in errs.items():
        assert e.max() < 1e-5, (k, e.max())
        assert np.mean(e > 1e-6) < 1e-3, (k, np.mean(e > 1e-6))


@pytest.mark.parametrize("f32", [0, 1])
def test_emu_multi_job_transfer(f32):
    """Three jobs of one trace group in one lane (transfer_fused_multi: the
    geometry of a step once, the fits, Stokes step and coherency tensor per
    job) against each job's own fused transfer on its rescaled records: the
    job at the trace's frequency bit for bit, the others (86 GHz trace reused
    at 230 and 345 GHz, other M_unit scales, one unpolarized with the
    counterjet cut, one Pandya) to the roundoff of the rescaling; then two
    jobs, and one."""
    h, prims = small_dump()
    _, fields = O.iharm_oracle(h, prims, M_unit=3e26)
    d = emu.pack_dump(fields)
    g, m = G.iharm_geom(h), G.iharm_model()
    kw = dict(nx=16, ny=16, thetacam=163., transfer_fp32=f32)
    im_t = G.make_image(g, m.L_unit, freqcgs=86e9, **kw)[0]
    jobs = [G.make_image(g, m.L_unit, freqcgs=86e9, **kw)[0],
            G.make_image(g, m.L_unit, freqcgs=230e9, only_unpolarized=1, counterjet=1, **kw)[0],
            G.make_image(g, m.L_unit, freqcgs=345e9, emission_type=1, **kw)[0]]
    scales = [1.0, 0.4, 2.5]
    for nj in (3, 2, 1):
        rc, out7, ns, fl = emu.render_multi(g, m, im_t, jobs[:nj], scales[:nj], d)
        assert rc == 0
        for j in range(nj):
            rc1, ref7, ns1, fl1 = emu.render_scaled(g, m, im_t, jobs[j], d, ne_scale=scales[j])
            assert rc1 == 0 and np.array_equal(ns, ns1) and np.array_equal(fl[:, j], fl1), (nj, j)
            got = out7[:, j, :]
            if j == 0:
                assert np.array_equal(got, ref7), (nj, j)
            else:
                scale = np.abs(ref7[:, :5]).max(axis=0)
                err = np.abs(got[:, :5] - ref7[:, :5]).max(axis=0) / np.where(scale > 0, scale, 1.)
                assert err.max() < 1e-11, (nj, j, err)
                assert np.allclose(got[:, 5:], ref7[:, 5:], rtol=1e-11, atol=0.), (nj, j)


def image_summary(I, Q, U, V):
    """Ftot-proportional sum, net LP and CP fractions and EVPA (rad) of an
    image, as print_image_stats forms them (src/main.c:1001-1042)."""
    sI, sQ, sU, sV = I.sum(), Q.sum(), U.sum(), V.sum()
    return sI, np.hypot(sQ, sU) / sI, sV / sI, 0.5 * np.arctan2(sU, sQ)


@pytest.mark.parametrize("thetacam,etype", [(163., 4), (60., 1)])
def test_emu_fp32_transfer(thetacam, etype):
    """FP32 coefficient shape functions and plasma-frame Stokes step (FP64
    geodesics, connection, tetrads and N), split and fused forms (both
    gather the FP32 copy of the cell records, as the GPU's mixed-precision
    kernels do), against the FP64 oracle: Ftot, LP, CP and EVPA well inside
    the 1e-4 bar of the mixed-precision mode (measured ~1e-7 here)."""
    h, prims = small_dump()
    orc, fields = O.iharm_oracle(h, prims, M_unit=3e26)
    p = O.default_params(nx=20, ny=20, thetacam=thetacam, freqcgs=230e9, emission_type=etype)
    ref = orc.render(p)
    S = ref["imageS"].reshape(-1, 5)
    rI, rLP, rCP, rEV = image_summary(S[:, 0], S[:, 1], S[:, 2], S[:, 3])
    g, m = G.iharm_geom(h), G.iharm_model()
    im, scale, _ = G.make_image(g, m.L_unit, nx=20, ny=20, thetacam=thetacam, freqcgs=230e9,
                                emission_type=etype, transfer_fp32=1)
    d = emu.pack_dump(fields)
    for rc, out7, ns, fl in (emu.render(g, m, im, d), emu.render_scaled(g, m, im, im, d)):
        assert rc == 0 and np.array_equal(ns, ref["nstep"].ravel())
        f3 = 230e9 ** 3
        gI, gLP, gCP, gEV = image_summary(out7[:, 1] * f3, -out7[:, 2] * f3, -out7[:, 3] * f3, out7[:, 4] * f3)
        errs = dict(Ftot=abs(gI / rI - 1), LP=abs(gLP - rLP), CP=abs(gCP - rCP), EVPA=abs(gEV - rEV),
                    Funpol=abs((out7[:, 0] * f3).sum() / ref["image"].sum() - 1))
        print(errs)
        assert errs["Ftot"] < 1e-5 and errs["Funpol"] < 1e-5
        assert errs["LP"] < 1e-5 and errs["CP"] < 1e-5 and errs["EVPA"] < 1e-5


def test_emu_adjoint_vs_finite_differences():
    """Reverse-mode gradient of the L2 Stokes-image loss (the GPU kernel's
    code on the CPU) against central finite differences of the same loss:
    d/dln(M_unit) and d/dR_high, d/dR_low through a re-derivation of the dump,
    and d/d(n_e scale) of single cells through their n_e.  The loss is
    strongly curved in ln M_unit (optical depth), so the differences use
    h = 1e-5, whose truncation error is ~4e-6 here (it falls as h^2: 4e-4 at
    h = 1e-4); bar 1e-4 relative (the adjoint's gradient-check bar is 1e-3)."""
    h, prims = small_dump(N=(32, 24, 24), seed=5)
    M, Rh, Rl = 3e26, 40., 1.
    g, m = G.iharm_geom(h), G.iharm_model()
    im, _, _ = G.make_image(g, m.L_unit, nx=12, ny=12, thetacam=60., freqcgs=230e9)

    def dump_at(Mu, rh, rl):
        _, f11 = O.iharm_oracle(h, prims, M_unit=Mu, trat_large=rh, trat_small=rl)
        return emu.pack_dump_adjoint(f11, G.iharm_units(Mu, trat_large=rh, trat_small=rl))

    _, _, St = emu.adjoint(g, m, im, dump_at(1.3 * M, Rh, Rl), np.zeros(5 * 144))
    target = np.zeros((144, 5))
    target[:, :4] = St
    w = np.array([1., 3., 3., 5.])
    d0 = dump_at(M, Rh, Rl)
    out, gcell, S = emu.adjoint(g, m, im, d0, target.ravel(), w)
    scale = out[0]
    assert scale > 0

    def loss(d):
        return emu.adjoint(g, m, im, d, target.ravel(), w)[0][0]

    hM = 1e-5
    fd_M = (loss(dump_at(M * np.exp(hM), Rh, Rl)) - loss(dump_at(M * np.exp(-hM), Rh, Rl))) / (2 * hM)
    fd_Rh = (loss(dump_at(M, Rh * (1 + hM), Rl)) - loss(dump_at(M, Rh * (1 - hM), Rl))) / (2 * hM * Rh)
    fd_Rl = (loss(dump_at(M, Rh, Rl * (1 + hM))) - loss(dump_at(M, Rh, Rl * (1 - hM)))) / (2 * hM * Rl)
    print("lnM", out[1], fd_M, "Rh", out[2], fd_Rh, "Rl", out[3], fd_Rl)
    # the reverse pass in segments (the GPU's path) gives the same gradient
    # to roundoff, whatever the segment length
    for L in (512, 64, 7):
        o2, g2, S2 = emu.adjoint(g, m, im, d0, target.ravel(), w, seg_len=L)
        assert np.allclose(S2, S, rtol=1e-10, atol=1e-12 * np.abs(S).max()), L
        for q in range(4):
            assert abs(o2[q] / out[q] - 1) < 1e-8, (L, q, o2[q], out[q])
        assert np.allclose(g2, gcell, rtol=1e-7, atol=1e-9 * np.abs(gcell).max()), L
    assert abs(out[1] / fd_M - 1) < 1e-4
    assert abs(out[2] / fd_Rh - 1) < 1e-4
    assert abs(out[3] / fd_Rl - 1) < 1e-4
    top = np.argsort(-np.abs(gcell))[:4]
    for c in top:
        dp, dm = d0.copy(), d0.copy()
        dp[16 * c + 6] *= 1 + hM
        dm[16 * c + 6] *= 1 - hM
        fd = (loss(dp) - loss(dm)) / (2 * hM)
        print("cell", c, gcell[c], fd)
        assert abs(gcell[c] / fd - 1) < 1e-4


@pytest.mark.parametrize("seg_len", [256, 37, 1])
def test_emu_segmented_chain(seg_len):
    """The composed-segment chain (used on the GPU for long rays) for every
    ray of an image: against the sequential chain of the same packets to
    roundoff (N, I, tau, tauF; the unpolarized map regroups approximate_solve
    and carries the cached emissivity across segments), and against the
    oracle at the plain bar.  Segment lengths down to 1 step put the cache
    carry at every segment boundary."""
    h, prims = small_dump()
    orc, fields = O.iharm_oracle(h, prims, M_unit=3e26)
    p = O.default_params(nx=16, ny=16, thetacam=60., freqcgs=230e9)
    ref = orc.render(p)
    g, m = G.iharm_geom(h), G.iharm_model()
    im, scale, _ = G.make_image(g, m.L_unit, nx=16, ny=16, thetacam=60., freqcgs=230e9)
    d = emu.pack_dump(fields)
    rc, seq, ns, fl = emu.render(g, m, im, d)
    rc2, seg, ns2, fl2 = emu.render_segmented(g, m, im, d, seg_len=seg_len)
    assert rc == 0 and rc2 == 0 and np.array_equal(ns, ns2) and np.array_equal(fl, fl2)
    assert ns.max() > 256   # some rays span several segments
    den = np.maximum(np.abs(seq[:, 1]), 1e-6 * np.abs(seq[:, 1]).max())
    assert (np.abs(seg[:, 1:5] - seq[:, 1:5]) / den[:, None]).max() < 1e-10
    den0 = np.maximum(np.abs(seq[:, 0]), 1e-6 * np.abs(seq[:, 0]).max())
    assert (np.abs(seg[:, 0] - seq[:, 0]) / den0).max() < 1e-10
    for c in (5, 6):
        assert np.allclose(seg[:, c], seq[:, c], rtol=1e-12, atol=1e-14 * np.abs(seq[:, c]).max())
    F, errs = image_errors(ref, seg, 230e9 ** 3, scale)
    assert F < 1e-6
    for k, e in errs.items():
        assert e.max() < 1e-5, (k, e.max())


@pytest.mark.parametrize("budget", [1, 50, 200])
def test_emu_trace_two_pass_bit_identical(budget):
    """A trajectory traced for `budget` steps and resumed from its last
    record (the streamed render) is bit-identical to the one-pass trace:
    same step count, same records."""
    h, _ = small_dump()
    g, m = G.iharm_geom(h), G.iharm_model()
    im, _, _ = G.make_image(g, m.L_unit, nx=16, ny=16, thetacam=60., freqcgs=230e9)
    paused_any = False
    for i, j in [(0, 0), (3, 11), (8, 8), (15, 2), (7, 9)]:
        ns, rec = emu.trace(g, im, i, j)
        ns2, rec2, paused = emu.trace_two_pass(g, im, i, j, budget)
        paused_any |= paused
        assert paused == (ns > budget)
        assert ns2 == ns
        assert np.array_equal(rec, rec2)
    assert paused_any


@pytest.mark.parametrize("cam", [dict(thetacam=60.), dict(thetacam=163., target_nturns=0),
                                 dict(thetacam=90., only_unpolarized=1)])
def test_emu_trace_pixel(cam):
    """The trace output's per-step records (trace_pixel_records, as
    k_trace_pixel runs it) against the oracle's or_trace_pixel, with the
    dump's rho, u and beta where the device derivation puts them."""
    h, prims = small_dump()
    orc, fields = O.iharm_oracle(h, prims, M_unit=3e26)
    p = O.default_params(nx=16, ny=16, freqcgs=230e9, **cam)
    oc = orc.camera(p)
    g, m = G.iharm_geom(h), G.iharm_model()
    im, _, _ = G.make_image(g, m.L_unit, nx=16, ny=16, freqcgs=230e9, **cam)
    nc = fields[0].size
    dump = emu.pack_dump_trace(fields, orc.rub[:2 * nc])
    for ij in [(8, 7), (3, 12), (15, 0)]:
        rc, ns, ref = orc.trace_pixel(p, oc, *ij)
        nse, got = emu.trace_pixel(g, m, im, dump, *ij)
        assert rc == 0 and nse == ns, (ij, nse, ns)
        trace_compare(ref, got, ij)
        nse2, got2 = emu.trace_pixel(g, m, im, dump, *ij, prims_ok=False)
        f2 = G.trace_fields(got2)
        assert not np.any(f2["prims"][:, :2]) and not np.any(f2["beta"])
        assert np.array_equal(np.delete(got2, [0, 1, 14], axis=1), np.delete(got, [0, 1, 14], axis=1))


def test_emu_trace_pixel_sphere():
    kw = dict(R0=50., tau0=2.0143e-5, Thetae0=10., beta0=30.)
    go, mo = O.sphere_model(**kw)
    orc = O.Oracle(go, mo)
    p = O.default_params(nx=16, ny=16, thetacam=60., dx=100., dy=100., freqcgs=230e9)
    oc = orc.camera(p)
    gs, ms = G.sphere_setup(**kw)
    im, _, _ = G.make_image(gs, ms.L_unit, nx=16, ny=16, thetacam=60., dx=100., dy=100., freqcgs=230e9)
    for ij in [(8, 8), (2, 9)]:
        rc, ns, ref = orc.trace_pixel(p, oc, *ij)
        nse, got = emu.trace_pixel(gs, ms, im, None, *ij)
        assert rc == 0 and nse == ns
        trace_compare(ref, got, ij)
