"""Pins of the oracle's thermal synchrotron fits (oracle/or_radiation.c, which
restates src/symphony/maxwell_juettner_fits.c:27-395 and the symphony wrappers
of src/symphony/fits.c) that do not rest on the fits themselves.

* Emissivities j_I, j_Q, j_V (Dexter, emission_type 4, and Pandya, emission_type
  1) against the exact thermal integrals of the ultrarelativistic single-electron
  kernels over a Maxwell-Juettner distribution:
      j_S = n sqrt(3) e^3 B sin(th) / (4 pi m c^2) int N(g) k_S(g, nu / nu_c(g)) dg,
      nu_c = 3/2 g^2 nu_B sin(th),
      k_I = F(x) = x int_x^inf K_5/3,   k_Q = G(x) = x K_2/3(x),
      k_V = (4 cot(th) / (3 g)) H(x),   H(x) = x K_1/3(x) + int_x^inf K_1/3.
  The kernels are derived here, not assumed: test_kernels_from_helical_orbit
  integrates the radiation amplitudes of a helical orbit (pitch angle alpha,
  observer at th = alpha + psi, Airy-function pulses) over an isotropic
  distribution of pitch angles and finds Q/I = G/F and V/I = (4/3) cot(th)/g
  H/F.  The V kernel comes from the first-order terms in psi: the pitch-angle
  dependence of the pulse and the t^2 term of the perpendicular velocity along
  the small circle the velocity traces in direction space.
* Faraday rotation and conversion at Theta_e >= 3 in the high-frequency limit
  (X -> 0), where they reduce to the relativistic thermal closed forms
  rho_V = rho_V,cold K_0(1/Theta)/K_2(1/Theta) (Quataert & Gruzinov 2000;
  Shcherbakov 2008) and rho_Q = rho_Q,cold (K_1/K_2 + 6 Theta) (Shcherbakov
  2008), with K_n from mpmath.
* The finite-X shapes of rho_Q's f(X) and rho_V's g(X) are fits (Shcherbakov
  2008, Dexter 2016) to numerical results that have no closed form and no table
  in the reference or offline; test_fit_bodies_match_reference_source pins the
  oracle's fit functions to the reference's own source text instead: every
  numeric literal and every arithmetic operator, in order, in each of the 12 fit
  functions.  (The reference build cannot compile these functions: they call
  GSL, which is not in this image.)
"""
import os
import re

import numpy as np
import pytest
from scipy import integrate, special
from scipy.interpolate import CubicSpline

import oracle as O
from test_oracle_math import coeffs_at, sphere_point_oracle

EE, CL, ME = O.EE, O.CL, O.ME
REF_FITS = "/root/reference/src/symphony/maxwell_juettner_fits.c"


# ------------------------------------------------------- synchrotron kernels

def _tail_int(nu, y):
    """int_y^inf K_nu(t) dt, as int K_nu(e^s) e^s ds (smooth in s)."""
    f = lambda s: special.kv(nu, np.exp(s)) * np.exp(s)
    lo = np.log(y)
    mid = max(lo, np.log(40.))
    a = integrate.quad(f, lo, mid, limit=400, epsabs=0, epsrel=1e-12)[0] if mid > lo else 0.
    return a + integrate.quad(f, mid, np.log(900.), limit=400, epsabs=0, epsrel=1e-12)[0]


def _tables():
    """F, G, H on a log grid, splined in log-log (relative error < 1e-7)."""
    y = np.logspace(-9, np.log10(250.), 1500)
    i13 = np.array([_tail_int(1. / 3., v) for v in y])
    k13, k23 = special.kv(1. / 3., y), special.kv(2. / 3., y)
    F = y * (2 * k23 - i13)          # int_y^inf K_5/3 = 2 K_2/3(y) - int_y^inf K_1/3
    G = y * k23
    H = y * k13 + i13
    ly = np.log(y)
    return {k: CubicSpline(ly, np.log(v)) for k, v in (("I", F), ("Q", G), ("V", H))}


_TAB = None


def kernel(kind, y):
    global _TAB
    if _TAB is None:
        _TAB = _tables()
    y = np.asarray(y, dtype=float)
    out = np.exp(_TAB[kind](np.log(np.clip(y, 1e-9, 250.))))
    return np.where(y > 250., 0., out)


def test_kernel_tables_vs_direct():
    for y in (1e-6, 0.03, 0.7, 4., 40.):
        F = y * _tail_int(5. / 3., y)
        G = y * special.kv(2. / 3., y)
        H = y * special.kv(1. / 3., y) + _tail_int(1. / 3., y)
        for k, v in (("I", F), ("Q", G), ("V", H)):
            assert abs(kernel(k, y) / v - 1) < 1e-6, (k, y)


def exact_thermal(kind, n, B, Th, theta, nu):
    """Exact thermal emissivity of Stokes I, Q or V (erg/s/cm^3/Hz/sr),
    ultrarelativistic kernels, Maxwell-Juettner N(g) = g^2 beta e^(-g/Th)/(Th K_2)."""
    nuB = EE * B / (2 * np.pi * ME * CL)
    K2 = special.kn(2, 1. / Th)
    xs = nu / (1.5 * nuB * np.sin(theta) * Th * Th)

    def integrand(g):
        beta = np.sqrt(1 - 1 / (g * g))
        N = g * g * beta * np.exp(-g / Th) / (Th * K2)
        k = kernel(kind, nu / (1.5 * g * g * nuB * np.sin(theta)))
        if kind == "V":
            k = k * 4. / (3. * g * np.tan(theta))
        return N * k

    zpk = (2 * xs) ** (1. / 3.)            # peak of z^2 e^-z k(xs/z^2), z = g/Th
    gmax = Th * (80. + 10. * zpk)
    val = integrate.quad(integrand, 1. + 1e-9, gmax, limit=600,
                         points=[Th, 3 * Th, 10 * Th, Th * zpk])[0]
    return n * np.sqrt(3) * EE ** 3 * B * np.sin(theta) / (4 * np.pi * ME * CL ** 2) * val


def test_kernels_from_helical_orbit():
    """Stokes kernels of one energy shell with isotropic pitch angles, from the
    far-field amplitudes of the helical orbit (no synchrotron formula assumed).
    Pitch angle a, observer at th = a + psi; along the pulse, with w the gyro
    frequency and q = om sin(th) sin(a) w^2 / 6, p = om (1/g^2 + psi^2) / 2:
      phase om (t - n.r/c) = p t + q t^3,
      beta.e1 = w sin(a) t,  beta.e2 = psi + cos(th) sin(a) w^2 t^2 / 2,
    so with I0 = int e^{i(pt + qt^3)} dt = 2 pi (3q)^(-1/3) Ai(p (3q)^(-1/3)):
      E1 = w sin(a) dI0/dp (times -i),  E2 = I0 (psi - cot(th)/2 (1/g^2 + psi^2))
    (t^2 -> -d^2/dp^2 and Ai'' = z Ai).  Each electron passes the observer once
    per gyration; electrons per d(alpha) ~ sin(a)."""
    th, g, w = 0.9, 2000., 1.
    u = np.linspace(-60., 60., 1200001)
    psi = u / g
    a = th - psi
    for x in (0.1, 1., 5.):
        om = x * 1.5 * g ** 3 * w * np.sin(th)
        p = 0.5 * om * (1 / g ** 2 + psi ** 2)
        q = om / 6. * w * w * np.sin(th) * np.sin(a)
        s = (3 * q) ** (-1. / 3.)
        ai, aip, _, _ = special.airy(p * s)
        E1 = w * np.sin(a) * 2 * np.pi * s * s * aip
        E2 = 2 * np.pi * s * ai * (psi - 0.5 / np.tan(th) * (1 / g ** 2 + psi ** 2))
        wt = np.sin(a)
        I = np.trapezoid(wt * (E1 * E1 + E2 * E2), psi)
        Q = np.trapezoid(wt * (E1 * E1 - E2 * E2), psi)
        V = np.trapezoid(wt * 2 * E1 * E2, psi)
        F, G, H = (float(kernel(k, x)) for k in "IQV")
        assert abs((Q / I) / (G / F) - 1) < 2e-4, x
        assert abs((V / I) / (4. / 3. / (g * np.tan(th)) * H / F) - 1) < 2e-4, x


GRID = [(Th, th, nf) for Th in (3., 10., 30.) for th in (0.5, 1.0, 2.0) for nf in (30., 1e2, 1e3, 1e4)]


def _coeff_ratios(dist, Th, th, nf):
    orc, g, m = sphere_point_oracle(Thetae=Th)
    nuB = EE * m.sph_B0 / (2 * np.pi * ME * CL)
    nu = nf * nuB * Th * Th
    c, X, K = coeffs_at(orc, m, th, nu, pol=1, dist=dist)
    out = {}
    for kind, idx in (("I", 0), ("Q", 1), ("V", 3)):
        out[kind] = c[idx] * nu * nu / exact_thermal(kind, m.sph_Ne0, m.sph_B0, Th, th, nu)
    return out


@pytest.mark.parametrize("Th,th,nf", GRID)
def test_dexter_emissivities_vs_exact(Th, th, nf):
    """Dexter (2016) fits, emission_type 4: j_I within 3.5%, j_Q within 3%,
    j_V within 3% of the exact integrals (signs included: the ratios are
    positive) for Theta_e 3-30, nu/(nu_B Theta_e^2) 30-1e4."""
    r = _coeff_ratios(4, Th, th, nf)
    assert abs(r["I"] - 1) < 0.035, r
    assert abs(r["Q"] - 1) < 0.03, r
    assert abs(r["V"] - 1) < 0.03, r


@pytest.mark.parametrize("Th,th,nf", GRID)
def test_pandya_emissivities_vs_exact(Th, th, nf):
    """Pandya et al. (2016) fits, emission_type 1 (fits to symphony's exact
    harmonic sums): j_I within 3%, j_Q within 10% of the exact integrals; their
    j_V fit is coarse in Theta_e and angle and lies within 0.45-1.7 of it."""
    r = _coeff_ratios(1, Th, th, nf)
    assert abs(r["I"] - 1) < 0.03, r
    assert abs(r["Q"] - 1) < 0.10, r
    assert 0.45 < r["V"] < 1.7, r


@pytest.mark.parametrize("Th", [3., 10., 30.])
def test_faraday_hot_high_frequency_limit(Th):
    """rho_V -> rho_V,cold K0/K2 and rho_Q -> rho_Q,cold (K1/K2 + 6 Theta) as
    X = Theta sqrt(sqrt2 sin(th) 1e3 nu_B / nu) -> 0, at temperatures where
    these factors are far from 1 (K0/K2 ~ 1e-3 - 0.2, K1/K2 + 6 Theta ~ 18-180)."""
    import mpmath as mp
    orc, g, m = sphere_point_oracle(Thetae=Th)
    n, B = m.sph_Ne0, m.sph_B0
    nuB = EE * B / (2 * np.pi * ME * CL)
    k0, k1, k2 = (float(mp.besselk(k, 1. / Th)) for k in (0, 1, 2))
    for th in (0.4, 1.1, 2.2):
        nu = 2e9 * np.sqrt(2) * np.sin(th) * nuB * Th * Th      # X ~ 7e-4
        c, X, K = coeffs_at(orc, m, th, nu)
        rQ, rV = c[8] / nu, c[10] / nu
        rV_cold = EE ** 3 * n * B * np.cos(th) / (np.pi * ME ** 2 * CL ** 2 * nu ** 2)
        rQ_cold = EE ** 4 * n * B ** 2 * np.sin(th) ** 2 / (4 * np.pi ** 2 * ME ** 3 * CL ** 3 * nu ** 3)
        assert abs(rV / (rV_cold * k0 / k2) - 1) < 1e-3, (th, rV / (rV_cold * k0 / k2))
        assert abs(rQ / (rQ_cold * (k1 / k2 + 6 * Th)) - 1) < 1e-3, (th, rQ / (rQ_cold * (k1 / k2 + 6 * Th)))


# ---------------------------------------------------- reference source text

def _strip_comments(s):
    s = re.sub(r"/\*.*?\*/", " ", s, flags=re.S)
    s = re.sub(r"//[^\n]*", " ", s)
    return re.sub(r"#if DEBUG.*?#endif", " ", s, flags=re.S)


def _body(src, name):
    m = re.search(r"\b" + name + r"\s*\([^;{)]*\)\s*\{", src)
    assert m, name
    i, depth = m.end(), 1
    while depth:
        depth += {"{": 1, "}": -1}.get(src[i], 0)
        i += 1
    return src[m.end():i - 1]


_TOK = re.compile(r"(\d+\.\d*|\.\d+|\d+)([eE][+-]?\d+)?|[A-Za-z_]\w*(?:->\w+)*|==|<=|>=|&&|\|\||[-+*/(),<>?:=!]")


def _literals_and_ops(body):
    """Numeric literal values and arithmetic operators, in order (a '*' right
    after '(' or ',' is a pointer dereference, not a product)."""
    toks = [m.group(0) for m in _TOK.finditer(body)]
    lits, ops = [], []
    for i, t in enumerate(toks):
        if t[0].isdigit() or t[0] == ".":
            lits.append(float(t))
        elif t in ("+", "-", "/") or (t == "*" and not (i > 0 and toks[i - 1] in ("(", ","))):
            ops.append(t)
    return lits, ops


FIT_PAIRS = [("I_I", "I_I"), ("I_Q", "I_Q"), ("I_V", "I_V"),
             ("maxwell_juettner_dexter_I", "dexter_I"), ("maxwell_juettner_dexter_Q", "dexter_Q"),
             ("maxwell_juettner_dexter_V", "dexter_V"), ("maxwell_juettner_leung_I", "leung_I"),
             ("maxwell_juettner_I", "mj_I"), ("maxwell_juettner_Q", "mj_Q"), ("maxwell_juettner_V", "mj_V"),
             ("maxwell_juettner_rho_Q", "mj_rho_Q"), ("maxwell_juettner_rho_V", "mj_rho_V")]


@pytest.mark.skipif(not os.path.exists(REF_FITS), reason="reference source not present")
@pytest.mark.parametrize("ref_name,oracle_name", FIT_PAIRS)
def test_fit_bodies_match_reference_source(ref_name, oracle_name):
    ref = _strip_comments(open(REF_FITS).read())
    orc = _strip_comments(open(os.path.join(os.path.dirname(O.__file__), "or_radiation.c")).read())
    rl, ro = _literals_and_ops(_body(ref, ref_name))
    ol, oo = _literals_and_ops(_body(orc, oracle_name))
    assert rl == ol, (ref_name, rl, ol)
    assert ro == oo, (ref_name, ro, oo)
