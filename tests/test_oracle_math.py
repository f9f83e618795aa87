"""Pins of the oracle's functions that the reference build cannot exercise
(the GSL-dependent thermal fits, K_n, the Stokes step of evolve_N) against
what the mathematics fixes: library Bessel functions, closed-form limits,
matrix exponentials of the transfer generators, Kirchhoff's law, and the
exact thermal synchrotron emissivity integral."""
import numpy as np
import pytest
from scipy import integrate, special
from scipy.linalg import expm

import oracle as O

EE, CL, ME = O.EE, O.CL, O.ME


def test_bessel_kn_vs_scipy_mpmath():
    import mpmath as mp
    for x in np.concatenate([10 ** np.linspace(-3, 2.8, 400), [1.999999, 2.0, 2.000001]]):
        for n in (0, 1, 2):
            ref = special.kn(n, x)
            got = O.bessel_kn(n, x)
            assert abs(got / ref - 1) < 2e-13, (n, x, got, ref)
    for x in (1e-3, 0.37, 2.0, 7.5, 91., 640.):
        for n in (0, 1, 2):
            ref = float(mp.besselk(n, x))
            assert abs(O.bessel_kn(n, x) / ref - 1) < 2e-13


# ---------------------------------------------------------------- Stokes step

def generator(c):
    """dS/dl = j - A S with A the absorption-rotation matrix in the sign
    convention of src/ipolarray.c:391-520."""
    jI, jQ, jU, jV, aI, aQ, aU, aV, rQ, rU, rV = c
    A = np.array([[aI, aQ, aU, aV],
                  [aQ, aI, rV, -rU],
                  [aU, -rV, aI, rQ],
                  [aV, rU, -rQ, aI]])
    return A, np.array([jI, jQ, jU, jV])


def exact_step(c, x, S0):
    A, j = generator(c)
    M = np.zeros((5, 5))
    M[:4, :4] = -A
    M[:4, 4] = j
    E = expm(M * x)
    return E[:4, :4] @ S0 + E[:4, 4]


def test_stokes_pure_rotation():
    rng = np.random.default_rng(0)
    for _ in range(50):
        c = np.zeros(11)
        c[8:] = rng.normal(size=3)
        S0 = rng.normal(size=4)
        x = rng.uniform(0.01, 3.)
        got = O.stokes_step(c, x, S0)
        assert np.allclose(got, exact_step(c, x, S0), rtol=1e-12, atol=1e-12)
        assert abs(np.dot(got[1:], got[1:]) - np.dot(S0[1:], S0[1:])) < 1e-12


def test_stokes_pure_emission_absorption():
    rng = np.random.default_rng(1)
    for _ in range(50):
        c = np.zeros(11)
        aI = rng.uniform(0.5, 2.)
        aP = rng.normal(size=3)
        aP *= rng.uniform(0.05, 0.9) * aI / np.linalg.norm(aP)
        c[4], c[5:8] = aI, aP
        c[0] = rng.uniform(0.1, 1.)
        jP = rng.normal(size=3)
        c[1:4] = jP * 0.5 * c[0] / np.linalg.norm(jP)
        S0 = np.array([1., 0.1, -0.2, 0.05])
        for x in (1e-7, 1e-3, 0.3, 4.):
            got = O.stokes_step(c, x, S0)
            assert np.allclose(got, exact_step(c, x, S0), rtol=1e-9, atol=1e-12), (x,)


def test_stokes_split_is_second_order():
    """With both rotation and absorption the step is a symmetric split
    (half rotation, full absorption/emission, half rotation): its error
    against the exact constant-coefficient solution falls as x^3."""
    c = np.array([0.6, 0.1, -0.05, 0.08, 1.2, 0.2, -0.1, 0.15, 0.7, -0.3, 1.1])
    S0 = np.array([1., 0.3, -0.2, 0.1])
    errs = []
    for x in (0.2, 0.1, 0.05, 0.025):
        errs.append(np.linalg.norm(O.stokes_step(c, x, S0) - exact_step(c, x, S0)))
    r = [errs[i] / errs[i + 1] for i in range(3)]
    assert all(7. < q < 9. for q in r), r


def test_approximate_solve_exact_constant():
    for j, k, x, I0 in [(0.3, 2.0, 0.5, 0.1), (1.0, 1e-5, 10., 0.), (0.2, 0.7, 1e-2, 3.)]:
        got, tau = O.approximate_solve(I0, j, k, j, k, x)
        exact = I0 * np.exp(-k * x) + j / k * (1 - np.exp(-k * x))
        assert abs(got - exact) <= 1e-9 * max(abs(exact), 1e-30) + 1e-12 * x * j
        assert abs(tau - k * x) < 1e-15 * k * x + 1e-300


# ------------------------------------------------------------ thermal fits

def sphere_point_oracle(Thetae=10., B=None, nu_over=None):
    """An oracle set to the isothermal sphere, and a point/wavevector inside
    it with a chosen angle to the field."""
    g, m = O.sphere_model(R0=50., tau0=1e-5, Thetae0=Thetae, beta0=20.)
    orc = O.Oracle(g, m)
    return orc, g, m


def coeffs_at(orc, m, theta_bk, nu_cgs, pol=1, dist=4):
    """Place a photon of fluid-frame frequency nu at angle theta_bk to B at
    r = 10 on the equator of the sphere model (static fluid: the fluid frame
    is the coordinate frame; B is along -e_theta there)."""
    X = np.array([0., np.log(10.), np.pi / 2, 0.])
    g = orc.gcov(X)
    # orthonormal spatial directions at (r, th) in eMinkowski: e_r, e_th, e_ph
    er = np.array([0, 1. / np.sqrt(g[1, 1]), 0, 0])
    eth = np.array([0, 0, 1. / np.sqrt(g[2, 2]), 0])
    nu = nu_cgs * O.HPL / (ME * CL * CL)
    khat = -np.cos(theta_bk) * eth + np.sin(theta_bk) * er   # B points along -e_th
    K = nu * (np.array([1., 0, 0, 0]) + khat)
    return orc.jar(dist, pol, X, K), X, K


def test_kirchhoff_and_polarization_bounds():
    orc, g, m = sphere_point_oracle()
    for th in (0.3, 1.0, 1.5, 2.5):
        for nu in (86e9, 230e9, 1e12):
            c, X, K = coeffs_at(orc, m, th, nu)
            jI, jQ, jU, jV, aI, aQ, aU, aV, rQ, rU, rV = c
            x = O.HPL * nu / (ME * CL * CL * m.sph_Thetae0)
            Bnu_inv = 2 * O.HPL / CL ** 2 / np.expm1(x)   # j/nu^2 over B_nu/nu^3
            for jS, aS in ((jI, aI), (jQ, aQ), (jV, aV)):
                assert abs(aS * Bnu_inv - jS) <= 1e-12 * abs(jS) + 1e-300
            assert np.hypot(np.hypot(jQ, jU), jV) <= 0.99 * jI * (1 + 1e-12)
            assert jU == 0 and rU == 0


def test_faraday_cold_plasma_limit():
    """rho_V -> e^3 n_e B cos(th) / (pi m^2 c^2 nu^2) and
    rho_Q -> e^4 n_e B^2 sin^2(th) / (4 pi^2 m^3 c^3 nu^3) as Theta_e -> 0
    (cold magnetized plasma; the relativistic factors of the fits go to 1)."""
    Th = 2e-3
    orc, g, m = sphere_point_oracle(Thetae=Th)
    n, B = m.sph_Ne0, m.sph_B0
    for th in (0.4, 1.1, 2.2):
        for nu in (230e9, 1e12):
            c, X, K = coeffs_at(orc, m, th, nu)
            rQ, rV = c[8], c[10]
            rV_cold = EE ** 3 * n * B * np.cos(th) / (np.pi * ME ** 2 * CL ** 2 * nu ** 2)
            rQ_cold = EE ** 4 * n * B ** 2 * np.sin(th) ** 2 / (4 * np.pi ** 2 * ME ** 3 * CL ** 3 * nu ** 3)
            # reported as invariants rho * nu
            assert abs(rV / nu / rV_cold - 1) < 0.01, (th, nu, rV / nu / rV_cold)
            assert abs(rQ / nu / rQ_cold - 1) < 0.02, (th, nu, rQ / nu / rQ_cold)


def exact_thermal_jI(n, B, Th, theta, nu):
    """Thermal synchrotron emissivity (erg/s/cm^3/Hz/sr) by direct integration
    of the single-particle spectrum over a Maxwell-Juettner distribution
    (ultra-relativistic, pitch angle = theta):
    j = n sqrt(3) e^3 B sin th / (4 pi m c^2) int N(g) F(nu / nu_c(g)) dg,
    nu_c = 3/2 g^2 nu_B sin th, F(x) = x int_x^inf K_5/3."""
    nuB = EE * B / (2 * np.pi * ME * CL)

    def F(x):
        if x > 50:
            return 0.
        return x * integrate.quad(lambda t: special.kv(5. / 3., t), x, np.inf, limit=200)[0]

    K2 = special.kn(2, 1. / Th)

    def integrand(g):
        beta = np.sqrt(1 - 1 / g ** 2)
        N = g ** 2 * beta * np.exp(-g / Th) / (Th * K2)
        return N * F(nu / (1.5 * g ** 2 * nuB * np.sin(theta)))

    gmax = 1 + 60 * Th
    val = integrate.quad(integrand, 1.0001, gmax, limit=400, points=[Th, 3 * Th, 10 * Th])[0]
    return n * np.sqrt(3) * EE ** 3 * B * np.sin(theta) / (4 * np.pi * ME * CL ** 2) * val


@pytest.mark.parametrize("Th,nu_fac", [(10., 1e3), (10., 1e5), (30., 1e4), (3., 1e3)])
def test_thermal_emissivity_vs_exact_integral(Th, nu_fac):
    """Dexter (emission_type 4) and Leung (unpolarized) fits against the
    exact integral; the fits are quoted to a few percent for Theta_e >~ 1."""
    orc, g, m = sphere_point_oracle(Thetae=Th)
    th = 1.0
    nuB = EE * m.sph_B0 / (2 * np.pi * ME * CL)
    nu = nu_fac * nuB
    exact = exact_thermal_jI(m.sph_Ne0, m.sph_B0, Th, th, nu)
    c_pol, X, K = coeffs_at(orc, m, th, nu, pol=1)
    c_un, _, _ = coeffs_at(orc, m, th, nu, pol=0)
    jI_dexter = c_pol[0] * nu ** 2
    jI_leung = c_un[0] * nu ** 2
    assert abs(jI_dexter / exact - 1) < 0.08, (jI_dexter / exact)
    assert abs(jI_leung / exact - 1) < 0.08, (jI_leung / exact)
