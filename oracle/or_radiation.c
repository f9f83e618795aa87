/*
 * Oracle (test infrastructure): thermal synchrotron transfer coefficients.
 * Restates src/model_radiation.c (jar_calc_dist, thermal branches only),
 * src/radiation.c and the Maxwell-Juettner fits of
 * src/symphony/maxwell_juettner_fits.c with the symphony wrappers of
 * src/symphony/fits.c.
 *
 * K_n(x): the reference calls GSL's gsl_sf_bessel_Kn, which is not in this
 * image.  It is restated from the mathematics: the power series of
 * Abramowitz & Stegun 9.6.11/9.6.13 for x <= 2 and, for x > 2, the
 * trapezoidal rule on the integral K_nu(x) = int_0^inf exp(-x cosh t)
 * cosh(nu t) dt, which converges geometrically for this entire integrand.
 * K_2 follows by the upward recurrence K_2 = K_0 + (2/x) K_1 as GSL does.
 * Pinned in tests against scipy.special.kn and mpmath.besselk.
 *
 * How the fits are pinned (tests/test_oracle_fits.py, test_oracle_math.py):
 * - j_I, j_Q, j_V (Dexter and Pandya) against the exact thermal integrals of
 *   the ultrarelativistic single-electron kernels F, G and (4/3)(cot th/g) H,
 *   which the tests derive from the radiation amplitudes of a helical orbit;
 * - rho_Q, rho_V at Theta_e 3-30 in the high-frequency limit against the
 *   relativistic thermal closed forms (K0/K2, K1/K2 + 6 Theta), and in the cold
 *   limit; Kirchhoff's law for the absorptivities;
 * - the finite-X shape functions f(X) of rho_Q and g(X) of rho_V, which are
 *   fits to numerical results with no closed form, only against the
 *   reference's source text (every literal and arithmetic operator of each
 *   fit function, in order); the reference build cannot compile them (GSL).
 */
#include "oracle.h"
#include <math.h>

extern or_geom_t OR_G;
extern or_model_t OR_M;

#define EULER_GAMMA 0.57721566490153286061

static void bessel_k01(double x, double *k0, double *k1)
{
  if (x <= 2.0) {
    /* I0, I1 and the K0/K1 series (A&S 9.6.10, 9.6.13, 9.6.11) */
    double y = 0.25 * x * x;
    double term0 = 1.0, term1 = 0.5 * x;   /* (x/2)^k / k!^2, (x/2)^(2k+1)/(k!(k+1)!) */
    double I0 = 0, I1 = 0, S0 = 0, S1 = 0;
    double Hk = 0.0;                         /* harmonic number H_k */
    double psik1 = -EULER_GAMMA, psik2 = 1.0 - EULER_GAMMA; /* psi(k+1), psi(k+2) */
    for (int k = 0; k < 40; k++) {
      if (k > 0) {
        term0 *= y / ((double)k * k);
        term1 *= y / ((double)k * (k + 1));
        Hk += 1.0 / k;
        psik1 += 1.0 / k;
        psik2 += 1.0 / (k + 1);
      }
      I0 += term0;
      I1 += term1;
      S0 += term0 * Hk;
      S1 += term1 * (psik1 + psik2);
      if (term0 < 1e-20 * I0 && term1 < 1e-20 * I1) break;
    }
    double lg = log(0.5 * x);
    *k0 = -(lg + EULER_GAMMA) * I0 + S0;
    *k1 = 1.0 / x + lg * I1 - 0.5 * S1;
    return;
  }
  /* exponentially scaled integrals, trapezoidal rule */
  double h = 0.5 / sqrt(x);
  if (h > 0.1) h = 0.1;
  double s0 = 0.5, s1 = 0.5;
  for (int m = 1; m < 100000; m++) {
    double t = m * h;
    double e = exp(-x * (cosh(t) - 1.0));
    s0 += e;
    s1 += e * cosh(t);
    if (e < 1e-19 * s0) break;
  }
  double ex = exp(-x);
  *k0 = ex * s0 * h;
  *k1 = ex * s1 * h;
}

double or_bessel_kn(int n, double x)
{
  double k0, k1;
  bessel_k01(x, &k0, &k1);
  if (n == 0) return k0;
  if (n == 1) return k1;
  double km = k0, kc = k1;
  for (int j = 1; j < n; j++) {
    double kp = km + (2.0 * j / x) * kc;
    km = kc;
    kc = kp;
  }
  return kc;
}

/* src/radiation.c:14-23 */
static double Bnu_inv(double nu, double Thetae)
{
  double x = OR_HPL * nu / (OR_ME * OR_CL * OR_CL * Thetae);
  if (x < 2.e-3)
    return ((2. * OR_HPL / (OR_CL * OR_CL)) / (x / 24. * (24. + x * (12. + x * (4. + x)))));
  else
    return ((2. * OR_HPL / (OR_CL * OR_CL)) / (exp(x) - 1.));
}

/* src/radiation.c:26-44 */
static double get_fluid_nu(const double K[4], const double Ucov[4])
{
  return -(K[0] * Ucov[0] + K[1] * Ucov[1] + K[2] * Ucov[2] + K[3] * Ucov[3]) * OR_ME * OR_CL
      * OR_CL / OR_HPL;
}

/* src/radiation.c:47-72 */
static double get_bk_angle(const double K[4], const double Ucov[4], const double Bcon[4],
                           const double Bcov[4])
{
  double B = sqrt(fabs(Bcon[0] * Bcov[0] + Bcon[1] * Bcov[1] + Bcon[2] * Bcov[2]
                       + Bcon[3] * Bcov[3]));
  if (B == 0.) return (M_PI / 2.);
  double k = fabs(K[0] * Ucov[0] + K[1] * Ucov[1] + K[2] * Ucov[2] + K[3] * Ucov[3]);
  double mu = (K[0] * Bcov[0] + K[1] * Bcov[1] + K[2] * Bcov[2] + K[3] * Bcov[3]) / (k * B);
  if (fabs(mu) > 1.) mu /= fabs(mu);
  return acos(mu);
}

typedef struct {
  double nu, B, Ne, theta, Thetae;
  int dexter_fit;
} fitpar;

/* constants of src/symphony/params.c:13-35 */
#define SY_PI 3.1415926535897932384
#define SY_ME 9.1093826e-28
#define SY_C 2.99792458e10
#define SY_E 4.80320680e-10

/* src/symphony/maxwell_juettner_fits.c:334-356 */
static double I_I(double x)
{
  return 2.5651 * (1 + 1.92 * pow(x, -1. / 3.) + 0.9977 * pow(x, -2. / 3.))
      * exp(-1.8899 * pow(x, 1. / 3.));
}
static double I_Q(double x)
{
  return 2.5651 * (1 + 0.93193 * pow(x, -1. / 3.) + 0.499873 * pow(x, -2. / 3.))
      * exp(-1.8899 * pow(x, 1. / 3.));
}
static double I_V(double x)
{
  return (1.81384 / x + 3.42319 * pow(x, -2. / 3.) + 0.0292545 * pow(x, -0.5)
          + 2.03773 * pow(x, -1. / 3.))
      * exp(-1.8899 * pow(x, 1. / 3.));
}

/* src/symphony/maxwell_juettner_fits.c:287-330 */
static double dexter_I(const fitpar *p)
{
  double nus = 3.0 * OR_EE * p->B * sin(p->theta) / 4.0 / M_PI / OR_ME / OR_CL * p->Thetae
      * p->Thetae + 1.0;
  double x = p->nu / nus;
  return p->Ne * OR_EE * OR_EE * p->nu / 2. / sqrt(3) / OR_CL / p->Thetae / p->Thetae * I_I(x);
}
static double dexter_Q(const fitpar *p)
{
  double nus = 3.0 * OR_EE * p->B * sin(p->theta) / 4.0 / M_PI / OR_ME / OR_CL * p->Thetae
      * p->Thetae + 1.0;
  double x = p->nu / nus;
  return -p->Ne * OR_EE * OR_EE * p->nu / 2. / sqrt(3) / OR_CL / p->Thetae / p->Thetae * I_Q(x);
}
static double dexter_V(const fitpar *p)
{
  double nus = 3.0 * OR_EE * p->B * sin(p->theta) / 4.0 / M_PI / OR_ME / OR_CL * p->Thetae
      * p->Thetae + 1.0;
  double x = p->nu / nus;
  return 2. * p->Ne * OR_EE * OR_EE * p->nu / tan(p->theta) / 3. / sqrt(3) / OR_CL / p->Thetae
      / p->Thetae / p->Thetae * I_V(x);
}

/* src/symphony/maxwell_juettner_fits.c:368-395 */
static double leung_I(const fitpar *p)
{
  double K2 = fmax(or_bessel_kn(2, 1. / p->Thetae), OR_SMALL);
  double nuc = OR_EE * p->B / (2. * M_PI * OR_ME * OR_CL);
  double nus = (2. / 9.) * nuc * p->Thetae * p->Thetae * sin(p->theta);
  if (p->nu > 1.e12 * nus) return 0.;
  double x = p->nu / nus;
  double f = pow(pow(x, 1. / 2.) + pow(2., 11. / 12.) * pow(x, 1. / 6.), 2);
  return (sqrt(2.) * M_PI * OR_EE * OR_EE * p->Ne * nus / (3. * OR_CL * K2)) * f
      * exp(-pow(x, 1. / 3.));
}

/* src/symphony/params.c:44-48 */
static double get_nu_c(const fitpar *p)
{
  return (SY_E * p->B) / (2. * SY_PI * SY_ME * SY_C);
}

/* src/symphony/maxwell_juettner_fits.c:27-129 (Pandya fits, dispatching to
 * Leung / Dexter by dexter_fit) */
static double mj_I(const fitpar *p)
{
  if (p->dexter_fit == 2) return leung_I(p);
  else if (p->dexter_fit) return dexter_I(p);
  double nu_c = get_nu_c(p);
  double nu_s = (2. / 9.) * nu_c * sin(p->theta) * p->Thetae * p->Thetae;
  double X = p->nu / nu_s;
  double prefactor = (p->Ne * pow(SY_E, 2.) * nu_c) / SY_C;
  double term1 = sqrt(2.) * SY_PI / 27. * sin(p->theta);
  double term2 = pow(pow(X, 0.5) + pow(2., 11. / 12.) * pow(X, 1. / 6.), 2.);
  double term3 = exp(-pow(X, 1. / 3.));
  return prefactor * term1 * term2 * term3;
}
static double mj_Q(const fitpar *p)
{
  if (p->dexter_fit) return dexter_Q(p);
  double nu_c = get_nu_c(p);
  double nu_s = (2. / 9.) * nu_c * sin(p->theta) * p->Thetae * p->Thetae;
  double X = p->nu / nu_s;
  double prefactor = (p->Ne * pow(SY_E, 2.) * nu_c) / SY_C;
  double term1 = sqrt(2.) * SY_PI / 27. * sin(p->theta);
  double term2 = (7. * pow(p->Thetae, 24. / 25.) + 35.) / (10. * pow(p->Thetae, 24. / 25.) + 75.);
  double term3 = pow(pow(X, 0.5) + term2 * pow(2., 11. / 12.) * pow(X, 1. / 6.), 2.);
  double ans = prefactor * term1 * term3 * exp(-pow(X, 1. / 3.));
  return -ans;
}
static double mj_V(const fitpar *p)
{
  if (p->dexter_fit) return dexter_V(p);
  double nu_c = get_nu_c(p);
  double nu_s = (2. / 9.) * nu_c * sin(p->theta) * p->Thetae * p->Thetae;
  double X = p->nu / nu_s;
  double prefactor = (p->Ne * pow(SY_E, 2.) * nu_c) / SY_C;
  double term1 = (37. - 87. * sin(p->theta - 28. / 25.)) / (100. * (p->Thetae + 1.));
  double term2 = pow(1. + (pow(p->Thetae, 3. / 5.) / 25. + 7. / 10.) * pow(X, 9. / 25.), 5. / 3.);
  return prefactor * term1 * term2 * exp(-pow(X, 1. / 3.));
}

/* src/symphony/maxwell_juettner_fits.c:207-240 */
static double mj_rho_Q(const fitpar *p)
{
  double omega0 = SY_E * p->B / (SY_ME * SY_C);
  double wp2 = 4. * SY_PI * p->Ne * pow(SY_E, 2.) / SY_ME;
  double x = p->Thetae * sqrt(sqrt(2.) * sin(p->theta) * (1.e3 * omega0 / (2. * SY_PI * p->nu)));
  double extraterm = (.011 * exp(-x / 47.2)
                      - pow(2., (-1. / 3.)) / pow(3., (23. / 6.)) * SY_PI * 1.e4
                          * pow((x + 1.e-16), (-8. / 3.)))
      * (0.5 + 0.5 * tanh((log(x) - log(120.)) / 0.1));
  double jffunc = 2.011 * exp(-pow(x, 1.035) / 4.7) - cos(x / 2.) * exp(-pow(x, 1.2) / 2.73)
      - .011 * exp(-x / 47.2) + extraterm;
  double k1 = or_bessel_kn(1, 1. / p->Thetae);
  double k2 = or_bessel_kn(2, 1. / p->Thetae);
  double k_ratio = (k2 > 0) ? k1 / k2 : 1;
  double eps11m22 = jffunc * wp2 * pow(omega0, 2.) / pow(2. * SY_PI * p->nu, 4.)
      * (k_ratio + 6. * p->Thetae) * pow(sin(p->theta), 2.);
  return 2. * SY_PI * p->nu / (2. * SY_C) * eps11m22;
}

/* src/symphony/maxwell_juettner_fits.c:251-284 */
static double mj_rho_V(const fitpar *p)
{
  double omega0 = SY_E * p->B / (SY_ME * SY_C);
  double wp2 = 4. * SY_PI * p->Ne * pow(SY_E, 2.) / SY_ME;
  double x = p->Thetae * sqrt(sqrt(2.) * sin(p->theta) * (1.e3 * omega0 / (2. * SY_PI * p->nu)));
  double k0 = or_bessel_kn(0, 1. / p->Thetae);
  double k2 = or_bessel_kn(2, 1. / p->Thetae);
  double fit_factor = 0;
  if (p->dexter_fit && k2 > 0) {
    double shgmfunc = 0.43793091 * log(1. + 0.00185777 * pow(x, 1.50316886));
    fit_factor = (k0 - shgmfunc) / k2;
  } else {
    double shgmfunc = 1 - 0.11 * log(1 + 0.035 * x);
    double k_ratio = (k2 > 0) ? k0 / k2 : 1;
    fit_factor = k_ratio * shgmfunc;
  }
  double eps12 = wp2 * omega0 / pow((2. * SY_PI * p->nu), 3.) * fit_factor * cos(p->theta);
  return 2. * SY_PI * p->nu / SY_C * eps12;
}

/* model/analytic/model.c:185-199 (emission_type 10 of the analytic model) */
static void analytic_jk(const double X[4], const double K[4], double *j, double *k)
{
  double n = or_ne(X);
  double Ucon[4], Ucov[4], Bcon[4], Bcov[4];
  or_fourv(X, K, Ucon, Ucov, Bcon, Bcov);
  double nu = get_fluid_nu(K, Ucov);
  *j = fmax(n * pow(nu / OR_M.an_freqcgs, -OR_M.an_alpha) / pow(nu, 2), 0);
  *k = fmax((OR_M.an_A * n * pow(nu / OR_M.an_freqcgs, -(2.5 + OR_M.an_alpha)) + 1.e-54) * nu, 0);
}

/* src/model_radiation.c:145-305.  out = jI jQ jU jV aI aQ aU aV rQ rU rV.
 * dist: 1 Pandya thermal, 4 Dexter thermal (default), 10 model-defined. */
void or_jar(int dist, int pol, const double X[4], const double K[4], double out[11])
{
  double *jI = out, *jQ = out + 1, *jU = out + 2, *jV = out + 3, *aI = out + 4, *aQ = out + 5,
         *aU = out + 6, *aV = out + 7, *rQ = out + 8, *rU = out + 9, *rV = out + 10;
  for (int n = 0; n < 11; n++) out[n] = 0.;
  double Ne = or_ne(X);
  if (Ne <= 0.) return;
  if (dist == 10) {
    double j, k;
    analytic_jk(X, K, &j, &k);
    *jI = j;
    *aI = k;
    return;
  }
  fitpar p;
  double Ucon[4], Ucov[4], Bcon[4], Bcov[4];
  or_fourv(X, K, Ucon, Ucov, Bcon, Bcov);
  double nu = get_fluid_nu(K, Ucov);
  double nusq = nu * nu;
  double theta = get_bk_angle(K, Ucov, Bcon, Bcov);
  p.Ne = Ne;
  p.nu = nu;
  p.theta = theta;
  p.B = or_b(X);
  p.Thetae = or_thetae(X);
  p.dexter_fit = (dist == 1) ? 0 : 1;

  if (theta <= 0.0 || theta >= M_PI) {
    if (pol) *rV = mj_rho_V(&p) * nu;
    return;
  }
  if (!pol) {
    p.dexter_fit = 2;
    *jI = mj_I(&p);
    *jI /= nusq;
    double Bnuinv = Bnu_inv(nu, p.Thetae);
    *aI = (Bnuinv > 0) ? *jI / Bnuinv : 0;
    return;
  }
  *jI = mj_I(&p);
  *jI /= nusq;
  *jQ = -mj_Q(&p) / nusq;
  *jU = -0. / nusq;
  *jV = mj_V(&p) / nusq;
  double jP = sqrt(*jQ * *jQ + *jU * *jU + *jV * *jV);
  const double max_pol_frac_e = 0.99;
  if (*jI < jP / max_pol_frac_e) {
    double pol_frac_e = *jI / jP * max_pol_frac_e;
    *jQ *= pol_frac_e;
    *jU *= pol_frac_e;
    *jV *= pol_frac_e;
  }
  double Bnuinv = Bnu_inv(nu, p.Thetae);
  if (Bnuinv > 0) {
    *aI = *jI / Bnuinv;
    *aQ = *jQ / Bnuinv;
    *aU = *jU / Bnuinv;
    *aV = *jV / Bnuinv;
  }
  p.dexter_fit = 0;
  *rQ = mj_rho_Q(&p) * nu;
  *rU = 0. * nu;
  *rV = mj_rho_V(&p) * nu;
}
