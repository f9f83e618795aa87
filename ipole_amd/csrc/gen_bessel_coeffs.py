"""Generate ipg_bessel_coeffs.h: Chebyshev expansions of the exponentially
scaled modified Bessel functions sqrt(x) e^x K_n(x), n = 0, 1, on x >= 2,
in the variable t = 4/x - 1 (x in [2, inf) -> t in (-1, 1]).

Method: Chebyshev interpolation at the Chebyshev-Gauss nodes in 60-digit
arithmetic (mpmath.besselk), coefficients by the discrete cosine sum,
truncated where |c_k| < 1e-18.  For x <= 2 the device code uses the power
series of Abramowitz & Stegun 9.6.10-9.6.13 instead.

Run:  python ipole_amd/csrc/gen_bessel_coeffs.py > ipole_amd/csrc/ipg_bessel_coeffs.h
"""
import mpmath as mp

mp.mp.dps = 60
NODES = 64


def cheb_coeffs(f):
    n = NODES
    vals = []
    for k in range(n):
        t = mp.cos(mp.pi * (k + mp.mpf("0.5")) / n)
        vals.append(f(t))
    c = []
    for j in range(n):
        s = mp.mpf(0)
        for k in range(n):
            s += vals[k] * mp.cos(mp.pi * j * (k + mp.mpf("0.5")) / n)
        c.append(2 * s / n)
    c[0] /= 2
    last = max(j for j in range(n) if abs(c[j]) > mp.mpf("1e-18"))
    return c[: last + 1]


def scaled_k(nu):
    def f(t):
        x = 4 / (t + 1)
        return mp.sqrt(x) * mp.exp(x) * mp.besselk(nu, x)
    return f


def main():
    print("/* Generated by gen_bessel_coeffs.py: Chebyshev coefficients of")
    print(" * sqrt(x) exp(x) K_n(x) on x >= 2 in t = 4/x - 1, f(t) = sum_k c_k T_k(t). */")
    print("#pragma once")
    for nu in (0, 1):
        c = cheb_coeffs(scaled_k(nu))
        print("#define IPG_K%d_NCHEB %d" % (nu, len(c)))
        print("static constexpr double ipg_k%d_cheb[%d] = {" % (nu, len(c)))
        for v in c:
            print("  %s," % mp.nstr(v, 20, min_fixed=-1, max_fixed=-1))
        print("};")


if __name__ == "__main__":
    main()
