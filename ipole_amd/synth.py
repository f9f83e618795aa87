"""Seeded synthetic iharm-format FMKS dumps.

The generator produces what an iharm3d dump holds (model/iharm/model.c:
860-1051 for the header, :1895-1942 for the primitives): FMKS grid
parameters and primitives rho, u, U1..U3 (normal-observer velocity in FMKS
coordinates) and B1..B3 (B^i in FMKS coordinates), [N1][N2][N3][8] float64.

Profile (all seeded):
  * rho: a torus-like ring, exp(-ln(r/r0)^2 / (2*0.8^2)) * exp(-cos^2(th) / (2*0.3^2)),
    plus a thin spherical atmosphere 1e-3 (r/2)^-1.5, times lognormal noise of
    width sigma (0.3 by default);
  * u = K rho^(13/9) with gam = 13/9;
  * U^i: Keplerian rotation U3 = 1/(r^1.5 + a) and a slow radial inflow, each
    with a multiplicative random perturbation (5% by default);
  * B^i = eps^{ijk} d_j A_k / sqrt(-g) from a seeded random vector potential
    (the ring profile times a few random Fourier modes) plus a weak uniform
    vertical field, scaled so that the median plasma beta inside the ring is
    beta_target (10 by default).
"""
import numpy as np

# iharm3d-typical FMKS parameters
FMKS_DEFAULTS = dict(hslope=0.3, poly_xt=0.82, poly_alpha=14.0, mks_smooth=0.5)


def fmks_header(N1, N2, N3, a=0.9375, Rin=None, Rout=1000., hslope=0.3, poly_xt=0.82,
                poly_alpha=14.0, mks_smooth=0.5, gam=13. / 9):
    """Header dict of an FMKS dump.  Rin defaults to 0.85 r_h (inside the horizon)."""
    rh = 1. + np.sqrt(1. - a * a)
    if Rin is None:
        Rin = 0.85 * rh
    x1s, x1e = np.log(Rin), np.log(Rout)
    return dict(metric="FMKS", N1=N1, N2=N2, N3=N3, n_prim=8, gam=gam, a=a, hslope=hslope,
                Rin=Rin, Rout=Rout, poly_xt=poly_xt, poly_alpha=poly_alpha, mks_smooth=mks_smooth,
                startx=[0., x1s, 0., 0.], dx=[0., (x1e - x1s) / N1, 1. / N2, 2 * np.pi / N3],
                t=0., dump_cadence=5.)


def poly_norm(h):
    # model/iharm/model.c:1041
    return 0.5 * np.pi * 1. / (1. + 1. / (h["poly_alpha"] + 1.) * 1. / h["poly_xt"] ** h["poly_alpha"])


def fmks_theta(X1, X2, h):
    """theta(X1, X2) and its X1, X2 derivatives for FMKS (src/coordinates.c:75-81, :284-304)."""
    hs, xt, al, ms = h["hslope"], h["poly_xt"], h["poly_alpha"], h["mks_smooth"]
    pn = poly_norm(h)
    x1s = h["startx"][1]
    thG = np.pi * X2 + ((1. - hs) / 2.) * np.sin(2. * np.pi * X2)
    y = 2 * X2 - 1.
    thJ = pn * y * (1. + (y / xt) ** al / (al + 1.)) + 0.5 * np.pi
    E = np.exp(ms * (x1s - X1))
    th = thG + E * (thJ - thG)
    dthG = np.pi + (1 - hs) * np.pi * np.cos(2 * np.pi * X2)
    dthJ = 2 * pn * (1. + (y / xt) ** al)
    return th, -ms * E * (thJ - thG), dthG + E * (dthJ - dthG)


def fmks_gcov(X1, X2, h):
    """Covariant FMKS metric, shape X1.shape + (4, 4)."""
    a = h["a"]
    X1, X2 = np.broadcast_arrays(X1, X2)
    r = np.exp(X1)
    th, d1, d2 = fmks_theta(X1, X2, h)
    cth, sth = np.cos(th), np.sin(th)
    s2 = sth * sth
    rho2 = r * r + a * a * cth * cth
    gks = np.zeros(X1.shape + (4, 4))
    gks[..., 0, 0] = -1. + 2. * r / rho2
    gks[..., 0, 1] = gks[..., 1, 0] = 2. * r / rho2
    gks[..., 0, 3] = gks[..., 3, 0] = -2. * a * r * s2 / rho2
    gks[..., 1, 1] = 1. + 2. * r / rho2
    gks[..., 1, 3] = gks[..., 3, 1] = -a * s2 * (1. + 2. * r / rho2)
    gks[..., 2, 2] = rho2
    gks[..., 3, 3] = s2 * (rho2 + a * a * s2 * (1. + 2. * r / rho2))
    J = np.zeros(X1.shape + (4, 4))
    J[..., 0, 0] = 1.
    J[..., 1, 1] = r
    J[..., 2, 1] = d1
    J[..., 2, 2] = d2
    J[..., 3, 3] = 1.
    return np.einsum("...lm,...lk,...kn->...mn", J, gks, J)


def zone_centers(h):
    N1, N2, N3 = h["N1"], h["N2"], h["N3"]
    X1 = h["startx"][1] + (np.arange(N1) + 0.5) * h["dx"][1]
    X2 = h["startx"][2] + (np.arange(N2) + 0.5) * h["dx"][2]
    X3 = h["startx"][3] + (np.arange(N3) + 0.5) * h["dx"][3]
    return X1, X2, X3


def _smooth_random(rng, shape, nmodes=6):
    """Sum of a few random low-order Fourier modes on the unit cube, periodic in x3."""
    N1, N2, N3 = shape
    x1 = (np.arange(N1) + 0.5) / N1
    x2 = (np.arange(N2) + 0.5) / N2
    x3 = (np.arange(N3) + 0.5) / N3
    f = np.zeros(shape)
    for _ in range(nmodes):
        k1, k2 = rng.uniform(0.5, 4.0, size=2)
        k3 = rng.integers(0, 4)
        p1, p2, p3 = rng.uniform(0, 2 * np.pi, size=3)
        amp = rng.normal()
        f += amp * (np.cos(2 * np.pi * k1 * x1 + p1)[:, None, None]
                    * np.cos(2 * np.pi * k2 * x2 + p2)[None, :, None]
                    * np.cos(2 * np.pi * k3 * x3 + p3)[None, None, :])
    return f / np.sqrt(nmodes)


def b_squared(prims, h):
    """b^mu b_mu per zone from the primitives, as model/iharm/model.c:1954-2009 does."""
    X1, X2, _ = zone_centers(h)
    g = fmks_gcov(X1[:, None], X2[None, :], h)            # (N1,N2,4,4)
    gc = np.linalg.inv(g)
    U = prims[..., 2:5]
    B = prims[..., 5:8]
    gs = g[:, :, None, 1:, 1:]
    UdotU = np.einsum("ijkab,ijka,ijkb->ijk", np.broadcast_to(gs, U.shape[:3] + (3, 3)), U, U)
    ufac = np.sqrt(-1. / gc[:, :, None, 0, 0] * (1 + np.abs(UdotU)))
    ucon = np.empty(U.shape[:3] + (4,))
    ucon[..., 0] = -ufac * gc[:, :, None, 0, 0]
    ucon[..., 1:] = U - ufac[..., None] * gc[:, :, None, 0, 1:]
    ucov = np.einsum("ijab,ijkb->ijka", g, ucon)
    udotB = np.einsum("ijka,ijka->ijk", ucov[..., 1:], B)
    bcon = np.empty_like(ucon)
    bcon[..., 0] = udotB
    bcon[..., 1:] = (B + ucon[..., 1:] * udotB[..., None]) / ucon[..., 0:1]
    bcov = np.einsum("ijab,ijkb->ijka", g, bcon)
    return np.einsum("ijka,ijka->ijk", bcon, bcov)


def make_dump(N1, N2, N3, a=0.9375, seed=1234, noise_sigma=0.3, vel_noise=0.05, beta_target=10.,
              Kent=0.1, r0=8., **hkw):
    """Returns (header, prims[N1][N2][N3][8])."""
    h = fmks_header(N1, N2, N3, a=a, **hkw)
    rng = np.random.default_rng(seed)
    X1, X2, X3 = zone_centers(h)
    r = np.exp(X1)[:, None, None]
    th, d1, d2 = fmks_theta(X1[:, None], X2[None, :], h)
    th = th[:, :, None]
    ring = np.exp(-np.log(r / r0) ** 2 / (2 * 0.8 ** 2)) * np.exp(-np.cos(th) ** 2 / (2 * 0.3 ** 2))
    atm = 1e-3 * (r / 2.) ** -1.5 * np.ones_like(th)
    shape = (N1, N2, N3)
    rho = (ring + atm) * np.exp(noise_sigma * rng.standard_normal(shape) - 0.5 * noise_sigma ** 2)
    uu = Kent * rho ** (13. / 9.)
    prims = np.zeros(shape + (8,))
    prims[..., 0] = rho
    prims[..., 1] = uu
    prims[..., 2] = -0.1 * r ** -1.5 * (1 + vel_noise * rng.standard_normal(shape))
    prims[..., 3] = 0.005 * r ** -1.5 * rng.standard_normal(shape)
    prims[..., 4] = 1. / (r ** 1.5 + a) * (1 + vel_noise * rng.standard_normal(shape))

    # vector potential: ring-weighted random modes + weak uniform vertical field
    ringA = np.maximum(ring - 0.05, 0.) * np.ones(shape)
    A1 = 0.3 * ringA * _smooth_random(rng, shape)
    A2 = 0.3 * ringA * _smooth_random(rng, shape)
    A3 = ringA * (1. + 0.3 * _smooth_random(rng, shape)) + 1e-4 * 0.5 * (r * np.sin(th)) ** 2
    gdet = np.sqrt(-np.linalg.det(fmks_gcov(X1[:, None], X2[None, :], h)))[:, :, None]
    dx1, dx2, dx3 = h["dx"][1], h["dx"][2], h["dx"][3]

    def d(A, axis, dx):
        if axis == 2:
            return (np.roll(A, -1, axis=2) - np.roll(A, 1, axis=2)) / (2 * dx)
        return np.gradient(A, dx, axis=axis, edge_order=2)

    prims[..., 5] = (d(A3, 1, dx2) - d(A2, 2, dx3)) / gdet
    prims[..., 6] = (d(A1, 2, dx3) - d(A3, 0, dx1)) / gdet
    prims[..., 7] = (d(A2, 0, dx1) - d(A1, 1, dx2)) / gdet

    bsq = b_squared(prims, h)
    beta = uu * (h["gam"] - 1.) / (0.5 * bsq)
    sel = ring * np.ones(shape) > 0.1
    scale = np.sqrt(np.median(beta[sel]) / beta_target)
    prims[..., 5:8] *= scale
    h["seed"] = seed
    return h, prims


def batch_jobs(seed=0, n=64):
    """The batching workload's job set on the 288x128x128 dump: n jobs with
    thetacam from {17, 50, 90, 130, 163}, frequency from {86, 230, 345} GHz,
    M_unit log-uniform over 1e24-1e26 and image side from {64, 96, 128, 192,
    256}, drawn from a seeded generator."""
    rng = np.random.default_rng(seed)
    th = rng.choice([17., 50., 90., 130., 163.], n)
    fr = rng.choice([86e9, 230e9, 345e9], n)
    mu = 10 ** rng.uniform(24., 26., n)
    sz = rng.choice([64, 96, 128, 192, 256], n)
    return [dict(thetacam=float(t), freqcgs=float(f), M_unit=float(m), n=int(z))
            for t, f, m, z in zip(th, fr, mu, sz)]
