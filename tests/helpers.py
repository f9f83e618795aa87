"""Shared set-up for the parity tests: synthetic dumps, oracle renders and the
error norm used throughout (per pixel |dS| / max(|I_px|, 1e-6 max I))."""
import numpy as np

import oracle as O
from ipole_amd import ipgpu as G
from ipole_amd import synth

F3 = lambda f: f ** 3  # noqa: E731


def small_dump(N=(48, 32, 32), seed=7, a=0.9375):
    return synth.make_dump(*N, seed=seed, a=a)


def stokes_errors(ref_imageS, ref_image, got_imageS, got_image):
    """Max over pixels of |dS|/max(|I|, 1e-6 max I) for I, Q, U, V, and of the
    unpolarized image, plus the fractions above 1e-6."""
    Iref = np.abs(ref_imageS[..., 0])
    den = np.maximum(Iref, 1e-6 * Iref.max())
    errs = {}
    for k, n in enumerate("IQUV"):
        e = np.abs(got_imageS[..., k] - ref_imageS[..., k]) / den
        errs[n] = (float(e.max()), float(np.mean(e > 1e-6)))
    Iu = np.abs(ref_image)
    e = np.abs(got_image - ref_image) / np.maximum(Iu, 1e-6 * Iu.max())
    errs["unpol"] = (float(e.max()), float(np.mean(e > 1e-6)))
    return errs


def oracle_iharm(h, prims, M_unit, **cam):
    orc, fields = O.iharm_oracle(h, prims, M_unit=M_unit)
    p = O.default_params(**cam)
    return orc, fields, p


def gpu_image_kwargs(p):
    """ipgpu.make_image keywords from oracle Params."""
    return dict(nx=p.nx, ny=p.ny, rcam=p.rcam, thetacam=p.thetacam, phicam=p.phicam, rotcam=p.rotcam,
                fovx_dsource=p.fovx_dsource, fovy_dsource=p.fovy_dsource, dx=p.dx, dy=p.dy,
                dsource=p.dsource, freqcgs=p.freqcgs, eps=p.eps, maxnstep=p.maxnstep,
                old_centering=p.old_centering, only_unpolarized=p.only_unpolarized, qu_conv=p.qu_conv,
                stokes_floors=p.stokes_floors, target_nturns=p.target_nturns,
                subring_dtheta=p.subring_dtheta, counterjet=p.counterjet,
                emission_type=p.emission_type, xoff=p.xoff, yoff=p.yoff)


TRACE_GEO = ("X", "Kcon", "dl", "r", "th", "phi")
TRACE_TETRAD = ("Econ", "Ecov", "e2Constructed", "e2ConstructedCov", "e2Transported", "e2TransportedCov",
                "ntetrad", "ncoord", "stokes")


def trace_compare(ref, got, tag, tol=1e-6, geo_tol=1e-7, tetrad_tol=1e-5):
    """Trace records (nstep + 1, TRACE_NF) against the oracle's, per field as
    |d| / max |ref| over the steps: the geodesic within geo_tol, the plasma
    tetrad and what is projected on it within tetrad_tol (Gram-Schmidt
    amplifies the positions' difference where K and B are nearly parallel),
    everything else within tol.  Returns the per-field error."""
    fr, fg = G.trace_fields(ref), G.trace_fields(got)
    worst = {}
    for name, _, _ in G.TRACE_SETS:
        a, b = fr[name], fg[name]
        sc = np.abs(a).max()
        if sc == 0:
            assert not np.any(b), (tag, name)
            continue
        e = float(np.abs(b - a).max() / sc)
        worst[name] = e
        t = geo_tol if name in TRACE_GEO else (tetrad_tol if name in TRACE_TETRAD else tol)
        assert e < t, (tag, name, e)
    return worst
