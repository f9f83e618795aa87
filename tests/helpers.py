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
