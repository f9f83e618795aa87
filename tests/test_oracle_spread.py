"""The reference algorithm's own sensitivity to rounding: the oracle (which
reproduces the reference's translation units bit for bit) built with and
without FMA contraction.  The finite-difference connection (DEL = 1e-7,
src/geometry.c:91) turns last-bit differences in the metric into ~1e-9
differences in Gamma, which the geodesics carry to the image.  This spread
is the floor any independent implementation meets; the GPU parity bar
(tests/test_gpu_parity.py) is set against it."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import small_dump, stokes_errors

FMA = os.path.join(os.path.dirname(O.LIB), "liboracle_fma.so")


@pytest.mark.skipif(not os.path.exists(FMA), reason="liboracle_fma.so not built")
def test_build_to_build_spread():
    h, prims = small_dump()
    L_fma = O.load(FMA)
    worst = {}
    for th in (60., 90.):
        p = O.default_params(nx=32, ny=32, thetacam=th, freqcgs=230e9)
        a = O.iharm_oracle(h, prims, M_unit=3e26)[0].render(p)
        b = O.iharm_oracle(h, prims, M_unit=3e26, L=L_fma)[0].render(p)
        assert np.array_equal(a["nstep"], b["nstep"])
        assert abs(b["imageS"][..., 0].sum() / a["imageS"][..., 0].sum() - 1) < 1e-7
        for k, (mx, frac) in stokes_errors(a["imageS"], a["image"], b["imageS"], b["image"]).items():
            worst[k] = max(worst.get(k, 0.), mx)
            assert mx < 1e-4 and frac < 0.01
    # the spread is real (not zero) and of the order the GPU bar allows
    assert max(worst["Q"], worst["U"]) > 1e-6
