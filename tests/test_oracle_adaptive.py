"""The oracle's adaptive refinement (or_render_adaptive: src/main.c:98-124,
505-792, 912-1000) pinned by its limits and by closed forms:

* every error estimate above its tolerance (refine_abs = refine_rel = 0,
  refine_cut < 0): every pixel is traced and the image equals the plain
  render bit for bit, with the per-level counts of the 2^N + 1 grids;
* no estimate above (tolerances 1e300): only the coarse grid is traced and
  the rest is the hierarchical interpolation of it, written out here level
  by level (the mean of two neighbours on coarser rows and columns, of four
  corners elsewhere; or the lower-left traced neighbour);
* Iavg is the coarse intensities weighted by the area each stands for,
  the product of 1-D weights sp/2 + 1, sp, ..., sp, sp/2 that sum to the
  internal side;
* at working tolerances each pixel is either traced (equal to the plain
  render) or interpolated from its level's neighbours in the result, and
  the choice follows the criterion evaluated on the result.
"""
import numpy as np
import pytest

import oracle as O
from helpers import small_dump

F3 = 230e9 ** 3


@pytest.fixture(scope="module")
def case():
    h, prims = small_dump()
    orc, _ = O.iharm_oracle(h, prims, M_unit=3e26)
    p = O.default_params(nx=32, ny=32, thetacam=163., freqcgs=230e9)
    ref = orc.render(p)
    return orc, p, ref


def level_of(i, j, sp):
    """Refinement level at which internal pixel (i, j) is decided."""
    for lev in range(0, 32):
        s = sp >> lev
        if s == 0:
            break
        if i % s == 0 and j % s == 0:
            return lev
    raise AssertionError


def test_all_traced_equals_plain(case):
    orc, p, ref = case
    a = orc.render_adaptive(p, 8, refine_abs=0., refine_rel=0., refine_cut=-1.)
    assert a["rc"] == 0 and a["levels"] == 3
    assert a["traced"] == [9 ** 2, 17 ** 2 - 9 ** 2, 33 ** 2 - 17 ** 2]
    assert a["interpolated"] == [0, 0, 0]
    assert not a["interp"].any()
    for k in ("image", "imageS", "taus"):
        assert np.array_equal(a[k], ref[k]), k


@pytest.mark.parametrize("nearest,unpol", [(0, 0), (1, 0), (1, 1)])
def test_nothing_refined_is_interpolation(case, nearest, unpol):
    orc, p0, _ = case
    p = O.default_params(nx=32, ny=32, thetacam=163., freqcgs=230e9, only_unpolarized=unpol)
    a = orc.render_adaptive(p, 8, refine_abs=1e300, refine_rel=1e300, nearest=nearest)
    assert a["traced"] == [81, 0, 0] and a["interpolated"] == [0, 208, 1008]
    # the coarse grid traced, on the 33 x 33 internal grid
    sp, n = 4, 33
    pix = np.array([(i, j) for i in range(0, n, sp) for j in range(0, n, sp)], np.int32)
    cam = orc.camera(p)
    rc, out7, _, _ = orc.render_pixels(p, cam, pix)
    assert rc == 0
    img = np.zeros((n, n))
    tau = np.zeros((n, n))
    S = np.zeros((n, n, 5))
    for q, (i, j) in enumerate(pix):
        img[i, j] = out7[q, 0] * F3
        tau[i, j] = out7[q, 5]
        S[i, j] = [out7[q, 1] * F3, -out7[q, 2] * F3, -out7[q, 3] * F3, out7[q, 4] * F3, out7[q, 6]]
    arrays = [img, tau] + ([] if unpol else [S])
    for lev in (1, 2):
        s = sp >> lev
        for i in range(0, n, s):
            for j in range(0, n, s):
                if i % (2 * s) == 0 and j % (2 * s) == 0:
                    continue
                if i % (2 * s) == 0:
                    nb = [(i, j - s), (i, j + s)]
                elif j % (2 * s) == 0:
                    nb = [(i - s, j), (i + s, j)]
                else:
                    nb = [(i - s, j - s), (i + s, j - s), (i - s, j + s), (i + s, j + s)]
                for k, A in enumerate(arrays):
                    if nearest:
                        if k == 0 or not unpol:      # save_pixelTransfer leaves tau alone with -unpol
                            A[i, j] = A[nb[0]]
                    elif len(nb) == 2:
                        A[i, j] = .5 * (A[nb[0]] + A[nb[1]])
                    else:
                        A[i, j] = .25 * (A[nb[0]] + A[nb[1]] + A[nb[2]] + A[nb[3]])
    assert np.array_equal(a["image"], img[:32, :32])
    assert np.array_equal(a["taus"], tau[:32, :32])
    if not unpol:
        assert np.array_equal(a["imageS"], S[:32, :32])
    assert a["interp"].sum() == 32 * 32 - 64


def test_iavg_is_area_weighted_mean(case):
    orc, p, ref = case
    a = orc.render_adaptive(p, 8, refine_abs=1e300, refine_rel=1e300)
    sp, n = 4, 33
    w1 = np.array([sp // 2 + 1] + [sp] * 7 + [sp // 2], float)
    assert w1.sum() == n
    cam = orc.camera(p)
    pix = np.array([(i, j) for i in range(0, n, sp) for j in range(0, n, sp)], np.int32)
    rc, out7, _, _ = orc.render_pixels(p, cam, pix)
    I = out7[:, 0].reshape(9, 9)
    want = (np.outer(w1, w1) * I).sum() * F3 / (n * n)
    assert a["Iavg"] == pytest.approx(want, rel=1e-14)


def test_working_tolerances(case):
    orc, p, ref = case
    ra, rr = 1e-6, 1e-2
    a = orc.render_adaptive(p, 8, refine_abs=ra, refine_rel=rr)
    assert 0 < a["traced"][1] < 208 and 0 < a["traced"][2] < 800
    img, interp = a["image"], a["interp"]
    traced = interp == 0
    for k in ("image", "taus"):
        assert np.array_equal(a[k][traced], ref[k][traced])
    assert np.array_equal(a["imageS"][traced], ref["imageS"][traced])
    # every decided pixel inside the visible 32 x 32 part, against the rule
    n, sp = 32, 4
    np.seterr(divide="ignore", invalid="ignore")
    for i in range(n):
        for j in range(n):
            lev = level_of(i, j, sp)
            if lev == 0:
                continue
            s = sp >> lev
            if i % (2 * s) == 0:
                nb = [(i, j - s), (i, j + s)]
            elif j % (2 * s) == 0:
                nb = [(i - s, j), (i + s, j)]
            else:
                nb = [(i - s, j - s), (i + s, j - s), (i - s, j + s), (i + s, j + s)]
            if max(max(q) for q in nb) >= n:
                continue                          # a neighbour on the internal extra row/column
            I1, I2 = img[nb[0]], img[nb[1]]
            if len(nb) == 2:
                ea, er = (I2 - I1) / 2 / a["Iavg"], (I2 - I1) / 2 / I1
                interp_val = .5 * (I1 + I2)
            else:
                I3 = img[nb[2]]
                ea, er = ((I2 + I3) / 2 - I1) / a["Iavg"], (I2 + I3) / (2 * I1) - 1.
                interp_val = .25 * (I1 + I2 + I3 + img[nb[3]])
            refine = abs(ea) > ra and abs(er) > rr and abs(I1) > 0.
            assert refine == (interp[i, j] == 0), (i, j)
            if not refine:
                assert img[i, j] == interp_val, (i, j)


def test_bad_sizes(case):
    orc, _, _ = case
    p = O.default_params(nx=30, ny=30, thetacam=163., freqcgs=230e9)
    assert orc.render_adaptive(p, 8)["rc"] == -2           # 31 - 1 is not 8 x 2^k
