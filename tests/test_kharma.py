"""The KHARMA dump reader of the host (host/h5io.c: model/kharma/model.c:
407-1010, 1251-1470): synthetic Parthenon files with shuffled meshblocks in
both layouts ("new"; "old", KHARMA v5), read back exactly, and read by the
reference's own KHARMA loader (oracle/_ref, where built) into the same
geometry and fluid as the oracle derives from this reader's primitives.
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import small_dump
from ipole_amd import h5

M_UNIT = 3e26
MB = (12, 8, 16)


def shuffled(h, mb, seed):
    nb = (h["N1"] // mb[0]) * (h["N2"] // mb[1]) * (h["N3"] // mb[2])
    return np.random.default_rng(seed).permutation(nb)


@pytest.mark.parametrize("fmt,electrons", [("new", None), ("old", None), ("new", (1.4, 1.6))])
def test_round_trip(tmp_path, fmt, electrons):
    h, prims = small_dump()
    path = str(tmp_path / "k.h5")
    h5.write_kharma(path, h, prims, MB, order=shuffled(h, MB, 1), fmt=fmt, electrons=electrons)
    hr, pr, info = h5.read_kharma(path, fmt)
    assert np.array_equal(pr, prims)
    assert info["num_meshblocks"] == (48 // 12) * (32 // 8) * (32 // 16) and info["mb"] == list(MB)
    assert (info["game"], info["gamp"]) == ((1.333333333333333, 1.666666666666667) if electrons is None
                                            else electrons)
    assert hr["metric"] == h["metric"] and (hr["N1"], hr["N2"], hr["N3"]) == (h["N1"], h["N2"], h["N3"])
    for k in ("gam", "a", "hslope", "Rin", "Rout", "poly_xt", "poly_alpha", "mks_smooth"):
        assert hr[k] == h[k], k
    for d in (1, 2, 3):
        assert hr["startx"][d] == h["startx"][d]
        assert hr["dx"][d] == pytest.approx(h["dx"][d], rel=1e-15)
    # the wrong layout does not read back the same field
    other = "old" if fmt == "new" else "new"
    _, pr2, _ = h5.read_kharma(path, other)
    assert not np.array_equal(pr2, prims)


def test_rejects_other_coordinates(tmp_path):
    h, prims = small_dump(N=(24, 16, 16))
    path = str(tmp_path / "k.h5")
    h5.write_kharma(path, h, prims, (12, 8, 16))
    # the parameter text with the eks transform
    import ctypes as C
    s = h5.header_struct(h)
    ki = h5.KharmaInfo()
    ki.format = b"new"
    ki.mb[0], ki.mb[1], ki.mb[2] = 12, 8, 16
    par = h5.kharma_par(h, (12, 8, 16)).replace("transform = fmks", "transform = eks")
    assert h5.lib().ipole_h5_write_kharma(path.encode(), C.byref(s), C.byref(ki), par.encode(), None,
                                          np.ascontiguousarray(prims).ctypes.data_as(h5.dp)) == 0
    r = h5.lib().ipole_h5_read_kharma_header(path.encode(), C.byref(h5.DumpHeader()), C.byref(h5.KharmaInfo()))
    assert r == -2


@pytest.mark.skipif(not O.ref_available("kharma"), reason="oracle/_ref not built")
@pytest.mark.parametrize("fmt,electrons", [("new", None), ("old", (1.4, 1.6))])
def test_reference_loader(tmp_path, fmt, electrons):
    """The reference's load_kharma_data on the synthetic file: the same
    geometry, and bit for bit the same four-vectors, n_e, Theta_e, B and
    sigma as the oracle derives from this reader's primitives."""
    h, prims = small_dump()
    path = str(tmp_path / "k.h5")
    h5.write_kharma(path, h, prims, MB, order=shuffled(h, MB, 2), fmt=fmt, electrons=electrons)
    hr, pr, info = h5.read_kharma(path, fmt)
    keys = dict(dump=path, M_unit=M_UNIT, MBH=6.2e9, trat_small=1, trat_large=40, kharma_format=fmt)
    g_ref, L_unit = O.ref_geom("kharma", keys)
    orc, _ = O.iharm_oracle(hr, pr, M_unit=M_UNIT, game=info["game"], gamp=info["gamp"])
    g = orc.geom
    assert g_ref.metric == g.metric and (g_ref.N1, g_ref.N2, g_ref.N3) == (g.N1, g.N2, g.N3)
    for name in ("a", "hslope", "poly_xt", "poly_alpha", "mks_smooth", "poly_norm", "Rin", "Rout", "Rh",
                 "rmax_geo", "rmin_geo"):
        assert getattr(g_ref, name) == getattr(g, name), name
    for name in ("startx", "stopx", "dx"):
        for mu in (1, 2, 3):
            assert getattr(g_ref, name)[mu] == getattr(g, name)[mu], (name, mu)
    orc.activate()
    rng = np.random.default_rng(0)
    n = 2000
    X = np.zeros((n, 4))
    X[:, 1] = rng.uniform(g.startx[1] - 0.05, g.stopx[1] + 0.05, n)
    X[:, 2] = rng.uniform(-0.02, 1.02, n)
    X[:, 3] = rng.uniform(-7, 13, n)
    K = rng.normal(size=(n, 4)) * 1e-9
    ref = O.ref_run("kharma", "fluid", np.hstack([X, K]), keys).reshape(n, 21)
    for q in range(n):
        mine = np.concatenate(orc.fourv(X[q], K[q]) + [np.array(orc.scalars(X[q]))])
        assert np.array_equal(mine, ref[q]), q
    assert ref[:, 16].max() > 0          # inside the grid: n_e > 0 somewhere
