"""The ipole host executable (ipole_amd/bin/ipole) end to end: par file and
--key=value parsing, dump loading from HDF5, the pixel loop on the GPU,
the stderr summary line the reference's Python harness parses
(scripts/ipole.py:127-130, printed at src/main.c:1034) and the HDF5 image
layout (src/io.c:229-273), against the oracle."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from helpers import small_dump
from ipole_amd import h5

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ipole_amd", "bin", "ipole")


def run(args, cwd):
    print("running ipole %s" % " ".join(args), flush=True)
    r = subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stderr


def test_host_iharm(tmp_path):
    h, prims = small_dump()
    dump = str(tmp_path / "dump.h5")
    h5.write_dump(dump, h, prims)
    par = tmp_path / "run.par"
    par.write_text("# test\nthetacam 163\nfreqcgs 230.e9\nnx 24\nny 24\nM_unit 3e26\nMBH 6.2e9\n"
                   "trat_small 1\ntrat_large 40\ndump %s\noutfile %s\n" % (dump, tmp_path / "img.h5"))
    rc, err = run(["-par", str(par), "--thetacam=150"], cwd=str(tmp_path))
    assert rc == 0, err[-3000:]
    m = re.search(r"freq: (\S+) Ftot: (\S+) Jy \((\S+) Jy unpol xfer\)", err)
    assert m, err[-2000:]
    img = h5.read_image(str(tmp_path / "img.h5"), 24, 24)
    orc, _ = O.iharm_oracle(h, prims, M_unit=3e26)
    ref = orc.render(O.default_params(nx=24, ny=24, thetacam=150., freqcgs=230e9))
    sc = ref["cam"].scale
    F_ref = ref["imageS"][..., 0].sum() * sc
    assert abs(img["Ftot"] / F_ref - 1) < 1e-6
    assert abs(float(m.group(2)) / F_ref - 1) < 1e-5   # printed with %g
    assert np.allclose(img["imageS"], ref["imageS"], rtol=1e-5, atol=1e-6 * np.abs(ref["imageS"][..., 0]).max())
    assert np.allclose(img["image"], ref["image"], rtol=1e-5, atol=1e-6 * ref["image"].max())


def test_host_iharm_mixed_precision(tmp_path):
    """The host with the library's mixed-precision transfer (--transfer_fp32=1:
    FP32 fits, Stokes step and cell records) against the oracle's FP64
    image: Ftot within 1e-5 and the Stokes images within the mode's 1e-4."""
    h, prims = small_dump()
    dump = str(tmp_path / "dump.h5")
    h5.write_dump(dump, h, prims)
    rc, err = run(["--dump=%s" % dump, "--M_unit=3e26", "--freqcgs=230e9", "--thetacam=150", "--nx=24", "--ny=24",
                   "--transfer_fp32=1", "--outfile=%s" % (tmp_path / "mixed.h5")], cwd=str(tmp_path))
    assert rc == 0, err[-3000:]
    img = h5.read_image(str(tmp_path / "mixed.h5"), 24, 24)
    orc, _ = O.iharm_oracle(h, prims, M_unit=3e26)
    ref = orc.render(O.default_params(nx=24, ny=24, thetacam=150., freqcgs=230e9))
    F_ref = ref["imageS"][..., 0].sum() * ref["cam"].scale
    assert abs(img["Ftot"] / F_ref - 1) < 1e-5
    assert np.allclose(img["imageS"], ref["imageS"], rtol=1e-4, atol=1e-4 * np.abs(ref["imageS"][..., 0]).max())


def test_host_unpol_quench_and_maxnstep(tmp_path):
    h, prims = small_dump()
    dump = str(tmp_path / "dump.h5")
    h5.write_dump(dump, h, prims)
    rc, err = run(["--dump=%s" % dump, "--M_unit=3e26", "--freqcgs=230e9", "--nx=16", "--ny=16",
                   "-unpol", "-quench", "--outfile=never.h5"], cwd=str(tmp_path))
    assert rc == 0, err[-2000:]
    assert not (tmp_path / "never.h5").exists()
    assert "Jy unpol xfer" in err
    rc, err = run(["--dump=%s" % dump, "--M_unit=3e26", "--freqcgs=230e9", "--nx=8", "--ny=8",
                   "--maxnstep=40", "-quench"], cwd=str(tmp_path))
    assert rc == 256 - 10, (rc, err[-1000:])          # exit(-10), src/main.c:841-845
    assert "MaxNStep exceeded in pixel" in err


def test_host_sphere(tmp_path):
    rc, err = run(["--model=sphere", "--R0=50", "--tau0=2.0143e-5", "--Thetae0=10", "--beta0=30",
                   "--dx=100", "--dy=100", "--thetacam=60", "--freqcgs=230e9", "--nx=16", "--ny=16",
                   "--outfile=%s" % (tmp_path / "s.h5")], cwd=str(tmp_path))
    assert rc == 0, err[-2000:]
    img = h5.read_image(str(tmp_path / "s.h5"), 16, 16)
    g, m = O.sphere_model(R0=50., tau0=2.0143e-5, Thetae0=10., beta0=30.)
    ref = O.Oracle(g, m).render(O.default_params(nx=16, ny=16, thetacam=60., dx=100., dy=100., freqcgs=230e9))
    F_ref = ref["imageS"][..., 0].sum() * ref["cam"].scale
    assert abs(img["Ftot"] / F_ref - 1) < 1e-6


@pytest.mark.parametrize("fmt,electrons", [("new", None), ("old", (1.4, 1.6))])
def test_host_kharma(tmp_path, fmt, electrons):
    """--model=kharma: a synthetic KHARMA file (shuffled meshblocks) through
    the host's reader and the device derivation, against the oracle on the
    primitives and indices the reader returns."""
    h, prims = small_dump()
    dump = str(tmp_path / "k.h5")
    order = np.random.default_rng(3).permutation(32)
    h5.write_kharma(dump, h, prims, (12, 8, 16), order=order, fmt=fmt, electrons=electrons)
    rc, err = run(["--model=kharma", "--kharma_format=%s" % fmt, "--dump=%s" % dump, "--M_unit=3e26",
                   "--freqcgs=230e9", "--nx=24", "--ny=24", "--thetacam=150",
                   "--outfile=%s" % (tmp_path / "img.h5")], cwd=str(tmp_path))
    assert rc == 0, err[-3000:]
    img = h5.read_image(str(tmp_path / "img.h5"), 24, 24)
    hr, pr, info = h5.read_kharma(dump, fmt)
    orc, _ = O.iharm_oracle(hr, pr, M_unit=3e26, game=info["game"], gamp=info["gamp"])
    ref = orc.render(O.default_params(nx=24, ny=24, thetacam=150., freqcgs=230e9))
    F_ref = ref["imageS"][..., 0].sum() * ref["cam"].scale
    assert abs(img["Ftot"] / F_ref - 1) < 1e-6
    assert np.allclose(img["imageS"], ref["imageS"], rtol=1e-5, atol=1e-6 * np.abs(ref["imageS"][..., 0]).max())
    assert h5.read_dataset(str(tmp_path / "img.h5"), "/fluid_header/geom/fmks/a") == h["a"]
    assert h5.read_dataset(str(tmp_path / "img.h5"), "/fluid_header/nx1_mb") == 12
