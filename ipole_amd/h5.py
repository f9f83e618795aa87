"""ctypes binding of the host HDF5 I/O (lib/libipole_h5.so, host/h5io.c):
iharm-format dumps and the reference's image file layout."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libipole_h5.so")
dp = C.POINTER(C.c_double)
D4 = C.c_double * 4


class DumpHeader(C.Structure):
    _fields_ = [("metric", C.c_char * 32), ("N1", C.c_int), ("N2", C.c_int), ("N3", C.c_int),
                ("n_prim", C.c_int), ("gam", C.c_double), ("a", C.c_double), ("hslope", C.c_double),
                ("Rin", C.c_double), ("Rout", C.c_double), ("poly_xt", C.c_double),
                ("poly_alpha", C.c_double), ("mks_smooth", C.c_double), ("startx", D4), ("dx", D4),
                ("t", C.c_double), ("dump_cadence", C.c_double)]


class ImageMeta(C.Structure):
    """ipole_image_meta of host/h5io.h: what the image and trace headers hold."""
    _fields_ = ([("version", C.c_char * 32), ("githash", C.c_char * 32)]
                + [(k, C.c_int) for k in ("nx", "ny", "only_unpolarized", "qu_conv", "model_kind",
                                          "reverse_field")]
                + [(k, C.c_double) for k in ("freqcgs", "scale", "dsource", "dx", "dy", "fovx_dsource",
                                             "fovy_dsource", "rcam", "thetacam", "phicam", "rotcam",
                                             "fovx", "fovy", "xoff", "yoff")]
                + [("Xcam", D4)]
                + [(k, C.c_double) for k in ("t", "sigma_cut", "trat_small", "trat_large", "beta_crit",
                                             "L_unit", "M_unit", "T_unit", "Thetae_unit", "Mdot",
                                             "MdotEdd", "Ladv")])


class KharmaInfo(C.Structure):
    """ipole_kharma_info of host/h5io.h."""
    _fields_ = [("format", C.c_char * 8), ("num_meshblocks", C.c_int), ("mb", C.c_int * 3), ("nghost", C.c_int),
                ("has_electrons", C.c_int), ("game", C.c_double), ("gamp", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("ipole_amd/lib/libipole_h5.so is not built (make -C ipole_amd)")
        L = C.CDLL(LIBPATH)
        L.ipole_h5_write_dump.argtypes = [C.c_char_p, C.POINTER(DumpHeader), dp]
        L.ipole_h5_read_dump_header.argtypes = [C.c_char_p, C.POINTER(DumpHeader)]
        L.ipole_h5_read_dump_prims.argtypes = [C.c_char_p, C.POINTER(DumpHeader), dp]
        L.ipole_h5_read_image.argtypes = [C.c_char_p, C.c_int, C.c_int, dp, dp, dp, dp, dp]
        L.ipole_h5_trace_open.argtypes = [C.c_char_p, C.POINTER(ImageMeta), C.c_int, C.c_int, C.c_int]
        L.ipole_h5_trace_open.restype = C.c_int64
        L.ipole_h5_trace_write.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, dp]
        L.ipole_h5_trace_close.argtypes = [C.c_int64]
        L.ipole_h5_dataset_shape.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int64)]
        L.ipole_h5_read_doubles.argtypes = [C.c_char_p, C.c_char_p, dp]
        L.ipole_h5_read_kharma_header.argtypes = [C.c_char_p, C.POINTER(DumpHeader), C.POINTER(KharmaInfo)]
        L.ipole_h5_read_kharma_prims.argtypes = [C.c_char_p, C.POINTER(DumpHeader), C.POINTER(KharmaInfo), dp]
        L.ipole_h5_write_kharma.argtypes = [C.c_char_p, C.POINTER(DumpHeader), C.POINTER(KharmaInfo), C.c_char_p,
                                            C.POINTER(C.c_int), dp]
        _lib = L
    return _lib


def header_struct(h):
    s = DumpHeader()
    s.metric = h["metric"].encode()
    s.N1, s.N2, s.N3, s.n_prim = h["N1"], h["N2"], h["N3"], h.get("n_prim", 8)
    for k in ("gam", "a", "hslope", "Rin", "Rout", "t", "dump_cadence"):
        setattr(s, k, h[k])
    for k in ("poly_xt", "poly_alpha", "mks_smooth"):
        setattr(s, k, h.get(k, 0.))
    for mu in range(4):
        s.startx[mu] = h["startx"][mu]
        s.dx[mu] = h["dx"][mu]
    return s


def write_dump(path, h, prims):
    pr = np.ascontiguousarray(prims, dtype=np.float64)
    s = header_struct(h)
    s.n_prim = pr.shape[-1]
    if lib().ipole_h5_write_dump(path.encode(), C.byref(s), pr.ctypes.data_as(dp)):
        raise IOError("could not write dump %s" % path)


def read_dump(path):
    s = DumpHeader()
    if lib().ipole_h5_read_dump_header(path.encode(), C.byref(s)):
        raise IOError("could not read dump header %s" % path)
    h = dict(metric=s.metric.decode(), N1=s.N1, N2=s.N2, N3=s.N3, n_prim=s.n_prim, gam=s.gam, a=s.a,
             hslope=s.hslope, Rin=s.Rin, Rout=s.Rout, poly_xt=s.poly_xt, poly_alpha=s.poly_alpha,
             mks_smooth=s.mks_smooth, startx=list(s.startx), dx=list(s.dx), t=s.t,
             dump_cadence=s.dump_cadence)
    prims = np.zeros((s.N1, s.N2, s.N3, s.n_prim))
    if lib().ipole_h5_read_dump_prims(path.encode(), C.byref(s), prims.ctypes.data_as(dp)):
        raise IOError("could not read prims %s" % path)
    return h, prims


def read_image(path, nx, ny, polarized=True):
    image, taus = np.zeros((nx, ny)), np.zeros((nx, ny))
    imageS = np.zeros((nx, ny, 5)) if polarized else None
    Ft, Fu = C.c_double(0.), C.c_double(0.)
    rc = lib().ipole_h5_read_image(path.encode(), nx, ny, image.ctypes.data_as(dp),
                                   imageS.ctypes.data_as(dp) if polarized else None,
                                   taus.ctypes.data_as(dp), C.byref(Ft), C.byref(Fu))
    if rc:
        raise IOError("could not read image %s" % path)
    return dict(image=image, imageS=imageS, taus=taus, Ftot=Ft.value, Ftot_unpol=Fu.value)


def dataset_shape(path, name):
    """Dims of a dataset, or None when it is missing."""
    d = (C.c_int64 * 16)()
    r = lib().ipole_h5_dataset_shape(path.encode(), name.encode(), d)
    return None if r < 0 else tuple(int(d[k]) for k in range(r))


def read_dataset(path, name):
    """Any numeric dataset as float64."""
    shape = dataset_shape(path, name)
    if shape is None:
        raise KeyError(name)
    out = np.zeros(shape)
    if lib().ipole_h5_read_doubles(path.encode(), name.encode(), out.ctypes.data_as(dp)):
        raise IOError("could not read %s from %s" % (name, path))
    return out


def write_trace(path, meta, nx, ny, maxnstep, pixels):
    """Trace file of src/io.c:279-543: pixels is a list of (i, j, nstep,
    records) with records as Context.trace_pixel returns them; meta an
    ImageMeta for the header (None: zeros).  An existing file is appended
    to, as the reference does."""
    L = lib()
    meta = meta if meta is not None else ImageMeta()
    f = L.ipole_h5_trace_open(path.encode(), C.byref(meta), nx, ny, maxnstep)
    if f < 0:
        raise IOError("could not open %s" % path)
    try:
        for i, j, ns, rec in pixels:
            r = np.ascontiguousarray(rec, dtype=np.float64)
            if r.shape[0] < ns + 1:
                raise ValueError("records shorter than nstep + 1")
            if L.ipole_h5_trace_write(f, i, j, ns, r.ctypes.data_as(dp)):
                raise IOError("could not write pixel %d %d of %s" % (i, j, path))
    finally:
        L.ipole_h5_trace_close(f)


def kharma_par(h, mb, electrons=None, nghost=4):
    """The parameter file text a KHARMA dump carries (<parthenon/mesh>,
    <parthenon/meshblock>, <coordinates>, <GRMHD>, <electrons> blocks) for
    a grid of the dump header h."""
    t = "fmks" if h["metric"] in ("FMKS", "MMKS") else "mks"
    N, x0, dx = (h["N1"], h["N2"], h["N3"]), h["startx"], h["dx"]
    lines = ["<parthenon/job>", "problem_id = torus", "", "<parthenon/mesh>", "refinement = none",
             "nghost = %d" % nghost]
    for d in range(3):
        lines += ["nx%d = %d" % (d + 1, N[d]), "x%dmin = %.17g" % (d + 1, x0[d + 1]),
                  "x%dmax = %.17g" % (d + 1, x0[d + 1] + N[d] * dx[d + 1])]
    lines += ["", "<parthenon/meshblock>"] + ["nx%d = %d" % (d + 1, mb[d]) for d in range(3)]
    lines += ["", "<coordinates>", "base = spherical_ks", "transform = %s" % t, "a = %.17g" % h["a"],
              "r_in = %.17g" % h["Rin"], "r_out = %.17g" % h["Rout"], "hslope = %.17g" % h["hslope"]]
    if t == "fmks":
        lines += ["mks_smooth = %.17g" % h["mks_smooth"], "poly_xt = %.17g" % h["poly_xt"],
                  "poly_alpha = %.17g" % h["poly_alpha"]]
    lines += ["", "<GRMHD>", "cfl = 0.9", "gamma = %.17g" % h["gam"], ""]
    if electrons is not None:
        lines += ["<electrons>", "on = true", "gamma_e = %.17g" % electrons[0], "gamma_p = %.17g" % electrons[1], ""]
    else:
        lines += ["<electrons>", "on = false", ""]
    return "\n".join(lines)


def write_kharma(path, h, prims, mb, order=None, fmt="new", electrons=None):
    """A synthetic KHARMA (Parthenon) dump of the layout load_kharma_data
    reads, from an iharm-style header h and primitives [N1][N2][N3][8];
    mb the meshblock size, order the stored blocks' grid positions."""
    pr = np.ascontiguousarray(prims, dtype=np.float64)
    s = header_struct(h)
    s.n_prim = 8
    ki = KharmaInfo()
    ki.format = fmt.encode()
    for d in range(3):
        ki.mb[d] = mb[d]
    oa = None
    if order is not None:
        oa = (C.c_int * len(order))(*[int(x) for x in order])
    if lib().ipole_h5_write_kharma(path.encode(), C.byref(s), C.byref(ki), kharma_par(h, mb, electrons).encode(),
                                   oa, pr.ctypes.data_as(dp)):
        raise IOError("could not write %s" % path)


def read_kharma(path, fmt="new"):
    """(header dict, primitives [N1][N2][N3][8], info dict) of a KHARMA dump."""
    s = DumpHeader()
    ki = KharmaInfo()
    ki.format = fmt.encode()
    L = lib()
    rc = L.ipole_h5_read_kharma_header(path.encode(), C.byref(s), C.byref(ki))
    if rc:
        raise IOError("could not read the KHARMA header of %s (%d)" % (path, rc))
    pr = np.zeros((s.N1, s.N2, s.N3, 8))
    if L.ipole_h5_read_kharma_prims(path.encode(), C.byref(s), C.byref(ki), pr.ctypes.data_as(dp)):
        raise IOError("could not read the KHARMA primitives of %s" % path)
    h = dict(metric=s.metric.decode(), N1=s.N1, N2=s.N2, N3=s.N3, n_prim=s.n_prim, gam=s.gam, a=s.a,
             hslope=s.hslope, Rin=s.Rin, Rout=s.Rout, poly_xt=s.poly_xt, poly_alpha=s.poly_alpha,
             mks_smooth=s.mks_smooth, startx=list(s.startx), dx=list(s.dx), t=s.t)
    info = dict(num_meshblocks=ki.num_meshblocks, mb=list(ki.mb), game=ki.game, gamp=ki.gamp,
                has_electrons=ki.has_electrons)
    return h, pr, info
