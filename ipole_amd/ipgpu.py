"""ctypes binding of libipgpu.so (include/ipgpu.h) and the host-side set-up
the reference's main.c does before its pixel loop.

The C-ABI is the product boundary; this module only marshals arguments.  It
raises if the HIP library is not built: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libipgpu.so")

D4 = C.c_double * 4
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)

METRIC_MKS, METRIC_FMKS, METRIC_EMINKOWSKI = 0, 2, 5
MODEL_IHARM, MODEL_SPHERE = 1, 2
NFIELDS = 10
ERR_ARG, ERR_HIP, ERR_NAN, ERR_NOMEM, ERR_STATE, ERR_MAXNSTEP = -1, -2, -3, -4, -5, -10

# constants of src/constants.h
CL, ME, MP, HPL, GNEWT = 2.99792458e10, 9.1093826e-28, 1.67262171e-24, 6.6260693e-27, 6.6742e-8
MSUN, PC, JY, MUAS_PER_RAD, SIGMA_THOMSON = 1.989e33, 3.085678e18, 1.e-23, 2.06265e11, 0.665245873e-24


class Geom(C.Structure):
    _fields_ = [("metric", C.c_int), ("N1", C.c_int), ("N2", C.c_int), ("N3", C.c_int),
                ("a", C.c_double), ("hslope", C.c_double), ("poly_xt", C.c_double),
                ("poly_alpha", C.c_double), ("mks_smooth", C.c_double), ("poly_norm", C.c_double),
                ("startx", D4), ("stopx", D4), ("dx", D4), ("cstartx", D4), ("cstopx", D4),
                ("Rin", C.c_double), ("Rout", C.c_double), ("Rh", C.c_double),
                ("rmin_geo", C.c_double), ("rmax_geo", C.c_double)]


class Model(C.Structure):
    _fields_ = [("kind", C.c_int), ("L_unit", C.c_double), ("th_beg", C.c_double),
                ("sigma_cut", C.c_double), ("sigma_cut_high", C.c_double), ("sph_R0", C.c_double),
                ("sph_Ne0", C.c_double), ("sph_Thetae0", C.c_double), ("sph_B0", C.c_double),
                ("sph_sigma", C.c_double), ("sph_beta", C.c_double)]


class Units(C.Structure):
    _fields_ = [("M_unit", C.c_double), ("MBH_solar", C.c_double), ("trat_small", C.c_double),
                ("trat_large", C.c_double), ("beta_crit", C.c_double), ("gam", C.c_double),
                ("game", C.c_double), ("gamp", C.c_double)]


class Image(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("Xcam", D4), ("fovx", C.c_double),
                ("fovy", C.c_double), ("freq", C.c_double), ("freqcgs", C.c_double),
                ("rotcam", C.c_double), ("xoff", C.c_double), ("yoff", C.c_double),
                ("eps", C.c_double), ("maxnstep", C.c_int), ("old_centering", C.c_int),
                ("only_unpolarized", C.c_int), ("qu_conv", C.c_int), ("stokes_floors", C.c_int),
                ("target_nturns", C.c_int), ("subring_dtheta", C.c_int), ("counterjet", C.c_int),
                ("emission_type", C.c_int), ("transfer_fp32", C.c_int)]


class Refine(C.Structure):
    """ipgpu_refine: adaptive refinement (src/par.c:77-84 defaults)."""
    _fields_ = [("nx_min", C.c_int), ("ny_min", C.c_int), ("refine_abs", C.c_double),
                ("refine_rel", C.c_double), ("refine_cut", C.c_double), ("nearest_neighbor", C.c_int)]


def make_refine(nx_min, ny_min=None, refine_abs=1e-5, refine_rel=1e-1, refine_cut=0., nearest_neighbor=0):
    r = Refine()
    r.nx_min, r.ny_min = nx_min, nx_min if ny_min is None else ny_min
    r.refine_abs, r.refine_rel, r.refine_cut, r.nearest_neighbor = refine_abs, refine_rel, refine_cut, nearest_neighbor
    return r


MAX_LEVELS = 16


class Job(C.Structure):
    _fields_ = [("image", Image), ("m_unit_scale", C.c_double), ("image_out", C.POINTER(C.c_double)),
                ("imageS_out", C.POINTER(C.c_double)), ("taus_out", C.POINTER(C.c_double)),
                ("nstep_out", C.POINTER(C.c_int)), ("flag_out", C.POINTER(C.c_int))]


# Trace output record (include/ipgpu.h IPGPU_TR_*): dataset name of the
# reference's trace file (src/io.c:445-520), offset in the record, per-step
# shape.
TRACE_NF = 126
TRACE_SETS = [
    ("prims", 0, (8,)), ("b", 8, ()), ("ne", 9, ()), ("thetae", 10, ()), ("nu", 11, ()),
    ("b_k_angle", 12, ()), ("sigma", 13, ()), ("beta", 14, ()), ("dl", 15, ()), ("r", 16, ()),
    ("th", 17, ()), ("phi", 18, ()), ("j_unpol", 19, ()), ("k_unpol", 20, ()), ("I_unpol", 21, ()),
    ("X", 22, (4,)), ("Kcon", 26, (4,)), ("j_inv", 30, (4,)), ("alpha_inv", 34, (4,)),
    ("rho_inv", 38, (4,)), ("stokes", 42, (4,)), ("e2Constructed", 46, (4,)),
    ("e2ConstructedCov", 50, (4,)), ("e2Transported", 54, (4,)), ("e2TransportedCov", 58, (4,)),
    ("ntetrad", 62, (4, 4)), ("ncoord", 78, (4, 4)), ("Econ", 94, (4, 4)), ("Ecov", 110, (4, 4)),
]


def trace_fields(rec):
    """Splits trace records (nsteps, TRACE_NF) into the named per-step arrays."""
    out = {}
    for name, off, shape in TRACE_SETS:
        w = int(np.prod(shape)) if shape else 1
        a = rec[:, off:off + w]
        out[name] = a.reshape((rec.shape[0],) + shape) if shape else a[:, 0]
    return out


class IpgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("ipgpu error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("ipole_amd/lib/libipgpu.so is not built: run "
                               "`python -c 'import __graft_entry__ as g; g.build()'` or make -C ipole_amd")
        L = C.CDLL(LIBPATH)
        vp = C.c_void_p
        L.ipgpu_create.argtypes = [C.POINTER(vp), C.c_int]
        L.ipgpu_destroy.argtypes = [vp]
        L.ipgpu_destroy.restype = None
        L.ipgpu_last_error.argtypes = [vp]
        L.ipgpu_last_error.restype = C.c_char_p
        L.ipgpu_set_geometry.argtypes = [vp, C.POINTER(Geom)]
        L.ipgpu_set_model.argtypes = [vp, C.POINTER(Model)]
        L.ipgpu_upload_fields.argtypes = [vp, C.POINTER(dp), C.c_size_t]
        L.ipgpu_upload_prims.argtypes = [vp, dp, C.c_int, C.POINTER(Units)]
        L.ipgpu_download_fields.argtypes = [vp, dp, C.c_size_t]
        L.ipgpu_render.argtypes = [vp, C.POINTER(Image), dp, dp, dp, ip, ip]
        L.ipgpu_fetch.argtypes = [vp, dp, dp, dp, ip, ip]
        L.ipgpu_bad_pixel.argtypes = [vp, ip, ip]
        L.ipgpu_render_batch.argtypes = [vp, C.c_int, C.POINTER(Job)]
        L.ipgpu_bad_job.argtypes = [vp, ip, ip, ip]
        L.ipgpu_set_batch_memory.argtypes = [vp, C.c_size_t]
        L.ipgpu_rerender_batch.argtypes = [vp, C.c_int, C.POINTER(Job)]
        L.ipgpu_set_batch_options.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ipgpu_autotune_batch.argtypes = [vp, C.c_int, C.POINTER(Job), ip]
        L.ipgpu_autotune_render.argtypes = [vp, C.POINTER(Image), ip]
        L.ipgpu_batch_counts.argtypes = [vp, C.POINTER(C.c_longlong)]
        L.ipgpu_set_render_memory.argtypes = [vp, C.c_size_t]
        L.ipgpu_set_gather_staging.argtypes = [vp, C.c_int]
        L.ipgpu_render_noalloc.argtypes = [vp, C.POINTER(Image), vp, vp, vp, vp, vp]
        L.ipgpu_render_info.argtypes = [vp, ip, ip, C.POINTER(C.c_size_t)]
        L.ipgpu_adjoint.argtypes = [vp, C.POINTER(Image), dp, dp, dp, dp, dp]
        L.ipgpu_backward.argtypes = [vp, C.POINTER(Image), dp, dp, dp, dp]
        L.ipgpu_scale_density.argtypes = [vp, dp]
        L.ipgpu_kernel_times.argtypes = [vp, dp, dp, dp]
        L.ipgpu_stage_times.argtypes = [vp, dp]
        L.ipgpu_set_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ipgpu_step_counts.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.ipgpu_stream.argtypes = [vp]
        L.ipgpu_stream.restype = vp
        L.ipgpu_native_coord.argtypes = [C.POINTER(Geom), C.c_double, C.c_double, C.c_double, dp]
        L.ipgpu_trace_pixel.argtypes = [vp, C.POINTER(Image), C.c_int, C.c_int, C.c_int, ip, dp]
        L.ipgpu_render_adaptive.argtypes = [vp, C.POINTER(Image), C.POINTER(Refine), dp, dp, dp, ip, ip, ip]
        L.ipgpu_adaptive_stats.argtypes = [vp, ip, dp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        _lib = L
    return _lib


def device_total_memory(device=0):
    """Total memory of a device in bytes, from the HIP runtime that
    libipgpu.so links (a process may also hold torch's own copy of the
    runtime; this asks the library's, whatever was initialised first)."""
    lib()
    hip = C.CDLL("libamdhip64.so.7")   # the copy the library loaded (same soname)
    hip.hipDeviceTotalMem.argtypes = [C.POINTER(C.c_size_t), C.c_int]
    n = C.c_size_t(0)
    rc = hip.hipDeviceTotalMem(C.byref(n), device)
    if rc != 0:
        raise RuntimeError("hipDeviceTotalMem failed (%d)" % rc)
    return n.value


def P(a):
    return a.ctypes.data_as(dp)


def IP(a):
    return a.ctypes.data_as(ip)


# ------------------------------------------------------------------ set-up

def iharm_geom(h, rmax_geo=100., rmin_geo=1.):
    """Coordinate globals of an iharm dump header, as init_iharm_grid sets them
    (model/iharm/model.c:979-1065) and init_model's horizon (:326)."""
    g = Geom()
    g.metric = METRIC_FMKS if h["metric"] in ("FMKS", "MMKS") else METRIC_MKS
    g.N1, g.N2, g.N3 = h["N1"], h["N2"], h["N3"]
    g.a, g.hslope = h["a"], h["hslope"]
    if g.metric == METRIC_FMKS:
        g.poly_xt, g.poly_alpha, g.mks_smooth = h["poly_xt"], h["poly_alpha"], h["mks_smooth"]
        g.poly_norm = 0.5 * np.pi * 1. / (1. + 1. / (g.poly_alpha + 1.) * 1. / g.poly_xt ** g.poly_alpha)
    g.Rin, g.Rout = h["Rin"], h["Rout"]
    for mu in range(1, 4):
        g.startx[mu] = h["startx"][mu]
        g.dx[mu] = h["dx"][mu]
    g.rmax_geo = min(rmax_geo, g.Rout)
    g.rmin_geo = max(rmin_geo, g.Rin)
    g.stopx[0] = 1.
    g.stopx[1] = g.startx[1] + g.N1 * g.dx[1]
    g.stopx[2] = g.startx[2] + g.N2 * g.dx[2]
    g.stopx[3] = g.startx[3] + g.N3 * g.dx[3]
    g.cstopx[1] = np.log(g.Rout)
    g.cstopx[2] = 1.0
    g.cstopx[3] = 2 * np.pi
    g.Rh = 1 + np.sqrt(1. - g.a * g.a)
    return g


def iharm_model(MBH=6.2e9, sigma_cut=1., sigma_cut_high=-1., th_beg=0.0174):
    m = Model()
    m.kind = MODEL_IHARM
    m.L_unit = GNEWT * MBH * MSUN / (CL * CL)
    m.th_beg, m.sigma_cut, m.sigma_cut_high = th_beg, sigma_cut, sigma_cut_high
    return m


def iharm_units(M_unit, MBH=6.2e9, trat_small=1., trat_large=40., beta_crit=1., gam=13. / 9):
    u = Units()
    u.M_unit, u.MBH_solar = M_unit, MBH
    u.trat_small, u.trat_large, u.beta_crit = trat_small, trat_large, beta_crit
    u.gam, u.game, u.gamp = gam, 1.333333333333333, 1.666666666666667
    return u


def sphere_setup(R0=100., tau0=1.e-5, Thetae0=10., beta0=20., gam=13. / 9, tp_over_te=3., MBH=4.1e6):
    """Isothermal sphere of model/sphere/model.c:49-117 (eMinkowski)."""
    m = Model()
    m.kind = MODEL_SPHERE
    m.L_unit = GNEWT * MBH * MSUN / (CL * CL)
    m.sph_R0 = R0
    m.sph_Ne0 = tau0 / SIGMA_THOMSON / R0 / m.L_unit
    theta_unit = MP / ME * (gam - 1.) / (1. + tp_over_te)
    m.sph_B0 = CL * np.sqrt(8 * np.pi * (gam - 1.) * (MP + ME) / beta0) * np.sqrt(m.sph_Ne0 * Thetae0) / np.sqrt(theta_unit)
    m.sph_Thetae0 = Thetae0
    m.sph_sigma = Thetae0 * tp_over_te / beta0
    m.sph_beta = beta0
    g = Geom()
    g.metric = METRIC_EMINKOWSKI
    g.N1, g.N2, g.N3 = 300, 100, 1
    g.dx[0], g.dx[1], g.dx[2], g.dx[3] = 0.1, R0 / 300, np.pi / 100, 2 * np.pi
    g.stopx[0] = 1.
    for mu in (1, 2, 3):
        g.stopx[mu] = g.startx[mu] + (g.N1, g.N2, g.N3)[mu - 1] * g.dx[mu]
    for mu in range(4):
        g.cstartx[mu] = g.startx[mu]
        g.cstopx[mu] = g.stopx[mu]
    g.rmin_geo = 0.
    g.rmax_geo = max(100., R0)
    return g, m


def native_coord(geom, r, th_deg, phi_deg):
    X = np.zeros(4)
    rc = lib().ipgpu_native_coord(C.byref(geom), r, th_deg, phi_deg, P(X))
    if rc:
        raise IpgpuError(rc, "native_coord")
    return X


def make_image(geom, L_unit, nx=160, ny=160, rcam=1000., thetacam=90., phicam=0., rotcam=0.,
               fovx_dsource=0., fovy_dsource=0., dx=0., dy=0., dsource=16.9e6, freqcgs=230e9,
               eps=0.01, maxnstep=50000, old_centering=0, only_unpolarized=0, qu_conv=0,
               stokes_floors=0, target_nturns=-1, subring_dtheta=0, counterjet=0, emission_type=4,
               transfer_fp32=0,
               xoff=0., yoff=0.):
    """Camera set-up of src/main.c:125-174.  Returns (Image, scale, info)."""
    im = Image()
    im.nx, im.ny = nx, ny
    X = native_coord(geom, rcam, thetacam, phicam)
    for mu in range(4):
        im.Xcam[mu] = X[mu]
    im.freq = freqcgs * HPL / (ME * CL * CL)
    im.freqcgs = freqcgs
    im.rotcam = rotcam * np.pi / 180.
    Dsource = dsource * PC
    fov_to_d = Dsource / L_unit / MUAS_PER_RAD
    if fovx_dsource != 0.0:
        fx, fy = fovx_dsource, fovy_dsource
    elif dx != 0.0:
        fx, fy = dx / fov_to_d, dy / fov_to_d
    else:
        fx = fy = 160.
    DX, DY = fx * fov_to_d, fy * fov_to_d
    im.fovx, im.fovy = DX / rcam, DY / rcam
    scale = (DX * L_unit / nx) * (DY * L_unit / ny) / (Dsource * Dsource) / JY
    im.xoff, im.yoff = xoff, yoff
    im.eps, im.maxnstep = eps, maxnstep
    im.old_centering, im.only_unpolarized, im.qu_conv = old_centering, only_unpolarized, qu_conv
    im.stokes_floors, im.target_nturns, im.subring_dtheta = stokes_floors, target_nturns, subring_dtheta
    im.counterjet, im.emission_type = counterjet, emission_type
    im.transfer_fp32 = transfer_fp32
    return im, scale, dict(DX=DX, DY=DY, fovx_dsource=fx, fovy_dsource=fy, Dsource=Dsource)


# ------------------------------------------------------------------ context

class Context:
    """One GPU context (ipgpu_create .. ipgpu_destroy)."""

    def __init__(self, device=0):
        self.L = lib()
        self.h = C.c_void_p()
        rc = self.L.ipgpu_create(C.byref(self.h), device)
        if rc:
            raise IpgpuError(rc, "ipgpu_create(device=%d)" % device)

    def _chk(self, rc):
        if rc:
            raise IpgpuError(rc, self.L.ipgpu_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.ipgpu_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_geometry(self, g):
        self.geom = g
        self._chk(self.L.ipgpu_set_geometry(self.h, C.byref(g)))

    def set_model(self, m):
        self.model = m
        self._chk(self.L.ipgpu_set_model(self.h, C.byref(m)))

    def upload_fields(self, fields10):
        """fields10: array-like of 10 ghosted (N1+2, N2+2, N3+2) float64 arrays."""
        arrs = [np.ascontiguousarray(f, dtype=np.float64) for f in fields10]
        ptrs = (dp * NFIELDS)(*[P(a) for a in arrs])
        self._chk(self.L.ipgpu_upload_fields(self.h, ptrs, arrs[0].size))

    def upload_prims(self, prims, units):
        pr = np.ascontiguousarray(prims, dtype=np.float64)
        self._chk(self.L.ipgpu_upload_prims(self.h, P(pr), pr.shape[-1], C.byref(units)))

    def download_fields(self):
        g = self.geom
        nc = (g.N1 + 2) * (g.N2 + 2) * (g.N3 + 2)
        out = np.zeros(11 * nc)
        self._chk(self.L.ipgpu_download_fields(self.h, P(out), nc))
        return out.reshape(11, g.N1 + 2, g.N2 + 2, g.N3 + 2)

    def render(self, image, fetch=True):
        """Renders one image.  With fetch, returns dict of host arrays in the
        reference layout; otherwise results stay on the device."""
        if not fetch:
            rc = self.L.ipgpu_render(self.h, C.byref(image), None, None, None, None, None)
            self._chk(rc)
            return None
        n = image.nx * image.ny
        out = dict(image=np.zeros(n), imageS=np.zeros(5 * n), taus=np.zeros(n),
                   nstep=np.zeros(n, np.int32), flag=np.zeros(n, np.int32))
        rc = self.L.ipgpu_render(self.h, C.byref(image), P(out["image"]), P(out["imageS"]),
                                 P(out["taus"]), IP(out["nstep"]), IP(out["flag"]))
        if rc == ERR_MAXNSTEP:
            i, j = C.c_int(), C.c_int()
            self.L.ipgpu_bad_pixel(self.h, C.byref(i), C.byref(j))
            raise IpgpuError(rc, "MaxNStep exceeded in pixel %d %d" % (i.value, j.value))
        self._chk(rc)
        nx, ny = image.nx, image.ny
        return dict(image=out["image"].reshape(nx, ny), imageS=out["imageS"].reshape(nx, ny, 5),
                    taus=out["taus"].reshape(nx, ny), nstep=out["nstep"].reshape(nx, ny),
                    flag=out["flag"].reshape(nx, ny))

    def _batch_call(self, fn, jobs, fetch):
        n = len(jobs)
        arr = (Job * n)()
        outs = []
        for k, (image, mscale) in enumerate(jobs):
            arr[k].image = image
            arr[k].m_unit_scale = mscale
            if fetch:
                npx = image.nx * image.ny
                o = dict(image=np.zeros(npx), imageS=np.zeros(5 * npx), taus=np.zeros(npx),
                         nstep=np.zeros(npx, np.int32), flag=np.zeros(npx, np.int32))
                arr[k].image_out, arr[k].imageS_out, arr[k].taus_out = P(o["image"]), P(o["imageS"]), P(o["taus"])
                arr[k].nstep_out, arr[k].flag_out = IP(o["nstep"]), IP(o["flag"])
                outs.append(o)
        rc = fn(self.h, n, arr)
        if rc == ERR_MAXNSTEP:
            jb, i, j = C.c_int(), C.c_int(), C.c_int()
            self.L.ipgpu_bad_job(self.h, C.byref(jb), C.byref(i), C.byref(j))
            raise IpgpuError(rc, "MaxNStep exceeded in job %d pixel %d %d" % (jb.value, i.value, j.value))
        self._chk(rc)
        if not fetch:
            return None
        res = []
        for (image, _), o in zip(jobs, outs):
            nx, ny = image.nx, image.ny
            res.append(dict(image=o["image"].reshape(nx, ny), imageS=o["imageS"].reshape(nx, ny, 5),
                            taus=o["taus"].reshape(nx, ny), nstep=o["nstep"].reshape(nx, ny),
                            flag=o["flag"].reshape(nx, ny)))
        return res

    def render_batch(self, jobs, fetch=True):
        """Batched render of [(Image, m_unit_scale), ...]; returns one dict per
        job as render() does (or None without fetch)."""
        return self._batch_call(self.L.ipgpu_render_batch, jobs, fetch)

    def rerender_batch(self, jobs, fetch=True):
        """The last batch again with new M_unit scales, [(Image, m_unit_scale),
        ...] with the same images in the same order: its geodesic records are
        reused and only the transfer runs."""
        return self._batch_call(self.L.ipgpu_rerender_batch, jobs, fetch)

    def autotune_batch(self, jobs):
        """Measure and keep the batch options for [(Image, m_unit_scale), ...]:
        returns dict(slot, queue, queue_waves, resident_waves_per_cu)."""
        n = len(jobs)
        arr = (Job * n)()
        for k, (image, mscale) in enumerate(jobs):
            arr[k].image = image
            arr[k].m_unit_scale = mscale
        ch = np.zeros(4, np.int32)
        self._chk(self.L.ipgpu_autotune_batch(self.h, n, arr, IP(ch)))
        return dict(slot=int(ch[0]), queue=int(ch[1]), queue_waves=int(ch[2]), resident_waves_per_cu=int(ch[3]))

    def batch_counts(self):
        """The last batch's transfer work: dict(items, item_steps,
        item_radiating_steps, job_items, job_steps)."""
        out = (C.c_longlong * 5)()
        self._chk(self.L.ipgpu_batch_counts(self.h, out))
        return dict(items=out[0], item_steps=out[1], item_radiating_steps=out[2], job_items=out[3],
                    job_steps=out[4])

    def autotune_render(self, image):
        """Measure and keep the streamed render's settings for this image:
        returns dict(trace_budget, trace_phases, seg_min_steps, seg_len)."""
        ch = np.zeros(4, np.int32)
        self._chk(self.L.ipgpu_autotune_render(self.h, C.byref(image), IP(ch)))
        return dict(trace_budget=int(ch[0]), trace_phases=int(ch[1]), seg_min_steps=int(ch[2]), seg_len=int(ch[3]))

    def set_batch_options(self, slot=-1, queue=-1, queue_waves=-1, rerender_graph=-1):
        """Batch launch options; -1 keeps the current value."""
        self._chk(self.L.ipgpu_set_batch_options(self.h, int(slot), int(queue), int(queue_waves),
                                                 int(rerender_graph)))

    def adjoint(self, image, target_imageS, weights=None, density=True):
        """Loss L = sum w_s (S_s - target_s)^2 over pixels and I Q U V, and its
        gradient: dict(loss, dM_unit, dR_high, dR_low, density (N1, N2, N3)
        dL/d(n_e scale) per interior cell or None, imageS)."""
        n = image.nx * image.ny
        tgt = np.ascontiguousarray(target_imageS, dtype=np.float64).reshape(-1)
        assert tgt.size == 5 * n
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
        out4 = np.zeros(4)
        g = self.geom
        dens = np.zeros(g.N1 * g.N2 * g.N3) if density else None
        S = np.zeros(5 * n)
        rc = self.L.ipgpu_adjoint(self.h, C.byref(image), P(tgt), P(w) if w is not None else None, P(out4),
                                  P(dens) if dens is not None else None, P(S))
        if rc == ERR_MAXNSTEP:
            raise IpgpuError(rc, "MaxNStep exceeded")
        self._chk(rc)
        return dict(loss=out4[0], dM_unit=out4[1], dR_high=out4[2], dR_low=out4[3],
                    density=dens.reshape(g.N1, g.N2, g.N3) if dens is not None else None,
                    imageS=S.reshape(image.nx, image.ny, 5))

    def backward(self, image, dL_dimageS, density=True):
        """Vector-Jacobian product: dL/dimageS (nx, ny, 5 or 4) of the image
        the current dump renders -> dict(dM_unit, dR_high, dR_low, density
        (N1, N2, N3) dL/d(n_e scale) at 1 or None, imageS)."""
        n = image.nx * image.ny
        cot = np.zeros((n, 5))
        c = np.asarray(dL_dimageS, dtype=np.float64).reshape(n, -1)
        cot[:, :4] = c[:, :4]
        cot = np.ascontiguousarray(cot).reshape(-1)
        out3 = np.zeros(3)
        g = self.geom
        dens = np.zeros(g.N1 * g.N2 * g.N3) if density else None
        S = np.zeros(5 * n)
        rc = self.L.ipgpu_backward(self.h, C.byref(image), P(cot), P(out3), P(dens) if dens is not None else None,
                                   P(S))
        if rc == ERR_MAXNSTEP:
            raise IpgpuError(rc, "MaxNStep exceeded")
        self._chk(rc)
        return dict(dM_unit=out3[0], dR_high=out3[1], dR_low=out3[2],
                    density=dens.reshape(g.N1, g.N2, g.N3) if dens is not None else None,
                    imageS=S.reshape(image.nx, image.ny, 5))

    def scale_density(self, scale):
        """n_e of each interior cell times scale (N1, N2, N3)."""
        s = np.ascontiguousarray(scale, dtype=np.float64).reshape(-1)
        self._chk(self.L.ipgpu_scale_density(self.h, P(s)))

    def set_batch_memory(self, nbytes):
        self._chk(self.L.ipgpu_set_batch_memory(self.h, int(nbytes)))

    def set_render_memory(self, nbytes):
        """Device memory a single-image render may use for ray records and
        step packets; larger images render in chunks of rays."""
        self._chk(self.L.ipgpu_set_render_memory(self.h, int(nbytes)))

    def set_gather_staging(self, on):
        """Packet kernels read the interpolation's dump cells staged in LDS
        per 8-step group (True) or straight from the dump (False, default)."""
        self._chk(self.L.ipgpu_set_gather_staging(self.h, int(bool(on))))

    def render_info(self):
        """(chunks, continued rays, device bytes held) of the last render."""
        a, b, c = C.c_int(), C.c_int(), C.c_size_t()
        self._chk(self.L.ipgpu_render_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def kernel_times(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self._chk(self.L.ipgpu_kernel_times(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def stage_times(self):
        """ms of the last render's stages: k_trace, k_scan + k_prep,
        k_seg + k_chain, k_finalize after a one-pass render; the trace
        launches, the span of the finished rays' transfer, the tail after the
        last trace launch, k_finalize after a streamed one."""
        return self.stage_times_mode()[0]

    def stage_times_mode(self):
        """(stage_times(), True if the last render was streamed)."""
        ms = (C.c_double * 4)()
        rc = self.L.ipgpu_stage_times(self.h, ms)
        self._chk(min(rc, 0))
        return tuple(ms), rc == 1

    def set_tuning(self, trace_budget=-1, trace_phases=-1, seg_min_steps=-1, seg_len=-1, record_steps=-1):
        """Streamed-render trace budget (0 = one pass) and number of bounded
        phases, the ray length above which polarized transfer is composed
        from segments, the segment length and the geodesic records held per
        ray; -1 keeps a setting."""
        self._chk(self.L.ipgpu_set_tuning(self.h, int(trace_budget), int(trace_phases), int(seg_min_steps),
                                          int(seg_len), int(record_steps)))

    def render_adaptive(self, image, refine):
        """Adaptive refinement (src/main.c:505-792): dict of the nx x ny
        image arrays as render() returns them plus interp (1 where a pixel
        was interpolated) and the stats of adaptive_stats()."""
        nx, ny = image.nx, image.ny
        n = nx * ny
        out = dict(image=np.zeros(n), imageS=np.zeros(5 * n), taus=np.zeros(n),
                   nstep=np.zeros(n, np.int32), flag=np.zeros(n, np.int32), interp=np.zeros(n, np.int32))
        rc = self.L.ipgpu_render_adaptive(self.h, C.byref(image), C.byref(refine), P(out["image"]),
                                          P(out["imageS"]), P(out["taus"]), IP(out["nstep"]), IP(out["flag"]),
                                          IP(out["interp"]))
        if rc == ERR_MAXNSTEP:
            i, j = C.c_int(), C.c_int()
            self.L.ipgpu_bad_pixel(self.h, C.byref(i), C.byref(j))
            raise IpgpuError(rc, "MaxNStep exceeded in pixel %d %d" % (i.value, j.value))
        self._chk(rc)
        res = {k: (v.reshape(nx, ny, 5) if k == "imageS" else v.reshape(nx, ny)) for k, v in out.items()}
        res.update(self.adaptive_stats())
        return res

    def adaptive_stats(self):
        lv, ia = C.c_int(), C.c_double()
        tr, it = (C.c_longlong * MAX_LEVELS)(), (C.c_longlong * MAX_LEVELS)()
        self._chk(self.L.ipgpu_adaptive_stats(self.h, C.byref(lv), C.byref(ia), tr, it))
        return dict(levels=lv.value, Iavg=ia.value, traced=list(tr)[:lv.value], interpolated=list(it)[:lv.value])

    def trace_pixel(self, image, i, j):
        """Trace output of pixel (i, j): (nstep, records) with records of
        shape (nstep + 1, TRACE_NF), the values the reference's trace file
        holds for the pixel (src/io.c:279-543); trace_fields names them."""
        cap = image.maxnstep
        rec = np.zeros((cap, TRACE_NF))
        ns = C.c_int()
        rc = self.L.ipgpu_trace_pixel(self.h, C.byref(image), int(i), int(j), cap, C.byref(ns), P(rec))
        self._chk(rc)
        return ns.value, rec[:ns.value + 1].copy()

    def step_counts(self):
        a, b = C.c_longlong(), C.c_longlong()
        self._chk(self.L.ipgpu_step_counts(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def stream(self):
        return self.L.ipgpu_stream(self.h)


def image_stats(res, scale, freqcgs, dsource_pc, only_unpolarized=False):
    """The summary print_image_stats computes (src/main.c:1001-1042)."""
    img, S = res["image"], res["imageS"]
    Ftot_unpol = float(np.sum(img * scale))
    out = dict(Ftot_unpol=Ftot_unpol)
    if not only_unpolarized:
        Ftot = float(np.sum(S[..., 0] * scale))
        Q, U, V = (float(np.sum(S[..., k] * scale)) for k in (1, 2, 3))
        out.update(Ftot=Ftot, Qtot=Q, Utot=U, Vtot=V,
                   LP=100. * np.sqrt(Q * Q + U * U) / Ftot, CP=100. * V / Ftot,
                   nuLnu=4. * np.pi * Ftot * (dsource_pc * PC) ** 2 * JY * freqcgs)
    return out
