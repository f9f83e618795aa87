"""ctypes wrapper of the CPU oracle (liboracle.so) and runner of the reference
driver build (oracle/_ref).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the ipole_amd package.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFDIR = os.path.join(HERE, "_ref")

D4 = C.c_double * 4
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)

METRIC_MKS, METRIC_FMKS, METRIC_MINKOWSKI, METRIC_EMINKOWSKI = 0, 2, 4, 5
MODEL_IHARM, MODEL_SPHERE, MODEL_ANALYTIC = 1, 2, 3

# constants of src/constants.h
EE, CL, ME, MP, HPL, GNEWT = 4.80320680e-10, 2.99792458e10, 9.1093826e-28, 1.67262171e-24, 6.6260693e-27, 6.6742e-8
MSUN, PC, JY, SIGMA_THOMSON = 1.989e33, 3.085678e18, 1.e-23, 0.665245873e-24


class Geom(C.Structure):
    _fields_ = [("metric", C.c_int), ("a", C.c_double), ("hslope", C.c_double),
                ("poly_xt", C.c_double), ("poly_alpha", C.c_double), ("mks_smooth", C.c_double),
                ("poly_norm", C.c_double), ("startx", D4), ("stopx", D4), ("dx", D4),
                ("cstartx", D4), ("cstopx", D4), ("Rin", C.c_double), ("Rout", C.c_double),
                ("Rh", C.c_double), ("rmin_geo", C.c_double), ("rmax_geo", C.c_double),
                ("N1", C.c_int), ("N2", C.c_int), ("N3", C.c_int)]


class Model(C.Structure):
    _fields_ = [("kind", C.c_int), ("L_unit", C.c_double), ("fields", dp),
                ("th_beg", C.c_double), ("sigma_cut", C.c_double), ("sigma_cut_high", C.c_double),
                ("sph_R0", C.c_double), ("sph_Ne0", C.c_double), ("sph_Thetae0", C.c_double),
                ("sph_B0", C.c_double), ("sph_sigma", C.c_double), ("sph_beta", C.c_double),
                ("an_model", C.c_int), ("an_A", C.c_double),
                ("an_alpha", C.c_double), ("an_height", C.c_double), ("an_l0", C.c_double),
                ("an_freqcgs", C.c_double), ("an_RHO_unit", C.c_double), ("an_B_unit", C.c_double)]


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("rcam", C.c_double), ("thetacam", C.c_double),
                ("phicam", C.c_double), ("rotcam", C.c_double), ("fovx_dsource", C.c_double),
                ("fovy_dsource", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("dsource", C.c_double), ("freqcgs", C.c_double), ("eps", C.c_double),
                ("maxnstep", C.c_int), ("old_centering", C.c_int), ("only_unpolarized", C.c_int),
                ("qu_conv", C.c_int), ("stokes_floors", C.c_int), ("target_nturns", C.c_int),
                ("subring_dtheta", C.c_int), ("counterjet", C.c_int), ("emission_type", C.c_int),
                ("xoff", C.c_double), ("yoff", C.c_double)]


class Camera(C.Structure):
    _fields_ = [("Xcam", D4), ("freq", C.c_double), ("fovx", C.c_double), ("fovy", C.c_double),
                ("DX", C.c_double), ("DY", C.c_double), ("scale", C.c_double),
                ("rotcam_rad", C.c_double), ("dsource_cm", C.c_double),
                ("bad_pixel_i", C.c_int), ("bad_pixel_j", C.c_int)]


class Units(C.Structure):
    _fields_ = [("M_unit", C.c_double), ("MBH_solar", C.c_double), ("trat_small", C.c_double),
                ("trat_large", C.c_double), ("beta_crit", C.c_double), ("gam", C.c_double),
                ("game", C.c_double), ("gamp", C.c_double)]


class Traj(C.Structure):
    _fields_ = [("nturns", C.c_int), ("dl", C.c_double), ("X", D4), ("Kcon", D4),
                ("Xhalf", D4), ("Kconhalf", D4)]


def default_params(**kw):
    """Defaults of src/par.c:43-101."""
    p = Params()
    p.nx = p.ny = 160
    p.rcam, p.thetacam, p.phicam, p.rotcam = 1000., 90., 0., 0.
    p.dsource = 16.9e6
    p.eps, p.maxnstep = 0.01, 50000
    p.emission_type = 4
    p.target_nturns = -1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load(LIB)
    return _lib


def load(path):
    """Loads and declares one build of the oracle library (each build keeps
    its own global state)."""
    if True:
        if not os.path.exists(path):
            raise RuntimeError("%s is not built (run make -C oracle)" % path)
        L = C.CDLL(path)
        L.or_set_geom.argtypes = [C.POINTER(Geom)]
        L.or_get_geom.argtypes = [C.POINTER(Geom)]
        L.or_set_model.argtypes = [C.POINTER(Model)]
        L.or_gcov.argtypes = [dp, dp]
        L.or_gcon.argtypes = [dp, dp]
        L.or_gcon.restype = C.c_int
        L.or_gdet.argtypes = [dp]
        L.or_gdet.restype = C.c_double
        L.or_connection.argtypes = [dp, dp]
        L.or_bl_coord.argtypes = [dp, dp, dp]
        L.or_native_coord.argtypes = [C.c_double] * 3 + [dp]
        L.or_trace.argtypes = [dp, dp, C.POINTER(Traj), C.c_double, C.c_int, dp, C.c_int]
        L.or_trace.restype = C.c_int
        L.or_init_XK.argtypes = [C.c_long, C.c_long, C.POINTER(Params), C.POINTER(Camera), dp, dp]
        L.or_plasma_tetrad.argtypes = [dp, dp, dp, dp, dp, dp]
        L.or_plasma_tetrad.restype = C.c_int
        L.or_camera_tetrad.argtypes = [dp, dp, dp]
        L.or_camera_tetrad.restype = C.c_int
        L.or_camera_tetrad_old.argtypes = [dp, dp, dp]
        L.or_camera_tetrad_old.restype = C.c_int
        L.or_radiating_region.argtypes = [dp]
        L.or_radiating_region.restype = C.c_int
        L.or_fourv.argtypes = [dp] * 6
        for f in ("or_ne", "or_thetae", "or_b", "or_sigma"):
            getattr(L, f).argtypes = [dp]
            getattr(L, f).restype = C.c_double
        L.or_iharm_derive.argtypes = [dp, C.c_int, C.POINTER(Units), dp]
        L.or_iharm_derive_ru.argtypes = [dp, C.c_int, C.POINTER(Units), dp, dp]
        L.or_trace_pixel.argtypes = [C.POINTER(Params), C.POINTER(Camera), C.c_long, C.c_long, dp, C.c_int,
                                     dp, ip]
        L.or_trace_pixel.restype = C.c_int
        L.or_integrate_emission.argtypes = [C.POINTER(Traj), C.c_int, dp, dp, dp, dp, C.POINTER(Params)]
        L.or_integrate_emission.restype = C.c_int
        L.or_render_adaptive.argtypes = [C.POINTER(Params), C.POINTER(Camera), C.c_int, C.c_int, C.c_double,
                                         C.c_double, C.c_double, C.c_int, dp, dp, dp, ip, dp,
                                         C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.c_int]
        L.or_render_adaptive.restype = C.c_int
        L.or_sphere_setup.argtypes = [C.c_double] * 7 + [C.POINTER(Model)]
        L.or_analytic_setup.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(Model), C.POINTER(Geom)]
        L.or_bessel_kn.argtypes = [C.c_int, C.c_double]
        L.or_bessel_kn.restype = C.c_double
        L.or_jar.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.or_approximate_solve.argtypes = [C.c_double] * 6 + [dp]
        L.or_approximate_solve.restype = C.c_double
        L.or_push_polar.argtypes = [dp, dp, dp, dp, dp, C.c_double]
        L.or_stokes_step.argtypes = [dp, C.c_double, dp]
        L.or_evolve_N.argtypes = [dp, dp, C.c_double, dp, dp, C.c_int, C.POINTER(Params)]
        L.or_evolve_N.restype = C.c_int
        L.or_project_N.argtypes = [dp, dp, dp, dp, dp, dp, C.c_double]
        L.or_setup_camera.argtypes = [C.POINTER(Params), C.POINTER(Camera)]
        L.or_setup_camera.restype = C.c_int
        L.or_render.argtypes = [C.POINTER(Params), C.POINTER(Camera), dp, dp, dp, ip, ip, C.c_int]
        L.or_render.restype = C.c_int
        L.or_render_pixels.argtypes = [C.POINTER(Params), C.POINTER(Camera), ip, C.c_int, dp, ip,
                                       ip, C.c_int]
        L.or_render_pixels.restype = C.c_int
        return L


def P(a):
    return a.ctypes.data_as(dp)


def IP(a):
    return a.ctypes.data_as(ip)


def f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class Oracle:
    """Holds the global geometry/model state of the oracle and keeps the
    field arrays alive."""

    def __init__(self, geom, model, fields=None, L=None):
        self.L = L or lib()
        self.geom = geom
        self.model = model
        self.fields = None if fields is None else f64(fields)
        self.rub = None  # ghosted rho, u, beta for the trace output
        if self.fields is not None:
            self.model.fields = P(self.fields)
        self.activate()

    def activate(self):
        self.L.or_set_geom(C.byref(self.geom))
        self.L.or_set_model(C.byref(self.model))

    def gcov(self, X):
        g = np.zeros((4, 4))
        self.L.or_gcov(P(f64(X)), P(g))
        return g

    def gcon(self, g):
        gc = np.zeros((4, 4))
        self.L.or_gcon(P(f64(g)), P(gc))
        return gc

    def connection(self, X):
        c = np.zeros((4, 4, 4))
        self.L.or_connection(P(f64(X)), P(c))
        return c

    def bl_coord(self, X):
        r, th = C.c_double(), C.c_double()
        self.L.or_bl_coord(P(f64(X)), C.byref(r), C.byref(th))
        return r.value, th.value

    def camera(self, params):
        cam = Camera()
        rc = self.L.or_setup_camera(C.byref(params), C.byref(cam))
        assert rc == 0
        return cam

    def init_XK(self, i, j, params, cam):
        X, K = np.zeros(4), np.zeros(4)
        self.L.or_init_XK(i, j, C.byref(params), C.byref(cam), P(X), P(K))
        return X, K * cam.freq

    def trace(self, X, K, eps, maxnstep, Xcam, subring_dtheta=0):
        traj = (Traj * maxnstep)()
        ns = self.L.or_trace(P(f64(X)), P(f64(K)), traj, eps, maxnstep, P(f64(Xcam)), subring_dtheta)
        arr = np.ctypeslib.as_array(C.cast(traj, C.POINTER(C.c_char)), shape=(C.sizeof(Traj) * (ns + 1),))
        out = np.zeros((ns + 1, 18))
        for s in range(ns + 1):
            t = traj[s]
            out[s] = [t.nturns, t.dl] + list(t.X) + list(t.Kcon) + list(t.Xhalf) + list(t.Kconhalf)
        del arr
        return ns, out

    def fourv(self, X, K):
        o = [np.zeros(4) for _ in range(4)]
        self.L.or_fourv(P(f64(X)), P(f64(K)), *[P(v) for v in o])
        return o

    def scalars(self, X):
        Xp = P(f64(X))
        return (self.L.or_ne(Xp), self.L.or_thetae(Xp), self.L.or_b(Xp), self.L.or_sigma(Xp),
                self.L.or_radiating_region(Xp))

    def jar(self, dist, pol, X, K):
        o = np.zeros(11)
        self.L.or_jar(dist, pol, P(f64(X)), P(f64(K)), P(o))
        return o

    def plasma_tetrad(self, U, K, B, g):
        Ec, El = np.zeros((4, 4)), np.zeros((4, 4))
        fl = self.L.or_plasma_tetrad(P(f64(U)), P(f64(K)), P(f64(B)), P(f64(g)), P(Ec), P(El))
        return Ec, El, fl

    def camera_tetrad(self, X, old=False):
        Ec, El = np.zeros((4, 4)), np.zeros((4, 4))
        f = self.L.or_camera_tetrad_old if old else self.L.or_camera_tetrad
        fl = f(P(f64(X)), P(Ec), P(El))
        return Ec, El, fl

    def push_polar(self, Xm, Km, Ni, Nm, dlam):
        Ni = np.ascontiguousarray(Ni, dtype=np.complex128)
        Nm = np.ascontiguousarray(Nm, dtype=np.complex128)
        Nf = np.zeros((4, 4), dtype=np.complex128)
        self.L.or_push_polar(P(f64(Xm)), P(f64(Km)), Ni.ctypes.data_as(dp), Nm.ctypes.data_as(dp),
                             Nf.ctypes.data_as(dp), dlam)
        return Nf

    def evolve_N(self, Xf, Kf, dlam, N, tauF, zero_emission, params):
        N = np.array(N, dtype=np.complex128)
        t = C.c_double(tauF)
        fl = self.L.or_evolve_N(P(f64(Xf)), P(f64(Kf)), dlam, N.ctypes.data_as(dp), C.byref(t),
                                zero_emission, C.byref(params))
        return N, t.value, fl

    def project_N(self, X, N, rotcam):
        N = np.ascontiguousarray(N, dtype=np.complex128)
        o = [C.c_double() for _ in range(4)]
        self.L.or_project_N(P(f64(X)), N.ctypes.data_as(dp), *[C.byref(v) for v in o], rotcam)
        return np.array([v.value for v in o])

    def integrate_emission(self, rows, nsteps, params, N=None):
        """or_integrate_emission over trajectory rows as trace() returns them
        (nturns, dl, X, Kcon, Xhalf, Kconhalf): steps rows[nsteps] -> rows[0].
        Returns (I, tau, tauF, N, flag)."""
        traj = (Traj * len(rows))()
        for s, r in enumerate(rows):
            traj[s].nturns = int(r[0])
            traj[s].dl = r[1]
            for mu in range(4):
                traj[s].X[mu], traj[s].Kcon[mu] = r[2 + mu], r[6 + mu]
                traj[s].Xhalf[mu], traj[s].Kconhalf[mu] = r[10 + mu], r[14 + mu]
        N = np.zeros((4, 4), np.complex128) if N is None else np.array(N, np.complex128)
        I_, tau, tauF = C.c_double(), C.c_double(), C.c_double()
        fl = self.L.or_integrate_emission(traj, nsteps, C.byref(I_), C.byref(tau), C.byref(tauF),
                                          N.ctypes.data_as(dp), C.byref(params))
        return I_.value, tau.value, tauF.value, N, fl

    def trace_pixel(self, params, cam, i, j):
        """Trace output of pixel (i, j) (src/io.c:279-543): (rc, nstep,
        records (nstep + 1, 126))."""
        cap = params.maxnstep
        out = np.zeros((cap, 126))
        ns = C.c_int()
        rub = None if self.rub is None else P(self.rub)
        rc = self.L.or_trace_pixel(C.byref(params), C.byref(cam), i, j, rub, cap, P(out), C.byref(ns))
        return rc, ns.value, out[:ns.value + 1].copy()

    def render(self, params, nthreads=0):
        n = params.nx * params.ny
        image, imageS, taus = np.zeros(n), np.zeros(5 * n), np.zeros(n)
        nsteps, flags = np.zeros(n, np.int32), np.zeros(n, np.int32)
        cam = Camera()
        rc = self.L.or_render(C.byref(params), C.byref(cam), P(image), P(imageS), P(taus),
                              IP(nsteps), IP(flags), nthreads)
        return dict(rc=rc, cam=cam, image=image.reshape(params.nx, params.ny),
                    imageS=imageS.reshape(params.nx, params.ny, 5), taus=taus.reshape(params.nx, params.ny),
                    nstep=nsteps.reshape(params.nx, params.ny), flag=flags.reshape(params.nx, params.ny))

    def render_adaptive(self, params, nx_min, ny_min=None, refine_abs=1e-5, refine_rel=1e-1, refine_cut=0.,
                        nearest=0, nthreads=0):
        """Adaptive refinement of src/main.c:505-792 (or_render_adaptive)."""
        nx, ny = params.nx, params.ny
        n = nx * ny
        image, imageS, taus = np.zeros(n), np.zeros(5 * n), np.zeros(n)
        interp = np.zeros(n, np.int32)
        Iavg = C.c_double()
        tr, it = (C.c_longlong * 16)(), (C.c_longlong * 16)()
        cam = Camera()
        rc = self.L.or_render_adaptive(C.byref(params), C.byref(cam), nx_min, nx_min if ny_min is None else ny_min,
                                       refine_abs, refine_rel, refine_cut, nearest, P(image), P(imageS), P(taus),
                                       IP(interp), C.byref(Iavg), tr, it, nthreads)
        lv = int(np.log2((nx - nx % 2) // (nx_min - nx_min % 2)) + 1)
        return dict(rc=rc, cam=cam, image=image.reshape(nx, ny), imageS=imageS.reshape(nx, ny, 5),
                    taus=taus.reshape(nx, ny), interp=interp.reshape(nx, ny), Iavg=Iavg.value, levels=lv,
                    traced=list(tr)[:lv], interpolated=list(it)[:lv])

    def render_pixels(self, params, cam, pix, nthreads=0):
        pix = np.ascontiguousarray(pix, dtype=np.int32)
        n = len(pix)
        out7 = np.zeros((n, 7))
        nsteps, flags = np.zeros(n, np.int32), np.zeros(n, np.int32)
        rc = self.L.or_render_pixels(C.byref(params), C.byref(cam), IP(pix), n, P(out7), IP(nsteps),
                                     IP(flags), nthreads)
        return rc, out7, nsteps, flags


def stokes_step(c11, dlam, S):
    S = np.array(S, dtype=np.float64)
    lib().or_stokes_step(P(f64(c11)), dlam, P(S))
    return S


def bessel_kn(n, x):
    return lib().or_bessel_kn(n, x)


def approximate_solve(Ii, ji, ki, jf, kf, dl):
    tau = C.c_double(0.)
    If = lib().or_approximate_solve(Ii, ji, ki, jf, kf, dl, C.byref(tau))
    return If, tau.value


def sphere_model(R0=50., tau0=2.0143e-5, Thetae0=10., beta0=30., gam=13. / 9, tp_over_te=3.,
                 MBH=4.1e6):
    """Isothermal sphere of model/sphere/model.c:49-117 in eMinkowski space."""
    m = Model()
    lib().or_sphere_setup(R0, tau0, Thetae0, beta0, gam, tp_over_te, MBH, C.byref(m))
    g = Geom()
    g.metric = METRIC_EMINKOWSKI
    g.N1, g.N2, g.N3 = 300, 100, 1
    g.dx[0], g.dx[1], g.dx[2], g.dx[3] = 0.1, R0 / 300, np.pi / 100, 2 * np.pi
    g.stopx[0] = 1.
    for mu in (1, 2, 3):  # model/sphere/model.c:103-106
        g.stopx[mu] = g.startx[mu] + (g.N1, g.N2, g.N3)[mu - 1] * g.dx[mu]
    for mu in range(4):
        g.cstartx[mu] = g.startx[mu]
        g.cstopx[mu] = g.stopx[mu]
    g.rmin_geo = 0.
    g.rmax_geo = max(100., R0)
    return g, m


def analytic_model(model, MBH=4.063e6, rho_unit=0., freqcgs=230e9):
    m, g = Model(), Geom()
    lib().or_analytic_setup(model, MBH, rho_unit, freqcgs, C.byref(m), C.byref(g))
    return g, m


# ---------------------------------------------------------------- reference driver

def ref_available(model="iharm"):
    return os.path.exists(os.path.join(REFDIR, "ref_driver_" + model))


def ref_run(model, cmd, records, keys=None, nout=None):
    """Run oracle/_ref/ref_driver_<model> <cmd> on float64 records (2-D array,
    one row per record).  Returns the flat float64 output."""
    exe = os.path.join(REFDIR, "ref_driver_" + model)
    records = np.atleast_2d(f64(records)) if records is not None and len(records) else np.zeros((0, 1))
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        np.concatenate([[float(len(records))], records.ravel()]).astype(np.float64).tofile(fi)
        args = [exe, cmd, fi, fo] + ["--%s=%s" % (k, v) for k, v in (keys or {}).items()]
        r = subprocess.run(args, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError("ref driver failed: %s\n%s" % (r.stderr[-2000:], r.stdout[-2000:]))
        out = np.fromfile(fo, dtype=np.float64)
    return out


def ref_geom(model, keys):
    o = ref_run(model, "geom", np.zeros((0, 1)), keys)
    g = Geom()
    g.metric = int(o[0])
    g.a, g.hslope, g.poly_xt, g.poly_alpha, g.mks_smooth, g.poly_norm = o[1:7]
    for name, off in (("startx", 7), ("stopx", 11), ("dx", 15), ("cstartx", 19), ("cstopx", 23)):
        arr = getattr(g, name)
        for mu in range(4):
            arr[mu] = o[off + mu]
    g.Rin, g.Rout, g.Rh, g.rmin_geo, g.rmax_geo = o[27:32]
    g.N1, g.N2, g.N3 = int(o[32]), int(o[33]), int(o[34])
    return g, o[35]


def iharm_geom(h, rmax_geo=100., rmin_geo=1.):
    """Geometry globals as model/iharm/model.c:979-1065 and :326 set them."""
    g = Geom()
    g.metric = METRIC_FMKS if h["metric"] in ("FMKS", "MMKS") else METRIC_MKS
    g.a, g.hslope = h["a"], h["hslope"]
    if g.metric == METRIC_FMKS:
        g.poly_xt, g.poly_alpha, g.mks_smooth = h["poly_xt"], h["poly_alpha"], h["mks_smooth"]
        g.poly_norm = 0.5 * np.pi * 1. / (1. + 1. / (g.poly_alpha + 1.) * 1. / g.poly_xt ** g.poly_alpha)
    g.Rin, g.Rout = h["Rin"], h["Rout"]
    g.N1, g.N2, g.N3 = h["N1"], h["N2"], h["N3"]
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


def iharm_units(M_unit, MBH=6.2e9, trat_small=1., trat_large=40., beta_crit=1., gam=13. / 9,
                game=1.333333333333333, gamp=1.666666666666667):
    u = Units()
    u.M_unit, u.MBH_solar = M_unit, MBH
    u.trat_small, u.trat_large, u.beta_crit = trat_small, trat_large, beta_crit
    u.gam, u.game, u.gamp = gam, game, gamp
    return u


def iharm_oracle(h, prims, M_unit, MBH=6.2e9, trat_small=1., trat_large=40., beta_crit=1.,
                 sigma_cut=1., th_beg=0.0174, rmax_geo=100., L=None, game=1.333333333333333,
                 gamp=1.666666666666667):
    """Oracle set up for an iharm dump: derives the ghosted fields as
    model/iharm/model.c does and returns (Oracle, fields11)."""
    g = iharm_geom(h, rmax_geo=rmax_geo)
    L = L or lib()
    L.or_set_geom(C.byref(g))
    u = iharm_units(M_unit, MBH, trat_small, trat_large, beta_crit, h["gam"], game, gamp)
    nc = (g.N1 + 2) * (g.N2 + 2) * (g.N3 + 2)
    out = np.zeros(11 * nc)
    ru = np.zeros(2 * nc)
    pr = f64(prims)
    L.or_iharm_derive_ru(P(pr), pr.shape[-1], C.byref(u), P(out), P(ru))
    m = Model()
    m.kind = MODEL_IHARM
    m.L_unit = GNEWT * MBH * MSUN / (CL * CL)
    m.th_beg, m.sigma_cut, m.sigma_cut_high = th_beg, sigma_cut, -1.
    fields = out[:10 * nc].copy()
    orc = Oracle(g, m, fields, L=L)
    orc.rub = np.concatenate([ru, out[10 * nc:]])
    return orc, out.reshape(11, g.N1 + 2, g.N2 + 2, g.N3 + 2)
