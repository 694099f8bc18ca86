"""ctypes front end of the CPU oracle and of the reference build (TEST INFRASTRUCTURE).

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import
this module, and only as the checker.  ``liboracle.so`` is the C restatement in ``oracle.c``;
``_ref/libref_*.so`` are the reference's own Fortran modules compiled in place by the Makefile
(present only where the reference tree was available at build time).

Arrays follow the reference layout: ghosted fields are numpy arrays of shape
``(nz+2, ny+2, nx+2)`` (C order, so x is fastest, as in Fortran ``q(0:nx+1,0:ny+1,0:nz+1)``).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = C.POINTER(C.c_double)


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


class MdoGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("re", C.c_double), ("forcing", C.c_double * 3),
                ("periodic", C.c_int * 3), ("bc_type", C.c_int * 24),
                ("bc_value", C.c_double * 24),
                ("ihc", _P), ("ihv", _P)]


class MdoGeometry(C.Structure):
    _fields_ = [("kind", C.c_int), ("second_order", C.c_int), ("n_wave_x", C.c_int),
                ("amp_x", C.c_double), ("phase_x", C.c_double),
                ("A", C.c_double), ("lam", C.c_double), ("half_height", C.c_double)]


class MdoSorStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int),
                ("div_inf", C.c_double), ("div0_inf", C.c_double)]


class MdoStepOpts(C.Structure):
    _fields_ = [("solver", C.c_int), ("niter", C.c_int), ("omega", C.c_double),
                ("tol", C.c_double), ("max_iter", C.c_int), ("check_every", C.c_int),
                ("flowrate", C.c_int), ("ub_target", C.c_double), ("height", C.c_double)]


def bci(d, side, var):
    """Flat index of bcType(d, side, var), d = 1..3 (Fortran column-major)."""
    return (d - 1) + 3 * side + 6 * var


_lib = None
_ref_rb = None
_ref_fft = None
_ref_io = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _lib = C.CDLL(path)
        _lib.mdo_get_cfl.restype = C.c_double
        _lib.mdo_div_inf.restype = C.c_double
        _lib.mdo_bulk_u.restype = C.c_double
        _lib.mdo_bulk_u_step.restype = C.c_double
        _lib.mdo_flowrate_shift.restype = C.c_double
        _lib.mdo_sor_residual_inf.restype = C.c_double
        _lib.mdo_uniform01.restype = C.c_double
        _lib.mdo_uniform01.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64]
        _lib.mdo_splitmix64.restype = C.c_uint64
        _lib.mdo_splitmix64.argtypes = [C.c_uint64]
        _lib.mdo_is_in_body.restype = C.c_int
    return _lib


def ref_rb():
    """Reference red-black build, or None when it was not built."""
    global _ref_rb
    if _ref_rb is None:
        path = os.path.join(HERE, "_ref", "libref_rb.so")
        if not os.path.exists(path):
            return None
        _ref_rb = C.CDLL(path)
        _ref_rb.ref_get_cfl.restype = C.c_double
    return _ref_rb


def ref_fft():
    global _ref_fft
    if _ref_fft is None:
        path = os.path.join(HERE, "_ref", "libref_fft.so")
        if not os.path.exists(path):
            return None
        _ref_fft = C.CDLL(path)
    return _ref_fft


def ref_io():
    """Reference output module (io.f90 + field_hdf5.c), or None when it was not built."""
    global _ref_io
    if _ref_io is None:
        path = os.path.join(HERE, "_ref", "libref_io.so")
        if not os.path.exists(path):
            return None
        _ref_io = C.CDLL(path)
    return _ref_io


class Grid:
    """Grid + BC table, mirroring grid_type (init.f90:5-12) and boundary_type (boundary.f90:6-10)."""

    def __init__(self, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0, re=100.0, forcing=(1.0, 0.0, 0.0),
                 periodic=(1, 0, 1), beta=None, p_neumann_y=True):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.lx, self.ly, self.lz = lx, ly, lz
        self.re = re
        self.forcing = tuple(forcing)
        self.periodic = tuple(int(p) for p in periodic)
        self.bc_type = np.zeros(24, np.int32)
        self.bc_value = np.zeros(24, np.float64)
        if p_neumann_y:
            self.bc_type[bci(2, 0, 0)] = 1
            self.bc_type[bci(2, 1, 0)] = 1
        self.beta = beta
        self.dx, self.dy, self.dz = lx / nx, ly / ny, lz / nz
        self.yc = np.zeros(ny + 2)
        self.yv = np.zeros(ny + 2)
        self.ihc = None
        self.ihv = None
        if beta:
            self.ihc = np.zeros(ny + 2)
            self.ihv = np.zeros(ny + 2)
            lib().mdo_stretched_metrics(C.c_int(ny), C.c_double(ly), C.c_double(beta),
                                        _ptr(self.yc), _ptr(self.yv), _ptr(self.ihc), _ptr(self.ihv))
        else:
            # same expressions as the IBM sample positions (ibm.f90:161-163)
            for j in range(ny + 2):
                self.yc[j] = ((j + 1) - 1.5) * self.dy
                self.yv[j] = ((j + 1) - 0.5 - 1.5) * self.dy

    def struct(self):
        g = MdoGrid()
        g.nx, g.ny, g.nz = self.nx, self.ny, self.nz
        g.lx, g.ly, g.lz = self.lx, self.ly, self.lz
        g.re = self.re
        for q in range(3):
            g.forcing[q] = self.forcing[q]
            g.periodic[q] = self.periodic[q]
        for q in range(24):
            g.bc_type[q] = int(self.bc_type[q])
            g.bc_value[q] = float(self.bc_value[q])
        lib().mdo_finalize_grid(C.byref(g))
        g.ihc = _ptr(self.ihc)
        g.ihv = _ptr(self.ihv)
        return g

    @property
    def gshape(self):
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    @property
    def ishape(self):
        return (self.nz, self.ny, self.nx)

    def zeros(self):
        return np.zeros(self.gshape)


class Geometry:
    def __init__(self, kind=1, second_order=1, n_wave_x=1, amp_x=0.0, phase_x=0.0,
                 A=0.1, lam=1.0, half_height=1.0):
        self.s = MdoGeometry(kind, second_order, n_wave_x, amp_x, phase_x, A, lam, half_height)


# ----------------------------------------------------------------------------- oracle calls
def momentum(g, f, coef, dta, dtb, dtg):
    """f: dict with un,vn,wn,pn,us,vs,ws,oru,orv,orw (modified in place)."""
    s = g.struct()
    cu, cv, cw = coef if coef is not None else (None, None, None)
    lib().mdo_momentum(C.byref(s), _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), _ptr(f["pn"]),
                       _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]),
                       _ptr(f["oru"]), _ptr(f["orv"]), _ptr(f["orw"]),
                       _ptr(cu), _ptr(cv), _ptr(cw),
                       C.c_double(dta), C.c_double(dtb), C.c_double(dtg))


def get_cfl(g, f):
    s = g.struct()
    return lib().mdo_get_cfl(C.byref(s), _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]))


def apply_bc_fft(g, f):
    s = g.struct()
    lib().mdo_apply_bc_fft(C.byref(s), _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]),
                           _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]), _ptr(f["pn"]))


def apply_bc_redblack(g, f):
    s = g.struct()
    lib().mdo_apply_bc_redblack(C.byref(s), _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), _ptr(f["pn"]))


def build_fft_rhs(g, f, dtg):
    s = g.struct()
    rhs = np.zeros((g.ny, g.nz, g.nx))
    lib().mdo_build_fft_rhs(C.byref(s), _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]),
                            C.c_double(dtg), _ptr(rhs))
    return rhs


def fft_r2c(g, rhs):
    s = g.struct()
    nxh = g.nx // 2 + 1
    hat = np.zeros((g.ny, g.nz, nxh), np.complex128)
    lib().mdo_fft_r2c_xz(C.byref(s), _ptr(np.ascontiguousarray(rhs)), hat.ctypes.data_as(_P))
    return hat


def fft_c2r(g, hat):
    s = g.struct()
    out = np.zeros((g.ny, g.nz, g.nx))
    h = np.ascontiguousarray(hat, dtype=np.complex128)
    lib().mdo_fft_c2r_xz(C.byref(s), h.ctypes.data_as(_P), _ptr(out))
    return out


def tridiag(g, hat):
    s = g.struct()
    nxh = g.nx // 2 + 1
    cp = np.zeros((g.nz, g.ny, nxh))
    di = np.zeros((g.nz, g.ny, nxh))
    ar = np.zeros(g.ny)
    lib().mdo_tridiag_coefficients(C.byref(s), _ptr(cp), _ptr(di), _ptr(ar))
    h = np.array(hat, dtype=np.complex128, order="C")
    dp = np.zeros((g.nz, g.ny, nxh), np.complex128)
    lib().mdo_tridiag_solve(C.byref(s), _ptr(cp), _ptr(di), _ptr(ar), h.ctypes.data_as(_P),
                            dp.ctypes.data_as(_P))
    return h, cp, di


def projection_fft(g, f, coef, dtg):
    s = g.struct()
    cu, cv, cw = coef if coef is not None else (None, None, None)
    lib().mdo_projection_fft(C.byref(s), C.c_double(dtg), _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]),
                             _ptr(cu), _ptr(cv), _ptr(cw),
                             _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), _ptr(f["pn"]))


def sola_redblack(g, f, coef, niter, sor, dtg):
    s = g.struct()
    cu, cv, cw = coef if coef is not None else (None, None, None)
    lib().mdo_sola_redblack(C.byref(s), C.c_int(niter), C.c_double(sor), C.c_double(dtg),
                            _ptr(cu), _ptr(cv), _ptr(cw),
                            _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]),
                            _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), _ptr(f["pn"]))


def projection_sor(g, f, coef, omega, tol, max_iter, check_every, dtg):
    s = g.struct()
    cu, cv, cw = coef if coef is not None else (None, None, None)
    phi = np.zeros(g.ishape)
    rhs = np.zeros(g.ishape)
    st = MdoSorStats()
    lib().mdo_projection_sor(C.byref(s), C.c_double(omega), C.c_double(tol), C.c_int(max_iter),
                             C.c_int(check_every), C.c_double(dtg), _ptr(cu), _ptr(cv), _ptr(cw),
                             _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]),
                             _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), _ptr(f["pn"]),
                             _ptr(phi), _ptr(rhs), C.byref(st))
    return st, phi, rhs


def sor_div_rhs(g, f):
    s = g.struct()
    rhs = np.zeros(g.ishape)
    lib().mdo_sor_div_rhs(C.byref(s), _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), _ptr(rhs))
    return rhs


def sor_sweep(g, coef, color, omega, dtg, rhs, phi):
    s = g.struct()
    cu, cv, cw = coef if coef is not None else (None, None, None)
    lib().mdo_sor_sweep(C.byref(s), C.c_int(color), C.c_double(omega), C.c_double(dtg),
                        _ptr(cu), _ptr(cv), _ptr(cw), _ptr(rhs), _ptr(phi))


def sor_residual_inf(g, coef, dtg, rhs, phi):
    s = g.struct()
    cu, cv, cw = coef if coef is not None else (None, None, None)
    return lib().mdo_sor_residual_inf(C.byref(s), C.c_double(dtg), _ptr(cu), _ptr(cv), _ptr(cw),
                                      _ptr(rhs), _ptr(phi))


def div_inf(g, f):
    s = g.struct()
    return lib().mdo_div_inf(C.byref(s), _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]))


def bulk_u(g, u, height):
    s = g.struct()
    return lib().mdo_bulk_u(C.byref(s), _ptr(u), C.c_double(height))


def bulk_u_step(g, u, height):
    """bulk_u with the sums in the order the time step's flow-rate control adds them."""
    s = g.struct()
    return lib().mdo_bulk_u_step(C.byref(s), _ptr(u), C.c_double(height))


def set_ibm_coeff(g, geo, dix, diy, diz):
    s = g.struct()
    c = g.zeros()
    lib().mdo_set_ibm_coeff(C.byref(s), C.byref(geo.s), C.c_int(dix), C.c_int(diy), C.c_int(diz), _ptr(c))
    return c


def ibm_coefs(g, geo):
    return (set_ibm_coeff(g, geo, 1, 0, 0), set_ibm_coeff(g, geo, 0, 1, 0), set_ibm_coeff(g, geo, 0, 0, 1))


def uniform01(seed, comp, index):
    return lib().mdo_uniform01(seed, comp, index)


def initial_field(g, f, seed, amp, geo=None):
    s = g.struct()
    gs = geo.s if geo is not None else MdoGeometry()
    lib().mdo_initial_field(C.byref(s), C.byref(gs), C.c_int(1 if geo is not None else 0),
                            _ptr(g.yc), _ptr(g.yv), C.c_uint32(seed), C.c_double(amp),
                            _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]))


def rk3_step(g, f, coef, dt, solver=0, niter=3, omega=1.8, tol=1e-8, max_iter=100000,
             check_every=10, flowrate=False, ub_target=1.0, height=2.0):
    s = g.struct()
    o = MdoStepOpts(solver, niter, omega, tol, max_iter, check_every, 1 if flowrate else 0,
                    ub_target, height)
    cu, cv, cw = coef if coef is not None else (None, None, None)
    phi = np.zeros(g.ishape)
    rhs = np.zeros(g.ishape)
    st = (MdoSorStats * 3)()
    lib().mdo_rk3_step(C.byref(s), C.byref(o), C.c_double(dt),
                       _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), _ptr(f["pn"]),
                       _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]),
                       _ptr(f["oru"]), _ptr(f["orv"]), _ptr(f["orw"]),
                       _ptr(cu), _ptr(cv), _ptr(cw), _ptr(phi), _ptr(rhs), st)
    return [(st[q].iterations, st[q].converged, st[q].div_inf, st[q].div0_inf) for q in range(3)]


def rk_constants():
    """(alpha, beta, gamma) of the oracle's RK3 stages (step.f90:21-23)."""
    L = lib()
    return tuple(tuple((C.c_double * 3).in_dll(L, name)) for name in ("mdo_rk_alpha", "mdo_rk_beta", "mdo_rk_gamma"))


def run_steps(g, f, coef, nsteps, dt, cflmax=0.0, dtmax=1e-3, **step_kw):
    """The time loop of main.f90:64-97 over rk3_step: each step runs with the current dt; after it,
    when cflmax > 0, cfl = get_cfl (step.f90:210) and dt = min(cflmax / cfl, dtmax) if cfl > 0.
    Returns (dts, cfls, stats): the dt of each step, the CFL after it (0 when not computed) and the
    per-step solver statistics."""
    dts, cfls, stats = [], [], []
    for _ in range(nsteps):
        dts.append(dt)
        stats.append(rk3_step(g, f, coef, dt, **step_kw))
        cfl = 0.0
        if cflmax > 0.0:
            cfl = get_cfl(g, f)
        cfls.append(cfl)
        if cflmax > 0.0 and cfl > 0.0:
            dt = min(cflmax / cfl, dtmax)
    return dts, cfls, stats


def sor_relax_cells(g, coef, omega, dtg, rhs, phi, cells):
    """Relax the listed (i, j, k) cells (1-based) one after another in the given order, with the
    relaxation of sor_sweep (test hook for the colouring)."""
    s = g.struct()
    cu, cv, cw = coef if coef is not None else (None, None, None)
    cl = np.ascontiguousarray(cells, dtype=np.int32)
    lib().mdo_sor_relax_cells(C.byref(s), C.c_double(omega), C.c_double(dtg), _ptr(cu), _ptr(cv), _ptr(cw),
                              _ptr(rhs), _ptr(phi), cl.ctypes.data_as(C.POINTER(C.c_int)), C.c_int(len(cl)))


def new_fields(g):
    f = {k: g.zeros() for k in ("un", "vn", "wn", "pn", "us", "vs", "ws")}
    for k in ("oru", "orv", "orw"):
        f[k] = np.zeros(g.ishape)
    return f


def random_fields(g, rng, amp=1.0):
    f = new_fields(g)
    for k in ("un", "vn", "wn", "pn", "us", "vs", "ws"):
        f[k][...] = amp * rng.standard_normal(g.gshape)
    for k in ("oru", "orv", "orw"):
        f[k][...] = amp * rng.standard_normal(g.ishape)
    return f


# ----------------------------------------------------------------------------- reference calls
def _ci(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def ref_momentum(g, f, coef, dta, dtb, dtg):
    r = ref_rb()
    forcing = np.array(g.forcing, np.float64)
    per = np.array(g.periodic, np.int32)
    if coef is None:
        coef = (g.zeros(), g.zeros(), g.zeros())
    r.ref_momentum(C.c_int(g.nx), C.c_int(g.ny), C.c_int(g.nz), C.c_double(g.lx), C.c_double(g.ly),
                   C.c_double(g.lz), C.c_double(g.re), _ptr(forcing), _ci(per),
                   _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), _ptr(f["pn"]),
                   _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]),
                   _ptr(f["oru"]), _ptr(f["orv"]), _ptr(f["orw"]),
                   _ptr(coef[0]), _ptr(coef[1]), _ptr(coef[2]),
                   C.c_double(dta), C.c_double(dtb), C.c_double(dtg))


def ref_get_cfl(g, f):
    return ref_rb().ref_get_cfl(C.c_int(g.nx), C.c_int(g.ny), C.c_int(g.nz), C.c_double(g.lx),
                                C.c_double(g.ly), C.c_double(g.lz), _ptr(f["un"]), _ptr(f["vn"]),
                                _ptr(f["wn"]))


def ref_apply_bc_redblack(g, f):
    per = np.array(g.periodic, np.int32)
    bt = np.array(g.bc_type, np.int32)
    ref_rb().ref_apply_bc_redblack(C.c_int(g.nx), C.c_int(g.ny), C.c_int(g.nz), C.c_double(g.lx),
                                   C.c_double(g.ly), C.c_double(g.lz), _ci(per), _ci(bt),
                                   _ptr(g.bc_value), _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]),
                                   _ptr(f["pn"]))


def ref_projection_redblack(g, f, coef, niter, sor, dtg):
    per = np.array(g.periodic, np.int32)
    bt = np.array(g.bc_type, np.int32)
    if coef is None:
        coef = (g.zeros(), g.zeros(), g.zeros())
    ref_rb().ref_projection_redblack(C.c_int(g.nx), C.c_int(g.ny), C.c_int(g.nz), C.c_double(g.lx),
                                     C.c_double(g.ly), C.c_double(g.lz), _ci(per), _ci(bt),
                                     _ptr(g.bc_value), C.c_int(niter), C.c_double(sor), C.c_double(dtg),
                                     _ptr(coef[0]), _ptr(coef[1]), _ptr(coef[2]),
                                     _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]),
                                     _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), _ptr(f["pn"]))


def ref_set_ibm_coeff(g, dix, diy, diz, amp_x=0.0, phase_x=0.0, n_wave_x=0):
    c = g.zeros()
    ref_rb().ref_set_ibm_coeff(C.c_int(g.nx), C.c_int(g.ny), C.c_int(g.nz), C.c_double(g.lx),
                               C.c_double(g.ly), C.c_double(g.lz), C.c_double(g.re),
                               C.c_double(amp_x), C.c_double(phase_x), C.c_int(n_wave_x),
                               C.c_int(dix), C.c_int(diy), C.c_int(diz), _ptr(c))
    return c


def ref_apply_bc_fft(g, f):
    ref_fft().ref_apply_bc_fft(C.c_int(g.nx), C.c_int(g.ny), C.c_int(g.nz),
                               _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]),
                               _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]), _ptr(f["pn"]))


def ref_tridiag(g, hat):
    nxh = g.nx // 2 + 1
    h = np.array(hat, dtype=np.complex128, order="C")
    cp = np.zeros((g.nz, g.ny, nxh), np.complex128)
    di = np.zeros((g.nz, g.ny, nxh), np.complex128)
    ref_fft().ref_tridiag(C.c_int(g.nx), C.c_int(g.ny), C.c_int(g.nz), C.c_double(g.lx),
                          C.c_double(g.ly), C.c_double(g.lz), h.ctypes.data_as(_P),
                          cp.ctypes.data_as(_P), di.ctypes.data_as(_P))
    return h, cp, di


def ref_run_redblack(g, f, coef, niter, sor, nsteps, dt, cflmax=0.0, dtmax=1e-3):
    """The reference's own time loop (main.f90:64-97, red-black build) run by ref_bridge_rb over its
    modules: per step three RK stages of momentum / pressure_projection / apply_bc_redblack with
    step.f90's rk_alpha, rk_beta, rk_gamma, then the CFL-driven dt update.  Returns (dts, cfls)."""
    per = np.array(g.periodic, np.int32)
    bt = np.array(g.bc_type, np.int32)
    forcing = np.array(g.forcing, np.float64)
    if coef is None:
        coef = (g.zeros(), g.zeros(), g.zeros())
    dts = np.zeros(nsteps)
    cfls = np.zeros(nsteps)
    ref_rb().ref_run_redblack(C.c_int(g.nx), C.c_int(g.ny), C.c_int(g.nz), C.c_double(g.lx),
                              C.c_double(g.ly), C.c_double(g.lz), C.c_double(g.re), _ptr(forcing), _ci(per),
                              _ci(bt), _ptr(g.bc_value), C.c_int(niter), C.c_double(sor), C.c_int(nsteps),
                              C.c_double(dt), C.c_double(cflmax), C.c_double(dtmax),
                              _ptr(coef[0]), _ptr(coef[1]), _ptr(coef[2]),
                              _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), _ptr(f["pn"]),
                              _ptr(f["us"]), _ptr(f["vs"]), _ptr(f["ws"]),
                              _ptr(f["oru"]), _ptr(f["orv"]), _ptr(f["orw"]), _ptr(dts), _ptr(cfls))
    return list(dts), list(cfls)


# ----------------------------------------------------------------------------- output
def center_vel(f):
    """io.f90:184-222: uc(i,j,k) = 0.5 (un(i,j,k) + un(i+1,j,k)), vc with j+1, wc with k+1, over
    interior cells (arrays of shape (nz, ny, nx))."""
    un, vn, wn = f["un"], f["vn"], f["wn"]
    uc = 0.5 * (un[1:-1, 1:-1, 1:-1] + un[1:-1, 1:-1, 2:])
    vc = 0.5 * (vn[1:-1, 1:-1, 1:-1] + vn[1:-1, 2:, 1:-1])
    wc = 0.5 * (wn[1:-1, 1:-1, 1:-1] + wn[2:, 1:-1, 1:-1])
    return uc, vc, wc


def ref_write_vtk(g, f, dt, t_current, cfl, step, prefix):
    """The reference's write_vtk (io.f90:168-182) on these fields: writes <prefix>_<step>.vtk."""
    ref_io().ref_write_vtk(C.c_int(g.nx), C.c_int(g.ny), C.c_int(g.nz), C.c_double(g.lx), C.c_double(g.ly),
                           C.c_double(g.lz), C.c_double(g.re), C.c_double(dt), C.c_double(t_current),
                           C.c_double(cfl), _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), C.c_int(step),
                           os.fsencode(prefix))


def ref_write_field(g, f, dt, t_current, step, prefix):
    """The reference's write_field (io.f90:105-137): writes <prefix>_<step>.h5 through field_hdf5.c."""
    ref_io().ref_write_field(C.c_int(g.nx), C.c_int(g.ny), C.c_int(g.nz), C.c_double(g.lx), C.c_double(g.ly),
                             C.c_double(g.lz), C.c_double(g.re), C.c_double(dt), C.c_double(t_current),
                             _ptr(f["un"]), _ptr(f["vn"]), _ptr(f["wn"]), _ptr(f["pn"]), C.c_int(step),
                             os.fsencode(prefix))


def ref_read_field(g, path):
    """The reference's read_field (io.f90:139-166)."""
    out = {k: g.zeros() for k in ("un", "vn", "wn", "pn")}
    ref_io().ref_read_field(C.c_int(g.nx), C.c_int(g.ny), C.c_int(g.nz), _ptr(out["un"]), _ptr(out["vn"]),
                            _ptr(out["wn"]), _ptr(out["pn"]), os.fsencode(path))
    return out
