"""mobydiff_amd -- MI355X (gfx950) time-step hot path of the mobydiff channel-flow solver.

The compute path is ``csrc/libmobydiff_hip.so`` (hand-written HIP), reached through its C ABI
(``include/mobydiff_hip.h``).  This module is the Python front end of that ABI: ctypes
structures mirroring the header and a ``Channel`` object whose methods mirror the reference's
module procedures (``momentum``, ``apply_bc``, ``apply_bc_redblack``, ``pressure_projection``,
``get_cfl``; step.f90, boundary.f90, pressure_solver.f90).  There is no CPU fallback: if the
library is not built, or no HIP device is present, loading fails with an error.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "csrc", "libmobydiff_hip.so")
IOLIBPATH = os.path.join(HERE, "csrc", "libmobydiff_io.so")

SOLVER_FFT, SOLVER_SOLA, SOLVER_SOR = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
GUESS_PREVIOUS, GUESS_LEARNED, GUESS_PHI_PREV = 0, 1, 2

_PD = C.POINTER(C.c_double)


class Config(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("re", C.c_double), ("forcing", C.c_double * 3),
                ("periodic", C.c_int32 * 3), ("bc_type", C.c_int32 * 24),
                ("bc_value", C.c_double * 24), ("stretch_beta", C.c_double),
                ("solver", C.c_int32), ("niter", C.c_int32), ("omega", C.c_double),
                ("tol", C.c_double), ("max_iter", C.c_int32), ("check_every", C.c_int32),
                ("precision", C.c_int32), ("inner_iters", C.c_int32), ("guess", C.c_int32),
                ("flowrate", C.c_int32), ("ub_target", C.c_double), ("flow_height", C.c_double),
                ("timing", C.c_int32)]


class Geometry(C.Structure):
    _fields_ = [("kind", C.c_int32), ("second_order", C.c_int32), ("n_wave_x", C.c_int32),
                ("amp_x", C.c_double), ("phase_x", C.c_double),
                ("A", C.c_double), ("lam", C.c_double), ("half_height", C.c_double)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("fallback", C.c_int32),
                ("checks", C.c_int32), ("div0_inf", C.c_double), ("div_inf", C.c_double),
                ("guess_div_inf", C.c_double), ("lds_sweeps", C.c_int32), ("guess_skipped", C.c_int32),
                ("lds_restarts", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepStats(C.Structure):
    _fields_ = [("stage", SolveStats * 3), ("ub", C.c_double)]


class PhaseTimes(C.Structure):
    _fields_ = [("stencil_ms", C.c_double), ("poisson_ms", C.c_double),
                ("projection_ms", C.c_double), ("guess_ms", C.c_double), ("stages", C.c_int64),
                ("sweep_ms", C.c_double), ("sweep_launches", C.c_int64), ("sweep_cells", C.c_int64),
                ("momentum_ms", C.c_double)]


def bci(d, side, var):
    """Flat index of bcType(d, side, var) (boundary.f90:8), d = 1..3."""
    return (d - 1) + 3 * side + 6 * var


_lib = None


class MdxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mobydiff_hip error {code}: {msg}")
        self.code = code


def lib():
    """Load libmobydiff_hip.so; raise if it is missing (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is not built: run __graft_entry__.build() "
                               "or make -C mobydiff_amd/csrc")
        _lib = C.CDLL(LIBPATH)
        _lib.mdx_last_error.restype = C.c_char_p
    return _lib


def _check(rc):
    if rc != 0:
        raise MdxError(rc, lib().mdx_last_error().decode())


def lds_read_rate():
    """(bytes/s, bytes per clock and CU) of the device's sustained LDS reads (mdx_lds_read_rate)"""
    b, c = C.c_double(), C.c_double()
    _check(lib().mdx_lds_read_rate(C.byref(b), C.byref(c)))
    return b.value, c.value


_iolib = None


def io_lib():
    """Load libmobydiff_io.so (HDF5 field files, include/mobydiff_io.h); raise if it is missing."""
    global _iolib
    if _iolib is None:
        if not os.path.exists(IOLIBPATH):
            raise RuntimeError(f"{IOLIBPATH} is not built: run make -C mobydiff_amd/csrc")
        _iolib = C.CDLL(IOLIBPATH)
        _iolib.mdx_io_last_error.restype = C.c_char_p
    return _iolib


def _io_check(rc):
    if rc != 0:
        raise MdxError(rc, io_lib().mdx_io_last_error().decode())


def write_field_h5(path, nx, ny, nz, lx, ly, lz, re, dt, t_current, un, vn, wn, pn):
    """write_field's file (io.f90:105-137, field_hdf5.c:149-184): ghosted un, vn, wn, pn of shape
    (nz+2, ny+2, nx+2) and the twelve attributes; dx = lx/nx etc. as finalize_grid sets them."""
    dx, dy, dz = lx / nx, ly / ny, lz / nz
    _io_check(io_lib().mdx_h5_write_field(
        os.fsencode(path), C.c_int(nx), C.c_int(ny), C.c_int(nz), C.c_double(lx), C.c_double(ly),
        C.c_double(lz), C.c_double(dx), C.c_double(dy), C.c_double(dz), C.c_double(re), C.c_double(dt),
        C.c_double(t_current), _p(un), _p(vn), _p(wn), _p(pn)))


def read_field_h5(path, nx, ny, nz):
    """read_field (io.f90:139-166, field_hdf5.c:186-204): the four ghosted fields of the file."""
    out = {k: np.zeros((nz + 2, ny + 2, nx + 2)) for k in ("un", "vn", "wn", "pn")}
    _io_check(io_lib().mdx_h5_read_field(os.fsencode(path), C.c_int(nx), C.c_int(ny), C.c_int(nz),
                                         _p(out["un"]), _p(out["vn"]), _p(out["wn"]), _p(out["pn"])))
    return out


def read_field_info(path):
    """(nx, ny, nz, dt, t_current) of a field file"""
    n = [C.c_int32() for _ in range(3)]
    dt, t = C.c_double(), C.c_double()
    _io_check(io_lib().mdx_h5_read_field_info(os.fsencode(path), *[C.byref(q) for q in n], C.byref(dt),
                                              C.byref(t)))
    return n[0].value, n[1].value, n[2].value, dt.value, t.value


def _p(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(_PD)


def channel_config(nx, ny, nz, lx, ly, lz, re, forcing=(0.0, 0.0, 0.0), periodic=(1, 0, 1),
                   solver=SOLVER_SOR, niter=3, omega=1.8, tol=1e-8, max_iter=200000,
                   check_every=20, precision=PREC_FP64, inner_iters=20, guess=GUESS_PREVIOUS,
                   flowrate=False, ub_target=1.0, flow_height=2.0, stretch_beta=0.0,
                   p_neumann_y=True, timing=False):
    """Config for a channel: x/z periodic, no-slip y walls (Dirichlet u,v,w = 0), Neumann p on the
    walls, as tutorials/wavychannel/input.ini sets up."""
    c = Config()
    c.nx, c.ny, c.nz = nx, ny, nz
    c.lx, c.ly, c.lz = lx, ly, lz
    c.re = re
    for q in range(3):
        c.forcing[q] = forcing[q]
        c.periodic[q] = int(periodic[q])
    if p_neumann_y:
        c.bc_type[bci(2, 0, 0)] = 1
        c.bc_type[bci(2, 1, 0)] = 1
    c.stretch_beta = stretch_beta
    c.solver, c.niter, c.omega, c.tol = solver, niter, omega, tol
    c.max_iter, c.check_every = max_iter, check_every
    c.precision, c.inner_iters, c.guess = precision, inner_iters, guess
    c.flowrate, c.ub_target, c.flow_height = int(flowrate), ub_target, flow_height
    c.timing = int(timing)
    return c


def wavy_geometry(A=0.1, lam=1.0, half_height=1.0, second_order=True):
    g = Geometry()
    g.kind, g.second_order, g.A, g.lam, g.half_height = 1, int(second_order), A, lam, half_height
    return g


def bump_geometry(amp_x=0.0, phase_x=0.0, n_wave_x=1, second_order=True):
    g = Geometry()
    g.kind, g.second_order, g.n_wave_x, g.amp_x, g.phase_x = 0, int(second_order), n_wave_x, amp_x, phase_x
    return g


def step_batch(channels, dt):
    """One RK3 step of independent channels, run concurrently by the library (mdx_step_batch);
    returns the StepStats of each channel."""
    n = len(channels)
    hs = (C.c_void_p * n)(*[ch.h.value for ch in channels])
    st = (StepStats * n)()
    _check(lib().mdx_step_batch(hs, C.c_int32(n), C.c_double(dt), st))
    return list(st)


RK_ALPHA = (64.0 / 120.0, 50.0 / 120.0, 90.0 / 120.0)   # step.f90:21-23
RK_BETA = (0.0, -34.0 / 120.0, -50.0 / 120.0)
RK_GAMMA = (64.0 / 120.0, 16.0 / 120.0, 40.0 / 120.0)

FIELD_NAMES = ("un", "vn", "wn", "pn", "us", "vs", "ws", "oru", "orv", "orw")


class Channel:
    """One device-resident flow state (field_type + ibm_type + pressure_solver_type)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.h = C.c_void_p()
        _check(lib().mdx_create(C.byref(cfg), C.byref(self.h)))
        self.nx, self.ny, self.nz = cfg.nx, cfg.ny, cfg.nz

    # shapes (C order, x fastest)
    @property
    def gshape(self):
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    @property
    def ishape(self):
        return (self.nz, self.ny, self.nx)

    def close(self):
        if self.h:
            lib().mdx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        # at interpreter shutdown the HIP runtime (and a profiler around it) may already be torn
        # down: leave the device memory to the process exit instead of calling into the library
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def upload(self, **f):
        args = [_p(f.get(k)) for k in FIELD_NAMES]
        _check(lib().mdx_upload_fields(self.h, *args))

    def download(self, names=FIELD_NAMES):
        out = {}
        args = []
        for k in FIELD_NAMES:
            if k in names:
                out[k] = np.zeros(self.ishape if k.startswith("or") else self.gshape)
                args.append(_p(out[k]))
            else:
                args.append(None)
        _check(lib().mdx_download_fields(self.h, *args))
        return out

    def set_ibm(self, cu, cv, cw):
        _check(lib().mdx_set_ibm(self.h, _p(cu), _p(cv), _p(cw)))

    def compute_ibm(self, geo):
        cu, cv, cw = (np.zeros(self.gshape) for _ in range(3))
        _check(lib().mdx_compute_ibm(self.h, C.byref(geo), _p(cu), _p(cv), _p(cw)))
        return cu, cv, cw

    def momentum(self, dta, dtb, dtg):
        _check(lib().mdx_momentum(self.h, C.c_double(dta), C.c_double(dtb), C.c_double(dtg)))

    def apply_bc(self):
        _check(lib().mdx_apply_bc(self.h))

    def apply_bc_redblack(self):
        _check(lib().mdx_apply_bc_redblack(self.h))

    def pressure_projection(self, dtg):
        st = SolveStats()
        _check(lib().mdx_pressure_projection(self.h, C.c_double(dtg), C.byref(st)))
        return st

    def get_cfl(self):
        v = C.c_double()
        _check(lib().mdx_get_cfl(self.h, C.byref(v)))
        return v.value

    def center_velocity(self):
        """io.f90:184-222 center_vel on the device: (uc, vc, wc), each of shape (nz, ny, nx)"""
        uc, vc, wc = (np.zeros(self.ishape) for _ in range(3))
        _check(lib().mdx_center_velocity(self.h, _p(uc), _p(vc), _p(wc)))
        return uc, vc, wc

    def write_field(self, path, dt, t_current):
        """write_field (io.f90:105-137): download un, vn, wn, pn and write the HDF5 field file"""
        f = self.download(("un", "vn", "wn", "pn"))
        c = self.cfg
        write_field_h5(path, c.nx, c.ny, c.nz, c.lx, c.ly, c.lz, c.re, dt, t_current,
                       f["un"], f["vn"], f["wn"], f["pn"])

    def read_field(self, path):
        """read_field (io.f90:139-166): read un, vn, wn, pn from the file and upload them"""
        f = read_field_h5(path, self.nx, self.ny, self.nz)
        self.upload(**f)

    def div_norm(self):
        v = C.c_double()
        _check(lib().mdx_div_norm(self.h, C.byref(v)))
        return v.value

    def flowrate_shift(self, dtg):
        v = C.c_double()
        _check(lib().mdx_flowrate_shift(self.h, C.c_double(dtg), C.byref(v)))
        return v.value

    def bulk_velocity(self):
        v = C.c_double()
        _check(lib().mdx_bulk_velocity(self.h, C.byref(v)))
        return v.value

    def init_fields(self, seed, amp, geo=None, dt=1e-3):
        _check(lib().mdx_init_fields(self.h, C.c_uint32(seed), C.c_double(amp),
                                     C.byref(geo) if geo is not None else None, C.c_double(dt)))

    def step(self, dt):
        st = StepStats()
        _check(lib().mdx_step(self.h, C.c_double(dt), C.byref(st)))
        return st

    def set_step_report(self, bulk_velocity):
        """Whether step() reports the bulk velocity in StepStats.ub (flow-rate control on; the
        default).  The report waits for the step's last kernel; without it, ub is NaN and step()
        returns once the step is queued (mdx_set_step_report)."""
        _check(lib().mdx_set_step_report(self.h, C.c_int32(1 if bulk_velocity else 0)))

    def synchronize(self):
        _check(lib().mdx_synchronize(self.h))

    def time_momentum(self, dta, dtb, dtg, reps=20):
        """device ms per launch of the momentum kernel (mdx_time_momentum; RK3 history kept)"""
        ms = C.c_float()
        _check(lib().mdx_time_momentum(self.h, C.c_double(dta), C.c_double(dtb), C.c_double(dtg), C.c_int32(reps),
                                       C.byref(ms)))
        return ms.value

    def phase_times(self):
        t = PhaseTimes()
        _check(lib().mdx_get_phase_times(self.h, C.byref(t)))
        return t

    def reset_phase_times(self):
        _check(lib().mdx_reset_phase_times(self.h))

    def set_guess(self, guess):
        _check(lib().mdx_set_guess(self.h, C.c_int32(guess)))

    def attach_net(self, net, capture=False):
        _check(lib().mdx_attach_net(self.h, net.h if net is not None else None, C.c_int32(1 if capture else 0)))
        self._net = net

    def info(self):
        buf = C.create_string_buffer(1024)
        _check(lib().mdx_get_info(self.h, buf, C.c_int32(1024)))
        return buf.value.decode()

    def y_coordinates(self):
        yc = np.zeros(self.ny + 2)
        yv = np.zeros(self.ny + 2)
        _check(lib().mdx_get_y_coordinates(self.h, _p(yc), _p(yv)))
        return yc, yv


_PF = C.POINTER(C.c_float)


def _pf(a):
    if a is None:
        return None
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise TypeError("arrays must be C-contiguous float32")
    return a.ctypes.data_as(_PF)


class Net:
    """The pressure-guess U-Net (BF16/MFMA HIP kernels); see mdx_net_* in include/mobydiff_hip.h."""

    def __init__(self, nx, ny, nz, seed=0, stride=1):
        """A network serving the nx x ny x nz solver grid; it runs on that grid coarsened by
        `stride` (1, 2 or 4): inputs restricted by the mean over stride^3 cells, the output
        prolonged trilinearly.  `shape` is the network grid (the host-side forward/gradient calls
        take inputs of that shape)."""
        self.h = C.c_void_p()
        _check(lib().mdx_net_create_strided(C.c_int32(nx), C.c_int32(ny), C.c_int32(nz), C.c_int32(stride),
                                            C.c_uint64(seed), C.byref(self.h)))
        self.stride = stride
        self.shape = (nz // stride, ny // stride, nx // stride)
        lib().mdx_net_param_count.restype = C.c_int64
        self.nparam = int(lib().mdx_net_param_count())

    def close(self):
        """Destroy the network; handles it is attached to are detached by the library."""
        if self.h:
            lib().mdx_net_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def restrict(self, fine):
        """solver-grid field (nz, ny, nx) float64 -> network grid, the mean over stride^3 cells"""
        fine = np.ascontiguousarray(fine, dtype=np.float64)
        out = np.zeros(self.shape, np.float32)
        _check(lib().mdx_net_restrict_host(self.h, _p(fine), _pf(out)))
        return out

    def prolong(self, coarse):
        """network-grid field -> solver grid, the trilinear interpolation of the learned guess"""
        coarse = np.ascontiguousarray(coarse, dtype=np.float32)
        nz, ny, nx = self.shape
        out = np.zeros((nz * self.stride, ny * self.stride, nx * self.stride))
        _check(lib().mdx_net_prolong_host(self.h, _pf(coarse), _p(out)))
        return out

    def get_params(self):
        p = np.zeros(self.nparam, np.float32)
        c0 = C.c_double()
        _check(lib().mdx_net_get_params(self.h, _pf(p), C.byref(c0)))
        return p, c0.value

    def set_params(self, p, c0=0.0):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().mdx_net_set_params(self.h, _pf(p), C.c_double(c0)))

    def get_grads(self):
        g = np.zeros(self.nparam, np.float32)
        _check(lib().mdx_net_get_grads(self.h, _pf(g)))
        return g

    def zero_grads(self):
        _check(lib().mdx_net_zero_grads(self.h))

    def forward(self, x2):
        """x2: (nz, ny, nx, 2) float32 -> (nz, ny, nx) float32"""
        x2 = np.ascontiguousarray(x2, dtype=np.float32)
        out = np.zeros(self.shape, np.float32)
        _check(lib().mdx_net_forward_host(self.h, _pf(x2), _pf(out)))
        return out

    def accumulate_grad(self, x2, target, batch=1):
        x2 = np.ascontiguousarray(x2, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        loss = C.c_float()
        _check(lib().mdx_net_accumulate_grad_host(self.h, _pf(x2), _pf(target), C.c_int32(batch), C.byref(loss)))
        return loss.value

    def adam_step(self, lr=1e-3):
        _check(lib().mdx_net_adam_step(self.h, C.c_float(lr)))

    def replay_reserve(self, capacity):
        _check(lib().mdx_net_replay_reserve(self.h, C.c_int32(capacity)))

    def replay_count(self):
        n = C.c_int32()
        _check(lib().mdx_net_replay_count(self.h, C.byref(n)))
        return n.value

    def train(self, steps, batch=4, lr=1e-3, seed=0):
        loss = C.c_float()
        _check(lib().mdx_net_train(self.h, C.c_int32(steps), C.c_int32(batch), C.c_float(lr), C.c_uint64(seed),
                                   C.byref(loss)))
        return loss.value

    def eval_replay(self, count):
        loss = C.c_float()
        _check(lib().mdx_net_eval_replay(self.h, C.c_int32(count), C.byref(loss)))
        return loss.value

    def time_forward(self, reps=10):
        """(ms per forward pass on the device, MFMA flops of one pass)"""
        ms, fl = C.c_float(), C.c_double()
        _check(lib().mdx_net_time_forward(self.h, C.c_int32(reps), C.byref(ms), C.byref(fl)))
        return ms.value, fl.value
