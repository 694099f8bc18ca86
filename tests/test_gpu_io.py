"""Output and restart on the GPU: center_vel on the device, the driver's VTK files against the
reference's writer, HDF5 checkpoint/restart, and the driver's adaptive time step against the oracle.

  * mdx_center_velocity equals the oracle's center_vel (io.f90:184-222) bit for bit;
  * the Fortran driver's VTK file equals, byte for byte, the file the reference's own write_vtk
    (io.f90:168-259, compiled into oracle/_ref/libref_io.so) writes from the same final fields;
  * a run that writes a field file at step k, restarts from it and continues equals the
    uninterrupted run bit for bit (C ABI and Fortran driver; FP64 SOR with IBM and flow-rate
    control, and the flat FFT path);
  * the driver with cflmax > 0 takes the oracle's dt sequence (main.f90:91-97, read back from the
    per-step field files) and ends with the oracle's fields.
"""
import os
import subprocess

import numpy as np
import pytest

import mobydiff_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mobydiff_amd", "fortran", "mobydiff")
needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
needs_ref = pytest.mark.skipif(O.ref_io() is None, reason="reference output module not built (oracle/_ref)")

NX, NY, NZ = 32, 24, 16


def _wavy_state(g, seed=3, amp=0.05):
    """Physical start: two-sided wavy-wall coefficients and the oracle's seeded initial field."""
    geo = O.Geometry(kind=1, A=0.1, lam=1.0, half_height=1.0)
    coef = O.ibm_coefs(g, geo)
    f = O.new_fields(g)
    O.initial_field(g, f, seed, amp, geo)
    for a, b in (("un", "us"), ("vn", "vs"), ("wn", "ws")):
        f[a][...] = f[b]
    O.apply_bc_redblack(g, f)
    return f, coef


def _run_driver(text, cwd):
    ini = cwd / "input.ini"
    ini.write_text(text)
    print(f"fortran driver: {ini}", flush=True)
    out = subprocess.run([DRIVER, str(ini)], capture_output=True, text=True, timeout=120, cwd=cwd)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


def _read_dump(path):
    raw = open(path, "rb").read()
    nx, ny, nz = np.frombuffer(raw[:12], np.int32)
    n = (nx + 2) * (ny + 2) * (nz + 2)
    data = np.frombuffer(raw[12:], np.float64)
    return {k: data[q * n:(q + 1) * n].reshape(nz + 2, ny + 2, nx + 2) for q, k in enumerate(("un", "vn", "wn", "pn"))}


def _wavy_ini(extra_time="", pressure="solver = sor\ntol = 1e-8\ncheck_every = 10\nomega = 1.8", flow="",
              output="interval = 0", init=""):
    return f"""[grid]
nx = {NX}
ny = {NY}
nz = {NZ}
lx = 4.0
ly = 2.2
lz = 2.0
[flow]
re = 2000.0
forcing_x = 0.0
{flow}
[time]
dt = 5e-4
{extra_time}
[pressure]
{pressure}
[boundary]
periodic_y = false
y_min_p_type = neumann
y_max_p_type = neumann
[ibm]
geometry = wavy
a = 0.1
[init]
{init}
[output]
{output}
"""


def test_center_velocity_bit_exact():
    print(f"center_vel {NX}x{NY}x{NZ}", flush=True)
    g = O.Grid(NX, NY, NZ, lx=4.0, ly=2.2, lz=2.0)
    rng = np.random.default_rng(5)
    f = {k: rng.standard_normal(g.gshape) for k in ("un", "vn", "wn")}
    ch = M.Channel(M.channel_config(NX, NY, NZ, 4.0, 2.2, 2.0, 2000.0))
    ch.upload(**f)
    got = ch.center_velocity()
    ch.close()
    for a, b in zip(got, O.center_vel(f)):
        assert np.array_equal(a, b)


@needs_driver
@needs_ref
def test_driver_vtk_equals_reference_file(tmp_path):
    """Tutorial setup (reference bump, SOLA nIter 3, body force 1) at 32x24x16, VTK every second
    step: the driver's data_2.vtk is the file the reference's write_vtk writes from the same fields."""
    text = f"""[grid]
nx = {NX}
ny = {NY}
nz = {NZ}
[flow]
re = 100.0
forcing_x = 1.0
[time]
dt = 1.0e-4
nsteps = 2
[boundary]
periodic_y = false
y_min_p_type = neumann
y_max_p_type = neumann
[output]
interval = 2
prefix = data
dump_final = final.bin
"""
    _run_driver(text, tmp_path)
    ours = (tmp_path / "data_2.vtk").read_bytes()
    f = _read_dump(tmp_path / "final.bin")
    g = O.Grid(NX, NY, NZ, lx=1.0, ly=1.0, lz=1.0, re=100.0)
    O.ref_write_vtk(g, f, 1e-4, 0.0 + 1e-4 + 1e-4, 0.0, 2, str(tmp_path / "ref"))
    theirs = (tmp_path / "ref_2.vtk").read_bytes()
    assert len(ours.splitlines()) == 12 + NX * NY * NZ
    assert ours == theirs


def _restart_pair(cfg_fn, prepare, tmp_path, steps=4, k=2, dt=5e-4):
    """(uninterrupted fields after `steps`, fields after k steps + write/read + steps - k)"""
    a = M.Channel(cfg_fn())
    prepare(a)
    for _ in range(steps):
        a.step(dt)
    fa = a.download(("un", "vn", "wn", "pn"))
    a.close()
    b = M.Channel(cfg_fn())
    coef = prepare(b)
    for _ in range(k):
        b.step(dt)
    path = str(tmp_path / "ck.h5")
    b.write_field(path, dt, k * dt)
    b.close()
    c = M.Channel(cfg_fn())
    if coef is not None:
        c.set_ibm(*coef)
    c.read_field(path)
    for _ in range(steps - k):
        c.step(dt)
    fc = c.download(("un", "vn", "wn", "pn"))
    c.close()
    return fa, fc


def test_restart_bit_identical_wavy_fp64_sor(tmp_path):
    print("restart: wavy channel, FP64 SOR to 1e-8, flow-rate control", flush=True)

    def cfg():
        return M.channel_config(NX, NY, NZ, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8,
                                check_every=10, flowrate=True, flow_height=2.0)

    def prepare(ch):
        geo = M.wavy_geometry(A=0.1)
        coef = ch.compute_ibm(geo)
        ch.init_fields(seed=4, amp=0.05, geo=geo, dt=5e-4)
        return coef

    fa, fc = _restart_pair(cfg, prepare, tmp_path)
    for q in fa:
        assert np.array_equal(fa[q], fc[q]), (q, np.abs(fa[q] - fc[q]).max())


def test_restart_bit_identical_flat_fft(tmp_path):
    print("restart: flat stretched channel, FFT solve, flow-rate control", flush=True)
    import math

    def cfg():
        return M.channel_config(NX, NY, NZ, 4 * math.pi, 2.0, 2 * math.pi, 2800.0, solver=M.SOLVER_FFT,
                                flowrate=True, flow_height=2.0, stretch_beta=1.5, p_neumann_y=False)

    def prepare(ch):
        ch.init_fields(seed=1, amp=0.1, geo=None, dt=1e-3)
        return None

    fa, fc = _restart_pair(cfg, prepare, tmp_path, dt=1e-3)
    for q in fa:
        assert np.array_equal(fa[q], fc[q]), (q, np.abs(fa[q] - fc[q]).max())


@needs_driver
def test_driver_restart_bit_identical(tmp_path):
    """Driver run A: 4 steps writing field files at steps 2 and 4.  Run B restarts from A's step-2
    file ([init] restart) and runs 2 steps: B's final fields equal A's, and A's step-4 file holds them."""
    a_dir, b_dir = tmp_path / "a", tmp_path / "b"
    a_dir.mkdir()
    b_dir.mkdir()
    init = "profile = poiseuille\nseed = 3\namplitude = 0.05"
    _run_driver(_wavy_ini("nsteps = 4", flow="flow_rate = true\nflow_height = 2.0", init=init,
                          output="interval = 0\nfield_interval = 2\nfield_prefix = field\ndump_final = final.bin"), a_dir)
    out = _run_driver(_wavy_ini("nsteps = 2", flow="flow_rate = true\nflow_height = 2.0",
                                init=f"restart = {a_dir / 'field_2.h5'}",
                                output="interval = 0\ndump_final = final.bin"), b_dir)
    assert "restart from" in out
    fa, fb = _read_dump(a_dir / "final.bin"), _read_dump(b_dir / "final.bin")
    f4 = M.read_field_h5(str(a_dir / "field_4.h5"), NX, NY, NZ)
    for q in fa:
        assert np.array_equal(fa[q], fb[q]), (q, np.abs(fa[q] - fb[q]).max())
        assert np.array_equal(fa[q], f4[q]), q
    t = 0.0
    for _ in range(4):
        t = t + 5e-4
    assert M.read_field_info(str(a_dir / "field_4.h5")) == (NX, NY, NZ, 5e-4, t)


@needs_driver
@pytest.mark.parametrize("cflmax,dtmax", [(0.05, 1e-2), (0.5, 2e-3)])
def test_driver_adaptive_dt_matches_oracle(tmp_path, cflmax, dtmax):
    """main.f90:91-97 in the driver: after each step get_cfl, then dt = min(cflmax / cfl, dtmax).
    The driver starts from a field file of the oracle's wavy start and writes a field file every
    step; the dt and t_current attributes of those files are the oracle's dt sequence and time, and
    the last file holds the oracle's fields (SOLA red-black, nIter 3, omega 1.8), bit for bit."""
    g = O.Grid(NX, NY, NZ, lx=4.0, ly=2.2, lz=2.0, re=2000.0, forcing=(0.0, 0.0, 0.0), periodic=(1, 0, 1))
    f, coef = _wavy_state(g)
    dt0 = 1e-3
    M.write_field_h5(str(tmp_path / "start.h5"), NX, NY, NZ, 4.0, 2.2, 2.0, 2000.0, dt0, 0.0,
                     f["un"], f["vn"], f["wn"], f["pn"])
    nsteps = 4
    _run_driver(_wavy_ini(f"nsteps = {nsteps}\ncflmax = {cflmax}\ndtmax = {dtmax}",
                          pressure="solver = sola\nniter = 3\nomega = 1.8", init=f"restart = {tmp_path / 'start.h5'}",
                          output="interval = 0\nfield_interval = 1\nfield_prefix = field"), tmp_path)
    dts, cfls, _ = O.run_steps(g, f, coef, nsteps, dt0, cflmax=cflmax, dtmax=dtmax, solver=1, niter=3, omega=1.8)
    next_dt = min(cflmax / cfls[-1], dtmax)
    t = 0.0
    for s in range(1, nsteps + 1):
        t = t + dts[s - 1]
        _, _, _, dt_file, t_file = M.read_field_info(str(tmp_path / f"field_{s}.h5"))
        assert t_file == t, (s, t_file, t)
        assert dt_file == (dts[s] if s < nsteps else next_dt), (s, dt_file)
    if cflmax == 0.05:
        assert dts[1] == cflmax / cfls[0] < dtmax
    else:
        assert dts[1] == dtmax
    last = M.read_field_h5(str(tmp_path / f"field_{nsteps}.h5"), NX, NY, NZ)
    for q in ("un", "vn", "wn", "pn"):
        assert np.array_equal(last[q], f[q]), (q, np.abs(last[q] - f[q]).max())
