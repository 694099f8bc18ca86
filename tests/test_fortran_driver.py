"""The Fortran drop-in driver (mobydiff_amd/fortran): input.ini grammar on the CPU, and on the GPU a
driver run against the oracle (reference setup and solver) and against the Python C-ABI path.

The grammar cases follow the reference reader (config.f90:10-431): `[section]` headers, `key =
value` and bare `key value`, `;`/`#` comments, case-insensitive keys and the same aliases,
t_final -> nsteps only when nsteps is absent, the boundary key grammar
<dir>_<side>_<var>_<type|value>, and unparsable values keeping the default.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mobydiff_amd", "fortran", "mobydiff")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER),
                                reason="Fortran driver not built (run __graft_entry__.build())")


def _parse(text, tmp_path):
    ini = tmp_path / "input.ini"
    ini.write_text(text)
    out = subprocess.run([DRIVER, str(ini), "--print-config"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    cfg, bc = {}, {}
    for line in out.stdout.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "bc":
            d, s, q, t = (int(x) for x in parts[1:5])
            bc[(d, s, q)] = (t, float(parts[5]))
        elif parts[0] in ("grid", "lengths", "re", "forcing", "dt", "nsteps", "t_final", "cfl",
                          "stretch_beta", "flow_rate", "pressure", "sor", "periodic", "ibm", "init",
                          "output", "dump_final"):
            cfg[parts[0]] = parts[1:]
    return cfg, bc, out.stdout


# the reference's tutorial input (tutorials/wavychannel/input.ini), restated
TUTORIAL = """[grid]
nx = 200          ; number of points in the streamwise direction
ny = 200
nz = 200
lx = 1.0
ly = 1.0
lz = 1.0

[flow]
re = 100.0
forcing_x = 1.0
forcing_y = 0.0
forcing_z = 0.0

[time]
dt = 1.0e-4       ; use 1.0e-3 for no-IBM runs if desired
nsteps = 3
cflmax = 0.0
dtmax = 1.0e-3

[pressure]
niter = 3          ; red-black SOR iterations, ignored by FFT backends

[boundary]
periodic_x = true
periodic_y = false
periodic_z = true

y_min_p_type = neumann
y_max_p_type = neumann

[output]
interval = 0       ; use 0 to disable VTK output
prefix = data
field_interval = 0 ; use >0 to write full staggered fields in HDF5
field_prefix = field
"""


def test_tutorial_input(tmp_path):
    cfg, bc, _ = _parse(TUTORIAL, tmp_path)
    assert cfg["grid"] == ["200", "200", "200"]
    assert [float(x) for x in cfg["lengths"]] == [1.0, 1.0, 1.0]
    assert float(cfg["re"][0]) == 100.0
    assert [float(x) for x in cfg["forcing"]] == [1.0, 0.0, 0.0]
    assert float(cfg["dt"][0]) == 1e-4 and cfg["nsteps"] == ["3"]
    assert cfg["pressure"][:2] == ["1", "3"] and float(cfg["pressure"][2]) == 1.5
    assert cfg["periodic"] == ["T", "F", "T"]
    for key, (t, v) in bc.items():
        expect = 1 if key in ((2, 0, 0), (2, 1, 0)) else 0
        assert t == expect and v == 0.0, (key, t, v)
    assert cfg["output"] == ["0", "data", "0", "field"]


def test_defaults_without_file(tmp_path):
    out = subprocess.run([DRIVER, str(tmp_path / "missing.ini"), "--print-config"], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 0
    assert "input file not found; using defaults" in out.stdout
    lines = dict((ln.split()[0], ln.split()[1:]) for ln in out.stdout.splitlines() if ln and not ln.startswith(" "))
    # init.f90:23-46 defaults
    assert lines["grid"] == ["200", "200", "200"] and lines["nsteps"] == ["10000"]
    assert float(lines["dt"][0]) == 1e-4 and float(lines["re"][0]) == 100.0
    assert lines["periodic"] == ["T", "T", "T"]


def test_grammar_aliases_comments_and_bare_values(tmp_path):
    text = """
# full-line comment
[GRID]
NX 48            # bare `key value`
ny=32
nz   =   16 ; trailing comment
[Flow]
meanpx = 2.5
body_force_z = -1.0
[time]
steps = 7
[pressure]
omega = 1.8
iterations = 5
[boundary]
x_periodic = no
isperiodic_y = .true.
streamwise_lo_un_type = neumann
i_high_u_value = 0.25
k_front_w_type = 1
wallnormal_top_vn_value = -3.5
"""
    cfg, bc, _ = _parse(text, tmp_path)
    assert cfg["grid"] == ["48", "32", "16"]
    assert [float(x) for x in cfg["forcing"]] == [2.5, 0.0, -1.0]
    assert cfg["nsteps"] == ["7"]
    assert cfg["pressure"][1] == "5" and float(cfg["pressure"][2]) == 1.8
    assert cfg["periodic"] == ["F", "T", "T"]
    assert bc[(1, 0, 1)][0] == 1
    assert bc[(1, 1, 1)][1] == 0.25
    assert bc[(3, 0, 3)][0] == 1
    assert bc[(2, 1, 2)][1] == -3.5


def test_t_final_sets_nsteps_only_without_nsteps(tmp_path):
    cfg, _, _ = _parse("[time]\ndt = 0.01\nt_final = 0.105\n", tmp_path)
    assert cfg["nsteps"] == ["11"]                      # ceiling(t_final/dt)
    assert abs(float(cfg["t_final"][0]) - 0.11) < 1e-15  # finalize_grid: dt * nsteps
    cfg, _, _ = _parse("[time]\ndt = 0.01\nt_final = 0.105\nnsteps = 4\n", tmp_path)
    assert cfg["nsteps"] == ["4"]


def test_bad_values_keep_defaults_with_warning(tmp_path):
    cfg, bc, out = _parse("[grid]\nnx = many\n[pressure]\nniter = 0\n[boundary]\nq_min_p_type = neumann\n"
                          "y_min_p_kind = 1\n", tmp_path)
    assert cfg["grid"][0] == "200"
    assert cfg["pressure"][1] == "3"
    assert "could not parse integer" in out
    assert "nIter must be positive" in out
    assert "unknown boundary key" in out
    assert "must end in _type or _value" in out


def test_extension_keys(tmp_path):
    text = """[grid]
stretch_beta = 1.5
[flow]
flow_rate = true
ub = 1.0
flow_height = 2.0
[pressure]
solver = sor
tol = 1e-9
max_iter = 5000
check_every = 10
precision = mixed
inner_iters = 16
[ibm]
geometry = wavy
a = 0.12
lambda = 2.0
half_height = 1.0
[init]
profile = poiseuille
seed = 7
amplitude = 0.05
[output]
dump_final = "final.bin"
"""
    cfg, _, _ = _parse(text, tmp_path)
    assert float(cfg["stretch_beta"][0]) == 1.5
    assert cfg["flow_rate"][0] == "T" and float(cfg["flow_rate"][2]) == 2.0
    assert cfg["pressure"][0] == "2"
    assert float(cfg["sor"][0]) == 1e-9 and cfg["sor"][1:] == ["5000", "10", "1", "16"]
    assert cfg["ibm"][:2] == ["1", "1"] and float(cfg["ibm"][3]) == 0.12 and float(cfg["ibm"][4]) == 2.0
    assert cfg["init"] == ["T", "7", cfg["init"][2]] and float(cfg["init"][2]) == 0.05
    assert cfg["dump_final"] == ["final.bin"]


# ------------------------------------------------------------------------------------------ GPU
def _read_dump(path):
    raw = open(path, "rb").read()
    nx, ny, nz = np.frombuffer(raw[:12], np.int32)
    n = (nx + 2) * (ny + 2) * (nz + 2)
    data = np.frombuffer(raw[12:], np.float64)
    assert data.size == 4 * n
    return {k: data[q * n:(q + 1) * n].reshape(nz + 2, ny + 2, nx + 2) for q, k in enumerate(("un", "vn", "wn", "pn"))}


def _run_driver(text, tmp_path):
    ini = tmp_path / "input.ini"
    ini.write_text(text)
    print(f"fortran driver: {ini}", flush=True)
    out = subprocess.run([DRIVER, str(ini)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "timing: nsteps" in out.stdout
    return out.stdout


@pytest.mark.gpu
def test_driver_tutorial_setup_matches_oracle(tmp_path):
    """The tutorial input at 32x24x16: reference bump geometry, SOLA red-black with nIter = 3, zero
    start, body force 1 -- the reference's own configuration -- bit-exact against the oracle."""
    from oracle import oracle as O
    text = TUTORIAL.replace("nx = 200          ; number of points in the streamwise direction", "nx = 32")
    text = text.replace("ny = 200", "ny = 24").replace("nz = 200", "nz = 16")
    text += "dump_final = final.bin\n"
    _run_driver(text, tmp_path)
    got = _read_dump(tmp_path / "final.bin")

    g = O.Grid(32, 24, 16, lx=1.0, ly=1.0, lz=1.0, re=100.0, forcing=(1.0, 0.0, 0.0), periodic=(1, 0, 1))
    # init_ibm (ibm.f90:47): bump amplitude 5 dy
    coef = O.ibm_coefs(g, O.Geometry(kind=0, amp_x=5 * (1.0 / 24)))
    f = O.new_fields(g)
    for _ in range(3):
        O.rk3_step(g, f, coef, 1e-4, solver=1, niter=3, omega=1.5)
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(got[k], f[k]), (k, np.abs(got[k] - f[k]).max())


@pytest.mark.gpu
def test_driver_wavy_mixed_sor_matches_library_path(tmp_path):
    """Wavy channel, Poiseuille + perturbation start, FP32 sweeps in FP64 defect correction to
    1e-8: the Fortran driver's fields equal the Python C-ABI path's, bit for bit."""
    import mobydiff_amd as M
    nx, ny, nz, dt = 32, 24, 16, 5e-4
    text = f"""[grid]
nx = {nx}
ny = {ny}
nz = {nz}
lx = 4.0
ly = 2.2
lz = 2.0
[flow]
re = 2000.0
forcing_x = 0.0
[time]
dt = {dt}
nsteps = 2
[pressure]
solver = sor
omega = 1.8
tol = 1e-8
check_every = 10
precision = mixed
inner_iters = 10
[boundary]
periodic_y = false
y_min_p_type = neumann
y_max_p_type = neumann
[ibm]
geometry = wavy
a = 0.1
[init]
profile = poiseuille
seed = 3
amplitude = 0.05
[output]
interval = 0
dump_final = final.bin
"""
    _run_driver(text, tmp_path)
    got = _read_dump(tmp_path / "final.bin")

    cfg = M.channel_config(nx, ny, nz, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8,
                           check_every=10, precision=M.PREC_MIXED, inner_iters=10)
    ch = M.Channel(cfg)
    geo = M.wavy_geometry(A=0.1)
    ch.compute_ibm(geo)
    ch.init_fields(seed=3, amp=0.05, geo=geo, dt=dt)
    for _ in range(2):
        ch.step(dt)
    ref = ch.download()
    ch.close()
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(got[k], ref[k]), (k, np.abs(got[k] - ref[k]).max())
