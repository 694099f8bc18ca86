"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on the same inputs.

Bar: bit-exact for the FP64 operators whose expression order is shared (momentum, boundary
conditions, SOLA sweeps, pressure-correction SOR, norms, IBM coefficients, the initial field);
FFT-based projections agree to 1e-12 relative (different, equally exact, FFT orderings); the
mixed-precision solve reaches the same ||div||_inf tolerance and agrees to 1e-6 relative.
"""
import numpy as np
import pytest

import mobydiff_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

KEYS = ("un", "vn", "wn", "pn", "us", "vs", "ws", "oru", "orv", "orw")


def _cfg_from(g, **kw):
    c = M.channel_config(g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, g.re, forcing=g.forcing,
                         periodic=g.periodic, p_neumann_y=False, stretch_beta=g.beta or 0.0, **kw)
    for q in range(24):
        c.bc_type[q] = int(g.bc_type[q])
        c.bc_value[q] = float(g.bc_value[q])
    return c


def _channel(g, f, coef=None, **kw):
    ch = M.Channel(_cfg_from(g, **kw))
    ch.upload(**f)
    if coef is not None:
        ch.set_ibm(*coef)
    return ch


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _bc_table(g, rng):
    for d in (1, 2, 3):
        for s in (0, 1):
            for v in range(4):
                g.bc_type[O.bci(d, s, v)] = int(rng.random() < 0.5)
                g.bc_value[O.bci(d, s, v)] = rng.standard_normal()


# ------------------------------------------------------------------------------ operators
@pytest.mark.parametrize("shape", [(16, 12, 8), (40, 20, 6)])
@pytest.mark.parametrize("per", [(1, 0, 1), (0, 1, 0)])
@pytest.mark.parametrize("ibm", [False, True])
@pytest.mark.parametrize("beta", [None, 1.5])
def test_momentum_bit_exact(shape, per, ibm, beta):
    rng = np.random.default_rng(1)
    g = O.Grid(*shape, lx=1.3, ly=2.0, lz=0.7, re=300.0, forcing=(1.0, 0.2, -0.1), periodic=per, beta=beta)
    f = O.random_fields(g, rng)
    coef = tuple(np.abs(rng.standard_normal(g.gshape)) * 30 for _ in range(3)) if ibm else None
    ch = _channel(g, f, coef, solver=M.SOLVER_SOLA)
    ch.momentum(0.1, -0.03, 0.05)
    out = ch.download()
    O.momentum(g, f, coef, 0.1, -0.03, 0.05)
    for k in KEYS:
        assert np.array_equal(out[k], f[k]), k


def test_momentum_bit_exact_wavy_production_shape():
    g = O.Grid(192, 128, 96, lx=4.0, ly=2.2, lz=2.0, re=2000.0, periodic=(1, 0, 1))
    geo = O.Geometry(kind=1, A=0.1, lam=1.0, half_height=1.0)
    coef = O.ibm_coefs(g, geo)
    rng = np.random.default_rng(2)
    f = O.random_fields(g, rng, amp=0.5)
    ch = _channel(g, f, coef, solver=M.SOLVER_SOR)
    ch.momentum(5e-4 * 8 / 15, 0.0, 5e-4 * 8 / 15)
    out = ch.download(("us", "vs", "ws", "oru"))
    O.momentum(g, f, coef, 5e-4 * 8 / 15, 0.0, 5e-4 * 8 / 15)
    for k in ("us", "vs", "ws", "oru"):
        assert np.array_equal(out[k], f[k]), k


def test_apply_bc_fft_bit_exact():
    rng = np.random.default_rng(3)
    g = O.Grid(16, 10, 8)
    f = O.random_fields(g, rng)
    ch = _channel(g, f, solver=M.SOLVER_FFT)
    ch.apply_bc()
    out = ch.download()
    O.apply_bc_fft(g, f)
    for k in KEYS:
        assert np.array_equal(out[k], f[k]), k


@pytest.mark.parametrize("per", [(1, 0, 1), (1, 1, 1), (0, 0, 0), (0, 1, 1)])
def test_apply_bc_redblack_bit_exact(per):
    rng = np.random.default_rng(4)
    g = O.Grid(12, 10, 6, lx=1.0, ly=2.0, lz=0.5, periodic=per)
    _bc_table(g, rng)
    f = O.random_fields(g, rng)
    ch = _channel(g, f, solver=M.SOLVER_SOLA)
    ch.apply_bc_redblack()
    out = ch.download()
    O.apply_bc_redblack(g, f)
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(out[k], f[k]), k


@pytest.mark.parametrize("per", [(1, 0, 1), (1, 1, 1), (0, 0, 0), (0, 1, 1)])
@pytest.mark.parametrize("niter", [1, 3])
def test_sola_projection_bit_exact(per, niter):
    rng = np.random.default_rng(5)
    g = O.Grid(12, 10, 6, lx=1.0, ly=1.5, lz=0.5, periodic=per)
    _bc_table(g, rng)
    f = O.random_fields(g, rng)
    coef = tuple(np.abs(rng.standard_normal(g.gshape)) * 20 for _ in range(3))
    ch = _channel(g, f, coef, solver=M.SOLVER_SOLA, niter=niter, omega=1.5)
    st = ch.pressure_projection(0.03)
    assert st.iterations == niter
    out = ch.download()
    O.sola_redblack(g, f, coef, niter, 1.5, 0.03)
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(out[k], f[k]), k


@pytest.mark.parametrize("per_y", [0, 1])
@pytest.mark.parametrize("ibm", [False, True])
@pytest.mark.parametrize("beta", [None, 1.5])
def test_sor_fixed_iterations_bit_exact(per_y, ibm, beta):
    if beta and per_y:
        pytest.skip("stretched grid is a wall-bounded grid")
    rng = np.random.default_rng(6)
    g = O.Grid(16, 12, 8, lx=1.0, ly=1.5, lz=0.5, periodic=(1, per_y, 1), beta=beta)
    g.bc_type[O.bci(2, 0, 0)] = 1
    g.bc_type[O.bci(2, 1, 0)] = 1
    f = O.random_fields(g, rng)
    coef = tuple(np.abs(rng.standard_normal(g.gshape)) * 20 for _ in range(3)) if ibm else None
    ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-300, max_iter=7, check_every=7)
    st = ch.pressure_projection(0.03)
    assert st.iterations == 7 and st.converged == 0
    out = ch.download()
    ost, phi, rhs = O.projection_sor(g, f, coef, 1.8, 0.0, 7, 7, 0.03)
    assert ost.iterations == 7
    # the reported residual is the residual of the final iterate
    assert st.div_inf == O.sor_residual_inf(g, coef, 0.03, rhs, phi)
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(out[k], f[k]), k


def test_sor_to_tolerance_bit_exact_wavy():
    g = O.Grid(48, 32, 24, lx=4.0, ly=2.2, lz=2.0, re=2000.0, periodic=(1, 0, 1))
    geo = O.Geometry(kind=1, A=0.1, lam=1.0, half_height=1.0)
    coef = O.ibm_coefs(g, geo)
    rng = np.random.default_rng(7)
    f = O.random_fields(g, rng, amp=0.05)
    ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8, max_iter=100000, check_every=10)
    st = ch.pressure_projection(5e-4 * 64 / 120)
    out = ch.download()
    ost, _, _ = O.projection_sor(g, f, coef, 1.8, 1e-8, 100000, 10, 5e-4 * 64 / 120)
    assert st.converged == 1 and ost.converged == 1
    assert st.iterations == ost.iterations
    assert st.div_inf < 1e-8
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(out[k], f[k]), k
    assert ch.div_norm() == O.div_inf(g, f)


def test_mixed_precision_sor_reaches_tolerance():
    g = O.Grid(48, 32, 24, lx=4.0, ly=2.2, lz=2.0, re=2000.0, periodic=(1, 0, 1))
    geo = O.Geometry(kind=1, A=0.1, lam=1.0, half_height=1.0)
    coef = O.ibm_coefs(g, geo)
    rng = np.random.default_rng(8)
    f = O.random_fields(g, rng, amp=0.05)
    ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8, max_iter=100000, check_every=10,
                  precision=M.PREC_MIXED, inner_iters=10)
    st = ch.pressure_projection(5e-4 * 64 / 120)
    out = ch.download()
    tight = {k: v.copy() for k, v in f.items()}
    ost, _, _ = O.projection_sor(g, f, coef, 1.8, 1e-8, 100000, 10, 5e-4 * 64 / 120)
    O.projection_sor(g, tight, coef, 1.8, 1e-13, 1000000, 10, 5e-4 * 64 / 120)
    assert st.converged == 1 and st.div_inf < 1e-8
    assert ch.div_norm() < 1.5e-8
    assert st.iterations <= 1.5 * ost.iterations + 10
    sl = (slice(1, -1),) * 3
    for k in ("un", "vn", "wn"):
        assert _rel(out[k][sl], f[k][sl]) < 1e-6, k

    def pm(a):   # p is fixed up to a constant by the Neumann solve
        return a[sl] - a[sl].mean()

    # both solves stop at ||div||_inf < 1e-8, so p differs by about the distance of either from the
    # converged p: the mixed p is no further from the converged (1e-13) p than 5x the FP64 oracle's
    e_oracle, e_mixed = _rel(pm(f["pn"]), pm(tight["pn"])), _rel(pm(out["pn"]), pm(tight["pn"]))
    print(f"p vs the converged solve: FP64 oracle {e_oracle:.3e}, mixed {e_mixed:.3e}; "
          f"mixed vs oracle {_rel(pm(out['pn']), pm(f['pn'])):.3e}", flush=True)
    assert e_mixed <= 5.0 * e_oracle
    assert _rel(pm(out["pn"]), pm(f["pn"])) < 1e-4


@pytest.mark.parametrize("beta", [None, 1.5])
@pytest.mark.parametrize("shape", [(16, 12, 8), (64, 24, 32)])
def test_fft_projection_matches_oracle(beta, shape):
    rng = np.random.default_rng(9)
    g = O.Grid(*shape, lx=2.0, ly=2.0, lz=1.0, periodic=(1, 0, 1), beta=beta)
    f = O.random_fields(g, rng)
    ch = _channel(g, f, solver=M.SOLVER_FFT)
    ch.apply_bc()
    ch.pressure_projection(0.07)
    ch.apply_bc()
    out = ch.download()
    O.apply_bc_fft(g, f)
    O.projection_fft(g, f, None, 0.07)
    O.apply_bc_fft(g, f)
    for k in ("un", "vn", "wn", "pn"):
        assert _rel(out[k], f[k]) < 1e-12, k
    assert ch.div_norm() < 1e-11


@pytest.mark.parametrize("solver", ["part", "wave", "thomas"])
@pytest.mark.parametrize("shape", [(16, 64, 8), (16, 160, 8), (8, 250, 8)])
def test_fft_tridiagonal_solvers_match_oracle(solver, shape, monkeypatch):
    """The three tridiagonal solvers of the FFT path on wall-normal sizes where the partitioned
    solve applies (64 <= ny <= 256): 16 segments of 4, of 12 with whole identity segments past ny,
    and of 16 with identity rows inside the last segment; stretched grid.  Each projection within
    1e-12 of the oracle's Thomas recurrence."""
    monkeypatch.setenv("MOBYDIFF_PART_TRIDIAG", "1" if solver == "part" else "0")
    monkeypatch.setenv("MOBYDIFF_WAVE_TRIDIAG", "0" if solver == "thomas" else "1")
    rng = np.random.default_rng(11)
    g = O.Grid(*shape, lx=2.0, ly=2.0, lz=1.0, periodic=(1, 0, 1), beta=1.5)
    f = O.random_fields(g, rng)
    ch = _channel(g, f, solver=M.SOLVER_FFT)
    ch.apply_bc()
    ch.pressure_projection(0.07)
    ch.apply_bc()
    out = ch.download()
    O.apply_bc_fft(g, f)
    O.projection_fft(g, f, None, 0.07)
    O.apply_bc_fft(g, f)
    for k in ("un", "vn", "wn", "pn"):
        assert _rel(out[k], f[k]) < 1e-12, (k, _rel(out[k], f[k]))
    assert ch.div_norm() < 1e-11


def test_cfl_and_div_norm_bit_exact():
    rng = np.random.default_rng(10)
    g = O.Grid(20, 14, 10, lx=2.0, ly=1.0, lz=0.5)
    f = O.random_fields(g, rng)
    ch = _channel(g, f, solver=M.SOLVER_SOLA)
    assert ch.get_cfl() == O.get_cfl(g, f)
    assert ch.div_norm() == O.div_inf(g, f)


@pytest.mark.parametrize("kind", [0, 1])
def test_ibm_coefficients_bit_exact(kind):
    if kind == 0:
        g = O.Grid(32, 24, 6, lx=1.0, ly=1.0, lz=0.25, re=100.0)
        geo = O.Geometry(kind=0, second_order=1, n_wave_x=1, amp_x=5 * g.dy)
        mgeo = M.bump_geometry()
    else:
        g = O.Grid(48, 40, 6, lx=4.0, ly=2.2, lz=0.5, re=2000.0)
        geo = O.Geometry(kind=1, second_order=1, A=0.13, lam=1.0, half_height=1.0)
        mgeo = M.wavy_geometry(A=0.13)
    ch = M.Channel(_cfg_from(g, solver=M.SOLVER_SOR))
    mine = ch.compute_ibm(mgeo)
    ref = O.ibm_coefs(g, geo)
    for a, b in zip(mine, ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("solver", [M.SOLVER_FFT, M.SOLVER_SOR])
def test_initial_field_matches_oracle(solver):
    dt = 1e-3
    if solver == M.SOLVER_FFT:
        g = O.Grid(32, 24, 16, lx=4 * np.pi, ly=2.0, lz=2 * np.pi, re=2800.0, periodic=(1, 0, 1), beta=1.5)
        geo = None
    else:
        g = O.Grid(48, 32, 24, lx=4.0, ly=2.2, lz=2.0, re=2000.0, periodic=(1, 0, 1))
        geo = O.Geometry(kind=1, A=0.1, lam=1.0, half_height=1.0)
    g.bc_type[O.bci(2, 0, 0)] = 1
    g.bc_type[O.bci(2, 1, 0)] = 1
    ch = M.Channel(_cfg_from(g, solver=solver, omega=1.8, tol=1e-8, check_every=10))
    coef = None
    if geo is not None:
        ch.compute_ibm(M.wavy_geometry(A=0.1))
        coef = O.ibm_coefs(g, geo)
    ch.init_fields(seed=3, amp=0.1, geo=M.wavy_geometry(A=0.1) if geo is not None else None, dt=dt)
    out = ch.download()
    f = O.new_fields(g)
    O.initial_field(g, f, 3, 0.1, geo)
    if solver == M.SOLVER_FFT:
        O.apply_bc_fft(g, f)
        O.projection_fft(g, f, None, dt * 64 / 120)
        O.apply_bc_fft(g, f)
        tol = 1e-13
    else:
        O.projection_sor(g, f, coef, 1.8, 1e-8, 1000000, 10, dt * 64 / 120)
        O.apply_bc_redblack(g, f)
        tol = 0.0
    for k in ("us", "vs", "ws", "un", "vn", "wn"):
        assert _rel(out[k], f[k]) <= tol, k
    assert not out["pn"].any()


# ------------------------------------------------------------------------------ LDS-resident sweeps
def _wavy(nx, ny, nz, per_y=0, ibm=True):
    g = O.Grid(nx, ny, nz, lx=4.0, ly=2.2, lz=2.0, re=2000.0, periodic=(1, per_y, 1))
    g.bc_type[O.bci(2, 0, 0)] = 1
    g.bc_type[O.bci(2, 1, 0)] = 1
    coef = O.ibm_coefs(g, O.Geometry(kind=1, A=0.1, lam=1.0, half_height=1.0)) if ibm else None
    return g, coef


@pytest.mark.parametrize("shape,per_y,ibm", [((48, 32, 24), 0, True), ((32, 16, 16), 1, False),
                                             ((96, 64, 48), 0, True)])
def test_lds_sweeps_bit_exact_fixed_iterations(shape, per_y, ibm, monkeypatch):
    monkeypatch.setenv("MOBYDIFF_LDS_SOR", "2")   # FP64 batches on the LDS-resident solver too
    g, coef = _wavy(*shape, per_y=per_y, ibm=ibm)
    rng = np.random.default_rng(21)
    f = O.random_fields(g, rng, amp=0.05)
    n = 9
    ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-300, max_iter=n, check_every=n)
    st = ch.pressure_projection(0.02)
    assert st.lds_sweeps == 1 and st.iterations == n, ch.info()
    out = ch.download()
    O.projection_sor(g, f, coef, 1.8, 0.0, n, n, 0.02)
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(out[k], f[k]), k


@pytest.mark.parametrize("kz", [1, 2, 4, 8])
@pytest.mark.parametrize("case", ["wavy", "stretched_random_mu", "periodic_y"])
def test_zmarch_sweeps_bit_exact_fixed_iterations(kz, case, monkeypatch):
    """Streaming sweeps with a thread marching kz planes (nz = 22: the last segment is short)
    against the oracle, bit for bit, for z-invariant (plane) and general 3-D mu tables."""
    monkeypatch.setenv("MOBYDIFF_LDS_SOR", "0")
    monkeypatch.setenv("MOBYDIFF_SWEEP_FUSED", "0")
    monkeypatch.setenv("MOBYDIFF_SWEEP_KZ", str(kz))
    rng = np.random.default_rng(23)
    if case == "wavy":
        g, coef = _wavy(48, 32, 22)
    elif case == "periodic_y":
        g, coef = _wavy(40, 16, 22, per_y=1, ibm=False)
    else:
        g = O.Grid(48, 24, 22, lx=4.0, ly=2.0, lz=2.0, re=2800.0, periodic=(1, 0, 1), beta=1.5)
        g.bc_type[O.bci(2, 0, 0)] = 1
        g.bc_type[O.bci(2, 1, 0)] = 1
        coef = tuple(np.where(rng.random(g.gshape) < 0.2, 0.0, 50.0 * rng.random(g.gshape)) for _ in range(3))
    f = O.random_fields(g, rng, amp=0.05)
    n = 7
    ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-300, max_iter=n, check_every=n)
    st = ch.pressure_projection(0.02)
    assert st.lds_sweeps == 0 and st.iterations == n, ch.info()
    out = ch.download()
    ch.close()
    O.projection_sor(g, f, coef, 1.8, 0.0, n, n, 0.02)
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(out[k], f[k]), (k, case, kz)


@pytest.mark.parametrize("kz", [2, 8])
def test_mixed_solve_same_for_sweep_kernels(kz, monkeypatch):
    """The FP32 defect-correction solve gives the same iterates with the one-cell-per-thread sweep
    (kz = 0) and the z-marching sweep."""
    g, coef = _wavy(48, 32, 24)
    rng = np.random.default_rng(24)
    f = O.random_fields(g, rng, amp=0.05)
    monkeypatch.setenv("MOBYDIFF_LDS_SOR", "0")
    monkeypatch.setenv("MOBYDIFF_SWEEP_FUSED", "0")
    outs = []
    for v in ("0", str(kz)):
        monkeypatch.setenv("MOBYDIFF_SWEEP_KZ", v)
        ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8, check_every=10,
                      precision=M.PREC_MIXED, inner_iters=10)
        st = ch.pressure_projection(5e-4 * 64 / 120)
        assert st.converged == 1
        outs.append(ch.download())
        ch.close()
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.parametrize("shape,case", [((48, 32, 22), "wavy"), ((48, 20, 22), "stretched_random_mu"),
                                        ((40, 16, 22), "periodic_y"), ((16, 12, 8), "wavy"),
                                        ((192, 128, 96), "wavy")])
def test_fused_iterations_bit_exact_fixed_iterations(shape, case, monkeypatch):
    """One launch per red+black iteration (tiles with a recomputed halo) against the oracle, bit
    for bit: partial tiles in x and z, walls and periodic y, z-invariant and 3-D mu, a grid smaller
    than one tile's halo region (16 x 12 x 8: the region wraps onto itself)."""
    monkeypatch.setenv("MOBYDIFF_LDS_SOR", "0")
    monkeypatch.setenv("MOBYDIFF_SWEEP_FUSED", "1")
    rng = np.random.default_rng(25)
    nx, ny, nz = shape
    if case == "wavy":
        g, coef = _wavy(nx, ny, nz)
    elif case == "periodic_y":
        g, coef = _wavy(nx, ny, nz, per_y=1, ibm=False)
    else:
        g = O.Grid(nx, ny, nz, lx=4.0, ly=2.0, lz=2.0, re=2800.0, periodic=(1, 0, 1), beta=1.5)
        g.bc_type[O.bci(2, 0, 0)] = 1
        g.bc_type[O.bci(2, 1, 0)] = 1
        coef = tuple(np.where(rng.random(g.gshape) < 0.2, 0.0, 50.0 * rng.random(g.gshape)) for _ in range(3))
    f = O.random_fields(g, rng, amp=0.05)
    n = 7
    ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-300, max_iter=n, check_every=n)
    st = ch.pressure_projection(0.02)
    assert st.iterations == n and "fused" in ch.info(), ch.info()
    out = ch.download()
    ch.close()
    O.projection_sor(g, f, coef, 1.8, 0.0, n, n, 0.02)
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(out[k], f[k]), (k, case, shape)


def test_fused_mixed_solve_equals_one_colour_sweeps(monkeypatch):
    g, coef = _wavy(48, 32, 24)
    rng = np.random.default_rng(26)
    f = O.random_fields(g, rng, amp=0.05)
    monkeypatch.setenv("MOBYDIFF_LDS_SOR", "0")
    outs = []
    for fused in ("0", "1"):
        monkeypatch.setenv("MOBYDIFF_SWEEP_FUSED", fused)
        ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8, check_every=10,
                      precision=M.PREC_MIXED, inner_iters=10)
        st = ch.pressure_projection(5e-4 * 64 / 120)
        assert st.converged == 1
        outs.append(ch.download())
        ch.close()
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_lds_fp32_batches_equal_streaming_at_production_shape(monkeypatch):
    """192 x 128 x 96 wavy channel, mixed precision, a fixed 60 iterations (three FP32 batches of
    20 between FP64 defect corrections): the LDS-resident batches (the default) and the streaming
    one-colour sweeps give bit-identical fields."""
    g, coef = _wavy(192, 128, 96)
    rng = np.random.default_rng(24)
    f = O.random_fields(g, rng, amp=0.05)
    outs, stats = [], []
    for mode in ("1", "0"):
        monkeypatch.setenv("MOBYDIFF_LDS_SOR", mode)
        ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-300, max_iter=60, check_every=20,
                      precision=M.PREC_MIXED, inner_iters=20)
        stats.append(ch.pressure_projection(0.02))
        outs.append(ch.download(("un", "vn", "wn", "pn")))
        ch.close()
    assert stats[0].lds_sweeps == 1 and stats[1].lds_sweeps == 0
    assert stats[0].iterations == stats[1].iterations == 60
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.parametrize("form", ["halves", "columns", "pairs"])
def test_lds_kernel_forms_equal_streaming_at_production_shape(form, monkeypatch):
    """Each thread layout of the LDS-resident kernel (half a z-column per thread, a whole column,
    x-pairs with an LDS coefficient table) at 192 x 128 x 96, mixed precision, a fixed 40
    iterations, including the kernel's defect epilogue: bit-identical to the streaming sweeps, and
    the plan reports the form asked for."""
    g, coef = _wavy(192, 128, 96)
    rng = np.random.default_rng(25)
    f = O.random_fields(g, rng, amp=0.05)
    outs, infos = [], []
    for mode in ("1", "0"):
        monkeypatch.setenv("MOBYDIFF_LDS_SOR", mode)
        monkeypatch.setenv("MOBYDIFF_LDS_FORM", form)
        ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-300, max_iter=40, check_every=20,
                      precision=M.PREC_MIXED, inner_iters=20)
        st = ch.pressure_projection(0.02)
        assert st.iterations == 40
        outs.append(ch.download(("un", "vn", "wn", "pn")))
        infos.append(ch.info())
        ch.close()
    want = {"halves": "(half a z-column per thread)", "columns": "(a z-column per thread)", "pairs": "(x-pairs)"}[form]
    assert want in infos[0], infos[0]
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_lds_overlapped_schedule_equals_plain_order(precision, monkeypatch):
    """The half-column form relaxes the interior columns' halo-free cells while the exchange is in
    flight (the rest of colour 1 after the publish, the next iteration's colour 0 after that);
    with MOBYDIFF_LDS_OVERLAP=0 every cell is relaxed in its colour pass.  At 192 x 128 x 96, in
    the FP32 batches of the mixed solve and in FP64 batches on the same kernel (MOBYDIFF_LDS_SOR=2),
    60 iterations give bit-identical fields."""
    g, coef = _wavy(192, 128, 96)
    rng = np.random.default_rng(27)
    f = O.random_fields(g, rng, amp=0.05)
    prec = M.PREC_MIXED if precision == "mixed" else M.PREC_FP64
    monkeypatch.setenv("MOBYDIFF_LDS_SOR", "1" if precision == "mixed" else "2")
    outs = []
    for ovl in ("1", "0"):
        monkeypatch.setenv("MOBYDIFF_LDS_OVERLAP", ovl)
        ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-300, max_iter=60, check_every=20,
                      precision=prec, inner_iters=20)
        st = ch.pressure_projection(0.02)
        assert st.lds_sweeps == 1 and st.iterations == 60, (st.lds_sweeps, st.iterations)
        outs.append(ch.download(("un", "vn", "wn", "pn")))
        ch.close()
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_lds_kernel_cooperative_launch_equals_ordinary(monkeypatch):
    """The LDS-resident kernel launched cooperatively (MOBYDIFF_LDS_COOP=1) computes exactly what
    the ordinary launch of the same co-resident grid computes."""
    g, coef = _wavy(96, 64, 48)
    rng = np.random.default_rng(26)
    f = O.random_fields(g, rng, amp=0.05)
    outs = []
    for coop in ("1", "0"):
        monkeypatch.setenv("MOBYDIFF_LDS_COOP", coop)
        ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8, check_every=20,
                      precision=M.PREC_MIXED, inner_iters=20)
        st = ch.pressure_projection(0.02)
        assert st.lds_sweeps == 1 and st.converged == 1
        outs.append((st.iterations, ch.download(("un", "vn", "wn", "pn"))))
        ch.close()
    assert outs[0][0] == outs[1][0]
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(outs[0][1][k], outs[1][1][k]), k


@pytest.mark.parametrize("lds", ["1", "0"])
def test_merged_batches_skip_only_failing_checks(lds, monkeypatch):
    """Mixed SOR with merged batches (up to three inner batches while the residual decay predicts
    their checks would fail) against every check kept (MOBYDIFF_MERGE_MAX=1): both reach
    the tolerance, the merged solve needs no more iterations than one extra inner batch and far
    fewer checks, and its fields stay within the tolerance's reach of the unmerged ones."""
    monkeypatch.setenv("MOBYDIFF_LDS_SOR", lds)
    g, coef = _wavy(96, 64, 48)
    rng = np.random.default_rng(27)
    f = O.random_fields(g, rng, amp=0.05)
    res = []
    for mm in ("1", "3"):
        monkeypatch.setenv("MOBYDIFF_MERGE_MAX", mm)
        ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8, check_every=20,
                      precision=M.PREC_MIXED, inner_iters=20)
        st = ch.pressure_projection(0.02)
        res.append((st, ch.download(("un", "vn", "wn"))))
        ch.close()
    (s1, f1), (s3, f3) = res
    assert s1.converged == 1 and s3.converged == 1
    assert s1.div_inf < 1e-8 and s3.div_inf < 1e-8
    assert s1.iterations <= s3.iterations <= s1.iterations + 20
    assert s3.checks < 0.6 * s1.checks, (s1.checks, s3.checks)
    for k in ("un", "vn", "wn"):
        assert np.max(np.abs(f1[k] - f3[k])) < 1e-6, k


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
def test_lds_sweeps_equal_streaming_sweeps(precision, monkeypatch):
    g, coef = _wavy(48, 32, 24)
    rng = np.random.default_rng(22)
    f = O.random_fields(g, rng, amp=0.05)
    outs, stats, infos = [], [], []
    for on in ("2", "0"):
        monkeypatch.setenv("MOBYDIFF_LDS_SOR", on)
        ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8, check_every=10,
                      precision=precision, inner_iters=10)
        stats.append(ch.pressure_projection(5e-4 * 64 / 120))
        outs.append(ch.download())
        infos.append(ch.info())
        ch.close()
    assert stats[0].lds_sweeps == 1 and stats[1].lds_sweeps == 0, infos
    assert stats[0].iterations == stats[1].iterations and stats[0].converged == 1
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
def test_lds_wait_timeout_restarts_on_streaming_sweeps(precision, monkeypatch):
    """A neighbour wait of the LDS-resident sweeps that gives up (forced from the host: a spin limit
    of 0 polls, so a wait that finds its neighbour not yet published gives up at once and the
    kernel runs to its end with the error word set) makes the solve restart from Phi = 0 on the
    streaming sweeps.  The step then equals a run on the streaming sweeps bit for bit, the solve
    reports the restart, the error word is cleared (later norms succeed), and after three restarts
    the handle stays on the streaming sweeps."""
    g, coef = _wavy(96, 64, 48)
    rng = np.random.default_rng(28)
    f = O.random_fields(g, rng, amp=0.05)
    outs, stats, infos = [], [], []
    for limit, lds in (("0", "2"), (None, "0")):
        if limit is None:
            monkeypatch.delenv("MOBYDIFF_LDS_SPIN_LIMIT", raising=False)
        else:
            monkeypatch.setenv("MOBYDIFF_LDS_SPIN_LIMIT", limit)
        monkeypatch.setenv("MOBYDIFF_LDS_SOR", lds)
        ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8, check_every=20,
                      precision=precision, inner_iters=20)
        st = ch.step(5e-4)
        stats.append([st.stage[q] for q in range(3)])
        assert ch.div_norm() < 1.5e-8 and ch.get_cfl() > 0.0
        outs.append(ch.download(("un", "vn", "wn", "pn")))
        infos.append(ch.info())
        ch.close()
    restarts = [s.lds_restarts for s in stats[0]]
    print("restarts per stage", restarts, infos[0], flush=True)
    assert sum(restarts) >= 1
    assert all(s.converged == 1 for s in stats[0])
    assert all(s.lds_sweeps == 0 for s in stats[0] if s.lds_restarts)
    assert [s.iterations for s in stats[0]] == [s.iterations for s in stats[1]]
    if sum(restarts) >= 3:
        assert "gave up 3 times" in infos[0], infos[0]
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.gpu
def test_time_momentum_keeps_rk3_history():
    """mdx_time_momentum (the bench's momentum-kernel timing) restores the old right-hand sides:
    a step after it is bit-identical to a step without it."""
    g, coef = _wavy(48, 32, 24)
    f = O.random_fields(g, np.random.default_rng(31), amp=0.05)
    res = []
    for timed in (False, True):
        ch = _channel(g, f, coef, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8, check_every=20)
        ch.step(5e-4)
        if timed:
            ms = ch.time_momentum(50 / 120 * 5e-4, -34 / 120 * 5e-4, 16 / 120 * 5e-4, reps=3)
            assert ms > 0.0
        ch.step(5e-4)
        res.append(ch.download(("un", "vn", "wn", "pn")))
        ch.close()
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(res[0][k], res[1][k]), k
