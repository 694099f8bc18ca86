"""Pin the C oracle to the reference's own Fortran modules (compiled in place into oracle/_ref).

Every comparison is bit-exact: both sides evaluate the same FP64 expressions in the same order.
Inputs are random fields (including ghosts) so that every branch and every ghost read of the
reference is exercised, not only physically consistent states.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.skipif(O.ref_rb() is None or O.ref_fft() is None,
                                reason="reference build oracle/_ref not present")

PERIODICITIES = [(1, 0, 1), (1, 1, 1), (0, 0, 0), (0, 1, 1)]


def _bc_table(g, rng, neumann_frac=0.5):
    for d in (1, 2, 3):
        for s in (0, 1):
            for v in range(4):
                g.bc_type[O.bci(d, s, v)] = int(rng.random() < neumann_frac)
                g.bc_value[O.bci(d, s, v)] = rng.standard_normal()


def _same(a, b, keys):
    for k in keys:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("per", PERIODICITIES)
@pytest.mark.parametrize("with_ibm", [False, True])
def test_momentum_bit_exact(per, with_ibm):
    rng = np.random.default_rng(11)
    g = O.Grid(12, 10, 8, lx=1.3, ly=1.1, lz=0.7, re=250.0, forcing=(1.0, 0.3, -0.2), periodic=per)
    f = O.random_fields(g, rng)
    coef = tuple(np.abs(rng.standard_normal(g.gshape)) * 50 for _ in range(3)) if with_ibm else None
    f2 = {k: v.copy() for k, v in f.items()}
    O.momentum(g, f, coef, 0.1, -0.05, 0.07)
    O.ref_momentum(g, f2, coef, 0.1, -0.05, 0.07)
    _same(f, f2, f.keys())


def test_get_cfl_matches():
    rng = np.random.default_rng(3)
    g = O.Grid(10, 8, 6, lx=2.0, ly=1.0, lz=0.5)
    f = O.random_fields(g, rng)
    assert O.get_cfl(g, f) == O.ref_get_cfl(g, f)


@pytest.mark.parametrize("per", PERIODICITIES)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_apply_bc_redblack_bit_exact(per, seed):
    rng = np.random.default_rng(100 + seed)
    g = O.Grid(8, 6, 4, lx=1.0, ly=2.0, lz=0.5, periodic=per)
    _bc_table(g, rng)
    f = O.random_fields(g, rng)
    f2 = {k: v.copy() for k, v in f.items()}
    O.apply_bc_redblack(g, f)
    O.ref_apply_bc_redblack(g, f2)
    _same(f, f2, ("un", "vn", "wn", "pn"))


def test_apply_bc_fft_bit_exact():
    rng = np.random.default_rng(5)
    g = O.Grid(8, 6, 4)
    f = O.random_fields(g, rng)
    f2 = {k: v.copy() for k, v in f.items()}
    O.apply_bc_fft(g, f)
    O.ref_apply_bc_fft(g, f2)
    _same(f, f2, ("un", "vn", "wn", "us", "vs", "ws", "pn"))


@pytest.mark.parametrize("per", PERIODICITIES)
@pytest.mark.parametrize("niter,sor", [(1, 1.5), (3, 1.8)])
def test_sola_redblack_bit_exact(per, niter, sor):
    rng = np.random.default_rng(7)
    g = O.Grid(8, 6, 4, lx=1.0, ly=1.5, lz=0.5, periodic=per)
    _bc_table(g, rng)
    f = O.random_fields(g, rng)
    coef = tuple(np.abs(rng.standard_normal(g.gshape)) * 20 for _ in range(3))
    f2 = {k: v.copy() for k, v in f.items()}
    O.sola_redblack(g, f, coef, niter, sor, 0.03)
    O.ref_projection_redblack(g, f2, coef, niter, sor, 0.03)
    _same(f, f2, ("un", "vn", "wn", "pn"))


@pytest.mark.parametrize("dirs", [(1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 0, 0)])
@pytest.mark.parametrize("amp,phase,nwave", [(0.0, 0.0, 0), (0.21, 0.4, 2)])
def test_set_ibm_coeff_bit_exact(dirs, amp, phase, nwave):
    g = O.Grid(24, 20, 4, lx=1.0, ly=1.0, lz=0.25, re=100.0)
    geo = O.Geometry(kind=0, second_order=1, n_wave_x=nwave if nwave else 1,
                     amp_x=amp if amp else 5 * g.dy, phase_x=phase)
    mine = O.set_ibm_coeff(g, geo, *dirs)
    ref = O.ref_set_ibm_coeff(g, *dirs, amp_x=amp, phase_x=phase, n_wave_x=nwave)
    assert np.array_equal(mine, ref)
    # the mask classes are non-trivial: solid, forced and fluid cells all occur
    assert (mine == 1e30 / g.re).any() and (mine == 0).any()
    assert ((mine > 0) & (mine < 1e30 / g.re)).any()


@pytest.mark.parametrize("shape", [(8, 6, 4), (16, 12, 8), (12, 9, 10)])
def test_tridiag_bit_exact(shape):
    rng = np.random.default_rng(9)
    nx, ny, nz = shape
    g = O.Grid(nx, ny, nz, lx=2.0, ly=1.5, lz=1.0)
    nxh = nx // 2 + 1
    hat = rng.standard_normal((ny, nz, nxh)) + 1j * rng.standard_normal((ny, nz, nxh))
    mine, cp, di = O.tridiag(g, hat)
    ref, rcp, rdi = O.ref_tridiag(g, hat)
    assert np.array_equal(cp, rcp.real) and not rcp.imag.any()
    assert np.array_equal(di, rdi.real) and not rdi.imag.any()
    assert np.array_equal(mine, ref)


def _wavy_state(g, seed=0, amp=0.05):
    """Physical start: two-sided wavy-wall coefficients and the oracle's seeded initial field."""
    geo = O.Geometry(kind=1, A=0.1, lam=1.0, half_height=1.0)
    coef = O.ibm_coefs(g, geo)
    f = O.new_fields(g)
    O.initial_field(g, f, seed, amp, geo)   # fills the predictor fields
    for a, b in (("un", "us"), ("vn", "vs"), ("wn", "ws")):
        f[a][...] = f[b]
    O.apply_bc_redblack(g, f)
    return f, coef


@pytest.mark.parametrize("per", [(1, 0, 1), (1, 1, 1), (0, 0, 0)])
@pytest.mark.parametrize("niter,sor", [(1, 1.5), (3, 1.8)])
def test_rk3_time_loop_bit_exact(per, niter, sor):
    """The oracle's RK3 stage composition (rk3_step, SOLA build) against the reference's own time loop
    (main.f90:64-89 over momentum / pressure_projection / apply_bc_redblack with step.f90's RK3
    constants), two steps from random fields, coefficients and BC tables."""
    rng = np.random.default_rng(21)
    g = O.Grid(8, 6, 4, lx=1.0, ly=1.5, lz=0.5, re=300.0, forcing=(0.7, 0.1, -0.2), periodic=per)
    _bc_table(g, rng)
    f = O.random_fields(g, rng)
    coef = tuple(np.abs(rng.standard_normal(g.gshape)) * 20 for _ in range(3))
    f2 = {k: v.copy() for k, v in f.items()}
    dts, cfls, _ = O.run_steps(g, f, coef, 2, 0.01, solver=1, niter=niter, omega=sor)
    rdts, rcfls = O.ref_run_redblack(g, f2, coef, niter, sor, 2, 0.01)
    assert dts == rdts and cfls == rcfls == [0.0, 0.0]
    _same(f, f2, f.keys())


def test_rk3_time_loop_wavy_channel_bit_exact():
    """Same on the wavy channel (IBM walls, seeded Poiseuille start), three steps."""
    g = O.Grid(16, 12, 8, lx=4.0, ly=2.2, lz=2.0, re=2000.0, forcing=(0.0, 0.0, 0.0), periodic=(1, 0, 1))
    f, coef = _wavy_state(g)
    f2 = {k: v.copy() for k, v in f.items()}
    O.run_steps(g, f, coef, 3, 5e-4, solver=1, niter=3, omega=1.8)
    O.ref_run_redblack(g, f2, coef, 3, 1.8, 3, 5e-4)
    _same(f, f2, f.keys())
    assert np.abs(f["un"]).max() > 0.5   # a real flow, not a zero field


@pytest.mark.parametrize("cflmax,dtmax", [(0.05, 1e-2), (0.5, 2e-3)])
def test_adaptive_dt_bit_exact(cflmax, dtmax):
    """main.f90:91-97: after each step get_cfl, then dt = min(cflmax / cfl, dtmax).  The dt sequence,
    the CFL values and the fields equal the reference loop's bit for bit; the first parameter set
    makes the CFL limit bind, the second the dtmax cap."""
    g = O.Grid(16, 12, 8, lx=4.0, ly=2.2, lz=2.0, re=2000.0, forcing=(0.0, 0.0, 0.0), periodic=(1, 0, 1))
    f, coef = _wavy_state(g, seed=3)
    f2 = {k: v.copy() for k, v in f.items()}
    dts, cfls, _ = O.run_steps(g, f, coef, 4, 1e-3, cflmax=cflmax, dtmax=dtmax, solver=1, niter=3, omega=1.8)
    rdts, rcfls = O.ref_run_redblack(g, f2, coef, 3, 1.8, 4, 1e-3, cflmax=cflmax, dtmax=dtmax)
    assert dts == rdts and cfls == rcfls
    assert all(c > 0 for c in cfls)
    _same(f, f2, f.keys())
    if cflmax == 0.05:
        assert dts[1] == cflmax / cfls[0] < dtmax   # the CFL limit sets dt
    else:
        assert dts[1] == dtmax                      # the cap sets dt
