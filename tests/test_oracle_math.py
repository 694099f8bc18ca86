"""Pin the oracle functions the reference build cannot run, by mathematics.

  * FFT r2c/c2r (FFTW in the reference, absent here): numpy.fft on the same layout.
  * FFT-path RHS / correction (pressure_solver.f90, needs FFTW to build): closed-form divergence
    of analytic fields and ||div||_inf after projection.
  * Momentum operator: second-order convergence to the analytic right-hand side, and the exact
    discrete Poiseuille steady state.
  * Pressure-correction SOR: equal to the reference SOLA iteration at fixed iteration counts.
  * New features (RNG, stretched grid, flow rate, wavy geometry): known-answer vectors, closed
    forms, invariants.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


# --------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(8, 3, 6), (16, 2, 8), (12, 2, 10), (6, 3, 5)])
def test_fft_r2c_matches_numpy(shape):
    nx, ny, nz = shape
    g = O.Grid(nx, ny, nz)
    rng = np.random.default_rng(0)
    rhs = rng.standard_normal((ny, nz, nx))
    hat = O.fft_r2c(g, rhs)
    ref = np.fft.rfft2(rhs, axes=(1, 2))
    assert np.max(np.abs(hat - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("shape", [(8, 3, 6), (16, 2, 8), (12, 2, 10)])
def test_fft_c2r_matches_numpy(shape):
    nx, ny, nz = shape
    g = O.Grid(nx, ny, nz)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((ny, nz, nx))
    hat = np.fft.rfft2(x, axes=(1, 2))
    back = O.fft_c2r(g, hat)
    # unnormalised, like FFTW's c2r
    assert np.max(np.abs(back - x * nx * nz)) < 1e-11 * nx * nz


# --------------------------------------------------------------------------- RHS / projection
def _analytic_div_fields(g):
    """u = sin(2 pi x/lx) cos(z), v = y^2 (x-independent), w = cos(3 z) + x: div known in closed form."""
    f = O.new_fields(g)
    nx, ny, nz = g.nx, g.ny, g.nz
    i = np.arange(nx + 2)[None, None, :]
    j = np.arange(ny + 2)[None, :, None]
    k = np.arange(nz + 2)[:, None, None]
    xf = (i - 1) * g.dx
    xc = (i - 0.5) * g.dx
    zf = (k - 1) * g.dz
    zc = (k - 0.5) * g.dz
    yv = g.yv[j]
    f["us"][...] = np.sin(2 * np.pi * xf / g.lx) * np.cos(zc) + 0 * yv
    f["vs"][...] = yv ** 2 + 0 * xc + 0 * zc
    f["ws"][...] = np.cos(3 * zf) + 0 * xc + 0 * yv
    return f


def test_build_fft_rhs_closed_form():
    g = O.Grid(16, 8, 12, lx=2.0, ly=1.0, lz=1.5)
    f = _analytic_div_fields(g)
    dtg = 0.25
    rhs = O.build_fft_rhs(g, f, dtg)  # (ny, nz, nx)
    i = np.arange(1, g.nx + 1)[None, None, :]
    j = np.arange(1, g.ny + 1)[:, None, None]
    k = np.arange(1, g.nz + 1)[None, :, None]
    a = 2 * np.pi / g.lx
    # exact discrete differences of the sampled analytic fields
    du = (np.sin(a * i * g.dx) - np.sin(a * (i - 1) * g.dx)) / g.dx * np.cos((k - 0.5) * g.dz)
    dv = (g.yv[j + 1] ** 2 - g.yv[j] ** 2) / g.dy
    dw = (np.cos(3 * k * g.dz) - np.cos(3 * (k - 1) * g.dz)) / g.dz
    assert np.max(np.abs(rhs - (du + dv + dw) / dtg)) < 1e-12


@pytest.mark.parametrize("beta", [None, 1.5])
def test_fft_projection_is_divergence_free(beta):
    g = O.Grid(16, 12, 8, lx=4 * np.pi / 4, ly=2.0, lz=np.pi / 2, beta=beta)
    rng = np.random.default_rng(4)
    f = O.random_fields(g, rng)
    f["pn"][...] = 0.0
    O.apply_bc_fft(g, f)
    O.projection_fft(g, f, None, 0.3)
    O.apply_bc_fft(g, f)
    assert O.div_inf(g, f) < 1e-11
    # pressure increment is the potential of the removed gradient: p = phi (p started at 0)
    assert np.isfinite(f["pn"]).all()


def test_fft_projection_idempotent():
    g = O.Grid(16, 12, 8, ly=2.0)
    rng = np.random.default_rng(5)
    f = O.random_fields(g, rng)
    O.apply_bc_fft(g, f)
    O.projection_fft(g, f, None, 0.3)
    O.apply_bc_fft(g, f)
    u1 = {k: f[k].copy() for k in ("un", "vn", "wn")}
    for k in ("un", "vn", "wn"):
        f[k.replace("n", "s")][...] = f[k]
    p0 = f["pn"].copy()
    O.projection_fft(g, f, None, 0.3)
    for k in ("un", "vn", "wn"):
        assert np.max(np.abs(f[k][1:-1, 1:-1, 1:-1] - u1[k][1:-1, 1:-1, 1:-1])) < 1e-11
    assert np.max(np.abs(f["pn"] - p0)) < 1e-9


# --------------------------------------------------------------------------- momentum
def _periodic_rhs_error(n):
    """Max error of the discrete momentum RHS against the analytic one, triply periodic box."""
    L = 2 * np.pi
    g = O.Grid(n, n, n, lx=L, ly=L, lz=L, re=10.0, forcing=(0.3, -0.2, 0.1), periodic=(1, 1, 1))
    f = O.new_fields(g)
    h = L / n
    i = np.arange(n + 2)[None, None, :]
    j = np.arange(n + 2)[None, :, None]
    k = np.arange(n + 2)[:, None, None]
    xf, xc = (i - 1) * h, (i - 0.5) * h
    yf, yc = (j - 1) * h, (j - 0.5) * h
    zf, zc = (k - 1) * h, (k - 0.5) * h

    def U(x, y, z): return np.sin(x) * np.cos(y) * np.cos(z) + 0.5
    def V(x, y, z): return np.cos(x) * np.sin(y) * np.cos(2 * z)
    def W(x, y, z): return np.cos(x) * np.cos(y) * np.sin(z) * 0.7
    f["un"][...] = U(xf, yc, zc)
    f["vn"][...] = V(xc, yf, zc)
    f["wn"][...] = W(xc, yc, zf)
    O.momentum(g, f, None, 1.0, 0.0, 0.0)
    rhs_u = (f["us"] - f["un"])[1:-1, 1:-1, 1:-1]

    # analytic: -d(uu)/dx - d(uv)/dy - d(uw)/dz + fx + nu lap u at u points
    x, y, z = xf[..., 1:-1], yc[:, 1:-1, :], zc[1:-1]
    e = 1e-6

    def d(fun, ax):
        dd = [0, 0, 0]
        dd[ax] = e
        return (fun(x + dd[0], y + dd[1], z + dd[2]) - fun(x - dd[0], y - dd[1], z - dd[2])) / (2 * e)
    conv = d(lambda a, b, c: U(a, b, c) ** 2, 0) + d(lambda a, b, c: U(a, b, c) * V(a, b, c), 1) \
        + d(lambda a, b, c: U(a, b, c) * W(a, b, c), 2)
    lap = -3 * (U(x, y, z) - 0.5)
    exact = -conv + 0.3 + lap / 10.0
    return np.max(np.abs(rhs_u - exact))


def test_momentum_second_order_convergence():
    e1 = _periodic_rhs_error(16)
    e2 = _periodic_rhs_error(32)
    assert e1 < 0.2
    assert 3.3 < e1 / e2 < 4.7, (e1, e2)


def test_momentum_discrete_poiseuille_is_steady():
    """u_j = (f Re/2)(1 - y_j^2) + f Re dy^2/8 is an exact steady state of the discrete operator
    (3-point Laplacian exact on quadratics; the ghost reflection puts the wall mean at zero)."""
    ny = 16
    g = O.Grid(8, ny, 4, lx=1.0, ly=2.0, lz=0.5, re=50.0, forcing=(0.04, 0.0, 0.0), periodic=(1, 0, 1))
    f = O.new_fields(g)
    yp = g.yc - 1.0
    prof = 0.04 * 50.0 / 2 * (1 - yp ** 2) + 0.04 * 50.0 * g.dy ** 2 / 8
    f["un"][...] = prof[None, :, None]
    O.apply_bc_fft(g, f)
    assert np.allclose(f["un"][:, 0, :], -f["un"][:, 1, :])
    O.momentum(g, f, None, 1e-2, 0.0, 1e-2)
    assert np.max(np.abs(f["us"] - f["un"])[1:-1, 1:-1, 1:-1]) < 1e-13
    assert np.max(np.abs(f["oru"])) < 1e-11


@pytest.mark.parametrize("beta", [1e-6])
def test_stretched_momentum_reduces_to_uniform(beta):
    rng = np.random.default_rng(2)
    gu = O.Grid(8, 10, 6, ly=2.0, periodic=(1, 0, 1))
    gs = O.Grid(8, 10, 6, ly=2.0, periodic=(1, 0, 1), beta=beta)
    assert np.max(np.abs(gs.ihc[1:-1] * gu.dy - 1)) < 1e-9
    f = O.random_fields(gu, rng)
    f2 = {k: v.copy() for k, v in f.items()}
    O.momentum(gu, f, None, 0.1, -0.02, 0.05)
    O.momentum(gs, f2, None, 0.1, -0.02, 0.05)
    assert np.max(np.abs(f["us"] - f2["us"])) < 1e-6 * np.max(np.abs(f["us"]))


def test_stretched_metrics_closed_form():
    g = O.Grid(4, 32, 4, ly=2.0, beta=1.5)
    j = np.arange(1, 34)
    eta = (j - 1) / 32
    yv = 1.0 * (1 + np.tanh(1.5 * (2 * eta - 1)) / np.tanh(1.5))
    assert np.max(np.abs(g.yv[1:33] - yv[:32])) < 1e-14
    h = 1 / g.ihc[1:-1]
    assert abs(h.sum() - 2.0) < 1e-13
    assert h[0] < h[16] / 3  # clustered at the walls


# --------------------------------------------------------------------------- SOR forms
@pytest.mark.parametrize("per_y", [0, 1])
@pytest.mark.parametrize("niter", [1, 4])
def test_pressure_correction_sor_equals_sola(per_y, niter):
    """The pressure-correction SOR and the reference SOLA sweep are the same iteration."""
    rng = np.random.default_rng(12)
    g = O.Grid(8, 6, 4, lx=1.0, ly=1.5, lz=0.5, periodic=(1, per_y, 1))
    f = O.random_fields(g, rng)
    coef = tuple(np.abs(rng.standard_normal(g.gshape)) * 20 for _ in range(3))
    f2 = {k: v.copy() for k, v in f.items()}
    O.sola_redblack(g, f, coef, niter, 1.8, 0.03)
    st, phi, rhs = O.projection_sor(g, f2, coef, 1.8, 0.0, niter, niter, 0.03)
    assert st.iterations == niter
    for k in ("un", "vn", "wn"):
        a, b = f[k][1:-1, 1:-1, 1:-1], f2[k][1:-1, 1:-1, 1:-1]
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(a)), k
    assert np.max(np.abs(f["pn"] - f2["pn"])[1:-1, 1:-1, 1:-1]) < 1e-10 * np.max(np.abs(f["pn"]))


def test_sor_to_tolerance_reaches_tolerance():
    rng = np.random.default_rng(13)
    g = O.Grid(16, 12, 8, lx=1.0, ly=1.0, lz=0.5, periodic=(1, 0, 1))
    f = O.random_fields(g, rng, amp=0.1)
    st, phi, rhs = O.projection_sor(g, f, None, 1.8, 1e-8, 20000, 10, 0.03)
    assert st.converged == 1
    assert st.div_inf < 1e-8
    assert O.div_inf(g, f) < 1e-8
    assert st.div0_inf > 1.0


# --------------------------------------------------------------------------- new features
def test_splitmix64_known_answers():
    # SplitMix64 reference sequence from state 0: outputs are mix(k * golden) for k = 1, 2, 3
    lib = O.lib()
    assert lib.mdo_splitmix64(0) == 0xE220A8397B1DCDAF
    assert lib.mdo_splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4
    assert lib.mdo_splitmix64((2 * 0x9E3779B97F4A7C15) % 2 ** 64) == 0x06C45D188009454F


def test_uniform01_range_and_mean():
    vals = np.array([O.uniform01(3, c, q) for c in range(3) for q in range(4000)])
    assert vals.min() >= 0 and vals.max() < 1
    assert abs(vals.mean() - 0.5) < 0.01
    assert O.uniform01(0, 0, 0) == (0xE220A8397B1DCDAF >> 11) / 2.0 ** 53


def test_initial_field_poiseuille_bulk():
    g = O.Grid(8, 32, 4, lx=1.0, ly=2.0, lz=0.5)
    f = O.new_fields(g)
    O.initial_field(g, f, seed=0, amp=0.0)
    # midpoint rule of 1.5(1-y^2) on [-1,1]: M = I - (h^2/24)(f'(1)-f'(-1)), so U_b = 1 + dy^2/8
    assert abs(O.bulk_u(g, f["us"], 2.0) - (1 + g.dy ** 2 / 8)) < 1e-14
    O.initial_field(g, f, seed=7, amp=0.1)
    pert = f["us"][1:-1, 1:-1, 1:-1] - (1.5 * (1 - (g.yc[1:-1] - 1) ** 2))[None, :, None]
    assert 0.09 < np.abs(pert).max() <= 0.1
    assert np.all(f["vs"][:, 1, :] == 0)


def test_flowrate_shift_sets_bulk_exactly():
    g = O.Grid(8, 16, 4, ly=2.0, beta=1.5)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(g.gshape)
    s = g.struct()
    import ctypes as C
    du = O.lib().mdo_flowrate_shift(C.byref(s), O._ptr(u), None, C.c_double(0.1), C.c_double(1.0), C.c_double(2.0))
    assert abs(O.bulk_u(g, u, 2.0) - 1.0) < 1e-14
    assert du != 0


@pytest.mark.parametrize("shape", [(8, 16, 4), (70, 13, 11), (192, 128, 96), (256, 128, 256)])
def test_step_order_bulk_sums_equal_exact_sum(shape):
    """The time step's order of the flow-rate sums (64-column tiles over 8 or 12 planes, 4 or 8 rows
    per tile; 192x128x96 takes 12 planes and 4 rows, 256x128x256 8 planes and 8 rows, 70x13x11
    ragged tiles in x, y and z) against the exactly rounded sum (math.fsum), and a uniform field
    against its closed form."""
    nx, ny, nz = shape
    g = O.Grid(nx, ny, nz, ly=2.0, beta=1.5 if nx < 100 else 0.0)
    rng = np.random.default_rng(nx + ny + nz)
    u = rng.standard_normal(g.gshape) + 1.0
    h = 1.0 / g.ihc[1:-1] if g.beta else np.full(ny, g.dy)
    terms = (u[1:-1, 1:-1, 1:-1] * h[None, :, None]).ravel()
    exact = math.fsum(terms.tolist()) / (nx * nz * 2.0)
    got = O.bulk_u_step(g, u, 2.0)
    assert abs(got - exact) <= 1e-13 * max(abs(exact), 1.0), (got, exact)
    assert abs(got - O.bulk_u(g, u, 2.0)) <= 1e-13 * max(abs(exact), 1.0)
    # u = 1: the sum of the cell heights over each column is ly
    one = np.ones(g.gshape)
    assert abs(O.bulk_u_step(g, one, 2.0) - 1.0) <= 1e-13


def test_wavy_geometry_bisection_matches_closed_form():
    """For y-neighbours the interface distance has a closed form: |y_wall(x) - y|."""
    g = O.Grid(32, 40, 2, lx=2.0, ly=2.2, lz=0.1, re=1.0)
    geo = O.Geometry(kind=1, second_order=1, A=0.1, lam=1.0, half_height=1.0)
    cv = O.set_ibm_coeff(g, geo, 0, 1, 0)
    checked = 0
    for i in range(0, g.nx + 2):
        x = ((i + 1) - 1.5) * g.dx
        h = 1.0 - 0.1 * math.cos(2 * math.pi * x / 1.0)
        for j in range(1, g.ny + 1):
            y = ((j + 1) - 0.5 - 1.5) * g.dy - 1.1
            if abs(y) > h:
                assert cv[1, j, i] == 1e30
                continue
            acc = 0.0
            for dj in (-1, 1):
                yn = y + dj * g.dy
                if abs(yn) > h:
                    d = h - abs(y)
                    acc += ((g.dy - d) / d) / g.dy ** 2
            xn = [x - g.dx, x + g.dx]
            if any(abs(y) > 1.0 - 0.1 * math.cos(2 * math.pi * q) for q in xn):
                continue  # x-neighbours need the bisection; not closed form
            if acc:
                assert abs(cv[1, j, i] - acc) < 1e-6 * acc
                checked += 1
            else:
                assert cv[1, j, i] == 0
    assert checked > 20


# --------------------------------------------------------------------------- RK3 and colouring
def test_rk3_constants_closed_form():
    """The low-storage RK3 of step.f90:21-23: gamma = alpha + beta per stage, sum(gamma) = 1 (the
    stages advance one full dt), and the exact rationals 8/15, 5/12, 3/4 / 0, -17/60, -5/12."""
    from fractions import Fraction as F
    a, b, c = O.rk_constants()
    want_a, want_b = (F(8, 15), F(5, 12), F(3, 4)), (F(0), F(-17, 60), F(-5, 12))
    for s in range(3):
        assert a[s] == float(want_a[s]) and b[s] == float(want_b[s])
        assert F(a[s]) + F(b[s]) - F(c[s]) == 0 or abs(a[s] + b[s] - c[s]) <= 2 ** -52
        assert c[s] == float(want_a[s] + want_b[s])
    assert abs(sum(c) - 1.0) <= 2 ** -52


def test_rk3_third_order_convergence():
    """The stage update of momentum (u <- u + dt alpha_s f(u) + dt beta_s f_prev, step.f90) with the
    oracle's constants integrates y' = -y^2 (y(0) = 1, exact 1/(1+t)) to third order: halving dt
    divides the error by 8 (a wrong constant drops the order to 1 or 2)."""
    a, b, _ = O.rk_constants()

    def integrate(n):
        dt, y = 1.0 / n, 1.0
        for _ in range(n):
            fprev = 0.0
            for s in range(3):
                fy = -y * y
                y = y + dt * a[s] * fy + dt * b[s] * fprev
                fprev = fy
        return abs(y - 0.5)

    e = [integrate(n) for n in (16, 32, 64)]
    for e0, e1 in zip(e, e[1:]):
        assert 7.0 < e0 / e1 < 9.0, e


def test_colouring_order_independence():
    """Red-black colouring: the cells of one colour (i+j+k parity) do not read each other, so
    relaxing them one by one in any order gives sor_sweep's result bit for bit.  Controls: both
    colours together, or a wrong parity (i+j), depend on the order."""
    rng = np.random.default_rng(17)
    g = O.Grid(8, 6, 4, lx=1.0, ly=1.5, lz=0.5, periodic=(1, 0, 1))
    coef = tuple(np.abs(rng.standard_normal(g.gshape)) * 20 for _ in range(3))
    rhs = rng.standard_normal(g.ishape)
    phi0 = rng.standard_normal(g.ishape)
    cells = np.array([(i, j, k) for k in range(1, 5) for j in range(1, 7) for i in range(1, 9)])
    par = cells.sum(axis=1) % 2

    def relax(sel, order_seed):
        phi = phi0.copy()
        c = cells[sel]
        c = c[np.random.default_rng(order_seed).permutation(len(c))]
        O.sor_relax_cells(g, coef, 1.8, 0.01, rhs, phi, c)
        return phi

    for color in (0, 1):
        swept = phi0.copy()
        O.sor_sweep(g, coef, color, 1.8, 0.01, rhs, swept)
        sel = par == (color + 1) % 2   # colour c: the cells with i + j + k + c odd
        for seed in range(4):
            assert np.array_equal(relax(sel, seed), swept)
        assert not np.array_equal(swept, phi0)
    everything = np.ones(len(cells), bool)
    assert not np.array_equal(relax(everything, 0), relax(everything, 1))
    wrong = (cells[:, 0] + cells[:, 1]) % 2 == 0
    assert not np.array_equal(relax(wrong, 0), relax(wrong, 1))
