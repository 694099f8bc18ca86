"""GPU runs of whole workloads against the oracle and against size-independent invariants.

* Flat channel, 64^3, tanh-stretched y (beta = 1.5), Lx x Ly x Lz = 4 pi x 2 x 2 pi, Re_b = 2800,
  constant flow rate, Poiseuille + 0.1 seeded perturbation, dt = 1e-3, 100 RK3 steps with the
  FFT + tridiagonal Poisson solve: u, v, w, p against the FP64 oracle to rel-L2 <= 1e-10, discrete
  divergence <= 1e-12, bulk velocity held at 1.
* Production flat channel 256 x 192 x 256 (Re_b = 5600, dt = 5e-4): a few steps on the GPU with
  the same invariants (divergence, bulk velocity) and two steps against the oracle.
* A batch of 8 wavy channels (A uniform in [0.05, 0.15], per-instance seeds) stepped
  concurrently: every instance equals its own single-channel run bit for bit, and the mixed-
  precision batch agrees with the FP64 oracle field by field.
"""
import math

import numpy as np
import pytest

import mobydiff_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _flat(nx, ny, nz, re):
    return M.channel_config(nx, ny, nz, 4 * math.pi, 2.0, 2 * math.pi, re, solver=M.SOLVER_FFT, flowrate=True,
                            ub_target=1.0, flow_height=2.0, stretch_beta=1.5, p_neumann_y=False)


def _oracle_grid(nx, ny, nz, re):
    return O.Grid(nx, ny, nz, lx=4 * math.pi, ly=2.0, lz=2 * math.pi, re=re, forcing=(0.0, 0.0, 0.0),
                  periodic=(1, 0, 1), beta=1.5, p_neumann_y=False)


def test_flat_channel_64_cubed_100_steps_matches_oracle():
    nx = ny = nz = 64
    dt = 1e-3
    print("flat 64^3, 100 steps", flush=True)
    ch = M.Channel(_flat(nx, ny, nz, 2800.0))
    ch.init_fields(seed=0, amp=0.1, geo=None, dt=dt)
    f = ch.download()
    assert ch.div_norm() <= 1e-12
    for _ in range(100):
        ch.step(dt)
    out = ch.download()
    div_gpu = ch.div_norm()
    ub = ch.bulk_velocity()
    ch.close()

    g = _oracle_grid(nx, ny, nz, 2800.0)
    for _ in range(100):
        O.rk3_step(g, f, None, dt, solver=0, flowrate=True, ub_target=1.0, height=2.0)
    for k in ("un", "vn", "wn", "pn"):
        sl = (slice(1, -1),) * 3
        assert _rel(out[k][sl], f[k][sl]) <= 1e-10, (k, _rel(out[k][sl], f[k][sl]))
    assert div_gpu <= 1e-12, div_gpu
    assert O.div_inf(g, f) <= 1e-12
    assert abs(ub - 1.0) < 1e-12, ub


@pytest.mark.parametrize("tol", [1e-12, 1e-8])
def test_flat_channel_mixed_precision_fft(tol):
    """FP32 transforms and Thomas solves in FP64 defect correction (FP64 mean mode and residual)
    against the FP64 FFT solve on the 64^3 stretched channel: the divergence meets tol and the
    fields agree with the FP64 path (tol 1e-12: to 1e-10 relative; tol 1e-8: to 1e-6)."""
    nx = ny = nz = 64
    dt = 1e-3
    outs, stats = [], []
    for prec in (M.PREC_FP64, M.PREC_MIXED):
        cfg = _flat(nx, ny, nz, 2800.0)
        cfg.precision, cfg.tol = prec, tol
        ch = M.Channel(cfg)
        ch.init_fields(seed=0, amp=0.1, geo=None, dt=dt)
        for _ in range(5):
            st = ch.step(dt)
        stats.append(st)
        outs.append(ch.download(("un", "vn", "wn", "pn")))
        if prec == M.PREC_MIXED:
            assert ch.div_norm() <= 1.5 * tol
        ch.close()
    for q in range(3):
        s = stats[1].stage[q]
        assert s.converged == 1 and s.div_inf < tol and 1 <= s.iterations <= 4, (s.iterations, s.div_inf)
    sl = (slice(1, -1),) * 3
    bar = 1e-10 if tol < 1e-10 else 1e-6
    for k in ("un", "vn", "wn", "pn"):
        r = _rel(outs[1][k][sl], outs[0][k][sl])
        assert r <= bar, (k, r)


def test_flat_channel_production_size():
    nx, ny, nz = 256, 192, 256
    dt = 5e-4
    print("flat 256x192x256", flush=True)
    ch = M.Channel(_flat(nx, ny, nz, 5600.0))
    ch.init_fields(seed=0, amp=0.1, geo=None, dt=dt)
    f = ch.download(("un", "vn", "wn", "pn", "us", "vs", "ws", "oru", "orv", "orw"))
    for _ in range(2):
        ch.step(dt)
    out = ch.download(("un", "vn", "wn", "pn"))
    for _ in range(8):
        ch.step(dt)
    assert ch.div_norm() <= 1e-12
    assert abs(ch.bulk_velocity() - 1.0) < 1e-12
    ch.close()
    g = _oracle_grid(nx, ny, nz, 5600.0)
    for _ in range(2):
        O.rk3_step(g, f, None, dt, solver=0, flowrate=True, ub_target=1.0, height=2.0)
    sl = (slice(1, -1),) * 3
    for k in ("un", "vn", "wn", "pn"):
        assert _rel(out[k][sl], f[k][sl]) <= 1e-10, (k, _rel(out[k][sl], f[k][sl]))


def _wavy_channel(nx, ny, nz, A, seed, dt, precision):
    cfg = M.channel_config(nx, ny, nz, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8,
                           check_every=10, precision=precision, inner_iters=10, flowrate=True, flow_height=2.0)
    ch = M.Channel(cfg)
    geo = M.wavy_geometry(A=A)
    coef = ch.compute_ibm(geo)
    ch.init_fields(seed=seed, amp=0.05, geo=geo, dt=dt)
    return ch, coef


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED])
def test_batch_of_8_wavy_channels_equals_single_runs(precision):
    nx, ny, nz, dt, steps = 48, 32, 24, 5e-4, 2
    amps = np.random.default_rng(2024).uniform(0.05, 0.15, 8)
    print("batch of 8 wavy channels", list(np.round(amps, 4)), flush=True)
    batch = [_wavy_channel(nx, ny, nz, float(a), q, dt, precision)[0] for q, a in enumerate(amps)]
    starts = [ch.download() for ch in batch]
    for _ in range(steps):
        stats = M.step_batch(batch, dt)
        for st in stats:
            assert all(st.stage[q].converged == 1 for q in range(3))
    results = [ch.download(("un", "vn", "wn", "pn")) for ch in batch]
    for ch in batch:
        ch.close()
    for q, a in enumerate(amps):
        single, coef = _wavy_channel(nx, ny, nz, float(a), q, dt, precision)
        for _ in range(steps):
            single.step(dt)
        ref = single.download(("un", "vn", "wn", "pn"))
        single.close()
        for k in ("un", "vn", "wn", "pn"):
            assert np.array_equal(results[q][k], ref[k]), (q, k)
        # against the FP64 oracle from the same start
        g = O.Grid(nx, ny, nz, lx=4.0, ly=2.2, lz=2.0, re=2000.0, forcing=(0.0, 0.0, 0.0), periodic=(1, 0, 1))
        f = {k: v.copy() for k, v in starts[q].items()}
        tight = {k: v.copy() for k, v in starts[q].items()}
        for _ in range(steps):
            O.rk3_step(g, f, coef, dt, solver=2, omega=1.8, tol=1e-8, max_iter=1000000, check_every=10,
                       flowrate=True, ub_target=1.0, height=2.0)
            O.rk3_step(g, tight, coef, dt, solver=2, omega=1.8, tol=1e-13, max_iter=1000000, check_every=10,
                       flowrate=True, ub_target=1.0, height=2.0)
        sl = (slice(1, -1),) * 3
        # FP64: round-off only (the bulk-velocity sums are ordered differently on the CPU)
        bar = 1e-10 if precision == M.PREC_FP64 else 1e-6
        for k in ("un", "vn", "wn"):
            r = _rel(results[q][k][sl], f[k][sl])
            assert r <= bar, (q, k, r)

        def pm(a):   # p with its mean removed (fixed up to a constant by the Neumann solve)
            return a[sl] - a[sl].mean()

        e_p = _rel(pm(results[q]["pn"]), pm(f["pn"]))
        if precision == M.PREC_FP64:
            assert e_p <= 1e-10, (q, e_p)
        else:
            # two solves stopped at ||div||_inf < 1e-8 differ in p by about the distance of either
            # from the converged p: the mixed p is no further from the converged (1e-13) solve than
            # 5x the FP64 oracle's p
            e_o, e_m = _rel(pm(f["pn"]), pm(tight["pn"])), _rel(pm(results[q]["pn"]), pm(tight["pn"]))
            print(f"instance {q}: p vs oracle {e_p:.3e}; vs the converged solve: oracle {e_o:.3e}, "
                  f"mixed {e_m:.3e}", flush=True)
            assert e_p <= 1e-4 and e_m <= 5.0 * e_o, (q, e_p, e_o, e_m)


def test_batch_rejects_repeated_handle():
    ch, _ = _wavy_channel(16, 12, 8, 0.1, 0, 5e-4, M.PREC_FP64)
    with pytest.raises(M.MdxError, match="appears twice"):
        M.step_batch([ch, ch], 5e-4)
    ch.close()
