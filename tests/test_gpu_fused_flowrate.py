"""The flow-rate correction formed inside the momentum kernel against its separate sum pass.

In a step, the tiled momentum kernel sums u* h_j and mu_u h_j over each of its tiles and one launch
folds those partial sums and shifts u*; MOBYDIFF_FUSED_FLOWRATE=0 keeps the separate pass
(bulk_partial_kernel over u*, then the same fold over its 512 partials).  The two differ only in
the order of the bulk sums, so from one state the steps agree to round-off: the bulk velocity each
step reports to 1e-13, u, v, w to 1e-12 relative, p (mean removed) to 1e-9 relative.  Cases: the
192x128x96 wavy channel (IBM, 4-row tiles), a stretched wall-normal grid, and walls in z with the
fixed-sweep SOLA solve (constant flow rate needs periodic x; SOR to tolerance periodic z).

The oracle's time step adds the same sums in the same order (mdo_rk3_step), so one RK3 step with
the fixed-sweep SOLA projection (momentum, boundary fills and SOLA being bit-exact each) equals
the oracle's step bit for bit, with 4-row tiles of 12 planes (the 192x128x96 wavy channel), 8-row
tiles of 8 planes (256x128x256) and a stretched grid.
"""
import numpy as np
import pytest

import mobydiff_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DT = 5e-4

CASES = {
    "wavy_192x128x96": dict(shape=(192, 128, 96), periodic=(1, 0, 1), A=0.1, beta=0.0, solver=M.SOLVER_SOR),
    "stretched_64x48x32": dict(shape=(64, 48, 32), periodic=(1, 0, 1), A=None, beta=1.5, solver=M.SOLVER_SOR),
    "walls_z_sola_64x48x32": dict(shape=(64, 48, 32), periodic=(1, 0, 0), A=None, beta=0.0,
                                  solver=M.SOLVER_SOLA),
}


def _run(case, fused, monkeypatch, steps=2):
    monkeypatch.setenv("MOBYDIFF_FUSED_FLOWRATE", "1" if fused else "0")
    nx, ny, nz = case["shape"]
    cfg = M.channel_config(nx, ny, nz, 4.0, 2.2, 2.0, 2000.0, periodic=case["periodic"], solver=case["solver"],
                           niter=20, omega=1.8, tol=1e-8, check_every=10, precision=M.PREC_FP64, flowrate=True,
                           flow_height=2.0, stretch_beta=case["beta"])
    ch = M.Channel(cfg)
    geo = None
    if case["A"] is not None:
        geo = M.wavy_geometry(A=case["A"])
        ch.compute_ibm(geo)
    ch.init_fields(seed=3, amp=0.05, geo=geo, dt=DT)
    ch.set_step_report(True)
    ub = [ch.step(DT).ub for _ in range(steps)]
    out = ch.download(("un", "vn", "wn", "pn"))
    ch.close()
    return ub, out


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name", list(CASES))
def test_fused_flowrate_matches_separate_pass(name, monkeypatch):
    case = CASES[name]
    print(f"fused flow rate: {name}", flush=True)
    ub_f, out_f = _run(case, True, monkeypatch)
    ub_s, out_s = _run(case, False, monkeypatch)
    for a, b in zip(ub_f, ub_s):
        print(f"  ub fused {a!r} separate {b!r}", flush=True)
        assert abs(a - b) <= 1e-13
    sl = (slice(1, -1),) * 3
    for k in ("un", "vn", "wn"):
        e = _rel(out_f[k][sl], out_s[k][sl])
        print(f"  {k}: rel-L2 {e:.3e}", flush=True)
        assert e <= 1e-12, (k, e)
    p, ps = out_f["pn"][sl], out_s["pn"][sl]
    e = _rel(p - p.mean(), ps - ps.mean())
    print(f"  p (mean removed): rel-L2 {e:.3e}", flush=True)
    assert e <= 1e-9


@pytest.mark.parametrize("shape,A,beta", [((192, 128, 96), 0.1, 0.0), ((256, 128, 256), None, 0.0),
                                          ((64, 48, 32), None, 1.5)])
def test_step_with_flow_rate_bit_exact_with_oracle(shape, A, beta):
    nx, ny, nz = shape
    print(f"SOLA step with constant flow rate {shape} A={A} beta={beta}", flush=True)
    cfg = M.channel_config(nx, ny, nz, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOLA, niter=3, omega=1.8,
                           flowrate=True, flow_height=2.0, stretch_beta=beta)
    ch = M.Channel(cfg)
    geo = M.wavy_geometry(A=A) if A else None
    coef = ch.compute_ibm(geo) if geo else None
    ch.init_fields(seed=5, amp=0.05, geo=geo, dt=DT)
    f = ch.download()
    ch.step(DT)
    out = ch.download(("un", "vn", "wn", "pn"))
    ch.close()
    g = O.Grid(nx, ny, nz, lx=4.0, ly=2.2, lz=2.0, re=2000.0, forcing=(0.0, 0.0, 0.0), periodic=(1, 0, 1),
               beta=beta or None)
    O.rk3_step(g, f, coef, DT, solver=1, niter=3, omega=1.8, flowrate=True, ub_target=1.0, height=2.0)
    sl = (slice(1, -1),) * 3
    for k in ("un", "vn", "wn", "pn"):
        d = float(np.max(np.abs(out[k][sl] - f[k][sl])))
        print(f"  {k}: max |GPU - oracle| {d:.3e}", flush=True)
        assert np.array_equal(out[k][sl], f[k][sl]), k


@pytest.mark.parametrize("kind", ["wavy_plane_tables", "random_full_tables"])
def test_tables_built_in_fold_equal_separate_build(kind, monkeypatch):
    """In a step with constant flow rate the fold-and-shift launch builds the stage's face
    coefficient tables (mu_entry, the arithmetic of build_mu_kernel); MOBYDIFF_FUSED_MU=0 builds
    them in their own launch.  Bit-identical fields and iterations: on the wavy channel (z-invariant
    coefficients, one plane per table, mixed precision) and with random per-cell coefficients (whole
    tables, FP64)."""
    print(f"face tables in the fold: {kind}", flush=True)
    outs = []
    for fused in (True, False):
        monkeypatch.setenv("MOBYDIFF_FUSED_MU", "1" if fused else "0")
        if kind == "wavy_plane_tables":
            nx, ny, nz, prec = 192, 128, 96, M.PREC_MIXED
        else:
            nx, ny, nz, prec = 32, 24, 16, M.PREC_FP64
        cfg = M.channel_config(nx, ny, nz, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8,
                               check_every=10, precision=prec, inner_iters=32, flowrate=True, flow_height=2.0)
        ch = M.Channel(cfg)
        if kind == "wavy_plane_tables":
            geo = M.wavy_geometry(A=0.1)
            ch.compute_ibm(geo)
        else:
            geo = None
            rng = np.random.default_rng(11)
            ch.set_ibm(*(np.abs(rng.standard_normal(ch.gshape)) * 20 for _ in range(3)))
        ch.init_fields(seed=2, amp=0.05, geo=geo, dt=DT)
        its = [[s.iterations for s in ch.step(DT).stage] for _ in range(2)]
        outs.append((its, ch.download(("un", "vn", "wn", "pn"))))
        ch.close()
    (its_f, f), (its_s, s) = outs
    print(f"  iterations {its_f} / {its_s}", flush=True)
    assert its_f == its_s
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(f[k], s[k]), k
