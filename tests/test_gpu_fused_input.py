"""The learned start's fused input pass against the separate kernels it replaces, and the per-step
bulk-velocity report switch, on the headline workload (192x128x96 wavy channel, A = 0.1, seed 0,
Re_b = 2000, constant flow rate, dt = 5e-4, FP32 sweeps in FP64 defect correction to 1e-8).

The fused pass (learned_input_kernel + input_from_sums_kernel) forms the right-hand side, the
residual of Phi_0 and the network's voxel sums in one sweep; the separate path (sor_prepare,
phi0_residual, restrict_input; MOBYDIFF_FUSED_INPUT=0) forms the same quantities with other
summation orders.  From one state, the same steps run both ways with the network evaluated for
every solve (after one step that fills the per-stage history): the residual of the network's start (what the network's input determines) agrees per
solve to 1e-3 relative, the solves take the same sweeps to within one batch, and the fields
agree to the solver tolerance.
"""
import os

import numpy as np
import pytest

import mobydiff_amd as M

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "mobydiff_amd", "weights", "unet_guess_s4_wavy_192x128x96.npz")
NX, NY, NZ, DT = 192, 128, 96, 5e-4


def _cfg():
    return M.channel_config(NX, NY, NZ, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8,
                            max_iter=400000, check_every=20, precision=M.PREC_MIXED, inner_iters=20,
                            flowrate=True, flow_height=2.0)


@pytest.fixture(scope="module")
def case():
    head = M.Channel(_cfg())
    geo = M.wavy_geometry(A=0.1)
    coef = head.compute_ibm(geo)
    head.init_fields(seed=0, amp=0.05, geo=geo, dt=DT)
    for _ in range(6):
        head.step(DT)
    state = head.download()
    head.close()
    return state, coef


def _run(state, coef, steps, net=None, report=True):
    ch = M.Channel(_cfg())
    ch.set_ibm(*coef)
    ch.upload(**state)
    if net is not None:
        ch.attach_net(net, capture=False)
        ch.set_guess(M.GUESS_LEARNED)
        ch.step(DT)   # the per-stage history the network's input needs
    ch.set_step_report(report)
    stats = [ch.step(DT) for _ in range(steps)]
    out = ch.download()
    if net is not None:
        ch.attach_net(None)
    ch.close()
    return out, stats


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="trained U-Net parameters not present")
def test_fused_input_pass_matches_separate_kernels(case, monkeypatch):
    from mobydiff_amd import train_guess
    state, coef = case
    net = M.Net(NX, NY, NZ, seed=0, stride=train_guess.stride_of(WEIGHTS))
    train_guess.load(WEIGHTS, net)
    monkeypatch.setenv("MOBYDIFF_NET_MARGIN", "1e30")   # keep every network start
    monkeypatch.setenv("MOBYDIFF_NET_TIME_GATE", "0")
    steps = 3
    monkeypatch.setenv("MOBYDIFF_FUSED_INPUT", "0")
    out_s, st_s = _run(state, coef, steps, net)
    monkeypatch.setenv("MOBYDIFF_FUSED_INPUT", "1")
    out_f, st_f = _run(state, coef, steps, net)
    net.close()
    sl = (slice(1, -1),) * 3
    for k in range(steps):
        for q in range(3):
            a, b = st_f[k].stage[q], st_s[k].stage[q]
            print(f"step {k} stage {q}: start residual fused {a.guess_div_inf:.6e} separate {b.guess_div_inf:.6e}, "
                  f"sweeps {a.iterations} / {b.iterations}", flush=True)
            assert a.fallback == 0 and b.fallback == 0
            assert a.guess_div_inf > 0.0
            assert abs(a.guess_div_inf - b.guess_div_inf) <= 1e-3 * b.guess_div_inf
            assert abs(a.iterations - b.iterations) <= 20
            assert a.div_inf < 1e-8 and b.div_inf < 1e-8
    for name in ("un", "vn", "wn"):
        e = _rel(out_f[name][sl], out_s[name][sl])
        print(f"{name}: rel-L2 fused vs separate {e:.3e}", flush=True)
        assert e < 1e-8
    p, ps = out_f["pn"][sl], out_s["pn"][sl]
    e = _rel(p - p.mean(), ps - ps.mean())
    print(f"p (mean removed): rel-L2 fused vs separate {e:.3e}", flush=True)
    assert e < 1e-4


def test_step_report_switch(case):
    """Without the bulk-velocity report the step computes the same fields; ub is NaN, and with the
    report it is the bulk velocity (the target, to the projection's change)."""
    state, coef = case
    out_on, st_on = _run(state, coef, 2, report=True)
    out_off, st_off = _run(state, coef, 2, report=False)
    for k in range(2):
        assert abs(st_on[k].ub - 1.0) < 1e-6
        assert np.isnan(st_off[k].ub)
    for name in ("un", "vn", "wn", "pn"):
        assert np.array_equal(out_on[name], out_off[name]), name
