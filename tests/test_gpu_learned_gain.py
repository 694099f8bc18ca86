"""What the trained network gains on the headline workload: the 192x128x96 wavy channel (A = 0.1,
seed 0, Re_b = 2000, constant flow rate, dt = 5e-4), FP32 red-black sweeps in FP64 defect
correction to ||div||_inf < 1e-8.

The network is the stride-4 U-Net trained by mobydiff_amd.train_guess on 20,100 pairs of 50 other
wavy channels (mobydiff_amd/weights/).  From one state (six steps after the seeded start) the same
steps run twice: from the plain start (the previous pressure, Phi = 0) and with the learned
guess, whose start the solver takes whenever its residual beats the plain start's (the time gate
switched off with MOBYDIFF_NET_TIME_GATE=0, so the network is evaluated for every solve).  At matched
steps and stages, the solves the network started need fewer sweeps, and the solve time per stage
(network evaluation included) and the wall time per step are less than the plain start's.
"""
import os
import time

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
                            flowrate=True, flow_height=2.0, timing=True)


def _run(state, coef, steps, net=None):
    ch = M.Channel(_cfg())
    ch.set_ibm(*coef)
    ch.upload(**state)
    if net is not None:
        ch.attach_net(net, capture=False)
        ch.set_guess(M.GUESS_LEARNED)
    its, started = [], []
    ch.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = ch.step(DT)
        its.append([st.stage[q].iterations for q in range(3)])
        started.append([st.stage[q].fallback == 0 for q in range(3)])
    ch.synchronize()
    el = time.perf_counter() - t0
    t = ch.phase_times()
    div = ch.div_norm()
    if net is not None:
        ch.attach_net(None)
    ch.close()
    return its, started, el, t, div


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="trained U-Net parameters not present")
def test_learned_guess_saves_sweeps_and_time(monkeypatch):
    from mobydiff_amd import train_guess
    print(f"learned-guess gain: wavy channel {NX}x{NY}x{NZ}, weights {os.path.basename(WEIGHTS)}", flush=True)
    head = M.Channel(_cfg())
    geo = M.wavy_geometry(A=0.1)
    coef = head.compute_ibm(geo)
    head.init_fields(seed=0, amp=0.05, geo=geo, dt=DT)
    for _ in range(6):
        head.step(DT)
    state = head.download()
    head.close()
    net = M.Net(NX, NY, NZ, seed=0, stride=train_guess.stride_of(WEIGHTS))
    train_guess.load(WEIGHTS, net)
    steps = 12
    its_p, _, el_p, t_p, div_p = _run(state, coef, steps)
    monkeypatch.setenv("MOBYDIFF_NET_MARGIN", "1.0")
    monkeypatch.setenv("MOBYDIFF_NET_TIME_GATE", "0")
    its_l, started, el_l, t_l, div_l = _run(state, coef, steps, net=net)
    net.close()
    pairs = [(its_l[k][q], its_p[k][q]) for k in range(steps) for q in range(3) if started[k][q]]
    print(f"plain start: {its_p}, {el_p * 1e3 / steps:.3f} ms/step", flush=True)
    print(f"learned:     {its_l}, {el_l * 1e3 / steps:.3f} ms/step, network started {len(pairs)} of {3 * steps}",
          flush=True)
    assert div_p < 1e-8 and div_l < 1e-8
    assert len(pairs) >= steps, pairs
    a, b = sum(p[0] for p in pairs), sum(p[1] for p in pairs)
    print(f"sweeps where the network started: {a / len(pairs):.1f} against {b / len(pairs):.1f} per solve "
          f"({1 - a / b:.1%} fewer)", flush=True)
    assert a < b
    # solve time per stage from the sampled stage events; the Poisson phase holds the network's
    # evaluation (guess_ms, over every evaluation)
    solve_p = t_p.poisson_ms / t_p.stages
    solve_l = t_l.poisson_ms / t_l.stages
    print(f"solve time per stage: learned {solve_l:.3f} ms (network {t_l.guess_ms / (3 * steps):.3f}) against "
          f"plain {solve_p:.3f} ms", flush=True)
    assert solve_l < solve_p
    assert el_l < el_p
