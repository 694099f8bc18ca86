"""Accuracy of every solver path at the headline size against the FP64 oracle.

Workload: the 192x128x96 wavy channel (A = 0.1, Re_b = 2000, constant flow rate, dt = 5e-4,
Poiseuille + 0.05 seeded perturbation), red-black SOR (omega 1.8) to ||div||_inf < 1e-8.  From one
state (four steps after the seeded start) each path runs three RK3 steps and is compared field by
field with the oracle's three steps from the same state:

  * plain FP64 (streaming sweeps, and the LDS-resident FP64 batches): round-off only, rel-L2 <= 1e-10;
  * FP32 sweeps in FP64 defect correction, without and with the learned initial guess (a network
    trained on this run's warm-up pairs, and the shipped stride-4 network): u, v, w within rel-L2 1e-6;
  * p is compared with its mean removed (a Neumann solve fixes p up to a constant).  Two solves that
    each stop at ||div||_inf < 1e-8 differ in p by about the distance of either from the converged
    solution, so the bar for p is set by the oracle's own distance from a solve converged to 1e-12:
    the mixed and learned paths' p must be no further from that converged p than 5x the oracle's p
    (and within 1e-4 of the oracle);
  * ||div||_inf after the third step < 1e-8 on every path.
The errors are printed so the run's log states them.
"""
import os

import numpy as np
import pytest

import mobydiff_amd as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NX, NY, NZ = 192, 128, 96
WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mobydiff_amd", "weights",
                       "unet_guess_s4_wavy_192x128x96.npz")
DT, TOL, STEPS, WARM = 5e-4, 1e-8, 3, 4
SL = (slice(1, -1),) * 3


def _cfg(precision, tol=TOL):
    return M.channel_config(NX, NY, NZ, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOR, omega=1.8, tol=tol,
                            max_iter=400000, check_every=20, precision=precision, inner_iters=20,
                            flowrate=True, flow_height=2.0)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _errors(out, ref):
    e = {k: _rel(out[k][SL], ref[k][SL]) for k in ("un", "vn", "wn")}
    p, pr = out["pn"][SL], ref["pn"][SL]
    e["p"] = _rel(p - p.mean(), pr - pr.mean())
    return e


@pytest.fixture(scope="module")
def case():
    print(f"accuracy case: wavy channel {NX}x{NY}x{NZ}, {WARM} warm-up steps, {STEPS} compared steps", flush=True)
    head = M.Channel(_cfg(M.PREC_MIXED))
    geo = M.wavy_geometry(A=0.1)
    coef = head.compute_ibm(geo)
    head.init_fields(seed=0, amp=0.05, geo=geo, dt=DT)
    net = M.Net(NX, NY, NZ, seed=0)
    net.replay_reserve(3 * WARM)
    head.attach_net(net, capture=True)
    for _ in range(WARM):
        head.step(DT)
    state = head.download()
    head.attach_net(None)
    head.close()
    net.train(100, batch=4, lr=1e-3)
    g = O.Grid(NX, NY, NZ, lx=4.0, ly=2.2, lz=2.0, re=2000.0, forcing=(0.0, 0.0, 0.0), periodic=(1, 0, 1))
    ref = {k: v.copy() for k, v in state.items()}
    its = [[s[0] for s in O.rk3_step(g, ref, coef, DT, solver=2, omega=1.8, tol=TOL, max_iter=400000,
                                     check_every=20, flowrate=True, ub_target=1.0, height=2.0)]
           for _ in range(STEPS)]
    print("oracle SOR iterations per stage", its, flush=True)
    tight = _run(state, coef, M.PREC_FP64, tol=1e-12)[0]
    p_oracle = _errors(ref, tight)["p"]
    print(f"oracle (tol 1e-8) p vs the solve converged to 1e-12: rel-L2 {p_oracle:.3e}", flush=True)
    yield {"state": state, "coef": coef, "ref": ref, "tight": tight, "net": net, "p_oracle": p_oracle}
    net.close()


def _run(state, coef, precision, tol=TOL, guess=M.GUESS_PREVIOUS, net=None, env=None):
    saved = {}
    for k, v in (env or {}).items():
        saved[k] = os.environ.get(k)
        os.environ[k] = v
    try:
        ch = M.Channel(_cfg(precision, tol))
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    ch.set_ibm(*coef)
    ch.upload(**state)
    if net is not None:
        ch.attach_net(net, capture=False)
    ch.set_guess(guess)
    stats = [ch.step(DT) for _ in range(STEPS)]
    out = ch.download(("un", "vn", "wn", "pn"))
    div = ch.div_norm()
    info = ch.info()
    if net is not None:
        ch.attach_net(None)
    ch.close()
    return out, stats, div, info


def _report(name, e, div, stats):
    its = [[s.stage[q].iterations for q in range(3)] for s in stats]
    print(f"{name}: rel-L2 u {e['un']:.3e} v {e['vn']:.3e} w {e['wn']:.3e} p(mean removed) {e['p']:.3e}; "
          f"||div||_inf {div:.3e}; SOR iterations {its}", flush=True)


@pytest.mark.parametrize("lds", ["1", "2"])
def test_plain_fp64_matches_oracle(case, lds):
    out, stats, div, info = _run(case["state"], case["coef"], M.PREC_FP64, env={"MOBYDIFF_LDS_SOR": lds})
    e = _errors(out, case["ref"])
    _report(f"plain FP64 ({info})", e, div, stats)
    for k, v in e.items():
        assert v <= 1e-10, (k, v)
    assert div < TOL


def test_mixed_matches_oracle(case):
    out, stats, div, _ = _run(case["state"], case["coef"], M.PREC_MIXED)
    e = _errors(out, case["ref"])
    _report("mixed (FP32 sweeps in FP64 defect correction)", e, div, stats)
    for k in ("un", "vn", "wn"):
        assert e[k] <= 1e-6, (k, e[k])
    p_tight = _errors(out, case["tight"])["p"]
    print(f"  p vs the 1e-12 solve {p_tight:.3e} (oracle's {case['p_oracle']:.3e})", flush=True)
    assert e["p"] <= 1e-4 and p_tight <= 5.0 * case["p_oracle"]
    assert div < TOL


@pytest.mark.parametrize("which", ["warm-up pairs", "shipped"])
def test_learned_matches_oracle(case, which):
    # the network's start is taken whenever it is evaluated (margin test off): the network trained
    # on this run's warm-up pairs, and the stride-4 network trained on 50 other wavy channels
    net = case["net"]
    if which == "shipped":
        from mobydiff_amd import train_guess
        if not os.path.exists(WEIGHTS):
            pytest.skip("trained U-Net parameters not present")
        net = M.Net(NX, NY, NZ, seed=0, stride=train_guess.stride_of(WEIGHTS))
        train_guess.load(WEIGHTS, net)
    try:
        out, stats, div, _ = _run(case["state"], case["coef"], M.PREC_MIXED, guess=M.GUESS_LEARNED, net=net,
                                  env={"MOBYDIFF_NET_MARGIN": "1e30", "MOBYDIFF_NET_TIME_GATE": "0"})
    finally:
        if net is not case["net"]:
            net.close()
    e = _errors(out, case["ref"])
    _report(f"learned guess ({which}) + mixed", e, div, stats)
    started = sum(1 for s in stats for q in range(3) if not s.stage[q].fallback)
    print(f"  network-started solves {started} of {3 * STEPS}", flush=True)
    assert started >= 1
    for k in ("un", "vn", "wn"):
        assert e[k] <= 1e-6, (k, e[k])
    p_tight = _errors(out, case["tight"])["p"]
    print(f"  p vs the 1e-12 solve {p_tight:.3e} (oracle's {case['p_oracle']:.3e})", flush=True)
    assert e["p"] <= 1e-4 and p_tight <= 5.0 * case["p_oracle"]
    assert div < TOL
