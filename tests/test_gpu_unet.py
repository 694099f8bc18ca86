"""GPU checks of the pressure-guess U-Net (BF16/MFMA HIP kernels) against the torch-CPU fp32
reference of the same network (oracle/unet_ref.py).

Bar (written per test): forward within 1e-2 relative of the bf16-emulating reference; parameter
gradients within 5e-2 relative (L2, per layer) of fp32 autograd on the same bf16-rounded inputs;
Adam bit-for-bit-close (1e-6) to the closed form applied to the GPU's own gradients.
"""
import json
import os

import numpy as np
import pytest
import torch

import mobydiff_amd as M
from oracle import unet_ref as R

pytestmark = pytest.mark.gpu

NX, NY, NZ = 64, 16, 16


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30))


def _inputs(seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    smooth = np.sin(2 * np.pi * x / NX) * np.cos(np.pi * y / NY) + 0.5 * np.cos(2 * np.pi * z / NZ)
    x2 = np.stack([smooth + 0.3 * rng.standard_normal(smooth.shape),
                   np.cos(2 * np.pi * x / NX + 0.3) + 0.3 * rng.standard_normal(smooth.shape)], axis=-1)
    target = (0.5 * smooth + 0.1 * rng.standard_normal(smooth.shape)).astype(np.float32)
    return x2.astype(np.float32), target


def _torch_x(x2):
    t = torch.zeros(1, 16, NZ, NY, NX)
    t[0, 0] = torch.as_tensor(x2[..., 0])
    t[0, 1] = torch.as_tensor(x2[..., 1])
    return t


def test_param_layout_matches_reference():
    net = M.Net(NX, NY, NZ, seed=3)
    p, _ = net.get_params()
    assert p.size == sum(k ** 3 * ci * co + co for _, ci, co, k in R.LAYERS)
    params = R.flat_to_params(p)
    assert np.array_equal(R.params_to_flat(params), p)
    # He-normal initialisation: per-layer std close to sqrt(2 / fan_in)
    for name, ci, co, k in R.LAYERS:
        w = params[name + ".w"].numpy()
        assert abs(w.std() / np.sqrt(2.0 / (k ** 3 * ci)) - 1) < 0.15, name


def test_forward_matches_bf16_reference():
    net = M.Net(NX, NY, NZ, seed=1)
    p, _ = net.get_params()
    params = R.flat_to_params(p)
    x2, _ = _inputs(0)
    out = net.forward(x2)
    ref = R.forward_bf16(params, _torch_x(x2))[0, 0].numpy()
    assert np.isfinite(out).all()
    assert _rel(out, ref) < 1e-2, _rel(out, ref)
    ref32 = R.forward(params, _torch_x(x2))[0, 0].detach().numpy()
    assert _rel(out, ref32) < 5e-2


def test_tiled_convolutions_match_direct_kernels(monkeypatch):
    """The LDS-tiled convolutions use the same fragments, K order and accumulation sequence as the
    direct kernels: forward outputs are bit-identical; gradients (FP32 atomics in the weight
    gradient) agree to rounding."""
    x2, target = _inputs(5)
    outs, grads = [], []
    for flag in ("1", "0"):
        monkeypatch.setenv("MOBYDIFF_NO_TILED_CONV", flag)
        net = M.Net(NX, NY, NZ, seed=11)
        outs.append(net.forward(x2))
        net.zero_grads()
        net.accumulate_grad(x2, target)
        grads.append(net.get_grads())
    assert np.array_equal(outs[0], outs[1])
    assert _rel(grads[1], grads[0]) < 1e-5


def test_split_k_convolutions_match_direct_kernels(monkeypatch):
    """On the coarse levels of the stride-4 network (24 x 16 x 12 and 12 x 8 x 6 voxels, which the
    tiled kernel cannot take) the split-K convolutions sum the direct kernels' K chunks in another
    order: forward outputs and gradients agree to rounding."""
    rng = np.random.default_rng(7)
    n0 = (48, 32, 24)
    x2 = rng.standard_normal(n0[::-1] + (2,)).astype(np.float32)
    target = rng.standard_normal(n0[::-1]).astype(np.float32)
    outs, grads = [], []
    for flag in ("1", "0"):
        monkeypatch.setenv("MOBYDIFF_NO_SPLIT_CONV", flag)
        net = M.Net(192, 128, 96, seed=3, stride=4)
        p, _ = net.get_params()
        outs.append(net.forward(x2))
        net.zero_grads()
        net.accumulate_grad(x2, target)
        grads.append(net.get_grads())
        net.close()
    t = torch.zeros(1, 16, *n0[::-1])
    t[0, 0] = torch.as_tensor(x2[..., 0])
    t[0, 1] = torch.as_tensor(x2[..., 1])
    ref = R.forward_bf16(R.flat_to_params(p), t)[0, 0].numpy()
    print(f"split-K against direct: forward rel {_rel(outs[1], outs[0]):.2e}, gradients rel "
          f"{_rel(grads[1], grads[0]):.2e}; against the reference: direct {_rel(outs[0], ref):.2e}, "
          f"split-K {_rel(outs[1], ref):.2e}", flush=True)
    assert _rel(outs[0], ref) < 1e-2 and _rel(outs[1], ref) < 1e-2
    assert _rel(outs[1], outs[0]) < 1e-2
    assert _rel(grads[1], grads[0]) < 1e-2


def test_gradients_match_autograd():
    net = M.Net(NX, NY, NZ, seed=2)
    p, _ = net.get_params()
    params = R.flat_to_params(p)
    x2, target = _inputs(1)
    net.zero_grads()
    loss = net.accumulate_grad(x2, target, batch=1)
    g = R.flat_to_params(net.get_grads())
    bparams = {k: (R._bf(v) if k.endswith(".w") and not k.startswith("head") else v) for k, v in params.items()}
    rl, rg = R.loss_and_grads(bparams, R._bf(_torch_x(x2)), torch.as_tensor(target)[None, None])
    assert abs(loss - float(rl)) < 2e-2 * float(rl)
    for name, ci, co, k in R.LAYERS:
        for suffix in (".w", ".b"):
            a, b = g[name + suffix].numpy(), rg[name + suffix].numpy()
            assert _rel(a, b) < 5e-2, (name + suffix, _rel(a, b))


def test_gradient_directional_finite_difference():
    """g . d against (L(p + eps d) - L(p - eps d)) / (2 eps) on the GPU network itself, per layer,
    with d the layer's own gradient: a random direction makes g . d a sum of cancelling terms
    whose size falls below the bf16 rounding noise of the loss."""
    net = M.Net(NX, NY, NZ, seed=4)
    p, _ = net.get_params()
    x2, target = _inputs(2)
    net.zero_grads()
    net.accumulate_grad(x2, target)
    g = net.get_grads().astype(np.float64)
    off = 0
    checked = 0
    for name, ci, co, k in R.LAYERS:
        n = k ** 3 * ci * co + co
        if name in ("e1a", "b3a", "d1b", "head"):
            d = np.zeros(p.size)
            d[off:off + n] = g[off:off + n]
            # step of 2% of the layer's weight norm along its gradient
            eps = 0.02 * np.linalg.norm(p[off:off + n]) / max(np.linalg.norm(d), 1e-30)
            losses = []
            for s in (+1, -1):
                net.set_params((p + s * eps * d).astype(np.float32))
                net.zero_grads()
                losses.append(net.accumulate_grad(x2, target))
            fd = (losses[0] - losses[1]) / (2 * eps)
            an = float(np.dot(g, d))
            assert abs(fd - an) < 0.1 * abs(an), (name, fd, an)
            checked += 1
        off += n
    assert off == p.size and checked == 4


def test_adam_matches_closed_form():
    net = M.Net(NX, NY, NZ, seed=6)
    p, _ = net.get_params()
    x2, target = _inputs(3)
    net.zero_grads()
    net.accumulate_grad(x2, target)
    g = net.get_grads()
    net.adam_step(1e-3)
    p1, _ = net.get_params()
    P = {"w": torch.as_tensor(p)}
    G = {"w": torch.as_tensor(g)}
    Z = {"w": torch.zeros_like(P["w"])}
    ref, _, _ = R.adam_step(P, G, Z, Z, 1, lr=1e-3)
    assert np.max(np.abs(p1 - ref["w"].numpy())) < 1e-6
    assert np.array_equal(net.get_grads(), np.zeros_like(g))


def test_training_reduces_loss():
    net = M.Net(NX, NY, NZ, seed=7)
    data = [_inputs(10 + q) for q in range(4)]
    first = None
    for step in range(60):
        loss = 0.0
        for x2, t in data:
            loss += net.accumulate_grad(x2, t, batch=len(data))
        net.adam_step(3e-3)
        first = loss if first is None else first
    assert loss < 0.5 * first, (first, loss)


def test_learned_guess_in_the_solver():
    """Capture (div u*, previous-step Phi of the stage) -> Phi pairs from the wavy-channel solver,
    train the network on them, then continue the same flow with the learned guess: the tolerance
    is still met, the fallback is reported, and the kept guesses start from a smaller residual
    than the previous-step guess (Phi = 0)."""
    nx, ny, nz, dt = 64, 32, 32, 5e-4
    geo = M.wavy_geometry(A=0.1)
    cfg = M.channel_config(nx, ny, nz, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8,
                           check_every=10, precision=M.PREC_MIXED, inner_iters=10, flowrate=True,
                           flow_height=2.0)
    ch = M.Channel(cfg)
    ch.compute_ibm(geo)
    ch.init_fields(seed=0, amp=0.05, geo=geo, dt=dt)
    net = M.Net(nx, ny, nz, seed=0)
    net.replay_reserve(64)
    ch.attach_net(net, capture=True)
    for _ in range(12):
        ch.step(dt)
    assert net.replay_count() == 36
    l0 = net.eval_replay(8)
    net.train(150, batch=4, lr=1e-3)
    l1 = net.eval_replay(8)
    assert l1 < 0.5 * l0, (l0, l1)
    ch.attach_net(net, capture=False)
    ch.set_guess(M.GUESS_LEARNED)
    its, fallbacks, kept = [], 0, 0
    for _ in range(3):
        st = ch.step(dt)
        for q in range(3):
            s = st.stage[q]
            assert s.converged == 1 and s.div_inf < 1e-8
            its.append(s.iterations)
            fallbacks += s.fallback
            if not s.fallback:
                kept += 1
                assert s.guess_div_inf == s.div0_inf
    assert ch.div_norm() < 1.5e-8
    assert kept >= 1, its
    print("learned-guess iterations", its, "fallbacks", fallbacks)
    ch.close()


def test_previous_phi_guess_in_the_solver():
    """The non-learned ablation: the previous-step Phi of the same stage, scaled by the ratio of the
    stages' dt*gamma, as the initial guess.  The same flow with the plain start (the previous
    pressure, Phi = 0) needs more sweeps after the first step; both meet the tolerance."""
    nx, ny, nz, dt = 64, 32, 32, 5e-4
    geo = M.wavy_geometry(A=0.1)
    its = {}
    for guess in (M.GUESS_PREVIOUS, M.GUESS_PHI_PREV):
        cfg = M.channel_config(nx, ny, nz, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8,
                               check_every=10, guess=guess)
        ch = M.Channel(cfg)
        ch.compute_ibm(geo)
        ch.init_fields(seed=0, amp=0.05, geo=geo, dt=dt)
        first = ch.step(dt)
        # no history yet: the plain start, nothing rejected
        assert all(first.stage[q].fallback == 0 for q in range(3))
        its[guess] = []
        for _ in range(4):
            st = ch.step(dt)
            for q in range(3):
                s = st.stage[q]
                assert s.converged == 1 and s.div_inf < 1e-8 and s.fallback == 0
                its[guess].append(s.iterations)
        assert ch.div_norm() < 1.5e-8
        ch.close()
    print("sweeps per solve: plain start", its[M.GUESS_PREVIOUS], "previous Phi", its[M.GUESS_PHI_PREV])
    assert sum(its[M.GUESS_PHI_PREV]) < sum(its[M.GUESS_PREVIOUS]), its


def test_train_guess_script_round_trip(tmp_path):
    """mobydiff_amd.train_guess end to end at its workload size, cut to one channel, two steps and
    two Adam steps: it captures three pairs per step, trains, and writes parameters that load into
    a fresh network unchanged (with c0 > 0, so the learned start can be used)."""
    from mobydiff_amd import train_guess
    out = str(tmp_path / "w.npz")
    assert train_guess.main(["--runs", "1", "--steps", "2", "--every", "1", "--adam-steps", "2", "--out", out]) == 0
    assert train_guess.stride_of(out) == 4
    net = M.Net(train_guess.NX, train_guess.NY, train_guess.NZ, seed=5, stride=4)
    meta = train_guess.load(out, net)
    assert meta["pairs"] == 6 and meta["adam_steps"] == 2
    p, c0 = net.get_params()
    with np.load(out, allow_pickle=False) as z:
        assert np.array_equal(p, z["params"]) and c0 == float(z["c0"]) and c0 > 0.0
    net.close()


SHIPPED = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mobydiff_amd", "weights",
                       "unet_guess_s4_wavy_192x128x96.npz")


@pytest.mark.skipif(not os.path.exists(SHIPPED), reason="trained U-Net parameters not present")
def test_train_guess_calibration_round_trip(tmp_path):
    """mobydiff_amd.train_guess --calibrate, cut to one held-out channel, one warm-up step, two
    measured steps and two factors: it runs the learned start at each factor, keeps the one with
    fewer iterations per solve (ties to the factor nearest 1), and writes the same parameters with
    the trained c0 times that factor and the sweep in the metadata."""
    from mobydiff_amd import train_guess
    out = str(tmp_path / "cal.npz")
    assert train_guess.main(["--calibrate", SHIPPED, "--out", out, "--scales", "0.5,1.0",
                             "--cal-channels", "0.08:60", "--cal-warmup", "1", "--cal-steps", "2"]) == 0
    with np.load(SHIPPED, allow_pickle=False) as a, np.load(out, allow_pickle=False) as b:
        assert np.array_equal(a["params"], b["params"])
        c0_out = float(b["c0"])
        meta_in = json.loads(str(a["meta"]))
        # the factors scale the trained c0 (the shipped file may be calibrated already)
        c0_in = float(meta_in["c0_calibration"]["c0_trained"]) if "c0_calibration" in meta_in else float(a["c0"])
    net = M.Net(train_guess.NX, train_guess.NY, train_guess.NZ, seed=0, stride=4)
    meta = train_guess.load(out, net)
    net.close()
    cal = meta["c0_calibration"]
    its = cal["mean_iterations_per_solve"]
    print(f"calibration: {its}, scale {cal['scale']}", flush=True)
    assert set(its) == {"0.5", "1"} and all(v > 0 for v in its.values())
    want = 0.5 if its["0.5"] < its["1"] else 1.0
    assert cal["scale"] == want and cal["c0_trained"] == c0_in
    assert abs(c0_out - c0_in * want) <= 1e-15 * c0_in


@pytest.mark.parametrize("fallback", ["prev", "zero"])
def test_learned_fallback_is_the_other_start(monkeypatch, fallback):
    """With the network's start never taken (MOBYDIFF_NET_MARGIN=0), every solve of the learned mode
    falls back: to the previous pressure (Phi = 0, the plain mode's start; the default) or, with
    MOBYDIFF_NET_FALLBACK=prev, to the previous-step guess Phi_0 (the previous-Phi mode's start).
    The fields after several steps are bit-identical to that mode's, and every solve reports the
    fallback."""
    monkeypatch.setenv("MOBYDIFF_NET_MARGIN", "0")
    monkeypatch.setenv("MOBYDIFF_NET_FALLBACK", fallback)
    other = M.GUESS_PHI_PREV if fallback == "prev" else M.GUESS_PREVIOUS
    nx, ny, nz, dt = 64, 32, 32, 5e-4
    geo = M.wavy_geometry(A=0.1)
    fields, fb = [], []
    for learned in (False, True):
        cfg = M.channel_config(nx, ny, nz, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8,
                               check_every=10, precision=M.PREC_MIXED, inner_iters=10, flowrate=True,
                               flow_height=2.0, guess=other)
        ch = M.Channel(cfg)
        ch.compute_ibm(geo)
        ch.init_fields(seed=0, amp=0.05, geo=geo, dt=dt)
        net = None
        if learned:
            net = M.Net(nx, ny, nz, seed=4)
            p, _ = net.get_params()
            net.set_params(p, c0=1.0)
            ch.attach_net(net, capture=False)
            ch.set_guess(M.GUESS_LEARNED)
        for _ in range(4):
            st = ch.step(dt)
            if learned:
                fb.extend(st.stage[q].fallback for q in range(3))
        fields.append(ch.download(("un", "vn", "wn", "pn")))
        ch.close()
        if net is not None:
            net.close()
    assert fb == [1] * 12, fb
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(fields[0][k], fields[1][k]), k


@pytest.mark.parametrize("s", [2, 4])
def test_strided_transfer_operators_match_oracle(s):
    """The strided network's restriction (mean over s^3 cells) and trilinear prolongation, run by
    the kernels of the solver coupling, against the oracle's separable-matrix restatement."""
    n = 64
    net = M.Net(n, n, n, seed=0, stride=s)
    assert net.shape == (n // s,) * 3
    rng = np.random.default_rng(s)
    fine = rng.standard_normal((n, n, n))
    r = net.restrict(fine)
    assert np.allclose(r, R.restrict_mean(fine, s), rtol=1e-6, atol=1e-7)
    coarse = rng.standard_normal(net.shape).astype(np.float32)
    p = net.prolong(coarse)
    ref = R.prolong_trilinear(coarse, s)
    assert np.abs(p - ref).max() <= 1e-12 * np.abs(ref).max()
    net.close()


def test_strided_learned_guess_in_the_solver():
    """A stride-2 network (its levels at 32x16x16, 16x8x8, 8x4x4 for a 64x32x32 channel): pairs
    captured from the solver, trained, then the learned start inside the solver meets the
    tolerance on every solve."""
    nx, ny, nz, dt = 64, 32, 32, 5e-4
    geo = M.wavy_geometry(A=0.1)
    cfg = M.channel_config(nx, ny, nz, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8,
                           check_every=10, precision=M.PREC_MIXED, inner_iters=10, flowrate=True,
                           flow_height=2.0)
    ch = M.Channel(cfg)
    ch.compute_ibm(geo)
    ch.init_fields(seed=0, amp=0.05, geo=geo, dt=dt)
    net = M.Net(nx, ny, nz, seed=0, stride=2)
    net.replay_reserve(64)
    ch.attach_net(net, capture=True)
    for _ in range(12):
        ch.step(dt)
    assert net.replay_count() == 36
    l0 = net.eval_replay(8)
    net.train(150, batch=4, lr=1e-3)
    l1 = net.eval_replay(8)
    assert l1 < 0.5 * l0, (l0, l1)
    ch.attach_net(net, capture=False)
    ch.set_guess(M.GUESS_LEARNED)
    for _ in range(3):
        st = ch.step(dt)
        for q in range(3):
            assert st.stage[q].converged == 1 and st.stage[q].div_inf < 1e-8
    assert ch.div_norm() < 1.5e-8
    ch.close()
    net.close()


@pytest.mark.parametrize("stride", [1, 4])
def test_inference_concat_read_in_place(stride, monkeypatch):
    """In inference the decoder convolutions read the up-sampled coarse tensor and the skip tensor
    where they lie (the loader forms the concat) instead of a materialised concat: the forward
    output is bit-identical to the run that stores the concat first (MOBYDIFF_NO_CAT_FUSION=1)."""
    rng = np.random.default_rng(13)
    grid = (NX, NY, NZ) if stride == 1 else (192, 128, 96)
    n0 = tuple(g // stride for g in grid)
    x2 = rng.standard_normal(n0[::-1] + (2,)).astype(np.float32)
    outs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("MOBYDIFF_NO_CAT_FUSION", flag)
        net = M.Net(*grid, seed=17, stride=stride)
        outs.append(net.forward(x2))
        net.close()
    assert np.isfinite(outs[0]).all()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("stride", [1, 4])
def test_inference_first_pooling_in_conv_epilogue(stride, monkeypatch):
    """In inference the convolution before the first pooling also writes the 2x2x2 averages of its
    output (pool_fwd_kernel's order of addition): the forward output is bit-identical to the run
    with the separate pooling launch (MOBYDIFF_NO_POOL_FUSION=1)."""
    rng = np.random.default_rng(29)
    grid = (NX, NY, NZ) if stride == 1 else (192, 128, 96)
    n0 = tuple(g // stride for g in grid)
    x2 = rng.standard_normal(n0[::-1] + (2,)).astype(np.float32)
    outs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("MOBYDIFF_NO_POOL_FUSION", flag)
        net = M.Net(*grid, seed=23, stride=stride)
        outs.append(net.forward(x2))
        net.close()
    assert np.isfinite(outs[0]).all()
    assert np.array_equal(outs[0], outs[1])
