"""Train the pressure-guess U-Net on pairs the wavy-channel solver generates, on one GPU.

Pairs: each of `runs` wavy channels (192 x 128 x 96, Lx x Ly x Lz = 4 x 2.2 x 2, Re_b = 2000 with
constant flow rate, dt = 5e-4, walls at y = +-(1 - A cos(2 pi x)) with A drawn uniformly from
[amin, amax], Poiseuille start with a 0.05 perturbation from seed `seed0 + run`) runs `steps` RK3
steps of the mixed-precision SOR solver (tolerance 1e-8, previous-step Phi guess); every `every`-th
step its three solves are captured into the network's replay buffer on the GPU as
(network input, converged Phi) pairs.  Then `adam_steps` Adam updates (batch 4, lr 1e-3, MSE) on
pairs drawn uniformly from the buffer.  The parameters (FP32), the target scale c0 and the stride
are written to an .npz file with plain arrays only (numpy.load with allow_pickle=False reads it).

The defaults are the recipe of the learned-guess workload: 50 channels (seeds 0..49, A uniform in
[0.05, 0.15]), 134 steps each with every solve captured: 20,100 pairs.  The network runs on the
solver grid coarsened by 4 (stride 4: its levels 48 x 32 x 24, 24 x 16 x 12, 12 x 8 x 6), so a
pair takes 0.3 MB of HBM and all of them stay resident (6 GB; 47 GB at stride 2).

    python -m mobydiff_amd.train_guess --out unet_guess_s4_wavy_192x128x96.npz

The network, its input and target, the capture and the Adam kernels are the library's own
(mdx_net_*, mdx_attach_net); this script only drives them.

Calibration of the output scale (`--calibrate WEIGHTS`, no training): the network is trained for
the mean-square error of its correction, but the SOR sweeps that follow remove the rough part of
an error quickly and its smooth part slowly, so the correction's best amplitude for the solve is
not the MSE-optimal one.  On held-out channels (seeds the training did not use, A across the
range) the script runs the learned start on every solve with the correction scaled by each
factor of `--scales` and keeps the factor with the fewest SOR iterations per solve; the weights
are written again with c0 times that factor, the sweep recorded in the metadata.

    python -m mobydiff_amd.train_guess --calibrate unet_guess_s4_wavy_192x128x96.npz --out calibrated.npz
"""
import argparse
import json
import math
import sys
import time

import numpy as np

NX, NY, NZ = 192, 128, 96
LX, LY, LZ = 4.0, 2.2, 2.0
RE, DT, AMP = 2000.0, 5e-4, 0.05
TOL, OMEGA = 1e-8, 1.8


def wavy_channel(M, A, seed, inner=20):
    """A fresh wavy channel (mixed precision, previous-step Phi guess) at its seeded start."""
    cfg = M.channel_config(NX, NY, NZ, LX, LY, LZ, RE, solver=M.SOLVER_SOR, omega=OMEGA, tol=TOL,
                           max_iter=200000, check_every=20, precision=M.PREC_MIXED, inner_iters=inner,
                           guess=M.GUESS_PHI_PREV, flowrate=True, flow_height=2.0)
    ch = M.Channel(cfg)
    geo = M.wavy_geometry(A=A)
    ch.compute_ibm(geo)
    ch.init_fields(seed=seed, amp=AMP, geo=geo, dt=DT)
    return ch


def save(path, net, meta):
    p, c0 = net.get_params()
    np.savez(path, params=p.astype(np.float32), c0=np.float64(c0), stride=np.int32(net.stride),
             meta=np.array(json.dumps(meta)))


def stride_of(path):
    """The stride of the network an .npz written by save() was trained with (1 for older files)."""
    with np.load(path, allow_pickle=False) as z:
        return int(z["stride"]) if "stride" in z.files else 1


def load(path, net):
    """Parameters and c0 from an .npz written by save() into `net`; returns the metadata dict."""
    with np.load(path, allow_pickle=False) as z:
        p, c0, meta = z["params"], float(z["c0"]), json.loads(str(z["meta"]))
        stride = int(z["stride"]) if "stride" in z.files else 1
    if p.size != net.nparam:
        raise ValueError(f"{path}: {p.size} parameters, the network has {net.nparam}")
    if stride != net.stride:
        raise ValueError(f"{path}: trained with stride {stride}, the network has stride {net.stride}")
    net.set_params(np.ascontiguousarray(p, dtype=np.float32), c0)
    return meta


def calibrate(M, path, out, scales, channels, warmup, steps, inner):
    """The output-scale sweep described in the module docstring, with `inner` FP32 iterations per
    defect-correction batch (the solver configuration the weights serve); the factors scale the
    trained c0 (a file calibrated before is recalibrated from its trained c0).  Returns the
    metadata written."""
    import os
    with np.load(path, allow_pickle=False) as z:
        stride = int(z["stride"]) if "stride" in z.files else 1
    net = M.Net(NX, NY, NZ, seed=0, stride=stride)
    meta = load(path, net)
    p, c0 = net.get_params()
    if "c0_calibration" in meta:
        c0 = float(meta["c0_calibration"]["c0_trained"])
    env = {"MOBYDIFF_NET_MARGIN": "1e30", "MOBYDIFF_NET_TIME_GATE": "0"}   # the network starts every solve
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    results = {f"{s:g}": [] for s in scales}
    try:
        for A, seed in channels:
            ch = wavy_channel(M, A, seed)
            geo = M.wavy_geometry(A=A)
            coef = ch.compute_ibm(geo)
            for _ in range(warmup):
                ch.step(DT)
            state = ch.download()
            ch.close()
            for s in scales:
                net.set_params(p, c0 * s)
                run = wavy_channel(M, A, seed, inner)
                run.set_ibm(*coef)
                run.upload(**state)
                run.attach_net(net, capture=False)
                run.set_guess(M.GUESS_LEARNED)
                run.step(DT)   # the per-stage history the network's input needs
                its = np.array([[st.stage[q].iterations for q in range(3)] for st in (run.step(DT) for _ in range(steps))])
                run.attach_net(None)
                run.close()
                results[f"{s:g}"].append(float(np.mean(its)))
                print(f"A={A} seed={seed} scale {s:g}: {np.mean(its):.1f} iterations per solve (stages "
                      f"{', '.join(f'{v:.1f}' for v in its.mean(axis=0))})", flush=True)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    mean = {k: float(np.mean(v)) for k, v in results.items()}
    # fewest iterations; among equals the factor nearest 1
    best = min(scales, key=lambda s: (round(mean[f"{s:g}"], 6), abs(s - 1.0)))
    net.set_params(p, c0 * best)
    meta = dict(meta, c0_calibration={"scale": best, "c0_trained": c0, "channels": [list(c) for c in channels],
                                      "warmup_steps": warmup, "steps": steps, "inner_iters": inner,
                                      "mean_iterations_per_solve": mean})
    save(out, net, meta)
    net.close()
    return meta


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--runs", type=int, default=50)
    ap.add_argument("--steps", type=int, default=134)
    ap.add_argument("--every", type=int, default=1)
    ap.add_argument("--amin", type=float, default=0.05)
    ap.add_argument("--amax", type=float, default=0.15)
    ap.add_argument("--seed0", type=int, default=0)
    ap.add_argument("--stride", type=int, default=4, help="network grid = solver grid / stride (1, 2 or 4)")
    ap.add_argument("--adam-steps", type=int, default=10000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-end", type=float, default=None,
                    help="cosine decay of the learning rate from --lr to this value over the Adam steps "
                         "(per chunk of 500 steps; default: constant --lr)")
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--out", required=True)
    ap.add_argument("--calibrate", default=None, metavar="WEIGHTS",
                    help="calibrate the output scale of trained weights instead of training")
    ap.add_argument("--scales", default="0.6,0.7,0.75,0.8,0.85,0.9,1.0")
    ap.add_argument("--cal-channels", default="0.06:50,0.10:51,0.14:52",
                    help="held-out channels A:seed of the calibration")
    ap.add_argument("--cal-warmup", type=int, default=10)
    ap.add_argument("--cal-steps", type=int, default=12)
    ap.add_argument("--cal-inner", type=int, default=32,
                    help="FP32 iterations per defect-correction batch of the calibration runs")
    args = ap.parse_args(argv)

    import mobydiff_amd as M

    if args.calibrate:
        scales = [float(x) for x in args.scales.split(",")]
        channels = [(float(a), int(s)) for a, s in (c.split(":") for c in args.cal_channels.split(","))]
        meta = calibrate(M, args.calibrate, args.out, scales, channels, args.cal_warmup, args.cal_steps,
                         args.cal_inner)
        print(json.dumps(meta["c0_calibration"]), flush=True)
        return 0

    rng = np.random.default_rng(args.seed0)
    amps = rng.uniform(args.amin, args.amax, args.runs)
    net = M.Net(NX, NY, NZ, seed=args.init_seed, stride=args.stride)
    per_run = 3 * ((args.steps + args.every - 1) // args.every)
    net.replay_reserve(args.runs * per_run)
    t0 = time.perf_counter()
    for r in range(args.runs):
        ch = wavy_channel(M, float(amps[r]), args.seed0 + r)
        for s in range(args.steps):
            ch.attach_net(net, capture=(s % args.every == 0))
            ch.step(DT)
        ch.synchronize()
        ch.attach_net(None)
        ch.close()
        print(f"run {r}: A={amps[r]:.4f} seed={args.seed0 + r} pairs={net.replay_count()} "
              f"({time.perf_counter() - t0:.1f} s)", flush=True)
    pairs = net.replay_count()
    t1 = time.perf_counter()
    l0 = net.eval_replay(min(pairs, 64))
    chunk, done, hist = 500, 0, []
    while done < args.adam_steps:
        n = min(chunk, args.adam_steps - done)
        lr = args.lr
        if args.lr_end is not None:
            lr = args.lr_end + (args.lr - args.lr_end) * 0.5 * (1.0 + math.cos(math.pi * done / args.adam_steps))
        hist.append(round(float(net.train(n, batch=args.batch, lr=lr, seed=done)), 6))
        done += n
        print(f"adam {done}/{args.adam_steps}: mean loss over the last {n} steps {hist[-1]:.5f} "
              f"({time.perf_counter() - t1:.1f} s)",
              flush=True)
    l1 = net.eval_replay(min(pairs, 64))
    meta = {"runs": args.runs, "steps": args.steps, "every": args.every, "A": [args.amin, args.amax],
            "seeds": [args.seed0, args.seed0 + args.runs - 1], "pairs": pairs, "adam_steps": args.adam_steps,
            "batch": args.batch, "lr": args.lr, "lr_end": args.lr_end, "mse_first64_before": float(l0),
            "mse_first64_after": float(l1),
            "loss_per_500_steps": hist, "capture_seconds": round(t1 - t0, 1),
            "train_seconds": round(time.perf_counter() - t1, 1), "grid": [NX, NY, NZ], "stride": args.stride}
    save(args.out, net, meta)
    print(json.dumps(meta), flush=True)
    net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
