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


def wavy_channel(M, A, seed):
    """A fresh wavy channel (mixed precision, previous-step Phi guess) at its seeded start."""
    cfg = M.channel_config(NX, NY, NZ, LX, LY, LZ, RE, solver=M.SOLVER_SOR, omega=OMEGA, tol=TOL,
                           max_iter=200000, check_every=20, precision=M.PREC_MIXED, inner_iters=20,
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
    args = ap.parse_args(argv)

    import mobydiff_amd as M

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
