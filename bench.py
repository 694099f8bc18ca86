"""Benchmark: wavy-wall channel time step on one MI355X.

Workload (the headline configuration): wavy channel, walls at y = +-(1 - A cos(2 pi x)), A = 0.1,
Lx x Ly x Lz = 4 x 2.2 x 2, uniform grid 192 x 128 x 96, Re_b = 2000 with constant flow rate,
dt = 5e-4, Poiseuille + 0.05 seeded perturbation.  Each step is one RK3 step: momentum + IBM,
red-black SOR pressure solve to ||div||_inf < 1e-8, projection, boundary conditions.

During the warm-up steps the solver's converged solves are captured as training pairs for the
U-Net initial guess, which is then trained on the GPU (Adam, batch 4) before the timed steps.
Timed configurations, all from the same state after the warm-up steps:
  * headline: learned initial guess + FP32 red-black sweeps inside an FP64 defect-correction loop;
  * plain: FP64 red-black SOR from the previous pressure (the comparison path);
  * mixed_no_guess: the headline's solver without the guess;
  * mixed_previous_phi_guess: the previous step's Phi of each stage as the guess, no network.
Prints one JSON line (see the "metric" field); value = Mcell-steps/s of the headline path.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s at ‖div‖∞<1e-8: learned+FP32 SOR vs plain FP64 solve"
NX, NY, NZ = 192, 128, 96
LX, LY, LZ = 4.0, 2.2, 2.0
RE, DT, A, AMP = 2000.0, 5e-4, 0.1, 0.05
TOL, OMEGA = 1e-8, 1.8
HBM_PEAK_GBS = 8000.0
# LDS read bandwidth of the chip: 256 CUs x 256 B per clock x 2.4 GHz (CDNA 4: 160 KB of LDS per CU
# with twice CDNA 3's 128 B per clock of read bandwidth); mdx_lds_read_rate measures the sustained
# rate on the device, reported beside it
LDS_PEAK_GBS = 256 * 256 * 2.4
# U-Net parameters (stride 4) trained by mobydiff_amd.train_guess on 20,100 solver pairs of 50 wavy
# channels (A uniform in [0.05, 0.15], seeds 0..49); the default network of the learned guess
WEIGHTS = os.path.join(ROOT, "mobydiff_amd", "weights", "unet_guess_s4_wavy_192x128x96.npz")
# canonical algorithmic bytes per cell for one full red+black iteration (FP64 88 B, FP32 44 B); a
# one-colour launch covers half of that, a fused launch all of it
BYTES_PER_CELL_ITER = {"fp64": 88.0, "fp32": 44.0}


def make_channel(M, precision, check_every, inner, tol=TOL):
    cfg = M.channel_config(NX, NY, NZ, LX, LY, LZ, RE, solver=M.SOLVER_SOR, omega=OMEGA, tol=tol,
                           max_iter=200000, check_every=check_every, precision=precision,
                           inner_iters=inner, flowrate=True, flow_height=2.0, timing=True)
    return M.Channel(cfg)


def run_steps(ch, k):
    # the timed steps queue back to back: no per-step bulk-velocity report (a wait for the step's
    # end that the reference's time loop does not have)
    ch.set_step_report(False)
    ch.synchronize()
    ch.reset_phase_times()
    its, fb, started = [], 0, []
    t0 = time.perf_counter()
    for _ in range(k):
        st = ch.step(DT)
        its.append([st.stage[q].iterations for q in range(3)])
        fb += sum(st.stage[q].fallback for q in range(3))
        started.append([not st.stage[q].fallback for q in range(3)])
    ch.synchronize()
    el = time.perf_counter() - t0
    div = ch.div_norm()
    run_steps.started = started   # which solves the network started (learned mode)
    return el, its, div, ch.phase_times(), fb


# HBM bytes per dispatch of the sweep kernels, from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this
# workload scaled by the counters' reading of a known-size copy (profiles/sweep_traffic_wavy_192x128x96.json)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sweep_traffic_wavy_192x128x96.json")


def measured_traffic(kind, kz, iters=None):
    """Measured HBM bytes per dispatch.  An LDS-resident dispatch is evaluated at `iters`, the
    sampled mean iterations per launch, from the affine fit of its counters over merged dispatches
    (or, without a fit, scaled from the fixed 20-iteration measurement)."""
    try:
        with open(TRAFFIC_FILE) as fh:
            table = json.load(fh)
    except (OSError, ValueError):
        return None
    entry = table.get("kernels", {}).get(f"{kind}_{kz}" if kz in ("fused", "lds") else f"{kind}_kz{kz}")
    if entry is None:
        return None
    b = entry.get("bytes_per_dispatch")
    per = entry.get("iterations_per_dispatch")
    if iters and "bytes_per_iteration" in entry:
        # affine fit over merged dispatches of 20 to 160 iterations: the defect epilogue and the
        # first loads are paid once per dispatch, the halo exchange once per iteration
        return round(entry["fixed_bytes"] + entry["bytes_per_iteration"] * iters)
    if b is not None and per and iters:
        b = round(b * iters / per)
    return b


def lds_plan(info):
    """(px, py, pz, bx, by, bz) of the LDS-resident solver's FP32 plan, from its info line
    ("... fp32 8x8x4 workgroups of 24x16x24 cells ...")."""
    import re
    m = re.search(r"fp32 (\d+)x(\d+)x(\d+) workgroups of (\d+)x(\d+)x(\d+) cells", info)
    return tuple(int(v) for v in m.groups()) if m else None


def lds_bytes_per_iteration(plan, elem=4):
    """LDS bytes one red+black iteration of the LDS-resident solver moves, over all workgroups:
    every relaxed cell takes six LDS loads and one store (the column forms: the z neighbour shared
    by two cells of one colour is loaded once); colour 0 relaxes the subdomain's half and the
    colour-0 cells of the face ring, colour 1 the subdomain's other half; and every exchanged value
    is read from LDS once to be published and written once into the receiver's halo (the colour-1
    half of the faces two layers deep and of the edges one cell thick; none across the walls)."""
    px, py, pz, bx, by, bz = plan
    total = 0
    for wy in range(py):
        has = {(0, -1, 0): wy > 0, (0, 1, 0): wy < py - 1}   # y neighbours (x, z periodic)
        nb = lambda dx, dy, dz: has.get((0, dy, 0), True) if dy else True   # noqa: E731
        ring = (by * bz) * (nb(-1, 0, 0) + nb(1, 0, 0)) // 2 + (bx * bz) * (nb(0, -1, 0) + nb(0, 1, 0)) // 2 \
            + (bx * by) * (nb(0, 0, -1) + nb(0, 0, 1)) // 2
        relaxed = bx * by * bz + ring
        faces = 2 * (by * bz * (nb(-1, 0, 0) + nb(1, 0, 0)) + bx * bz * (nb(0, -1, 0) + nb(0, 1, 0))
                     + bx * by * (nb(0, 0, -1) + nb(0, 0, 1)))
        edges = sum(bz for sx in (-1, 1) for sy in (-1, 1) if nb(0, sy, 0)) \
            + sum(by for sx in (-1, 1) for sz in (-1, 1)) + sum(bx for sy in (-1, 1) for sz in (-1, 1) if nb(0, sy, 0))
        total += px * pz * (7 * elem * relaxed + 2 * elem * (faces + edges) // 2)
    return total


def roofline(times, kind, info=""):
    """Roofline entry of the red-black sweep kernel over its event-timed launches.
    * LDS-resident batches (rb_lds_kernel): the iterations run from LDS, so the bound is LDS
      bandwidth: LDS bytes per launch (lds_bytes_per_iteration x the launch's iterations) over the
      launch time, against the LDS peak.  The HBM side is the counter-measured traffic per launch
      (FETCH_SIZE + WRITE_SIZE, scaled by a known-size copy) over the same time.
    * streaming one-colour sweeps: canonical bytes (44 B per cell and iteration in FP32, 88 in FP64,
      half per colour) over the launch time against the HBM peak; at 192x128x96 the working set sits
      in the 256 MB MALL, so the counter-measured HBM bytes are reported beside it and the fraction
      is taken of those."""
    if times.sweep_launches == 0:
        return None
    avg_s = times.sweep_ms / times.sweep_launches * 1e-3
    cells_per_launch = times.sweep_cells / times.sweep_launches
    canonical = BYTES_PER_CELL_ITER[kind] / 2.0 * cells_per_launch
    iters = cells_per_launch / (2.0 * NX * NY * NZ)
    lds = info.startswith("LDS-resident") and (kind == "fp32" or "; fp64 " in info)
    if lds:
        plan = lds_plan(info)
        lds_b = lds_bytes_per_iteration(plan) * iters if plan else None
        traffic = measured_traffic(kind, "lds", iters)
        out = {"kernel": f"rb_lds_kernel: LDS-resident red-black SOR batch ({kind}, {iters:.1f} iterations per "
                         "launch on average)",
               "bound": "lds", "unit": "GB/s", "peak": LDS_PEAK_GBS,
               "achieved": round(lds_b / avg_s / 1e9, 1) if lds_b else None,
               "frac": round(lds_b / avg_s / 1e9 / LDS_PEAK_GBS, 4) if lds_b else None,
               "lds_bytes_per_launch": lds_b, "traffic": traffic, "avg_launch_us": round(avg_s * 1e6, 2),
               "us_per_iteration": round(avg_s * 1e6 / iters, 3) if iters else None,
               "canonical_bytes_if_streamed": canonical}
        if traffic:
            out["hbm"] = {"achieved": round(traffic / avg_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                          "frac": round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4), "source": "rocprofv3 counters"}
        return out
    if iters > 0.75:
        kz, label = "fused", f"red-black SOR iteration ({kind}, both colours in one launch)"
    else:
        kz = os.environ.get("MOBYDIFF_SWEEP_KZ", "4")
        label = f"sor_sweep_zm_kernel: red-black SOR sweep ({kind}, one colour)"
    traffic = measured_traffic(kind, kz, None)
    out = {"kernel": label, "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS, "bytes_per_launch": canonical,
           "canonical_gbs": round(canonical / avg_s / 1e9, 1), "traffic": traffic,
           "avg_launch_us": round(avg_s * 1e6, 2)}
    if traffic:
        out["achieved"] = round(traffic / avg_s / 1e9, 1)
        out["frac"] = round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4)
    return out


def momentum_roofline(ch, cells, ibm, dt, workload):
    """The momentum kernel alone: device time from HIP events on the channel's stream around 20
    launches of the second RK3 stage's predictor (mdx_time_momentum; the RK3 history is restored).
    Canonical algorithmic traffic 128 B per cell with IBM (10 FP64 reads + 6 writes), 104 B without."""
    ms = ch.time_momentum(50.0 / 120.0 * dt, -34.0 / 120.0 * dt, 16.0 / 120.0 * dt, reps=20)
    b = (128.0 if ibm else 104.0) * cells
    ach = b / (ms * 1e-3) / 1e9
    return {"kernel": "momentum_tiled_kernel", "workload": workload, "bound": "hbm",
            "avg_launch_us": round(ms * 1e3, 2), "bytes_per_launch": b, "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4)}


def momentum_in_step(t, cells, ibm, workload):
    """The momentum kernel inside the timed steps: HIP events on the channel's stream around its
    launch in every fourth RK stage (mdx_phase_times.momentum_ms over `stages`).  Canonical
    algorithmic traffic 128 B per cell with IBM, 104 B without."""
    if t.stages == 0:
        return None
    avg_s = t.momentum_ms / t.stages * 1e-3
    b = (128.0 if ibm else 104.0) * cells
    ach = b / avg_s / 1e9
    return {"kernel": "momentum_tiled_kernel", "workload": workload,
            "measured": f"in the timed steps ({t.stages} sampled launches)", "bound": "hbm",
            "avg_launch_us": round(avg_s * 1e6, 2), "bytes_per_launch": b, "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4)}


def network_gain(its, started, its_plain):
    """Sweeps of the solves the network started against the same step and stage of the run from the
    plain start (mixed_no_guess, from the same state): the guess's gain where it was used."""
    pairs = [(its[k][q], its_plain[k][q]) for k in range(len(its)) for q in range(3) if started[k][q]]
    if not pairs:
        return {"count": 0}
    a, b = sum(p[0] for p in pairs), sum(p[1] for p in pairs)
    return {"count": len(pairs), "iterations": a / len(pairs), "plain_start_iterations": b / len(pairs),
            "sweeps_saved": round(1.0 - a / b, 4)}


def flat_fft_production(M, steps, precision=None):
    """Secondary line: the flat channel at production size, 256 x 192 x 256, tanh-stretched y
    (beta 1.5), 4 pi x 2 x 2 pi, Re_b = 5600, constant flow rate, dt = 5e-4, FFT + tridiagonal
    Poisson solve.  Canonical algorithmic traffic 768 B per cell-step (3 stages x 256 B)."""
    import math
    nx, ny, nz = 256, 192, 256
    cfg = M.channel_config(nx, ny, nz, 4 * math.pi, 2.0, 2 * math.pi, 5600.0, solver=M.SOLVER_FFT,
                           flowrate=True, ub_target=1.0, flow_height=2.0, stretch_beta=1.5, p_neumann_y=False,
                           timing=True)
    if precision is not None:
        cfg.precision = precision   # mixed: FP32 transforms/Thomas in FP64 defect correction to tol
    ch = M.Channel(cfg)
    ch.init_fields(seed=0, amp=0.1, geo=None, dt=5e-4)
    for _ in range(2):
        ch.step(5e-4)
    ch.synchronize()
    ch.reset_phase_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        ch.step(5e-4)
    ch.synchronize()
    el = time.perf_counter() - t0
    t = ch.phase_times()
    div = ch.div_norm()
    cells = nx * ny * nz
    mom = momentum_in_step(t, cells, False, "flat channel 256x192x256") if precision is None else None
    mom_rep = momentum_roofline(ch, cells, False, 5e-4, "flat channel 256x192x256") if precision is None else None
    ch.close()
    gbs = 768.0 * cells * steps / el / 1e9
    label = "FP64" if precision in (None, M.PREC_FP64) else "FP32 transforms/Thomas in FP64 defect correction to 1e-8"
    return {"workload": "flat channel 256x192x256 (beta=1.5, Re_b=5600, dt=5e-4), FFT+tridiagonal Poisson, " + label,
            "value": round(cells * steps / el / 1e6, 3), "unit": "Mcell-steps/s", "steps": steps,
            "ms_per_step": round(el / steps * 1e3, 3),
            "phase_ms_per_step": {"stencil": round(t.stencil_ms * 3.0 / max(t.stages, 1), 3),
                                  "poisson": round(t.poisson_ms * 3.0 / max(t.stages, 1), 3),
                                  "projection": round(t.projection_ms * 3.0 / max(t.stages, 1), 3)},
            "div_inf_final": div,
            "step_roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(gbs / HBM_PEAK_GBS, 4),
                              "bytes_per_step": 768.0 * cells},
            "stencil_roofline": mom, "stencil_roofline_repeated_launches": mom_rep}


def cpu_baseline(f_host, coef, cores, steps):
    """The oracle (C restatement of the reference CPU path, FP64, OpenMP) on the same state.  Returns
    the baseline entry and the oracle's fields after the steps (the reference of the accuracy check)."""
    from oracle import oracle as O
    g = O.Grid(NX, NY, NZ, lx=LX, ly=LY, lz=LZ, re=RE, forcing=(0.0, 0.0, 0.0), periodic=(1, 0, 1))
    f = {k: v.copy() for k, v in f_host.items()}
    t0 = time.perf_counter()
    its = []
    for _ in range(steps):
        its.append([s[0] for s in O.rk3_step(g, f, coef, DT, solver=2, omega=OMEGA, tol=TOL, max_iter=200000,
                                              check_every=20, flowrate=True, ub_target=1.0, height=2.0)])
    el = time.perf_counter() - t0
    entry = {"value": round(NX * NY * NZ * steps / el / 1e6, 3), "unit": "Mcell-steps/s", "cores": cores,
             "kind": "port", "sample": f"{steps} RK3 steps of the same 192x128x96 wavy workload from the "
             f"post-warm-up state, FP64 SOR to 1e-8 (iterations per stage {its})",
             "seconds": round(el, 3)}
    return entry, f


def field_errors(out, ref):
    """rel-L2 of u, v, w and of p with its mean removed (the Neumann solve fixes p up to a constant)
    over the interior cells."""
    import numpy as np
    sl = (slice(1, -1),) * 3
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))   # noqa: E731
    e = {k[0]: rel(out[k][sl], ref[k][sl]) for k in ("un", "vn", "wn")}
    p, pr = out["pn"][sl], ref["pn"][sl]
    e["p"] = rel(p - p.mean(), pr - pr.mean())
    return e


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="oracle steps timed as the CPU baseline; the GPU paths run the same steps from the "
                         "same state for the accuracy check")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--check-every", type=int, default=20)
    # FP32 iterations per defect-correction batch: the calibrated network's starts reach the
    # tolerance in about 25-32 iterations, so batches of 32 end most of them with one check
    # (20: two batches and two checks); measured on the bench channel and a held-out one
    ap.add_argument("--inner", type=int, default=32)
    ap.add_argument("--train-steps", type=int, default=200)
    ap.add_argument("--no-learned", action="store_true", help="headline without the learned guess")
    ap.add_argument("--net-weights", default=WEIGHTS if os.path.exists(WEIGHTS) else "",
                    help="U-Net parameters written by mobydiff_amd.train_guess (.npz); empty: train on the "
                         "warm-up steps' pairs instead")
    ap.add_argument("--stride", type=int, default=4,
                    help="stride of a network trained on the warm-up steps (network grid = grid / stride)")
    ap.add_argument("--flat-steps", type=int, default=10, help="steps of the flat FFT production line (0: skip)")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device once (the known-size kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import mobydiff_amd as M

    if args.calibration_copy_mib > 0:
        import torch
        n = args.calibration_copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    geo = M.wavy_geometry(A=A)
    head = make_channel(M, M.PREC_MIXED, args.check_every, args.inner)
    coef = head.compute_ibm(geo)
    head.init_fields(seed=0, amp=AMP, geo=geo, dt=DT)
    learned = not args.no_learned
    training = None
    net = None
    pretrained = learned and bool(args.net_weights)
    if pretrained:
        # parameters trained beforehand on other wavy channels (mobydiff_amd.train_guess)
        from mobydiff_amd import train_guess
        net = M.Net(NX, NY, NZ, seed=0, stride=train_guess.stride_of(args.net_weights))
        meta = train_guess.load(args.net_weights, net)
        training = dict(meta, weights=os.path.basename(args.net_weights),
                        source="(div u*, previous-step Phi) -> Phi pairs from the mixed SOR solves of "
                               f"{meta['runs']} wavy channels (A uniform in {meta['A']}, seeds "
                               f"{meta['seeds'][0]}..{meta['seeds'][1]}), trained on the GPU by this library")
        head.attach_net(net, capture=False)
    elif learned:
        # the network learns from the solver's own converged solves during the warm-up steps
        net = M.Net(NX, NY, NZ, seed=0, stride=args.stride)
        net.replay_reserve(3 * args.warmup)
        head.attach_net(net, capture=True)
    for _ in range(args.warmup):
        head.step(DT)
    state = head.download()
    if pretrained:
        head.set_guess(M.GUESS_LEARNED)
    elif learned:
        t0 = time.perf_counter()
        pairs = net.replay_count()
        l0 = net.eval_replay(pairs)
        net.train(args.train_steps, batch=4, lr=1e-3)
        l1 = net.eval_replay(pairs)
        head.synchronize()
        training = {"pairs": pairs, "adam_steps": args.train_steps, "batch": 4, "lr": 1e-3,
                    "mse_before": l0, "mse_after": l1, "seconds": round(time.perf_counter() - t0, 2),
                    "source": f"(div u*, previous-step Phi) -> Phi pairs captured from the {args.warmup} "
                              "warm-up steps of this workload"}
        head.attach_net(net, capture=False)
        head.set_guess(M.GUESS_LEARNED)

    def fresh(precision, guess=M.GUESS_PREVIOUS, prime=False, lds_fp64=False, with_net=False, tol=TOL):
        # MOBYDIFF_LDS_SOR is read when a handle is created: 2 puts the FP64 batches on the
        # LDS-resident kernel too
        saved = os.environ.get("MOBYDIFF_LDS_SOR")
        if lds_fp64:
            os.environ["MOBYDIFF_LDS_SOR"] = "2"
        try:
            ch = make_channel(M, precision, args.check_every, args.inner, tol)
        finally:
            if lds_fp64:
                if saved is None:
                    del os.environ["MOBYDIFF_LDS_SOR"]
                else:
                    os.environ["MOBYDIFF_LDS_SOR"] = saved
        ch.set_ibm(*coef)
        ch.upload(**state)
        if with_net:
            ch.attach_net(net, capture=False)
        ch.set_guess(guess)
        if prime:
            ch.step(DT)   # fills the per-stage history the guess uses
        return ch

    el_h, its_h, div_h, t_h, fb_h = run_steps(head, args.steps)
    started_h = run_steps.started
    plain = fresh(M.PREC_FP64)
    el_p, its_p, div_p, t_p, _ = run_steps(plain, args.steps)
    info_p = plain.info()
    plain.close()
    plain_lds = fresh(M.PREC_FP64, lds_fp64=True)
    el_pl, its_pl, div_pl, t_pl, _ = run_steps(plain_lds, args.steps)
    info_pl = plain_lds.info()
    plain_lds.close()
    mixed = fresh(M.PREC_MIXED)
    el_m, its_m, div_m, t_m, _ = run_steps(mixed, args.steps)
    mixed.close()
    prevphi = fresh(M.PREC_MIXED, M.GUESS_PHI_PREV, prime=True)
    el_q, its_q, div_q, t_q, fb_q = run_steps(prevphi, args.steps)
    prevphi.close()
    network = None
    if learned:
        # MFMA utilisation of the network: device time of one forward pass (HIP events on its stream)
        ms_f, fl = net.time_forward(10)
        ach = fl / (ms_f * 1e-3) / 1e12
        network = {"kernel": "U-Net forward pass (BF16 MFMA, FP32 accumulate), one 192x128x96 input",
                   "bound": "mfma", "forward_ms": round(ms_f, 3), "flops": fl, "achieved": round(ach, 1),
                   "peak": 2500.0, "unit": "TFLOP/s", "frac": round(ach / 2500.0, 4)}
    cells = NX * NY * NZ
    rate = lambda el: cells * args.steps / el / 1e6   # noqa: E731
    per_stage = lambda its: sum(map(sum, its)) / (3.0 * args.steps)   # noqa: E731

    def split(t, k):
        # stencil / Poisson / projection come from a sample of timed stages (t.stages of them):
        # three stages make a step; the network's time covers every start
        st = 3.0 / max(t.stages, 1)
        return {"stencil": round(t.stencil_ms * st, 3), "poisson": round(t.poisson_ms * st, 3),
                "projection": round(t.projection_ms * st, 3), "guess": round(t.guess_ms / k, 3)}

    def side(el, its, div, t):
        return {"value": round(rate(el), 3), "ms_per_step": round(el / args.steps * 1e3, 3),
                "phase_ms_per_step": split(t, args.steps), "sor_iterations_per_stage": per_stage(its),
                "div_inf_final": div}

    v_h, v_p, v_pl = rate(el_h), rate(el_p), rate(el_pl)
    out = {
        "metric": METRIC, "value": round(v_h, 3), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el_h / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": "wavy-wall channel 192x128x96 (A=0.1, Re_b=2000, dt=5e-4), RK3 step with "
                               "red-black SOR to ||div||_inf<1e-8: " +
                               ("learned initial guess + " if learned else "") +
                               "FP32 sweeps in FP64 defect correction",
                   "grid": [NX, NY, NZ], "learned_guess": learned, "omega": OMEGA, "tol": TOL,
                   "check_every": args.check_every, "inner_iters": args.inner, "parallelism": "1 GPU",
                   "sor_kernel": head.info()},
        "phase_ms_per_step": split(t_h, args.steps),
        "sor_iterations_per_stage": per_stage(its_h),
        "div_inf_final": div_h,
        "fallback_rate": round(fb_h / (3.0 * args.steps), 4) if learned else None,
        "network_started_solves": network_gain(its_h, started_h, its_m) if learned else None,
        "training": training,
        "network": network,
        "plain_fp64": dict(side(el_p, its_p, div_p, t_p), kernel="streaming one-colour sweeps",
                           roofline=roofline(t_p, "fp64", info_p)),
        "plain_fp64_lds": dict(side(el_pl, its_pl, div_pl, t_pl), kernel="LDS-resident FP64 batches",
                               sor_kernel=info_pl),
        "mixed_no_guess": side(el_m, its_m, div_m, t_m),
        "mixed_previous_phi_guess": dict(side(el_q, its_q, div_q, t_q),
                                         fallback_rate=round(fb_q / (3.0 * args.steps), 4)),
        "sweeps_saved_vs_no_guess": round(1.0 - per_stage(its_h) / per_stage(its_m), 4),
        # against the faster of the two plain FP64 kernels
        "speedup_vs_plain_fp64": round(v_h / max(v_p, v_pl), 3),
        "roofline": roofline(t_h, "fp32", head.info()),
    }
    if out["roofline"] and out["roofline"].get("bound") == "lds":
        rate, per_clk = M.lds_read_rate()
        out["roofline"]["lds_read_probe"] = {"achieved": round(rate / 1e9, 1), "bytes_per_clock_per_cu": round(per_clk, 1),
                                             "source": "mdx_lds_read_rate: 16-byte LDS reads, 8 workgroups per CU"}
    out["stencil_roofline"] = momentum_in_step(t_h, cells, True, "wavy channel 192x128x96 (IBM)")
    # the same kernel launched 20 times back to back from one state (its working set then partly
    # served by the 256 MB MALL)
    out["stencil_roofline_repeated_launches"] = momentum_roofline(head, cells, True, DT,
                                                                  "wavy channel 192x128x96 (IBM)")
    if args.flat_steps > 0:
        out["flat_fft_production"] = flat_fft_production(M, args.flat_steps)
        out["flat_fft_production_mixed"] = flat_fft_production(M, args.flat_steps, M.PREC_MIXED)
    if not args.no_cpu_baseline:
        aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        cores = int(os.environ.get("OMP_NUM_THREADS", aff))   # the oracle's OpenMP thread count
        out["cpu_baseline"], ref = cpu_baseline(state, coef, cores, args.cpu_steps)
        # accuracy: each GPU path runs the oracle's steps from the same state
        paths = [("headline", M.PREC_MIXED, M.GUESS_LEARNED if learned else M.GUESS_PREVIOUS, False),
                 ("mixed_no_guess", M.PREC_MIXED, M.GUESS_PREVIOUS, False),
                 ("plain_fp64", M.PREC_FP64, M.GUESS_PREVIOUS, False),
                 ("plain_fp64_lds", M.PREC_FP64, M.GUESS_PREVIOUS, True)]
        acc = {"steps": args.cpu_steps, "reference": "the oracle (FP64 SOR to 1e-8) from the same state",
               "metric": "rel-L2 over interior cells; p with its mean removed",
               "p_vs_converged": "p against FP64 SOR to ||div||_inf < 1e-12 from the same state: two solves "
                                 "that each stop at 1e-8 differ in p by about their distance from it"}
        # the converged solve: plain FP64 to 1e-12 (the sweeps' fixed point, reached on the GPU)
        ch = fresh(M.PREC_FP64, tol=1e-12)
        for _ in range(args.cpu_steps):
            ch.step(DT)
        tight = ch.download(("un", "vn", "wn", "pn"))
        ch.close()
        acc["oracle_p_vs_converged"] = float(f"{field_errors(ref, tight)['p']:.3e}")
        for name, prec, guess, lds64 in paths:
            ch = fresh(prec, guess, lds_fp64=lds64, with_net=learned and name == "headline")
            for _ in range(args.cpu_steps):
                ch.step(DT)
            got = ch.download(("un", "vn", "wn", "pn"))
            e = field_errors(got, ref)
            e["div_inf"] = ch.div_norm()
            e["p_vs_converged"] = field_errors(got, tight)["p"]
            ch.close()
            acc[name] = {k: float(f"{v:.3e}") for k, v in e.items()}
        out["error_vs_fp64_oracle"] = acc
    else:
        out["cpu_baseline"] = None
    head.close()
    if net is not None:
        net.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
