"""Timing calibration of the CPU oracle against the reference's own parallel CPU path (TEST
INFRASTRUCTURE, run in the build container where the reference compiles).

The oracle is the timed CPU baseline of the benchmark, so it must not be a strawman: on the same
case and the same threads its time step must be no slower than the reference's red-black build
with host OpenMP (the reference modules compiled from their sources with -fopenmp by
oracle/Makefile into _ref/libref_rb_omp.so, run through the bridge's time loop, main.f90:64-97).

Case: 64^3, lx = ly = lz = 1, Re = 100, body force 1, dt = 1e-4, the reference's bottom bump
(init_ibm, second order), fixed nIter = 3 SOLA sweeps with omega 1.5, zero start; one warm-up step,
then `steps` timed steps.  Both codes start from the same fields and must end with the same
fields (bit for bit).  Each code runs in a process of its own with OMP_NUM_THREADS threads.

    python -m oracle.calibrate --threads 8
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
N, DT, NITER, OMEGA = 64, 1e-4, 3, 1.5


def _case():
    sys.path.insert(0, os.path.dirname(HERE))
    from oracle import oracle as O
    g = O.Grid(N, N, N, lx=1.0, ly=1.0, lz=1.0, re=100.0, forcing=(1.0, 0.0, 0.0), periodic=(1, 0, 1))
    coef = O.ibm_coefs(g, O.Geometry(kind=0, amp_x=5 * (1.0 / N)))
    return O, g, coef


def run_one(which, steps):
    O, g, coef = _case()
    f = O.new_fields(g)
    if which == "oracle":
        step = lambda n: O.run_steps(g, f, coef, n, DT, solver=1, niter=NITER, omega=OMEGA)   # noqa: E731
    else:
        lib = C.CDLL(os.path.join(HERE, "_ref", "libref_rb_omp.so"))
        O._ref_rb = lib   # the bridge entry points of the OpenMP build
        lib.ref_get_cfl.restype = C.c_double
        step = lambda n: O.ref_run_redblack(g, f, coef, NITER, OMEGA, n, DT)   # noqa: E731
    step(1)   # warm-up
    t0 = time.perf_counter()
    step(steps)
    el = time.perf_counter() - t0
    out = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"calibrate_{which}.npy")
    np.save(out, np.stack([f[k] for k in ("un", "vn", "wn", "pn")]))
    return {"which": which, "ms_per_step": 1e3 * el / steps, "fields": out}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--one", choices=("oracle", "reference"))
    args = ap.parse_args(argv)
    if args.one:
        print(json.dumps(run_one(args.one, args.steps)))
        return 0
    env = dict(os.environ, OMP_NUM_THREADS=str(args.threads))
    for k in ("OMP_PROC_BIND", "OMP_PLACES"):   # "cores" places serialised the reference build here
        env.pop(k, None)
    res = {}
    for which in ("reference", "oracle"):
        r = subprocess.run([sys.executable, "-m", "oracle.calibrate", "--one", which, "--steps", str(args.steps)],
                           capture_output=True, text=True, env=env, cwd=os.path.dirname(HERE), check=True)
        res[which] = json.loads(r.stdout.strip().splitlines()[-1])
    a, b = (np.load(res[w]["fields"]) for w in ("reference", "oracle"))
    out = {"case": f"{N}^3, dt {DT}, SOLA nIter {NITER}, omega {OMEGA}, bump IBM, {args.steps} steps",
           "threads": args.threads, "reference_ms_per_step": round(res["reference"]["ms_per_step"], 2),
           "oracle_ms_per_step": round(res["oracle"]["ms_per_step"], 2),
           "oracle_over_reference": round(res["oracle"]["ms_per_step"] / res["reference"]["ms_per_step"], 3),
           "fields_bit_identical": bool(np.array_equal(a, b))}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
