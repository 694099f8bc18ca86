"""The channel's boundary conditions written by the SOR projection kernel against the separate fills.

With x and z periodic and Dirichlet u, v, w on both y walls, the projection kernel (sor_apply)
also writes the ghost values the fills bc_rb_y, bc_rb_x and bc_rb_z then set, each from the one
interior value it depends on, and the step runs no fill after the projection;
MOBYDIFF_FUSED_BC=0 keeps the three fills.  From one state the same steps must give the same
arrays bit for bit, ghost shell included (edges and corners keep the projection's u* values either
way): on the headline workload (192x128x96 wavy channel, mixed precision, 32-iteration batches),
on a stretched grid in FP64, and with moving walls (u = 0.5 on the top wall, w = -0.25 on the
bottom one), where the y ghosts are 2 V minus the wall-adjacent value.  The projection called on
its own (mdx_pressure_projection) writes no boundary condition, as the reference's projection:
tests/test_gpu_parity.py compares its whole arrays with the oracle's.
"""
import numpy as np
import pytest

import mobydiff_amd as M

pytestmark = pytest.mark.gpu

DT = 5e-4

CASES = {
    "wavy_192x128x96_mixed": dict(shape=(192, 128, 96), A=0.1, beta=0.0, prec=M.PREC_MIXED, walls=None),
    "stretched_64x48x32_fp64": dict(shape=(64, 48, 32), A=None, beta=1.5, prec=M.PREC_FP64, walls=None),
    "moving_walls_64x48x32": dict(shape=(64, 48, 32), A=None, beta=0.0, prec=M.PREC_FP64,
                                  walls={(1, 1): 0.5, (0, 3): -0.25}),
}


def _run(case, fused, monkeypatch, steps=2):
    monkeypatch.setenv("MOBYDIFF_FUSED_BC", "1" if fused else "0")
    nx, ny, nz = case["shape"]
    cfg = M.channel_config(nx, ny, nz, 4.0, 2.2, 2.0, 2000.0, solver=M.SOLVER_SOR, omega=1.8, tol=1e-8,
                           check_every=10, precision=case["prec"], inner_iters=32, flowrate=True,
                           flow_height=2.0, stretch_beta=case["beta"])
    for (side, var), v in (case["walls"] or {}).items():
        cfg.bc_value[M.bci(2, side, var)] = v
    ch = M.Channel(cfg)
    geo = None
    if case["A"] is not None:
        geo = M.wavy_geometry(A=case["A"])
        ch.compute_ibm(geo)
    ch.init_fields(seed=4, amp=0.05, geo=geo, dt=DT)
    its = [[s.iterations for s in ch.step(DT).stage] for _ in range(steps)]
    out = ch.download(("un", "vn", "wn", "pn"))
    ch.close()
    return its, out


@pytest.mark.parametrize("name", list(CASES))
def test_projection_bc_equal_separate_fills(name, monkeypatch):
    case = CASES[name]
    print(f"boundary conditions in the projection: {name}", flush=True)
    its_f, out_f = _run(case, True, monkeypatch)
    its_s, out_s = _run(case, False, monkeypatch)
    print(f"  iterations {its_f} / {its_s}", flush=True)
    assert its_f == its_s
    for k in ("un", "vn", "wn", "pn"):
        assert np.array_equal(out_f[k], out_s[k]), k
    if case["walls"]:
        # the y ghosts of u on the top wall: 2 V - u(j = ny); of w on the bottom: 2 V - w(j = 1)
        u, w = out_f["un"], out_f["wn"]
        ny = case["shape"][1]
        assert np.allclose(u[1:-1, ny + 1, 1:-1] + u[1:-1, ny, 1:-1], 1.0, atol=1e-14, rtol=0)
        assert np.allclose(w[1:-1, 0, 1:-1] + w[1:-1, 1, 1:-1], -0.5, atol=1e-14, rtol=0)
