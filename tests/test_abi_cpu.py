"""C-ABI checks that need no GPU: the library loads, exports every entry point declared in
include/mobydiff_hip.h, and rejects invalid configurations before touching the device."""
import os
import re

import pytest

import mobydiff_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mobydiff_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|int64_t|const char \*)\s*(mdx_\w+)\(", text, re.M)))


def test_declaration_scan_sees_every_entry_point():
    text = open(os.path.join(ROOT, "include", "mobydiff_hip.h")).read()
    every = set(re.findall(r"\b(mdx_\w+)\(", text))
    assert every <= set(_declared()), sorted(every - set(_declared()))


def test_library_exports_every_declared_symbol():
    lib = M.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def _expect_error(cfg, code, text):
    h = M.C.c_void_p()
    rc = M.lib().mdx_create(M.C.byref(cfg), M.C.byref(h))
    assert rc == code
    assert text in M.lib().mdx_last_error().decode()
    assert not h.value


def test_redblack_rejects_odd_nx():
    # pressure_redblack.f90:16-18
    cfg = M.channel_config(9, 8, 8, 1, 1, 1, 100, solver=M.SOLVER_SOLA)
    _expect_error(cfg, 1, "even nx and nz")


def test_fft_requires_powers_of_two():
    cfg = M.channel_config(12, 8, 8, 1, 1, 1, 100, solver=M.SOLVER_FFT)
    _expect_error(cfg, 3, "powers of two")


def test_sor_rejects_unsupported_bc():
    cfg = M.channel_config(8, 8, 8, 1, 1, 1, 100, solver=M.SOLVER_SOR, p_neumann_y=False)
    _expect_error(cfg, 3, "Neumann p")


def test_sor_requires_periodic_xz():
    cfg = M.channel_config(8, 8, 8, 1, 1, 1, 100, solver=M.SOLVER_SOR, periodic=(0, 0, 1))
    _expect_error(cfg, 3, "periodic")


def test_bad_omega_and_sizes():
    _expect_error(M.channel_config(8, 8, 8, 1, 1, 1, 100, omega=2.5), 1, "omega")
    _expect_error(M.channel_config(1, 8, 8, 1, 1, 1, 100), 1, ">= 2")
    _expect_error(M.channel_config(8, 8, 8, 1, 1, 1, -1), 1, "positive")
    _expect_error(M.channel_config(8, 8, 8, 1, 1, 1, 100, precision=M.PREC_MIXED, inner_iters=0), 1,
                  "inner_iters")


def test_null_arguments():
    assert M.lib().mdx_create(None, None) == 1
    assert M.lib().mdx_momentum(None, M.C.c_double(0), M.C.c_double(0), M.C.c_double(0)) == 1
    assert M.lib().mdx_destroy(None) == 0


def test_guess_weights_round_trip(tmp_path):
    """mobydiff_amd.train_guess writes plain arrays (no pickled objects) and load() restores the
    parameters and c0 exactly, refusing a parameter count that does not match the network."""
    import numpy as np
    from mobydiff_amd import train_guess

    class FakeNet:
        def __init__(self, n, stride=2):
            self.nparam, self.p, self.c0, self.stride = n, None, None, stride

        def get_params(self):
            return np.arange(self.nparam, dtype=np.float32) * 0.5, 1.25

        def set_params(self, p, c0):
            self.p, self.c0 = p, c0

    src = FakeNet(10)
    path = str(tmp_path / "w.npz")
    train_guess.save(path, src, {"runs": 2, "pairs": 12})
    dst = FakeNet(10)
    meta = train_guess.load(path, dst)
    assert meta == {"runs": 2, "pairs": 12}
    assert dst.c0 == 1.25 and np.array_equal(dst.p, np.arange(10, dtype=np.float32) * 0.5)
    with np.load(path, allow_pickle=False) as z:
        assert set(z.files) == {"params", "c0", "stride", "meta"}
    assert train_guess.stride_of(path) == 2
    import pytest
    with pytest.raises(ValueError):
        train_guess.load(path, FakeNet(11))
    with pytest.raises(ValueError, match="stride"):
        train_guess.load(path, FakeNet(10, stride=1))
