"""The strided U-Net's transfer operators in the oracle (oracle/unet_ref.py), pinned by what the
mathematics fixes: the restriction of a field constant on each block is that constant; the rows
of each axis' interpolation matrix are convex weights; the prolongation reproduces constants
everywhere and fields linear in the index away from the clamped outermost cells, and a smooth
field to second order in the voxel size.  (The GPU transfer kernels are compared with these in
tests/test_gpu_unet.py.)"""
import numpy as np
import pytest

from oracle import unet_ref as R


@pytest.mark.parametrize("s", [2, 4])
def test_restriction_of_blockwise_constants(s):
    rng = np.random.default_rng(s)
    coarse = rng.standard_normal((3, 4, 5))
    fine = np.repeat(np.repeat(np.repeat(coarse, s, 0), s, 1), s, 2)
    assert np.allclose(R.restrict_mean(fine, s), coarse, rtol=0, atol=1e-15)


@pytest.mark.parametrize("s", [2, 4])
def test_prolongation_weights_and_linear_fields(s):
    P = R.prolong_matrix(8 * s, s)
    assert np.all(P >= 0) and np.allclose(P.sum(axis=1), 1.0)
    assert np.count_nonzero(P, axis=1).max() <= 2
    cz, cy, cx = 4, 5, 6
    z, y, x = np.meshgrid(np.arange(cz), np.arange(cy), np.arange(cx), indexing="ij")
    lin = 0.5 + 2.0 * x - 0.75 * y + 0.25 * z
    fine = R.prolong_trilinear(lin, s)
    assert np.allclose(R.prolong_trilinear(np.full((cz, cy, cx), 3.5), s), 3.5, atol=1e-14)
    fz, fy, fx = np.meshgrid(*(((np.arange(n * s) + 0.5) / s - 0.5) for n in (cz, cy, cx)), indexing="ij")
    exact = 0.5 + 2.0 * fx - 0.75 * fy + 0.25 * fz
    h = s // 2   # cells within half a voxel of the ends are clamped
    inner = (slice(h, cz * s - h), slice(h, cy * s - h), slice(h, cx * s - h))
    assert np.allclose(fine[inner], exact[inner], atol=1e-12)


def test_prolongation_second_order():
    errs = []
    for n in (8, 16, 32):
        c = (np.arange(n) + 0.5) / n
        coarse = np.sin(np.pi * c)[None, None, :] * np.ones((2, 2, 1))
        f = (np.arange(2 * n) + 0.5) / (2 * n)
        fine = R.prolong_trilinear(coarse, 2)[0, 0]
        inner = slice(1, 2 * n - 1)
        errs.append(np.abs(fine[inner] - np.sin(np.pi * f[inner])).max())
    assert errs[1] / errs[2] > 3.5 and errs[0] / errs[1] > 3.5, errs
