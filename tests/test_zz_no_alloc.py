"""mdx_div_norm_noalloc: the no-allocation variant of mdx_div_norm (partial maxima in a static device
array of the library, not in a handle's buffers).  Called through the C ABI on device arrays the
test allocates with the HIP runtime the library itself uses; checked against mdx_div_norm of a
handle holding the same fields, bit for bit, on two calls in a row (the second on a smaller grid,
so the static array holds the first call's partials when it starts).  Sizes the static array cannot
hold are rejected before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

import mobydiff_amd as M

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    M.lib()   # loads the HIP runtime the library links (libamdhip64.so.7)
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _dev(hip, a):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
    assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def _case(hip, nx, ny, nz, lx, ly, lz, seed):
    rng = np.random.default_rng(seed)
    f = {k: rng.standard_normal((nz + 2, ny + 2, nx + 2)) for k in ("un", "vn", "wn")}
    ch = M.Channel(M.channel_config(nx, ny, nz, lx, ly, lz, 1000.0))
    ch.upload(**f)
    expect = ch.div_norm()
    ch.close()
    ptrs = [_dev(hip, f[k]) for k in ("un", "vn", "wn")]
    out = _dev(hip, np.full(1, -1.0))
    per = (C.c_int32 * 3)(1, 0, 1)
    print(f"mdx_div_norm_noalloc {nx}x{ny}x{nz}", flush=True)
    rc = M.lib().mdx_div_norm_noalloc(C.c_int32(nx), C.c_int32(ny), C.c_int32(nz), C.c_double(lx), C.c_double(ly),
                                      C.c_double(lz), per, *ptrs, out, None)
    got = np.zeros(1)
    assert hip.hipMemcpy(got.ctypes.data, out, 8, D2H) == 0   # the null stream: after the kernels
    for p in ptrs + [out]:
        hip.hipFree(p)
    assert rc == 0, M.lib().mdx_last_error().decode()
    return got[0], expect


def test_noalloc_div_norm_equals_handle_norm_twice():
    hip = _hip()
    g1, e1 = _case(hip, 96, 48, 40, 4.0, 2.2, 2.0, 1)
    g2, e2 = _case(hip, 32, 24, 16, 1.0, 1.0, 1.0, 2)
    assert g1 == e1 and g2 == e2, (g1, e1, g2, e2)
    assert e1 != e2


def test_noalloc_div_norm_rejects_sizes_beyond_the_static_buffer():
    per = (C.c_int32 * 3)(1, 0, 1)
    dummy = C.c_void_p(256)   # never dereferenced: the size check fails before any launch
    rc = M.lib().mdx_div_norm_noalloc(C.c_int32(4096), C.c_int32(4096), C.c_int32(64), C.c_double(1.0), C.c_double(1.0),
                                      C.c_double(1.0), per, dummy, dummy, dummy, dummy, None)
    assert rc == 1
    assert "too large for the static partial-max buffer" in M.lib().mdx_last_error().decode()
