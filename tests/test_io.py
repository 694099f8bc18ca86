"""Field files (libmobydiff_io, include/mobydiff_io.h) against the reference's own writer and reader.

The reference's io.f90 and field_hdf5.c are compiled from their sources into oracle/_ref/libref_io.so
(oracle/Makefile).  The HDF5 files this build writes must be the reference's files: the same
datasets in the same [i][j][k] (k fastest) order and the same twelve attributes (h5diff finds no
difference, attributes included), and each reader must read the other's file.  These run on the
CPU: the field files are host code.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import mobydiff_amd as M
from oracle import oracle as O

H5DIFF = shutil.which("h5diff") or "/opt/conda/bin/h5diff"

needs_ref = pytest.mark.skipif(O.ref_io() is None, reason="reference output module not built (oracle/_ref)")
needs_io = pytest.mark.skipif(not os.path.exists(M.IOLIBPATH), reason="libmobydiff_io not built")


def _fields(g, seed):
    rng = np.random.default_rng(seed)
    return {k: rng.standard_normal(g.gshape) for k in ("un", "vn", "wn", "pn")}


GRIDS = [(6, 5, 4), (9, 3, 7), (33, 17, 40)]


@needs_io
@needs_ref
@pytest.mark.parametrize("shape", GRIDS)
def test_h5_file_equals_reference_file(tmp_path, shape):
    nx, ny, nz = shape
    g = O.Grid(nx, ny, nz, lx=4.0, ly=2.2, lz=2.0, re=2000.0)
    f = _fields(g, nx)
    dt, t = 5e-4, 0.0375
    O.ref_write_field(g, f, dt, t, 7, str(tmp_path / "ref"))
    ours = tmp_path / "ours_7.h5"
    M.write_field_h5(str(ours), nx, ny, nz, 4.0, 2.2, 2.0, 2000.0, dt, t, f["un"], f["vn"], f["wn"], f["pn"])
    r = subprocess.run([H5DIFF, "-v", str(tmp_path / "ref_7.h5"), str(ours)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    # every attribute and dataset was compared, and none differs
    names = ("nx", "ny", "nz", "lx", "ly", "lz", "dx", "dy", "dz", "re", "dt", "t_current")
    for a in names:
        assert f"attribute: <{a} of </>> and <{a} of </>>" in r.stdout, (a, r.stdout)
    for d in ("un", "vn", "wn", "pn"):
        assert f"dataset: </{d}> and </{d}>" in r.stdout, (d, r.stdout)
    # the root group, its twelve attributes and the four datasets
    assert r.stdout.count("0 differences found") == 1 + len(names) + 4, r.stdout
    # each reader reads the other's file
    back = M.read_field_h5(str(tmp_path / "ref_7.h5"), nx, ny, nz)
    theirs = O.ref_read_field(g, str(ours))
    for k in f:
        assert np.array_equal(back[k], f[k]) and np.array_equal(theirs[k], f[k]), k
    assert M.read_field_info(str(ours)) == (nx, ny, nz, dt, t)


@needs_io
def test_h5_layout_is_k_fastest(tmp_path):
    """The dataset stores q(i,j,k) at ((i*(ny+2) + j)*(nz+2) + k) (field_hdf5.c:9-12): a field that
    encodes its own indices reads back, through h5dump's flat listing, in that order."""
    nx, ny, nz = 3, 2, 4
    g = O.Grid(nx, ny, nz)
    code = np.zeros(g.gshape)
    for k in range(nz + 2):
        for j in range(ny + 2):
            for i in range(nx + 2):
                code[k, j, i] = 100 * i + 10 * j + k
    z = np.zeros(g.gshape)
    path = tmp_path / "code.h5"
    M.write_field_h5(str(path), nx, ny, nz, 1.0, 1.0, 1.0, 1.0, 1e-3, 0.0, code, z, z, z)
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    r = subprocess.run([h5dump, "-d", "un", "-y", "-w", "0", "-m", "%g", str(path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    body = r.stdout.split("DATA {", 1)[1].split("}", 1)[0]
    vals = [float(x) for x in body.replace(",", " ").split()]
    expect = [100 * i + 10 * j + k for i in range(nx + 2) for j in range(ny + 2) for k in range(nz + 2)]
    assert vals == expect
    assert "DATASPACE  SIMPLE { ( 5, 4, 6 ) / ( 5, 4, 6 ) }" in r.stdout


@needs_io
def test_h5_reader_rejects_other_grid_and_missing_file(tmp_path):
    g = O.Grid(6, 5, 4)
    f = _fields(g, 1)
    path = tmp_path / "f.h5"
    M.write_field_h5(str(path), 6, 5, 4, 1.0, 1.0, 1.0, 1.0, 1e-3, 0.0, f["un"], f["vn"], f["wn"], f["pn"])
    with pytest.raises(M.MdxError, match=r"dataset un: shape \[8\]\[7\]\[6\], expected \[9\]\[7\]\[6\]"):
        M.read_field_h5(str(path), 7, 5, 4)
    with pytest.raises(M.MdxError, match="cannot open"):
        M.read_field_h5(str(tmp_path / "absent.h5"), 6, 5, 4)
    with pytest.raises(M.MdxError, match="cannot create"):
        M.write_field_h5(str(tmp_path / "no_dir" / "x.h5"), 6, 5, 4, 1.0, 1.0, 1.0, 1.0, 1e-3, 0.0,
                         f["un"], f["vn"], f["wn"], f["pn"])


def test_center_vel_restatement_against_definition():
    """The oracle's center_vel (io.f90:184-222) on a field whose face values are affine in the
    index: the mean of the two faces is the value half a cell further on."""
    g = O.Grid(5, 4, 3)
    f = {k: g.zeros() for k in ("un", "vn", "wn")}
    kk, jj, ii = np.meshgrid(np.arange(5), np.arange(6), np.arange(7), indexing="ij")
    f["un"][...] = 2.0 * ii + 0.25 * jj
    f["vn"][...] = 4.0 * jj - kk
    f["wn"][...] = 8.0 * kk + ii
    uc, vc, wc = O.center_vel(f)
    I, J, K = ii[1:-1, 1:-1, 1:-1], jj[1:-1, 1:-1, 1:-1], kk[1:-1, 1:-1, 1:-1]
    assert np.array_equal(uc, 2.0 * (I + 0.5) + 0.25 * J)
    assert np.array_equal(vc, 4.0 * (J + 0.5) - K)
    assert np.array_equal(wc, 8.0 * (K + 0.5) + I)
