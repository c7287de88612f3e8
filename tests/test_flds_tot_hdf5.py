"""The flds.tot.NNN writer (output_tot, code/output.F90:2494-3875) on the
host: files written through tmpg_flds_tot_write are read back through the
HDF5 C library and compared with

* closed forms: uniform fields give their value at every node; one particle
  gives ch / 25 (2D) or ch / 125 (3D) in its box of +-2 cells and its own
  3-velocity u / gamma there, 0 outside; a field linear along its staggered
  axis gives the line's value at the node; the current datasets are the
  current arrays at the nodes;
* the restatement oracle/flds_tot_ref.py, written the reference's way
  (per-particle boxes, the full interpolfld expression), on random states
  near the ghost layers, in 2D (both dataset lists) and 3D, orders 1 and 3,
  istep 1 and 2: E, B, J bit-equal; the moments to 1e-6 relative (FP64 sums
  in another order).
The reference ships no flds.tot file, so its own output is not compared
(parity of the file format and sampling rests on the restatement)."""
import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tristan-mp-pu-master_densdecomp_amd", "python"))
import flds_tot_ref as ref  # noqa: E402
import tmpg  # noqa: E402

H5LIB = "/opt/conda/lib/libhdf5.so"

pytestmark = pytest.mark.skipif(not os.path.exists(H5LIB), reason="no HDF5 C library")


class H5Reader:
    """the datasets of one file through the HDF5 1.10 C API"""

    def __init__(self):
        h = ctypes.CDLL(H5LIB)
        h.H5open()
        h.H5Fopen.restype = ctypes.c_int64
        h.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, ctypes.c_int64]
        h.H5Dopen2.restype = ctypes.c_int64
        h.H5Dopen2.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]
        h.H5Dget_space.restype = ctypes.c_int64
        h.H5Dget_space.argtypes = [ctypes.c_int64]
        h.H5Dget_type.restype = ctypes.c_int64
        h.H5Dget_type.argtypes = [ctypes.c_int64]
        h.H5Tequal.argtypes = [ctypes.c_int64, ctypes.c_int64]
        h.H5Sget_simple_extent_ndims.argtypes = [ctypes.c_int64]
        h.H5Sget_simple_extent_dims.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        h.H5Dread.argtypes = [ctypes.c_int64] * 5 + [ctypes.c_void_p]
        for f in ("H5Dclose", "H5Sclose", "H5Fclose", "H5Tclose"):
            getattr(h, f).argtypes = [ctypes.c_int64]
        h.H5Lexists.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]

        class GInfo(ctypes.Structure):
            _fields_ = [("storage_type", ctypes.c_int), ("nlinks", ctypes.c_uint64),
                        ("max_corder", ctypes.c_int64), ("mounted", ctypes.c_bool)]
        self.GInfo = GInfo
        h.H5Gget_info.argtypes = [ctypes.c_int64, ctypes.POINTER(GInfo)]
        self.h = h
        self.native_float = ctypes.c_int64.in_dll(h, "H5T_NATIVE_FLOAT_g").value

    def read(self, path):
        h = self.h
        f = h.H5Fopen(os.fsencode(path), 0, 0)   # H5F_ACC_RDONLY, H5P_DEFAULT
        assert f >= 0
        info = self.GInfo()
        assert h.H5Gget_info(f, ctypes.byref(info)) >= 0
        out = {}
        for name in ref.DATASETS_3D:
            if h.H5Lexists(f, name.encode(), 0) <= 0:
                continue
            d = h.H5Dopen2(f, name.encode(), 0)
            s = h.H5Dget_space(d)
            t = h.H5Dget_type(d)
            assert h.H5Tequal(t, self.native_float) > 0, f"{name} is not FP32"
            assert h.H5Sget_simple_extent_ndims(s) == 3
            dims = (ctypes.c_uint64 * 3)()
            h.H5Sget_simple_extent_dims(s, dims, None)
            a = np.zeros(tuple(dims), np.float32)
            assert h.H5Dread(d, self.native_float, 0, 0, 0, a.ctypes.data) >= 0
            out[name] = a
            h.H5Tclose(t)
            h.H5Sclose(s)
            h.H5Dclose(d)
        h.H5Fclose(f)
        assert info.nlinks == len(out)
        return out


@pytest.fixture(scope="module")
def h5():
    return H5Reader()


def _state(dim, order, nx, ny, nz, npart, seed):
    mx, my, mz, ng, ngz = tmpg.extent(dim, order, nx, ny, nz)
    rng = np.random.default_rng(seed)
    fields = rng.standard_normal((6, mz, my, mx)).astype(np.float32)
    cur = rng.standard_normal((3, mz, my, mx)).astype(np.float32)
    h, hz = ng // 2, ngz // 2

    def prt(n):
        p = np.zeros((7, n), np.float32)
        # half of them within two cells of the ghost layers
        edge = rng.random(n) < 0.5
        x = rng.uniform(h + 1, h + 1 + nx, n)
        x[edge] = np.where(rng.random(edge.sum()) < 0.5, rng.uniform(h + 1, h + 3, edge.sum()),
                           rng.uniform(h + nx - 1, h + 1 + nx - 1e-3, edge.sum()))
        p[0] = x
        p[1] = rng.uniform(h + 1, h + 1 + ny, n)
        p[2] = rng.uniform(hz + 1, hz + 1 + nz, n) if dim == 3 else hz + 1 + rng.random(n)
        p[3:6] = rng.standard_normal((3, n)) * 0.5
        p[6] = rng.uniform(0.5, 1.5, n)
        return p

    return fields, cur, prt(npart), prt(npart + 7)


@pytest.mark.parametrize("dim,order,shape,istep,in_plane", [
    (2, 1, (14, 11, 1), 1, True),
    (2, 3, (16, 12, 1), 2, False),
    (3, 1, (8, 7, 6), 1, False),
    (3, 3, (9, 8, 6), 2, False),
])
def test_flds_tot_matches_restatement(tmp_path, h5, dim, order, shape, istep, in_plane):
    nx, ny, nz = shape
    fields, cur, ions, lecs = _state(dim, order, nx, ny, nz, 60 if dim == 3 else 150, seed=dim * 10 + order)
    path = tmp_path / "flds.tot.001"
    tmpg.flds_tot_write(str(path), dim, order, nx, ny, nz, istep, fields, cur, ions, lecs, in_plane=in_plane)
    got = h5.read(path)
    want = ref.flds_tot(dim, order, nx, ny, nz, istep, fields, cur, ions, lecs, in_plane=in_plane)
    names = ref.DATASETS_2D_IN_PLANE if (dim == 2 and in_plane) else ref.DATASETS_3D
    assert list(got) == names
    mx, my, mz, _, _ = tmpg.extent(dim, order, nx, ny, nz)
    for name in names:
        assert got[name].shape == ((mz // istep if dim == 3 else 1), my // istep, mx // istep), name
        if name[0] in "ebj":
            np.testing.assert_array_equal(got[name], want[name], err_msg=name)
        else:
            np.testing.assert_allclose(got[name], want[name], rtol=1e-6, atol=1e-6, err_msg=name)
    assert np.abs(got["dens"]).max() > 0 and np.abs(got["v3yi"]).max() > 0


def _one_particle(x, y, z, u, v, w, ch):
    return np.array([[x], [y], [z], [u], [v], [w], [ch]], np.float32), np.zeros((7, 0), np.float32)


@pytest.mark.parametrize("dim", [2, 3])
def test_flds_tot_closed_forms(tmp_path, h5, dim):
    order, nx, ny, nz = 2, 12, 10, (8 if dim == 3 else 1)
    mx, my, mz, ng, ngz = tmpg.extent(dim, order, nx, ny, nz)
    vals = np.array([0.5, -1.25, 2.0, 3.0, -0.75, 1.5], np.float32)
    fields = np.broadcast_to(vals[:, None, None, None], (6, mz, my, mx)).copy()
    cur = np.random.default_rng(0).standard_normal((3, mz, my, mx)).astype(np.float32)
    # one ion well inside (its box stays off the ghost layers)
    i0, j0, k0 = 8, 7, (7 if dim == 3 else 1)
    u, v, w, ch = 0.6, -0.3, 0.2, 1.5
    ions, lecs = _one_particle(i0 + 0.3, j0 + 0.8, k0 + 0.5, u, v, w, ch)
    path = tmp_path / "flds.tot.002"
    tmpg.flds_tot_write(str(path), dim, order, nx, ny, nz, 1, fields, cur, ions, lecs)
    got = h5.read(path)
    for q, name in enumerate(["ex", "ey", "ez", "bx", "by", "bz"]):
        assert np.all(got[name] == vals[q]), name
    for q, name in enumerate(["jx", "jy", "jz"]):
        np.testing.assert_array_equal(got[name], cur[q])
    box = np.zeros((mz, my, mx), bool)
    kz = slice(k0 - 3, k0 + 2) if dim == 3 else slice(0, 1)
    box[kz, j0 - 3:j0 + 2, i0 - 3:i0 + 2] = True
    vol = 125.0 if dim == 3 else 25.0
    for name in ("dens", "densi"):
        np.testing.assert_allclose(got[name][box], np.float32(ch / vol), rtol=1e-7)
        assert np.all(got[name][~box] == 0.0)
    gam = np.sqrt(1.0 + u * u + v * v + w * w)
    for name, comp in (("v3x", u), ("v3xi", u), ("v3y", v), ("v3yi", v), ("v3z", w), ("v3zi", w)):
        np.testing.assert_allclose(got[name][box], comp / gam, rtol=1e-6)
        assert np.all(got[name][~box] == 0.0)
    # electrons count in dens and v3x, not in densi and v3xi
    lecs = np.array([[i0 + 0.5], [j0 + 0.5], [k0 + 0.5], [-u], [0.0], [0.0], [ch]], np.float32)
    tmpg.flds_tot_write(str(path), dim, order, nx, ny, nz, 1, fields, cur, ions, lecs)
    got = h5.read(path)
    np.testing.assert_allclose(got["dens"][box], 2 * ch / vol, rtol=1e-6)
    np.testing.assert_allclose(got["densi"][box], ch / vol, rtol=1e-6)
    ge = np.sqrt(1.0 + u * u)
    np.testing.assert_allclose(got["v3x"][box], (u / gam * ch - u / ge * ch) / (2 * ch), rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(got["v3xi"][box], u / gam, rtol=1e-6)


def test_flds_tot_linear_field_at_nodes(tmp_path, h5):
    """E_x staggered at i + 1/2: a field linear in x, ex(i) = a (i + 1/2),
    gives a i at node i (the average of its two neighbours), except at the
    clamped nodes 1 and mx, which read nodes 2 and mx - 1"""
    dim, order, nx, ny, nz = 3, 1, 9, 6, 6
    mx, my, mz, ng, ngz = tmpg.extent(dim, order, nx, ny, nz)
    i = np.arange(1, mx + 1, dtype=np.float32)
    fields = np.zeros((6, mz, my, mx), np.float32)
    fields[0] = 0.25 * (i + 0.5)
    cur = np.zeros((3, mz, my, mx), np.float32)
    empty = np.zeros((7, 0), np.float32)
    path = tmp_path / "flds.tot.003"
    tmpg.flds_tot_write(str(path), dim, order, nx, ny, nz, 1, fields, cur, empty, empty)
    got = h5.read(path)["ex"]
    want = 0.25 * np.clip(i, 2, mx - 1)
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), rtol=0, atol=1e-6)
    assert np.all(h5.read(path)["dens"] == 0.0) and np.all(h5.read(path)["v3x"] == 0.0)


def test_flds_tot_rejects_bad_input(tmp_path):
    dim, order, nx, ny, nz = 2, 1, 8, 8, 1
    mx, my, mz, _, _ = tmpg.extent(dim, order, nx, ny, nz)
    f = np.zeros((6, mz, my, mx), np.float32)
    c = np.zeros((3, mz, my, mx), np.float32)
    outside = np.array([[mx + 2.0], [5.0], [3.5], [0], [0], [0], [1]], np.float32)
    with pytest.raises(tmpg.TmpgError, match="outside"):
        tmpg.flds_tot_write(str(tmp_path / "x"), dim, order, nx, ny, nz, 1, f, c, outside,
                            np.zeros((7, 0), np.float32))
    with pytest.raises(tmpg.TmpgError, match="istep"):
        tmpg.flds_tot_write(str(tmp_path / "x"), dim, order, nx, ny, nz, mx + 1, f, c,
                            np.zeros((7, 0), np.float32), np.zeros((7, 0), np.float32))
    with pytest.raises(tmpg.TmpgError, match="H5Fcreate"):
        tmpg.flds_tot_write(str(tmp_path / "no" / "such" / "dir"), dim, order, nx, ny, nz, 1, f, c,
                            np.zeros((7, 0), np.float32), np.zeros((7, 0), np.float32))
