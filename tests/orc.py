"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use this module, as the checker or the CPU baseline.

Arrays follow the reference layout: fields are Fortran column-major
(mx, my, mz) arrays including ghosts (code/fields.F90:330-351), here held as
numpy arrays of shape (ncomp, mz, my, mx) in C order, which is the same
memory; particles are 7 rows (x, y, z, u, v, w, ch) of absolute 1-based
coordinates (code/particles.F90:51-55).
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

Q_GATHER2 = 1
Q_BXY_Z = 2
Q_CARRY = 4
COMPAT_ALL = 7


class Cfg(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("order", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("ntimes", ctypes.c_int),
        ("compat", ctypes.c_int),
        ("c", ctypes.c_double),
        ("corr", ctypes.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"])
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _scal(dtype, v):
    return ctypes.c_float(v) if np.dtype(dtype) == np.float32 else ctypes.c_double(v)


def make_cfg(dim, order, nx, ny, nz=1, ntimes=4, compat=COMPAT_ALL, c=0.45, corr=1.025):
    return Cfg(dim, order, nx, ny, nz if dim == 3 else 1, ntimes, compat, c, corr)


def extent(cfg):
    mx, my, mz, ng, ngz = (ctypes.c_int() for _ in range(5))
    lib().orc_extent(ctypes.byref(cfg), *(ctypes.byref(v) for v in (mx, my, mz, ng, ngz)))
    return mx.value, my.value, mz.value, ng.value, ngz.value


def interior_slices(cfg):
    """numpy slices (z, y, x) of the interior cells of a padded array."""
    mx, my, mz, ng, ngz = extent(cfg)
    h, hz = ng // 2, ngz // 2
    if cfg.dim == 3:
        return (slice(hz, hz + cfg.nz), slice(h, h + cfg.ny), slice(h, h + cfg.nx))
    return (slice(0, 1), slice(h, h + cfg.ny), slice(h, h + cfg.nx))


def zeros_grid(cfg, ncomp, dtype):
    mx, my, mz, _, _ = extent(cfg)
    return np.zeros((ncomp, mz, my, mx), dtype=dtype)


def normalize(dtype, c, c_omp, gamma0, ppc0, me, mi):
    """{qe, qi, me, mi, qme, qmi} as the reference computes them in `dtype`."""
    out = np.zeros(6, dtype=dtype)
    fn = getattr(lib(), "orc_normalize_" + _sfx(dtype))
    fn(*(_scal(dtype, v) for v in (c, c_omp, gamma0, ppc0, me, mi)), _ptr(out))
    return dict(zip(("qe", "qi", "me", "mi", "qme", "qmi"), (out[i] for i in range(6))))


def ghost_fill(cfg, a3):
    assert a3.flags.c_contiguous and a3.shape[0] == 3
    getattr(lib(), "orc_ghost_fill_" + _sfx(a3.dtype))(ctypes.byref(cfg), _ptr(a3))


def bhalf(cfg, f6):
    getattr(lib(), "orc_bhalf_" + _sfx(f6.dtype))(ctypes.byref(cfg), _ptr(f6))


def efull(cfg, f6):
    getattr(lib(), "orc_efull_" + _sfx(f6.dtype))(ctypes.byref(cfg), _ptr(f6))


def fold(cfg, j3):
    getattr(lib(), "orc_fold_" + _sfx(j3.dtype))(ctypes.byref(cfg), _ptr(j3))


def filt(cfg, j3, ntimes):
    getattr(lib(), "orc_filter_" + _sfx(j3.dtype))(ctypes.byref(cfg), _ptr(j3), ctypes.c_int(ntimes))


def mover(cfg, f6, prt, qm):
    assert prt.flags.c_contiguous and prt.shape[0] == 7
    getattr(lib(), "orc_mover_" + _sfx(prt.dtype))(
        ctypes.byref(cfg), _ptr(f6), _ptr(prt), ctypes.c_int64(prt.shape[1]), _scal(prt.dtype, qm))


def deposit(cfg, j3, prt, q):
    getattr(lib(), "orc_deposit_" + _sfx(prt.dtype))(
        ctypes.byref(cfg), _ptr(j3), _ptr(prt), ctypes.c_int64(prt.shape[1]), _scal(prt.dtype, q))


def rho(cfg, prt, q):
    """charge density (float64) on the interior grid, shape (nz, ny, nx)."""
    out = np.zeros((cfg.nz, cfg.ny, cfg.nx), dtype=np.float64)
    getattr(lib(), "orc_rho_" + _sfx(prt.dtype))(
        ctypes.byref(cfg), _ptr(prt), ctypes.c_int64(prt.shape[1]), _scal(prt.dtype, q), _ptr(out))
    return out


def step(cfg, f6, j3, pi, pe, norm, nsteps=1):
    dt = f6.dtype
    getattr(lib(), "orc_step_" + _sfx(dt))(
        ctypes.byref(cfg), _ptr(f6), _ptr(j3), _ptr(pi), ctypes.c_int64(pi.shape[1]),
        _ptr(pe), ctypes.c_int64(pe.shape[1]),
        _scal(dt, norm["qmi"]), _scal(dt, norm["qme"]), _scal(dt, norm["qi"]), _scal(dt, norm["qe"]),
        ctypes.c_int(nsteps))


def fill_from_interior(cfg, interior, dtype):
    """padded (ncomp, mz, my, mx) array from interior values (ncomp, nz, ny, nx),
    ghosts filled periodically the way bc_b1/bc_e1 fill them."""
    ncomp = interior.shape[0]
    out = zeros_grid(cfg, ncomp, dtype)
    sl = interior_slices(cfg)
    for q in range(ncomp):
        out[(q,) + sl] = interior[q]
    for q0 in range(0, ncomp, 3):
        blk = np.ascontiguousarray(out[q0:q0 + 3])
        if blk.shape[0] == 3:
            ghost_fill(cfg, blk)
            out[q0:q0 + 3] = blk
    return out
