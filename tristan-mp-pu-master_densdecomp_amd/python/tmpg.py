"""Python binding of libtmpg.so (the C ABI of include/tmpg.h), via ctypes.

The names follow the reference's per-step routines: advance_b_half
(advance_Bhalfstep), move (move_particles), deposit (deposit_particles),
filter (apply_filter1_opt), advance_e (advance_Efield), add_current,
step (one main-loop iteration, code/tristanmainloop.F90:107-274).

Grid arrays at this boundary use the reference layout: Fortran column-major
(mx, my, mz) with ghosts, held here as numpy float32 arrays of shape
(mz, my, mx) (same memory).  Particles are (7, n) float32 rows x, y, z, u, v,
w, ch in the reference's absolute 1-based coordinates.

There is no CPU fallback: a missing or unloadable libtmpg.so raises.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libtmpg.so")

Q_GATHER2 = 1
Q_BXY_Z = 2
COMPAT_DEFAULT = 3

EXPORTS = [
    "tmpg_create", "tmpg_destroy", "tmpg_last_error", "tmpg_sync",
    "tmpg_set_fields", "tmpg_get_fields", "tmpg_set_currents", "tmpg_get_currents",
    "tmpg_set_particles", "tmpg_add_particles", "tmpg_count_particles", "tmpg_inject_plasma_region", "tmpg_get_particles", "tmpg_sort_particles",
    "tmpg_advance_b_half", "tmpg_advance_e", "tmpg_reset_currents", "tmpg_move", "tmpg_deposit",
    "tmpg_push_deposit", "tmpg_fold_currents", "tmpg_filter", "tmpg_add_current", "tmpg_step",
    "tmpg_diag", "tmpg_gauss_residual", "tmpg_gauss_residual_noalloc", "tmpg_charge_density", "tmpg_track_continuity", "tmpg_load_uniform",
    "tmpg_last_timings", "tmpg_phase_cycles", "tmpg_input_open", "tmpg_input_close", "tmpg_input_get_int",
    "tmpg_input_get_double",
    "tmpg_model_create", "tmpg_model_destroy", "tmpg_model_last_error", "tmpg_model_num_params",
    "tmpg_model_get_params", "tmpg_model_set_params", "tmpg_model_set_scale", "tmpg_model_get_scale",
    "tmpg_model_set_base_passes", "tmpg_model_get_base_passes", "tmpg_model_set_amplitude_weight",
    "tmpg_model_train_batch", "tmpg_model_eval", "tmpg_model_add_sample", "tmpg_model_train",
    "tmpg_denoise", "tmpg_set_denoiser", "tmpg_flds_tot_write", "tmpg_write_flds_tot",
]


class Config(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("order", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("ntimes", ctypes.c_int32),
        ("compat", ctypes.c_int32),
        ("track_ids", ctypes.c_int32),
        ("c", ctypes.c_float),
        ("corr", ctypes.c_float),
        ("qm", ctypes.c_float * 2),
        ("q", ctypes.c_float * 2),
        ("mass", ctypes.c_float * 2),
        ("capacity", ctypes.c_int64 * 2),
        ("tile", ctypes.c_int32 * 3),
        ("sort_every", ctypes.c_int32),
        ("highorder", ctypes.c_int32),
        ("pusher", ctypes.c_int32),
        ("scheme", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 1),
    ]


class PlasmaRegion(ctypes.Structure):
    """tmpg_plasma_region: the arguments of inject_plasma_region
    (code/particles.F90:2549-2938) and the domain they refer to"""
    _fields_ = [
        ("dim", ctypes.c_int32), ("order", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("c", ctypes.c_float),
        ("x1", ctypes.c_float), ("x2", ctypes.c_float), ("y1", ctypes.c_float), ("y2", ctypes.c_float),
        ("z1", ctypes.c_float), ("z2", ctypes.c_float),
        ("ppc", ctypes.c_float), ("gamma_drift", ctypes.c_float),
        ("delgam_i", ctypes.c_float), ("delgam_e", ctypes.c_float), ("weight", ctypes.c_float),
        ("direction", ctypes.c_int32), ("pcosthmult", ctypes.c_int32), ("sigma", ctypes.c_float),
        ("upsamp_e", ctypes.c_int32),
    ]


class TmpgError(RuntimeError):
    pass


class FldsTotState(ctypes.Structure):
    """tmpg_flds_tot_state: what output_tot (code/output.F90:2494-3875)
    samples, on the padded reference layout"""
    _fields_ = [
        ("dim", ctypes.c_int32), ("order", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("istep", ctypes.c_int32), ("in_plane", ctypes.c_int32),
    ] + [(k, ctypes.c_void_p) for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz")] + [
        ("ni", ctypes.c_int64), ("ne", ctypes.c_int64),
        ("ions", ctypes.c_void_p), ("lecs", ctypes.c_void_p),
    ]


FLDS_TOT_DATASETS = ["dens", "densi", "v3x", "v3xi", "v3y", "v3yi", "v3z", "v3zi",
                     "ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"]


def flds_tot_write(path, dim, order, nx, ny, nz, istep, fields, currents, ions, lecs, in_plane=False):
    """write flds.tot from host arrays (no GPU): fields (6, ...) and
    currents (3, ...) on the padded grid (mz, my, mx) C order, particles
    (7, n) rows x, y, z, u, v, w, ch"""
    lib = load_library()
    mx, my, mz, _, _ = extent(dim, order, nx, ny, nz)
    arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(mz, my, mx) for a in list(fields) + list(currents)]
    prt = [np.ascontiguousarray(a, dtype=np.float32) for a in (ions, lecs)]
    st = FldsTotState(dim=dim, order=order, nx=nx, ny=ny, nz=nz, istep=istep, in_plane=int(bool(in_plane)))
    for k, a in zip(("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"), arrs):
        setattr(st, k, a.ctypes.data)
    st.ni, st.ne = prt[0].shape[1], prt[1].shape[1]
    st.ions = prt[0].ctypes.data if st.ni else None
    st.lecs = prt[1].ctypes.data if st.ne else None
    err = ctypes.create_string_buffer(256)
    rc = lib.tmpg_flds_tot_write(os.fsencode(path), ctypes.byref(st), err, ctypes.c_int(256))
    if rc != 0:
        raise TmpgError(f"tmpg_flds_tot_write: {err.value.decode()} ({rc})")


def inject_plasma_region(region, dseed):
    """(ions (7, n), electrons (7, n*upsamp_e), new seed): the reference's
    plasma loader on the host (no GPU)"""
    lib = load_library()
    seed = ctypes.c_double(dseed)
    n = ctypes.c_int64()
    rc = lib.tmpg_inject_plasma_region(ctypes.byref(region), ctypes.byref(seed), ctypes.c_int64(0),
                                       ctypes.byref(n), None, None)
    if rc not in (0, -4):
        raise TmpgError(f"tmpg_inject_plasma_region: {rc}")
    cap = max(1, n.value * max(1, region.upsamp_e))
    ions = np.zeros((7, cap), np.float32)
    lecs = np.zeros((7, cap), np.float32)
    seed = ctypes.c_double(dseed)
    rc = lib.tmpg_inject_plasma_region(ctypes.byref(region), ctypes.byref(seed), ctypes.c_int64(cap),
                                       ctypes.byref(n), _p(ions), _p(lecs))
    if rc != 0:
        raise TmpgError(f"tmpg_inject_plasma_region: {rc}")
    ne = n.value * max(1, region.upsamp_e)
    return ions[:, :n.value].copy(), lecs[:, :ne].copy(), seed.value


_lib = None


def load_library(path=LIB_PATH):
    """Load libtmpg.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise TmpgError(f"{path} is not built (run the package build first)")
        lib = ctypes.CDLL(path)
        lib.tmpg_last_error.restype = ctypes.c_char_p
        lib.tmpg_last_error.argtypes = [ctypes.c_void_p]
        lib.tmpg_model_last_error.restype = ctypes.c_char_p
        lib.tmpg_model_last_error.argtypes = [ctypes.c_void_p]
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def extent(dim, order, nx, ny, nz):
    ng = 5 if order == 1 else 7
    ngz = 5 if dim == 2 else ng
    return nx + ng, ny + ng, (nz + ngz if dim == 3 else 1), ng, ngz


class Context:
    """One simulation domain on the current GPU (tmpg_ctx)."""

    def __init__(self, dim, order, nx, ny, nz=1, ntimes=4, compat=COMPAT_DEFAULT, c=0.45, corr=1.025,
                 qm=(1.0, -1.0), q=(1.0, -1.0), mass=(1.0, 1.0), capacity=(0, 0), track_ids=False,
                 tile=(0, 0, 0), sort_every=0, highorder=0, pusher=0, scheme=0):
        self.lib = load_library()
        cfg = Config()
        cfg.dim, cfg.order, cfg.nx, cfg.ny, cfg.nz = dim, order, nx, ny, (nz if dim == 3 else 1)
        cfg.ntimes, cfg.compat, cfg.track_ids = ntimes, compat, int(bool(track_ids))
        cfg.c, cfg.corr = c, corr
        for s in range(2):
            cfg.qm[s], cfg.q[s], cfg.mass[s], cfg.capacity[s] = qm[s], q[s], mass[s], capacity[s]
        for a in range(3):
            cfg.tile[a] = tile[a]
        cfg.sort_every = sort_every
        cfg.highorder = int(bool(highorder))
        cfg.pusher = int(pusher)
        cfg.scheme = int(scheme)
        self.cfg = cfg
        self.dim, self.order = dim, order
        self.nx, self.ny, self.nz = nx, ny, (nz if dim == 3 else 1)
        self.mx, self.my, self.mz, self.ng, self.ngz = extent(dim, order, nx, ny, self.nz)
        h = ctypes.c_void_p()
        rc = self.lib.tmpg_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise TmpgError(f"tmpg_create: {rc}: {self.lib.tmpg_last_error(None).decode()}")
        self.h = h

    # -- helpers
    def _chk(self, rc, what):
        if rc != 0:
            raise TmpgError(f"{what}: {rc}: {self.lib.tmpg_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.tmpg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid_shape(self):
        return (self.mz, self.my, self.mx)

    def _grid(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.shape == self.grid_shape(), (a.shape, self.grid_shape())
        return a

    # -- state
    def set_fields(self, f6):
        arr = [self._grid(f6[q]) for q in range(6)]
        self._chk(self.lib.tmpg_set_fields(self.h, *(_p(a) for a in arr)), "set_fields")

    def get_fields(self):
        out = np.zeros((6,) + self.grid_shape(), dtype=np.float32)
        self._chk(self.lib.tmpg_get_fields(self.h, *(_p(out[q]) for q in range(6))), "get_fields")
        return out

    def set_currents(self, j3):
        arr = [self._grid(j3[q]) for q in range(3)]
        self._chk(self.lib.tmpg_set_currents(self.h, *(_p(a) for a in arr)), "set_currents")

    def get_currents(self):
        out = np.zeros((3,) + self.grid_shape(), dtype=np.float32)
        self._chk(self.lib.tmpg_get_currents(self.h, *(_p(out[q]) for q in range(3))), "get_currents")
        return out

    def set_particles(self, species, prt, ids=None):
        prt = np.ascontiguousarray(prt, dtype=np.float32)
        n = prt.shape[1]
        rows = [np.ascontiguousarray(prt[r]) for r in range(7)]
        idp = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        self._chk(self.lib.tmpg_set_particles(self.h, species, ctypes.c_int64(n), *(_p(r) for r in rows),
                                              _p(idp)), "set_particles")

    def add_particles(self, species, prt, ids=None):
        """append (7, n) particles to the species (tmpg_add_particles)"""
        prt = np.ascontiguousarray(prt, dtype=np.float32)
        n = prt.shape[1]
        rows = [np.ascontiguousarray(prt[r]) for r in range(7)]
        idp = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        self._chk(self.lib.tmpg_add_particles(self.h, species, ctypes.c_int64(n), *(_p(r) for r in rows),
                                              _p(idp)), "add_particles")

    def count(self, species):
        n = ctypes.c_int64()
        self._chk(self.lib.tmpg_count_particles(self.h, species, ctypes.byref(n)), "count_particles")
        return n.value

    def get_particles(self, species, with_ids=False):
        n = self.count(species)
        out = np.zeros((7, n), dtype=np.float32)
        ids = np.zeros(n, dtype=np.int32) if with_ids else None
        got = ctypes.c_int64()
        self._chk(self.lib.tmpg_get_particles(self.h, species, ctypes.c_int64(n), ctypes.byref(got),
                                              *(_p(out[r]) for r in range(7)), _p(ids)), "get_particles")
        return (out, ids) if with_ids else out

    def sort_particles(self):
        self._chk(self.lib.tmpg_sort_particles(self.h), "sort_particles")

    def load_uniform(self, species, ppc, drift_gb, theta, seed):
        self._chk(self.lib.tmpg_load_uniform(self.h, species, ppc, ctypes.c_float(drift_gb),
                                             ctypes.c_float(theta), ctypes.c_uint64(seed)), "load_uniform")

    # -- per-step kernels
    def advance_b_half(self):
        self._chk(self.lib.tmpg_advance_b_half(self.h), "advance_b_half")

    def advance_e(self):
        self._chk(self.lib.tmpg_advance_e(self.h), "advance_e")

    def reset_currents(self):
        self._chk(self.lib.tmpg_reset_currents(self.h), "reset_currents")

    def move(self):
        self._chk(self.lib.tmpg_move(self.h), "move")

    def deposit(self):
        self._chk(self.lib.tmpg_deposit(self.h), "deposit")

    def push_deposit(self):
        self._chk(self.lib.tmpg_push_deposit(self.h), "push_deposit")

    def fold_currents(self):
        self._chk(self.lib.tmpg_fold_currents(self.h), "fold_currents")

    def filter(self, npasses):
        self._chk(self.lib.tmpg_filter(self.h, npasses), "filter")

    def add_current(self):
        self._chk(self.lib.tmpg_add_current(self.h), "add_current")

    def step(self, nsteps=1):
        self._chk(self.lib.tmpg_step(self.h, nsteps), "step")

    def sync(self):
        self._chk(self.lib.tmpg_sync(self.h), "sync")

    # -- diagnostics
    def diag(self):
        out = np.zeros(8, dtype=np.float64)
        self._chk(self.lib.tmpg_diag(self.h, _p(out)), "diag")
        return out

    def gauss_residual(self):
        """[max |div_b E - F^n rho|, rms F^n rho, max |F^n rho|, max |div_b E|]"""
        out = np.zeros(4, dtype=np.float64)
        self._chk(self.lib.tmpg_gauss_residual(self.h, _p(out)), "gauss_residual")
        return out

    def gauss_residual_noalloc(self):
        """gauss_residual on the library's static scratch (no device allocation)"""
        out = np.zeros(4, dtype=np.float64)
        self._chk(self.lib.tmpg_gauss_residual_noalloc(self.h, _p(out)), "gauss_residual_noalloc")
        return out

    def write_flds_tot(self, path, istep=1, in_plane=False):
        """the reference's flds.tot file of the current state"""
        self._chk(self.lib.tmpg_write_flds_tot(self.h, os.fsencode(path), ctypes.c_int(istep),
                                               ctypes.c_int(int(bool(in_plane)))), "write_flds_tot")

    def charge_density(self):
        out = np.zeros((self.nz, self.ny, self.nx), dtype=np.float64)
        self._chk(self.lib.tmpg_charge_density(self.h, _p(out)), "charge_density")
        return out

    def track_continuity(self, on=True):
        self._chk(self.lib.tmpg_track_continuity(self.h, int(bool(on))), "track_continuity")

    def last_timings(self):
        out = np.zeros(6, dtype=np.float64)
        self._chk(self.lib.tmpg_last_timings(self.h, _p(out)), "last_timings")
        return out


def phase_cycles(reset=True):
    """Per-phase shader-cycle sums and stamp counts of the 3D cell-sum push
    (tmpg_phase_cycles); raises unless the loaded library is the
    phase-stamped build (lib/libtmpg_phases.so)."""
    lib = load_library()
    cyc = np.zeros(8, dtype=np.uint64)
    cnt = np.zeros(8, dtype=np.uint64)
    rc = lib.tmpg_phase_cycles(_p(cyc), _p(cnt), int(reset))
    if rc != 0:
        raise TmpgError(f"tmpg_phase_cycles failed ({rc}): the library has no phase stamps" if rc == -5
                        else f"tmpg_phase_cycles failed ({rc})")
    return cyc, cnt


class Model:
    """The learned current denoiser (tmpg_model): a periodic CNN whose output
    A gives the correction J += s * curl_b(A).  Current arrays here are the
    periodic interior, shape (batch, 3, nz, ny, nx) or (3, nz, ny, nx)."""

    def __init__(self, dim, seed=0):
        self.lib = load_library()
        self.dim = dim
        h = ctypes.c_void_p()
        rc = self.lib.tmpg_model_create(dim, ctypes.c_uint64(seed), ctypes.byref(h))
        if rc != 0:
            raise TmpgError(f"tmpg_model_create: {rc}")
        self.h = h

    def _chk(self, rc, what):
        if rc != 0:
            raise TmpgError(f"{what}: {rc}: {self.lib.tmpg_model_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.tmpg_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nparams(self):
        n = ctypes.c_int64()
        self._chk(self.lib.tmpg_model_num_params(self.h, ctypes.byref(n)), "num_params")
        return n.value

    def get_params(self):
        out = np.zeros(self.nparams(), dtype=np.float32)
        self._chk(self.lib.tmpg_model_get_params(self.h, _p(out)), "get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams()
        self._chk(self.lib.tmpg_model_set_params(self.h, _p(p)), "set_params")

    def set_scale(self, s):
        self._chk(self.lib.tmpg_model_set_scale(self.h, ctypes.c_float(s)), "set_scale")

    def scale(self):
        s = ctypes.c_float()
        self._chk(self.lib.tmpg_model_get_scale(self.h, ctypes.byref(s)), "get_scale")
        return s.value

    def set_amplitude_weight(self, lam):
        self._chk(self.lib.tmpg_model_set_amplitude_weight(self.h, ctypes.c_float(lam)), "set_amplitude_weight")

    def set_base_passes(self, n):
        self._chk(self.lib.tmpg_model_set_base_passes(self.h, int(n)), "set_base_passes")

    def base_passes(self):
        n = ctypes.c_int()
        self._chk(self.lib.tmpg_model_get_base_passes(self.h, ctypes.byref(n)), "get_base_passes")
        return n.value

    @staticmethod
    def _batch(j):
        j = np.ascontiguousarray(j, dtype=np.float32)
        if j.ndim == 4:
            j = j[None]
        return j

    def train_batch(self, jlo, jt, lr=0.0, update=False, want_grad=True):
        a, b = self._batch(jlo), self._batch(jt)
        bsz, _, nz, ny, nx = a.shape
        loss = ctypes.c_double()
        g = np.zeros(self.nparams(), dtype=np.float32) if want_grad else None
        self._chk(self.lib.tmpg_model_train_batch(self.h, nx, ny, nz, bsz, _p(a), _p(b), ctypes.c_float(lr),
                                                  int(bool(update)), ctypes.byref(loss), _p(g)), "train_batch")
        return loss.value, g

    def eval(self, jlo, jt, want_j=False):
        a, b = self._batch(jlo), self._batch(jt)
        bsz, _, nz, ny, nx = a.shape
        loss = ctypes.c_double()
        out = np.zeros_like(a) if want_j else None
        self._chk(self.lib.tmpg_model_eval(self.h, nx, ny, nz, bsz, _p(a), _p(b), ctypes.byref(loss), _p(out)),
                  "eval")
        return (loss.value, out) if want_j else loss.value

    def add_sample(self, jlo, jt):
        a = np.ascontiguousarray(jlo, dtype=np.float32)
        b = np.ascontiguousarray(jt, dtype=np.float32)
        _, nz, ny, nx = a.shape
        self._chk(self.lib.tmpg_model_add_sample(self.h, nx, ny, nz, _p(a), _p(b)), "add_sample")

    def train(self, iters, lr=1e-3, seed=0):
        losses = np.zeros(iters, dtype=np.float64)
        self._chk(self.lib.tmpg_model_train(self.h, iters, ctypes.c_float(lr), ctypes.c_uint64(seed), _p(losses)),
                  "train")
        return losses


def _ctx_denoise(self, model):
    self._chk(self.lib.tmpg_denoise(self.h, model.h), "denoise")


def _ctx_set_denoiser(self, model):
    self._chk(self.lib.tmpg_set_denoiser(self.h, None if model is None else model.h), "set_denoiser")
    self._denoiser = model   # keep the model alive while attached


Context.denoise = _ctx_denoise
Context.set_denoiser = _ctx_set_denoiser
