"""Shared helpers of the GPU parity tests: build matching GPU / oracle states."""
import ast
import os

import numpy as np

import orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def load_fixture(dim, order, suffix=""):
    fx = np.load(os.path.join(GOLD, f"ref_{dim}d_ord{order}{suffix}.npz"))
    p = ast.literal_eval(str(fx["params"]))
    return fx, p


def norm32(p):
    return orc.normalize(np.float32, p["c"], p["c_omp"], p["gamma0"], p["ppc0"], p["me"], p["mi"])


def norm64(p):
    return orc.normalize(np.float64, p["c"], p["c_omp"], p["gamma0"], p["ppc0"], p["me"], p["mi"])


def make_context(tmpg, dim, order, nx, ny, nz, p, ntimes=4, compat=None, track_ids=True, capacity=(0, 0),
                 sort_every=0, highorder=0, pusher=0, tile=(0, 0, 0), scheme=0):
    nm = norm32(p)
    kw = {}
    if compat is not None:
        kw["compat"] = compat
    return tmpg.Context(dim, order, nx, ny, nz, ntimes=ntimes, c=p["c"], corr=p["corr"],
                        qm=(float(nm["qmi"]), float(nm["qme"])), q=(float(nm["qi"]), float(nm["qe"])),
                        mass=(float(nm["mi"]), float(nm["me"])), track_ids=track_ids, capacity=capacity,
                        sort_every=sort_every, highorder=highorder, pusher=pusher, tile=tile, scheme=scheme,
                        **kw)


def sorted_by_id(prt, ids):
    out = np.empty_like(prt)
    out[:, ids] = prt
    return out


def periodic_diff(a, b, lo, n):
    """a - b on a periodic axis [lo, lo+n)."""
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return (d + 0.5 * n) % n - 0.5 * n


def random_state(dim, order, nx, ny, nz, npart, seed, emax=0.05, bmax=0.1, umax=1.0):
    """interior E,B (6, nz, ny, nx) float32 and two species (7, n) float32."""
    rng = np.random.default_rng(seed)
    cfg = orc.make_cfg(dim, order, nx, ny, nz)
    mx, my, mz, ng, ngz = orc.extent(cfg)
    h, hz = ng // 2, ngz // 2
    nzz = nz if dim == 3 else 1
    f = np.empty((6, nzz, ny, nx), dtype=np.float32)
    f[:3] = rng.uniform(-emax, emax, f[:3].shape)
    f[3:] = rng.uniform(-bmax, bmax, f[3:].shape)
    out = []
    for s in range(2):
        n = npart + 7 * s
        pos = np.empty((3, n))
        pos[0] = h + 1 + rng.random(n) * nx
        pos[1] = h + 1 + rng.random(n) * ny
        pos[2] = (hz + 1 + rng.random(n) * nz) if dim == 3 else hz + 1.5
        mom = rng.uniform(-umax, umax, (3, n))
        out.append(np.concatenate([pos, mom, np.ones((1, n))]).astype(np.float32))
    return f, out[0], out[1]
