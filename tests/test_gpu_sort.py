"""The periodic in-tile cell sort of the particles only reorders them: a run
that sorts before every step and one that never sorts hold the same particles
(matched by id) and the same fields after several main-loop steps, up to the
order of floating-point accumulation of the current."""
import numpy as np
import pytest

from gpu_util import make_context, periodic_diff, random_state, rel, sorted_by_id

pytestmark = pytest.mark.gpu

PHYS = dict(c=0.45, corr=1.025, c_omp=5.0, gamma0=0.5, ppc0=8.0, me=1.0, mi=1.0)


def _run(tmpg, dim, order, shape, f, pi, pe, sort_every, steps):
    ctx = make_context(tmpg, dim, order, *shape, PHYS, sort_every=sort_every)
    ctx.set_fields(f)
    ctx.set_particles(0, pi)
    ctx.set_particles(1, pe)
    ctx.step(steps)
    out = [ctx.get_particles(s, with_ids=True) for s in range(2)]
    fields = ctx.get_fields()
    n = [ctx.count(s) for s in range(2)]
    ctx.close()
    return [sorted_by_id(p, i) for p, i in out], fields, n


# ppc sets the tile segment length and so the particles per sorting thread:
# 3 -> 4, 6 -> 8, 14 -> 16 (tiles of 256 cells)
@pytest.mark.parametrize("dim,order,shape,ppc,weights", [
    (3, 3, (16, 16, 16), 6, False),
    (3, 1, (24, 16, 8), 3, True),
    (2, 2, (48, 32, 1), 6, False),
    (2, 1, (32, 32, 1), 14, True),
])
def test_sorted_run_matches_unsorted(dim, order, shape, ppc, weights):
    import orc
    import tmpg
    nx, ny, nz = shape
    ncell = nx * ny * (nz if dim == 3 else 1)
    f, pi, pe = random_state(dim, order, nx, ny, nz, ppc * ncell, seed=40 + order + dim, emax=0.02, bmax=0.05,
                             umax=0.8)
    cfg = orc.make_cfg(dim, order, nx, ny, nz)
    mx, my, mz, ng, ngz = orc.extent(cfg)
    full = np.zeros((6, mz, my, mx), np.float32)
    h, hz = ng // 2, ngz // 2
    if dim == 3:
        full[:, hz + 1: hz + 1 + nz, h + 1: h + 1 + ny, h + 1: h + 1 + nx] = f
    else:
        full[:, :, h + 1: h + 1 + ny, h + 1: h + 1 + nx] = f
    if weights:
        rng = np.random.default_rng(5)
        pi[6] = rng.uniform(0.5, 1.5, pi.shape[1])
        pe[6] = rng.uniform(0.5, 1.5, pe.shape[1])
    ps, fs, ns = _run(tmpg, dim, order, shape, full, pi, pe, 1, 6)
    pu, fu, nu = _run(tmpg, dim, order, shape, full, pi, pe, -1, 6)
    assert ns == nu == [pi.shape[1], pe.shape[1]]
    n3 = (nx, ny, nz)
    for s in range(2):
        assert np.array_equal(ps[s][6], pu[s][6])
        for r in range(dim):
            assert np.abs(periodic_diff(ps[s][r], pu[s][r], 0, n3[r])).max() < 1e-4, (s, r)
        for r in range(3, 6):
            assert rel(ps[s][r], pu[s][r]) < 1e-6, (s, r)
    for q in range(6):
        assert rel(fs[q], fu[q]) < 1e-5, q


def _key(p, lo, dim, shape, tile, eps):
    """tile index (x fastest) * cells per tile + cell within the tile"""
    n = p.shape[1]
    c = [np.floor(p[a].astype(np.float64) - eps).astype(np.int64) - lo[a] for a in range(dim)]
    c = [x % shape[a] for a, x in enumerate(c)] + ([] if dim == 3 else [np.zeros(n, np.int64)])
    t = [c[a] // tile[a] for a in range(3)]
    nt = [-(-shape[a] // tile[a]) if a < dim else 1 for a in range(3)]
    tid = t[0] + nt[0] * (t[1] + nt[1] * t[2])
    loc = (c[0] - t[0] * tile[0]) + tile[0] * ((c[1] - t[1] * tile[1]) + tile[1] * (c[2] - t[2] * tile[2]))
    return tid * (tile[0] * tile[1] * tile[2]) + loc


@pytest.mark.parametrize("dim,shape,ppc", [(3, (32, 16, 16), 3), (3, (16, 16, 8), 7), (2, (64, 48, 1), 14)])
def test_sort_orders_each_tile_by_cell(dim, shape, ppc):
    """after tmpg_sort_particles the library's order runs tile by tile
    (x fastest) and, inside each tile, by cell (x fastest); the particles
    themselves are unchanged"""
    import orc
    import tmpg
    nx, ny, nz = shape
    ncell = nx * ny * (nz if dim == 3 else 1)
    f, pi, pe = random_state(dim, 2, nx, ny, nz, ppc * ncell, seed=3 + ppc)
    tile = (8, 8, 4) if dim == 3 else (16, 16, 1)
    ctx = make_context(tmpg, dim, 2, nx, ny, nz, PHYS, sort_every=-1, tile=tile)
    ctx.set_particles(0, pi)
    ctx.set_particles(1, pe)
    before = [ctx.get_particles(s, with_ids=True) for s in range(2)]
    ctx.sort_particles()
    after = [ctx.get_particles(s, with_ids=True) for s in range(2)]
    ctx.close()
    cfg = orc.make_cfg(dim, 2, nx, ny, nz)
    _, _, _, ng, ngz = orc.extent(cfg)
    lo = (ng // 2 + 1, ng // 2 + 1, ngz // 2 + 1)
    for s in range(2):
        (p0, i0), (p1, i1) = before[s], after[s]
        assert np.array_equal(sorted_by_id(p0, i0), sorted_by_id(p1, i1))
        # a position a few ulp above an integer may be a cell-plus-offset
        # just below it, rounded up when the two were summed: such a
        # particle may sort with either cell
        k_hi, k_lo = _key(p1, lo, dim, shape, tile, 0.0), _key(p1, lo, dim, shape, tile, 4e-6)
        prev = -1
        for a, b in zip(k_lo.tolist(), k_hi.tolist()):
            if a >= prev:
                prev = a
            else:
                assert b >= prev
                prev = b
        # the order before the sort was tile by tile but not by cell
        assert np.any(np.diff(_key(p0, lo, dim, shape, tile, 0.0)) < 0)
