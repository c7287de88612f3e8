"""Training pairs and training for the learned current denoiser.

Pairs come from the method itself: one high-particles-per-cell run of a setup
(the plain method: deposit + binomial filter).  At a capture step the step is
issued routine by routine, so the raw high-ppc deposit and its filtered
version (the training target) are both available; a throw-away context
holding every k-th particle of the same state (charge scaled by k, so the
charge density is unchanged) deposits the same motion, which gives the
exactly paired low-ppc current.

The model corrects the current in curl form, J += curl_b(A), which leaves
the longitudinal part (div_b J) as it is.  With base_passes = 0 it acts on
the raw deposit and replaces the filter entirely; then the longitudinal part
keeps the raw deposit's shot noise.  With base_passes = n it acts on the
current after n binomial passes, so the longitudinal part is the plain
method's and the network improves the transverse part.

    python train_denoiser.py --dim 2 --order 2 --nx 256 --ny 256 --ppc-hi 64 --ppc-lo 4 \
        --steps 400 --every 20 --seeds 0 1 2 3 --iters 3000 --out model.npz

The recipe that reproduces a 32-ppc pair Weibel run from 2 ppc per species
(8 filter passes under the correction, amplitude weight 8, pairs from three
96x96 runs), compared on 1024x1024 cells:

    python train_denoiser.py --nx 96 --ny 96 --ppc-hi 32 --ppc-lo 2 --steps 600 --every 30 \
        --seeds 0 1 2 --iters 3000 --study 1024
"""
import argparse
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import tmpg  # noqa: E402


def weibel_norm(c=0.45, c_omp=5.0, gb=0.5, ppc_species=8, me=1.0, mi=1.0):
    """code/particles.F90:206-235 with ppc0 = 2 * ppc per species"""
    g0 = np.sqrt(1.0 + gb * gb)
    ppc0 = 2.0 * ppc_species
    omp = c / c_omp
    qe = -(omp * omp * g0) / ((ppc0 * 0.5) * (1.0 + me / mi))
    qi = -qe
    return dict(qe=qe, qi=qi, qme=qe / (me * abs(qi)), qmi=qi / (mi * abs(qi)), mi=mi * abs(qi), me=me * abs(qi))


def make_weibel(dim, order, nx, ny, nz, ppc, gb, theta, seed, c=0.45, c_omp=5.0, ntimes=4, q_scale=1.0,
                load=True):
    """two counter-streaming pair plasma regions (user/user_weibel.F90:252-292)"""
    nm = weibel_norm(c, c_omp, gb, ppc)
    # a subsampled run carries q_scale times the charge and the mass per
    # particle: same q/m, same charge and mass density
    ctx = tmpg.Context(dim, order, nx, ny, nz, ntimes=ntimes, c=c, qm=(nm["qmi"], nm["qme"]),
                       q=(nm["qi"] * q_scale, nm["qe"] * q_scale), mass=(nm["mi"] * q_scale, nm["me"] * q_scale))
    if load:
        half = ppc // 2
        ctx.load_uniform(0, half, -gb, theta, 4 * seed + 1)
        ctx.load_uniform(1, half, -gb, theta, 4 * seed + 2)
        ctx.load_uniform(0, half, gb, theta, 4 * seed + 3)
        ctx.load_uniform(1, half, gb, theta, 4 * seed + 4)
    return ctx, nm


def interior(ctx, a):
    h, hz = ctx.ng // 2, ctx.ngz // 2
    if ctx.dim == 3:
        return np.ascontiguousarray(a[..., hz:hz + ctx.nz, h:h + ctx.ny, h:h + ctx.nx])
    return np.ascontiguousarray(a[..., :, h:h + ctx.ny, h:h + ctx.nx])


def capture_step(hi, lo_factor, make_lo):
    """one main-loop step of `hi` issued routine by routine; returns
    (raw low-ppc current, filtered high-ppc current, raw high-ppc current)
    as interior arrays (3, nz, ny, nx)"""
    f = hi.get_fields()
    lo = make_lo()
    lo.set_fields(f)
    for s in range(2):
        p = hi.get_particles(s)
        lo.set_particles(s, p[:, ::lo_factor])
    # the same step on both: B half, push + deposit (raw current)
    for c in (hi, lo):
        c.advance_b_half()
        c.reset_currents()
        c.push_deposit()
    j_lo = interior(lo, lo.get_currents())
    j_hi_raw = interior(hi, hi.get_currents())
    hi.advance_b_half()
    hi.filter(hi.cfg.ntimes)
    j_hi = interior(hi, hi.get_currents())
    hi.advance_e()
    hi.add_current()
    lo.close()
    return j_lo, j_hi, j_hi_raw


def make_pairs(dim=2, order=2, nx=128, ny=128, nz=1, ppc_hi=64, ppc_lo=4, gb=0.5, theta=2.5e-3, steps=150,
               every=10, first=10, seed=0, c_omp=5.0, ntimes=4):
    k = ppc_hi // ppc_lo
    hi, _ = make_weibel(dim, order, nx, ny, nz, ppc_hi, gb, theta, seed, c_omp=c_omp, ntimes=ntimes)

    def make_lo():
        ctx, _ = make_weibel(dim, order, nx, ny, nz, ppc_hi, gb, theta, seed, c_omp=c_omp, ntimes=ntimes,
                             q_scale=float(k), load=False)
        return ctx

    pairs = []
    step = 0
    while step < steps:
        if step >= first and (step - first) % every == 0:
            jl, jh, _ = capture_step(hi, k, make_lo)
            pairs.append((jl, jh))
        else:
            hi.step(1)
        step += 1
    hi.close()
    return pairs


def make_pairs_onpolicy(model, dim=2, order=2, nx=96, ny=96, nz=1, ppc_hi=32, ppc_lo=4, gb=0.5, theta=2.5e-3,
                        steps=600, every=30, first=10, seed=0, c_omp=5.0, ntimes=4):
    """pairs from the states a denoised low-ppc run visits.  The low-ppc
    particles (every k-th particle of a high-ppc load, charge and mass x k)
    drive the fields with their denoised current (the filtered one when
    model is None); the high-ppc set moves in the same fields and its
    filtered current is the target.  Both contexts hold identical fields at
    every step, so each low-ppc particle follows its high-ppc copy."""
    k = ppc_hi // ppc_lo
    hi, _ = make_weibel(dim, order, nx, ny, nz, ppc_hi, gb, theta, seed, c_omp=c_omp, ntimes=ntimes)
    lo, _ = make_weibel(dim, order, nx, ny, nz, ppc_hi, gb, theta, seed, c_omp=c_omp, ntimes=ntimes,
                        q_scale=float(k), load=False)
    lo.set_fields(hi.get_fields())
    for s in range(2):
        lo.set_particles(s, hi.get_particles(s)[:, ::k])
    pairs = []
    for step in range(steps):
        cap = step >= first and (step - first) % every == 0
        for c in (hi, lo):
            c.advance_b_half()
            c.reset_currents()
            c.push_deposit()
        if cap:
            j_lo = interior(lo, lo.get_currents())
        if model is not None:
            lo.denoise(model)
        else:
            lo.filter(ntimes)
        j_used = lo.get_currents()
        if cap:
            hi.filter(ntimes)
            pairs.append((j_lo, interior(hi, hi.get_currents())))
        hi.set_currents(j_used)
        for c in (hi, lo):
            c.advance_b_half()
            c.advance_e()
            c.add_current()
    hi.close()
    lo.close()
    return pairs


def helmholtz(j, dim):
    """split a periodic Yee-grid current (3, nz, ny, nx) into its
    longitudinal part grad_f(phi), with div_b grad_f phi = div_b J, and its
    transverse part J - grad_f(phi) (div_b of it is zero).  div_b is the
    backward difference Gauss's law and the charge update use; a curl-form
    correction changes only the transverse part."""
    j = np.asarray(j, np.float64)
    nz, ny, nx = j.shape[1:]
    axes = (2, 1, 0) if dim == 3 else (2, 1)
    ks = [2 * np.pi * np.fft.fftfreq(n) for n in (nz, ny, nx)]
    kz, ky, kx = np.meshgrid(*ks, indexing="ij")
    kk = (kx, ky, kz)
    jh = [np.fft.fftn(j[q]) for q in range(3)]
    div = sum((1 - np.exp(-1j * kk[q])) * jh[q] for q in range(len(axes)))
    lap = sum(-4 * np.sin(kk[q] / 2) ** 2 for q in range(len(axes)))
    lap[0, 0, 0] = 1.0
    phi = div / lap
    phi[0, 0, 0] = 0.0
    jl = np.zeros_like(j)
    for q in range(len(axes)):
        jl[q] = np.real(np.fft.ifftn((np.exp(1j * kk[q]) - 1) * phi))
    return jl, j - jl


def filtered(jl, ntimes, dim, order):
    """the plain method's current: ntimes binomial passes of the library"""
    nz, ny, nx = jl.shape[1:]
    ctx = tmpg.Context(dim, order, nx, ny, nz, ntimes=ntimes)
    pad = np.zeros((3,) + ctx.grid_shape(), np.float32)
    h, hz = ctx.ng // 2, ctx.ngz // 2
    if dim == 3:
        pad[:, hz:hz + nz, h:h + ny, h:h + nx] = jl
    else:
        pad[:, :, h:h + ny, h:h + nx] = jl
    ctx.set_currents(pad)
    ctx.filter(ntimes)
    out = interior(ctx, ctx.get_currents())
    ctx.close()
    return out


def errors(model, pairs, ntimes, dim, order):
    """norm-wise relative errors vs the target (filtered high-ppc current),
    averaged over the pairs, for the raw low-ppc current, the filtered
    low-ppc current (the plain method) and the denoised one: 'all' over the
    whole current, 'T' over its transverse part (what a curl-form correction
    can change), 'L' over the longitudinal part"""
    out = {k: {"all": [], "T": [], "L": []} for k in ("raw", "filtered", "denoised")}
    nb = model.base_passes()
    for jl, jt in pairs:
        jf = filtered(jl, ntimes, dim, order)
        jin = filtered(jl, nb, dim, order) if nb > 0 else jl
        _, jd = model.eval(jin, jt, want_j=True)
        tl, tt = helmholtz(jt, dim)
        nt, ntt, ntl = np.linalg.norm(jt), np.linalg.norm(tt), np.linalg.norm(tl)
        for name, j in (("raw", jl), ("filtered", jf), ("denoised", jd[0])):
            l_, t_ = helmholtz(j, dim)
            out[name]["all"].append(np.linalg.norm(j - jt) / nt)
            out[name]["T"].append(np.linalg.norm(t_ - tt) / ntt)
            out[name]["L"].append(np.linalg.norm(l_ - tl) / max(ntl, 1e-30))
    return {k: {p: float(np.mean(v)) for p, v in d.items()} for k, d in out.items()}


def energy_history(ctx, nsteps, every):
    """(laps, field E energy, field B energy, kinetic energies) every `every` laps"""
    laps, hist = [], []
    for lap in range(every, nsteps + 1, every):
        ctx.step(every)
        d = ctx.diag()
        laps.append(lap)
        hist.append(d[:4].copy())
    return np.array(laps), np.array(hist)


def growth_rate(laps, eb, lo_frac=1e-3, hi_frac=0.1):
    """exponential growth rate (per lap) of the B energy: least-squares slope
    of log(E_B) over the laps where E_B is between lo_frac and hi_frac of its
    maximum, before the maximum"""
    imax = int(np.argmax(eb))
    emax = eb[imax]
    sel = [i for i in range(imax + 1) if lo_frac * emax <= eb[i] <= hi_frac * emax]
    if len(sel) < 3:
        return float("nan")
    a = np.polyfit(laps[sel], np.log(eb[sel]), 1)
    return float(a[0])


def compare_runs(model, dim=2, order=2, nx=96, ny=96, nz=1, ppc_hi=32, ppc_lo=4, gb=0.5, theta=2.5e-3,
                 steps=600, every=10, seed=11, c_omp=5.0, ntimes=4):
    """Weibel runs of one setup: high-ppc plain (the yardstick), low-ppc plain and
    low-ppc with the denoiser in place of the filter.  Returns per run the
    energy history, B-energy growth rate, saturated B energy and GPU time
    per lap (the library's own event timings)."""
    import time
    runs = {}
    for name, ppc, dn in (("hi_plain", ppc_hi, False), ("lo_plain", ppc_lo, False), ("lo_denoised", ppc_lo, True)):
        ctx, _ = make_weibel(dim, order, nx, ny, nz, ppc_hi, gb, theta, seed, c_omp=c_omp, ntimes=ntimes,
                             q_scale=ppc_hi / ppc, load=False)
        half = ppc // 2
        for k, (s, d) in enumerate(((0, -gb), (1, -gb), (0, gb), (1, gb))):
            ctx.load_uniform(s, half, d, theta, 4 * seed + k + 1)
        if dn:
            ctx.set_denoiser(model)
        ctx.last_timings()
        t0 = time.perf_counter()
        laps, hist = energy_history(ctx, steps, every)
        wall = time.perf_counter() - t0
        ms = ctx.last_timings()
        runs[name] = dict(laps=laps, hist=hist, growth=growth_rate(laps, hist[:, 1]), eb_max=float(hist[:, 1].max()),
                          ms_per_lap=float(ms[4] / max(ms[3], 1)), ms_split=(ms[:3] / max(ms[3], 1)).tolist(),
                          wall_s=wall, particles=ctx.count(0) + ctx.count(1))
        ctx.close()
    return runs


def weibel_run(dim, order, n, ppc, seed, model=None, laps=600, every=10, ppc_norm=32, gb=0.5, theta=2.5e-3,
               c_omp=5.0, ntimes=4):
    """one pair Weibel run on n^dim cells at `ppc` particles per cell per
    species (charge and mass scaled to the density of `ppc_norm`), with the
    denoiser in place of the filter when `model` is given.  Returns the
    B-energy growth rate, the peak B energy and the GPU time per lap."""
    ctx, _ = make_weibel(dim, order, n, n, n if dim == 3 else 1, ppc_norm, gb, theta, seed, c_omp=c_omp,
                         ntimes=ntimes, q_scale=ppc_norm / ppc, load=False)
    for k, (s, d) in enumerate(((0, -gb), (1, -gb), (0, gb), (1, gb))):
        ctx.load_uniform(s, ppc // 2, d, theta, 4 * seed + k + 1)
    if model is not None:
        ctx.set_denoiser(model)
    ctx.step(1)
    ctx.sync()
    ctx.last_timings()
    laps_, hist = energy_history(ctx, laps, every)
    ms = ctx.last_timings()
    ctx.close()
    eb = hist[:, 1]
    return growth_rate(laps_, eb, lo_frac=1e-2, hi_frac=0.3), float(eb.max()), float(ms[4] / max(ms[3], 1))


def parity_study(model, n=192, seeds=range(11, 19), ppc_hi=32, ppc_lo=2, laps=600, order=2, dim=2, c_omp=5.0,
                 log=None):
    """the denoised low-ppc run against the plain high- and low-ppc runs of
    the same seeds: mean growth rate, peak B energy (and its standard error)
    and GPU ms per lap of each"""
    out = {}
    for name, ppc, m in (("hi_plain", ppc_hi, None), ("lo_plain", ppc_lo, None), ("lo_denoised", ppc_lo, model)):
        rs = []
        for s in seeds:
            rs.append(weibel_run(dim, order, n, ppc, s, m, laps=laps, ppc_norm=ppc_hi, c_omp=c_omp))
            if log:
                log(f"{name} seed {s}: growth {rs[-1][0]:.4f} peak B {rs[-1][1]:.4g} {rs[-1][2]:.3f} ms/lap")
        rs = np.array(rs)
        out[name] = dict(growth=float(rs[:, 0].mean()), eb_max=float(rs[:, 1].mean()),
                         eb_max_se=float(rs[:, 1].std() / np.sqrt(len(rs))), ms_per_lap=float(rs[:, 2].mean()))
    return out


RECIPE = dict(base_passes=8, amplitude_weight=8.0, iters=3000, lr=2e-3)


def train(pairs, dim, iters=2000, lr=2e-3, seed=0, base_passes=0, order=2, amplitude_weight=0.0):
    """a model trained on (raw low-ppc, target) pairs; with base_passes > 0
    it learns the correction of the current after that many filter passes;
    amplitude_weight > 0 adds the loss term on the error along the target"""
    m = tmpg.Model(dim, seed=seed)
    m.set_base_passes(base_passes)
    m.set_amplitude_weight(amplitude_weight)
    for jl, jt in pairs:
        jin = filtered(jl, base_passes, dim, order) if base_passes > 0 else jl
        m.add_sample(jin, jt)
    losses = m.train(iters, lr=lr, seed=seed + 1)
    return m, losses


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dim", type=int, default=2)
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=256)
    ap.add_argument("--nz", type=int, default=1)
    ap.add_argument("--ppc-hi", type=int, default=64)
    ap.add_argument("--ppc-lo", type=int, default=4)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--seeds", type=int, nargs="+", default=[0, 1])
    ap.add_argument("--test-seed", type=int, default=99)
    ap.add_argument("--iters", type=int, default=2000)
    ap.add_argument("--base-passes", type=int, default=RECIPE["base_passes"])
    ap.add_argument("--amplitude-weight", type=float, default=RECIPE["amplitude_weight"])
    ap.add_argument("--study", type=int, default=0, metavar="N",
                    help="then compare the model on N x N cells (8 seeds): growth rate, peak B energy and GPU time "
                         "of the denoised low-ppc run against the plain high- and low-ppc runs")
    ap.add_argument("--out", default="")
    a = ap.parse_args()
    pairs = []
    for s in a.seeds:
        pairs += make_pairs(a.dim, a.order, a.nx, a.ny, a.nz, a.ppc_hi, a.ppc_lo, steps=a.steps, every=a.every,
                            seed=s)
    test = make_pairs(a.dim, a.order, a.nx, a.ny, a.nz, a.ppc_hi, a.ppc_lo, steps=a.steps, every=a.every,
                      seed=a.test_seed)
    m, losses = train(pairs, a.dim, a.iters, base_passes=a.base_passes, order=a.order,
                      amplitude_weight=a.amplitude_weight)
    e = errors(m, test, 4, a.dim, a.order)
    print(f"loss {losses[:50].mean():.4g} -> {losses[-50:].mean():.4g}")
    for k, d in e.items():
        print(f"held-out relative error vs filtered high-ppc current, {k:9s}: all {d['all']:.4f}  "
              f"transverse {d['T']:.4f}  longitudinal {d['L']:.4f}")
    if a.study:
        r = parity_study(m, n=a.study, ppc_hi=a.ppc_hi, ppc_lo=a.ppc_lo, order=a.order)
        h = r["hi_plain"]
        for k, d in r.items():
            print(f"{k:12s} growth {d['growth']:.4f} ({100 * (d['growth'] / h['growth'] - 1):+.1f}%)  peak B energy "
                  f"{d['eb_max']:.4g} ({100 * (d['eb_max'] / h['eb_max'] - 1):+.1f}% +- "
                  f"{100 * d['eb_max_se'] / h['eb_max']:.1f})  GPU {d['ms_per_lap']:.3f} ms/lap "
                  f"({d['ms_per_lap'] / h['ms_per_lap']:.3f} of {a.ppc_hi} ppc)")
    if a.out:
        np.savez(a.out, params=m.get_params(), scale=m.scale(), dim=a.dim, base_passes=m.base_passes())


if __name__ == "__main__":
    main()
