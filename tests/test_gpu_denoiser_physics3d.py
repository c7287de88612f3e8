"""The trained denoiser in the loop in 3D at order 3: pair-Weibel runs at 2
particles per cell per species with the denoiser against the 32-ppc runs
(the plain method, the yardstick) and the plain 2-ppc runs, on 96^3 cells
over 6 seeds (11-16) that training never saw.

The model is trained on the GPU with the library's kernels: pairs from three
64^3 runs at 32 ppc (target: the filtered current) and their 1-in-16
subsamples (input: the raw current), the correction on the current after 4
binomial passes, amplitude weight 8, 3000 Adam steps; 700 laps per judged
run (the B energy peaks near lap 400).

Measured when the test was written: the plain 2-ppc runs grow 3.6% slower
and peak 6.7% lower than the 32-ppc runs; the denoised runs grow 0.3%
slower and peak 5.4% lower, at 0.10 of the 32-ppc runs' GPU time per lap.
Bars: growth rate within 5% and closer to the 32-ppc runs' than the plain
2-ppc runs'; peak B energy within 10%; GPU time per lap at most 1/8 of the
32-ppc runs'."""
import pytest

pytestmark = pytest.mark.gpu


def test_denoised_2ppc_run_tracks_the_32ppc_run_3d_order3():
    import train_denoiser as td
    kw = dict(dim=3, order=3, nx=64, ny=64, nz=64, ppc_hi=32, ppc_lo=2, steps=700, every=30, first=10, c_omp=8.0)
    pairs = []
    for s in (0, 1, 2):
        pairs += td.make_pairs(seed=s, **kw)
    assert len(pairs) == 69
    m, losses = td.train(pairs, 3, iters=3000, lr=2e-3, base_passes=4, order=3, amplitude_weight=8.0)
    assert losses[-100:].mean() < losses[:100].mean()
    r = td.parity_study(m, n=96, seeds=range(11, 17), ppc_hi=32, ppc_lo=2, laps=700, order=3, dim=3, c_omp=8.0)
    h = r["hi_plain"]
    dev = {k: (d["growth"] / h["growth"] - 1.0, d["eb_max"] / h["eb_max"] - 1.0) for k, d in r.items()}
    for k, (g, e) in dev.items():
        print(f"{k:12s} growth {100 * g:+.1f}%  peak B energy {100 * e:+.1f}%  GPU {r[k]['ms_per_lap']:.3f} ms/lap")
    g_dn, e_dn = dev["lo_denoised"]
    g_lo, _ = dev["lo_plain"]
    assert abs(g_dn) < 0.05
    assert abs(g_dn) < abs(g_lo)
    assert abs(e_dn) < 0.10
    assert r["lo_denoised"]["ms_per_lap"] <= h["ms_per_lap"] / 8
