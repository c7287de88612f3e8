"""What the trained denoiser gains over the method without it.

Pairs from 2D pair-Weibel runs at 32 particles per cell per species
(2nd-order shapes): the target is the run's filtered current (the plain
method at high ppc); the input is the raw current of a 1-in-8 subsample of
the same particles (4 per cell per species, charge x8).  The model is trained
on the GPU with the library's own kernels and evaluated on held-out snapshots
of a run with another seed.

A curl-form correction J += curl_b(A) leaves div_b J, and so the
longitudinal part of J, exactly as deposited (that is what keeps charge
continuity at round-off).  The gain is therefore measured on the transverse
part, the only part the model can change: there the denoised low-ppc current
must be closer to the target than the plain method's filtered low-ppc
current.  The longitudinal part of the denoised current must equal the raw
deposit's."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tristan-mp-pu-master_densdecomp_amd", "python"))

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def weibel_pairs():
    import train_denoiser as td
    kw = dict(dim=2, order=2, nx=96, ny=96, ppc_hi=32, ppc_lo=4, steps=250, every=20, first=10, c_omp=5.0)
    pairs = td.make_pairs(seed=0, **kw) + td.make_pairs(seed=1, **kw)
    test = td.make_pairs(seed=7, **kw)
    assert len(pairs) == 24 and len(test) == 12
    return td, pairs, test


@pytest.mark.parametrize("base_passes", [4, 0])
def test_denoiser_beats_filter_on_held_out_weibel_snapshots(weibel_pairs, base_passes):
    """base_passes = 4: the correction acts on the plain method's filtered
    current; base_passes = 0: it replaces the filter"""
    td, pairs, test = weibel_pairs
    m, losses = td.train(pairs, 2, iters=2000, lr=2e-3, base_passes=base_passes)
    e = td.errors(m, test, 4, 2, 2)
    for k, d in e.items():
        print(f"{k:9s} all {d['all']:.4f}  transverse {d['T']:.4f}  longitudinal {d['L']:.4f}")
    assert np.all(np.isfinite(losses))
    assert losses[-100:].mean() < losses[:100].mean()
    assert e["denoised"]["T"] < e["raw"]["T"]
    assert e["denoised"]["T"] < e["filtered"]["T"], (e["denoised"]["T"], e["filtered"]["T"])
    # the correction is a discrete curl: the longitudinal error is that of
    # the current it acts on
    ref = e["filtered"]["L"] if base_passes else e["raw"]["L"]
    assert abs(e["denoised"]["L"] - ref) < 1e-3 * ref
    if base_passes:
        assert e["denoised"]["all"] < e["filtered"]["all"]
