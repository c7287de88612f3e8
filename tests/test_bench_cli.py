"""bench.py's contract on the CPU: one GPU only (a multi-GPU request exits
with a message), the workload table, and the normalisation of the
reference (code/particles.F90:206-235) that sizes charge and mass."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


def test_workloads_load_their_stated_particles_per_cell():
    for name, w in bench.WORKLOADS.items():
        per_species = [0, 0]
        for sp, ppc, _, _ in w["load"]:
            per_species[sp] += ppc
        # every species is loaded; the heavier-loaded one matches "ppc"
        assert max(per_species) == w["ppc"], name
        assert min(per_species) > 0, name


def test_normalisation_keeps_density_independent_of_ppc():
    """charge and mass per macro-particle scale as 1/ppc, so the 2-ppc
    denoised workloads describe the plasma of their 32-ppc references"""
    hi = bench.normalisation(bench.WORKLOADS["weibel3d_ord3_256cube_ppc32"])
    lo = bench.normalisation(bench.WORKLOADS["weibel3d_ord3_256cube_ppc2_denoised"])
    for k in ("qe", "qi", "me", "mi"):
        assert np.isclose(lo[k] * 2, hi[k] * 32), k
    for k in ("qme", "qmi"):
        assert np.isclose(lo[k], hi[k]), k
    # plasma frequency: wp^2 = n q^2 / m with n = 2 ppc per cell, c/wp = c_omp
    w = bench.WORKLOADS["weibel3d_ord3_256cube_ppc32"]
    n = 2.0 * w["ppc"]
    wp2 = n * hi["qe"] ** 2 / hi["me"] / 2 + n * hi["qi"] ** 2 / hi["mi"] / 2
    g0 = np.sqrt(1.0 + w["gb"] ** 2)
    assert np.isclose(wp2 / g0, (w["c"] / w["c_omp"]) ** 2, rtol=1e-6)
