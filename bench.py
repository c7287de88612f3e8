#!/usr/bin/env python3
"""Throughput of the per-step PIC loop on one MI355X.

Workload (default): 3D pair Weibel, 3rd-order shape, 512 x 256 x 256 periodic
grid, 4 particles per cell per species (268,435,456 particles), counter-
streaming beams +-x at gamma*beta = 1, thermal spread 0.1c (theta = 0.01),
c = 0.45, c/wp = 6 cells, 4 filter passes.  Synthetic initial state from the
library's loader; fields start at zero.

A step is one full main-loop iteration (B half step, push+deposit, B half
step, 4 filter passes, E step + current).  value = particles x steps / time.

roofline: the tiled push+deposit kernel (k_push_cells in 3D), priced at 56 B
per particle (read+write of the packed cell, 3 offsets, 3 momenta) + 48 B per
cell (6 field components read, 3 current components read-modify-write) per
launch (one launch per step), divided by the kernel's average duration
measured with HIP events around its launch on the library's stream.
traffic: HBM bytes per launch from FETCH_SIZE + WRITE_SIZE counter runs,
scaled by the same counters on a plain copy of known size
(profiles/traffic_<workload>.json).
roofline.counters: VALU busy, LDS active and bank-conflicted fractions of
the push kernel and its VALU lane-instructions per particle, from a
rocprofv3 --pmc pass over a short run of this script in a child process.
cpu_baseline: the reference's own Fortran routines (oracle/_ref, compiled
from its sources at -O3) on the host cores, one single-rank process per core
on its own 64 x 32 x 32-cell sample of the same plasma (core count and CPU
model stated).
denoised_vs_plain (default workload): step time of the 2-ppc run with the
learned denoiser against the 32-ppc plain run (3D pair Weibel, 3rd order,
256^3 cells), the second half of the metric.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "tristan-mp-pu-master_densdecomp_amd")
sys.path.insert(0, os.path.join(PKG, "python"))

HBM_PEAK_GBS = 8000.0

# "load": (species, ppc, drift gamma*beta along x, theta) per loader call.
# Weibel: like user/user_weibel.F90:252-292, two regions drifting -x and +x,
# each holding both species.
WORKLOADS = {
    # 3D pair Weibel precision study, 3rd order, full size
    "weibel3d_ord3_512x256x256_ppc4": dict(dim=3, order=3, nx=512, ny=256, nz=256, ppc=4, gb=1.0, theta=0.01,
                                           c=0.45, c_omp=6.0, me=1.0, mi=1.0, ntimes=4,
                                           load=[(0, 2, -1.0, 0.01), (1, 2, -1.0, 0.01),
                                                 (0, 2, 1.0, 0.01), (1, 2, 1.0, 0.01)]),
    # the same run with the Vay pusher (the reference's -Dvay build)
    "weibel3d_ord3_512x256x256_ppc4_vay": dict(dim=3, order=3, nx=512, ny=256, nz=256, ppc=4, gb=1.0, theta=0.01,
                                               c=0.45, c_omp=6.0, me=1.0, mi=1.0, ntimes=4, pusher=1,
                                               load=[(0, 2, -1.0, 0.01), (1, 2, -1.0, 0.01),
                                                     (0, 2, 1.0, 0.01), (1, 2, 1.0, 0.01)]),
    # 3D electron-ion two-stream, 2nd order: electron beams +-0.2c over cold ions
    "twostream3d_ord2_192cube_ppc16": dict(dim=3, order=2, nx=192, ny=192, nz=192, ppc=16, gb=0.2041, theta=1e-4,
                                           c=0.45, c_omp=8.0, me=1.0, mi=16.0, ntimes=4,
                                           load=[(0, 16, 0.0, 0.0), (1, 8, -0.2041, 1e-4), (1, 8, 0.2041, 1e-4)]),
    # 2D pair Weibel, 1st order
    "weibel2d_ord1_256x256_ppc8": dict(dim=2, order=1, nx=256, ny=256, nz=1, ppc=8, gb=0.5, theta=2.5e-3,
                                       c=0.45, c_omp=10.0, me=1.0, mi=1.0, ntimes=4,
                                       load=[(0, 4, -0.5, 2.5e-3), (1, 4, -0.5, 2.5e-3),
                                             (0, 4, 0.5, 2.5e-3), (1, 4, 0.5, 2.5e-3)]),
    # 2D pair Weibel, 2nd order, 1024^2: the plain run at 32 ppc per species
    # and the 2-ppc run with the learned denoiser on 8 filter passes (its
    # weights do not change the time; the physics comparison is in
    # train_denoiser.parity_study)
    "weibel2d_ord2_1024x1024_ppc32": dict(dim=2, order=2, nx=1024, ny=1024, nz=1, ppc=32, gb=0.5, theta=2.5e-3,
                                          c=0.45, c_omp=5.0, me=1.0, mi=1.0, ntimes=4,
                                          load=[(0, 16, -0.5, 2.5e-3), (1, 16, -0.5, 2.5e-3),
                                                (0, 16, 0.5, 2.5e-3), (1, 16, 0.5, 2.5e-3)]),
    # 3D pair Weibel at scale, 3rd order, 256^3: the plain run at 32 ppc per
    # species (1.07e9 particles) and the 2-ppc run with the learned denoiser
    # on 4 filter passes
    "weibel3d_ord3_256cube_ppc32": dict(dim=3, order=3, nx=256, ny=256, nz=256, ppc=32, gb=0.5, theta=2.5e-3,
                                        c=0.45, c_omp=8.0, me=1.0, mi=1.0, ntimes=4,
                                        load=[(0, 16, -0.5, 2.5e-3), (1, 16, -0.5, 2.5e-3),
                                              (0, 16, 0.5, 2.5e-3), (1, 16, 0.5, 2.5e-3)]),
    "weibel3d_ord3_256cube_ppc2_denoised": dict(dim=3, order=3, nx=256, ny=256, nz=256, ppc=2, gb=0.5,
                                                theta=2.5e-3, c=0.45, c_omp=8.0, me=1.0, mi=1.0, ntimes=4,
                                                denoise=dict(base_passes=4),
                                                load=[(0, 1, -0.5, 2.5e-3), (1, 1, -0.5, 2.5e-3),
                                                      (0, 1, 0.5, 2.5e-3), (1, 1, 0.5, 2.5e-3)]),
    "weibel2d_ord2_1024x1024_ppc2_denoised": dict(dim=2, order=2, nx=1024, ny=1024, nz=1, ppc=2, gb=0.5,
                                                  theta=2.5e-3, c=0.45, c_omp=5.0, me=1.0, mi=1.0, ntimes=4,
                                                  denoise=dict(base_passes=8),
                                                  load=[(0, 1, -0.5, 2.5e-3), (1, 1, -0.5, 2.5e-3),
                                                        (0, 1, 0.5, 2.5e-3), (1, 1, 0.5, 2.5e-3)]),
}


def normalisation(w):
    """code/particles.F90:206-235 for ppc0 = 2*ppc (both species)."""
    import numpy as np
    g0 = np.sqrt(1.0 + w["gb"] ** 2) if w["gb"] >= 0 else 1.0
    ppc0 = 2.0 * w["ppc"]
    omp = w["c"] / w["c_omp"]
    qe = -(omp * omp * g0) / ((ppc0 * 0.5) * (1.0 + w["me"] / w["mi"]))
    qi = -qe
    me = w["me"] * abs(qi)
    mi = w["mi"] * abs(qi)
    return dict(qe=qe, qi=qi, me=me, mi=mi, qme=qe / me, qmi=qi / mi, g0=g0, ppc0=ppc0)


def make_context(w, sort_every=0):
    import tmpg
    nm = normalisation(w)
    ncell = w["nx"] * w["ny"] * (w["nz"] if w["dim"] == 3 else 1)
    cap = ncell * w["ppc"]
    ctx = tmpg.Context(w["dim"], w["order"], w["nx"], w["ny"], w["nz"], ntimes=w["ntimes"], c=w["c"],
                       qm=(nm["qmi"], nm["qme"]), q=(nm["qi"], nm["qe"]), mass=(nm["mi"], nm["me"]),
                       capacity=(cap, cap), sort_every=sort_every, pusher=w.get("pusher", 0))
    for sp, ppc, gb, theta in w["load"]:
        ctx.load_uniform(sp, ppc, gb, theta, 7 + 2 * sp + (gb > 0))
    if w.get("denoise"):
        m = tmpg.Model(w["dim"], seed=0)
        m.set_base_passes(w["denoise"]["base_passes"])
        m.set_scale(1.0)
        ctx.set_denoiser(m)
    return ctx, ncell


def host_cpu():
    """(CPU model, cores this process may use, capped at the 16 host cores
    that go with one GPU)"""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return model, max(1, min(n, 16))


def cpu_reference_baseline(w, steps=30, max_cores=0):
    """The reference's own Fortran routines (oracle/_ref, built from its
    sources at -O3, FP32) on the host cores: one single-rank process per core,
    each on its own 64x32x32 sample of the workload's plasma (disjoint seeds),
    all started together.  Throughput = all processes' particle-steps / the
    slowest process's time, start-up (a 0-step run) subtracted."""
    import numpy as np
    dim, order = w["dim"], w["order"]
    exe = os.path.join(ROOT, "oracle", "_ref", f"tristan_ref_{dim}d_dd{order}_f32_o3")
    if not os.path.exists(exe):
        return None
    model, ncore = host_cpu()
    if max_cores > 0:
        ncore = min(ncore, max_cores)
    sx, sy, sz = (64, 32, 32) if dim == 3 else (256, 256, 1)
    nm = normalisation(w)
    ng = 5 if order == 1 else 7
    ngz = 5 if dim == 2 else ng
    mx, my, mz = sx + ng, sy + ng, (sz + ngz if dim == 3 else 1)
    h, hz = ng // 2, ngz // 2
    ncell = sx * sy * sz
    n = ncell * w["ppc"]

    def species(rng):
        pos = np.empty((3, n))
        pos[0] = h + 1 + rng.random(n) * sx
        pos[1] = h + 1 + rng.random(n) * sy
        pos[2] = (hz + 1 + rng.random(n) * sz) if dim == 3 else hz + 1.5
        th = np.sqrt(w["theta"])
        mom = rng.standard_normal((3, n)) * th
        mom[0, : n // 2] -= w["gb"]          # half of each species per beam
        mom[0, n // 2:] += w["gb"]
        return np.concatenate([pos, mom, np.ones((1, n))]).astype(np.float32)

    inp = f"""<node_configuration>
sizex = 1
sizey = 1
<time>
c = {w['c']}
<grid>
mx0 = {sx}
my0 = {sy}
mz0 = {sz}
<algorithm>
highorder = 0
Corr = 1.025
ntimes = {w['ntimes']}
<boundaries>
periodicx = 1
periodicy = 1
periodicz = 1
<particles>
maxptl0 = {4 * n + 100000}
ppc0 = {nm['ppc0']}
delgam = {w['theta']}
me = {w['me']}
mi = {w['mi']}
gamma0 = {nm['g0']}
c_omp = {w['c_omp']}
<problem>
distr_dim = 3
external_fields = 0
user_part_bcs = 0
"""
    with tempfile.TemporaryDirectory() as td:
        open(os.path.join(td, "input"), "w").write(inp)
        states = []
        for r in range(ncore):
            rng = np.random.default_rng(1000 + r)
            st = os.path.join(td, f"state{r}.bin")
            with open(st, "wb") as fh:
                fh.write(np.zeros(6 * mx * my * mz, np.float32).tobytes())
                fh.write(np.array([n, n], np.int32).tobytes())
                for _ in (1, -1):
                    fh.write(species(rng).tobytes())
            states.append(st)

        def run_all(nsteps):
            t0 = time.perf_counter()
            procs = [subprocess.Popen([exe, "input", st, os.path.join(td, f"out{r}.bin"), str(nsteps), "0"], cwd=td,
                                      stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                     for r, st in enumerate(states)]
            for pr in procs:
                if pr.wait(timeout=600) != 0:
                    raise RuntimeError(f"{exe} failed")
            return time.perf_counter() - t0

        t_start = run_all(0)
        t_run = run_all(steps)
        per_step = max(t_run - t_start, 1e-6) / steps
    return dict(value=ncore * 2 * n / per_step, unit="particle-steps/s", cores=ncore, kind="reference",
                sample=f"reference Fortran (amdflang -O3, FP32), {ncore} single-rank processes started together on "
                       f"{ncore} host cores ({model}), each on its own {sx}x{sy}x{sz}-cell sample of this plasma "
                       f"({w['ppc']} ppc/species, {2 * n} particles), {steps} full main-loop steps; "
                       f"start-up subtracted")


def read_traffic(workload):
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(p):
        try:
            return json.load(open(p)).get("bytes_per_step")
        except Exception:
            return None
    return None


def time_steps(w, warmup, steps):
    """ms per full step of workload w (the library's own timing events)"""
    ctx, _ = make_context(w)
    try:
        ctx.step(warmup)
        ctx.sync()
        ctx.last_timings()
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.sync()
        t1 = time.perf_counter()
        npart = ctx.count(0) + ctx.count(1)
    finally:
        ctx.close()
    return 1e3 * (t1 - t0) / steps, npart


def denoise_ratio(warmup=2, steps=4):
    """The learned-denoiser part of the metric: step time of the 2-ppc run with
    the denoiser against the 32-ppc plain run it is judged against (3D pair
    Weibel, 3rd order, 256^3 cells).  The network's weights do not change the
    time; the physics at parity is checked by the denoiser tests and the
    parity study of python/train_denoiser.py."""
    plain_ms, n_plain = time_steps(WORKLOADS["weibel3d_ord3_256cube_ppc32"], warmup, steps)
    den_ms, n_den = time_steps(WORKLOADS["weibel3d_ord3_256cube_ppc2_denoised"], warmup, steps)
    return {"plain": "weibel3d_ord3_256cube_ppc32", "denoised": "weibel3d_ord3_256cube_ppc2_denoised",
            "plain_ms_per_step": plain_ms, "denoised_ms_per_step": den_ms, "ratio": den_ms / plain_ms,
            "target_ratio": 0.125, "particles": [n_plain, n_den], "steps": steps, "warmup": warmup}


# counters of one profiled pass over the push kernel (limits of one rocprofv3
# pass: at most 8 SQ and 2 GRBM counters)
PUSH_COUNTERS = ["SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_LDS_IDX_ACTIVE", "SQ_LDS_BANK_CONFLICT",
                 "SQ_WAIT_INST_LDS", "GRBM_GUI_ACTIVE"]
N_XCD, N_CU, N_SIMD = 8, 256, 1024


def push_counters(workload, kernel, npart, timeout=240):
    """VALU and LDS utilisation of the push kernel, from a rocprofv3 --pmc
    pass over a short run of this script (a child process; 2 steps after 2
    warm-up steps).  VALU busy = SQ_INSTS_VALU x 4 cycles (a wave64 VALU
    instruction on a 16-lane SIMD) / 1024 SIMDs / the launch's cycles
    (GRBM_GUI_ACTIVE / 8 XCDs); LDS active and bank-conflicted:
    SQ_LDS_IDX_ACTIVE and SQ_LDS_BANK_CONFLICT / 256 CUs / cycles.  Mean over
    the profiled launches.  None when rocprofv3 is absent or fails."""
    import csv
    import glob
    import shutil
    prof = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3")
                                         else None)
    if not prof:
        return None
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        cmd = [prof, "--pmc", *PUSH_COUNTERS, "--output-format", "csv", "-d", td, "-o", "pc", "--",
               sys.executable, os.path.abspath(__file__), "--workload", workload, "--steps", "2", "--warmup", "2",
               "--no-cpu", "--no-ratio", "--no-counters"]
        env = dict(os.environ, TMPDIR="/tmp")
        r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True,
                           timeout=timeout)
        files = glob.glob(os.path.join(td, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            print(f"push counters: no counter file (rocprofv3 rc {r.returncode}): {r.stderr[-400:]}",
                  file=sys.stderr)
            return None
        per = {}
        for f in files:
            for row in csv.DictReader(open(f)):
                if kernel not in row.get("Kernel_Name", ""):
                    continue
                d = per.setdefault(row["Dispatch_Id"], {})
                d[row["Counter_Name"]] = d.get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    launches = [d for d in per.values() if all(c in d for c in PUSH_COUNTERS)]
    if not launches:
        return None
    mean = {c: sum(d[c] for d in launches) / len(launches) for c in PUSH_COUNTERS}
    cycles = mean["GRBM_GUI_ACTIVE"] / N_XCD
    return {"source": f"rocprofv3 --pmc {' '.join(PUSH_COUNTERS)}, {len(launches)} launches of {kernel}",
            "valu_busy": mean["SQ_INSTS_VALU"] * 4.0 / N_SIMD / cycles,
            "lds_active": mean["SQ_LDS_IDX_ACTIVE"] / N_CU / cycles,
            "lds_bank_conflicted": mean["SQ_LDS_BANK_CONFLICT"] / N_CU / cycles,
            "valu_lane_insts_per_particle": mean["SQ_INSTS_VALU"] * 64.0 / npart,
            "lds_lane_insts_per_particle": mean["SQ_INSTS_LDS"] * 64.0 / npart,
            "waves_per_launch": mean["SQ_WAVES"], "cycles_per_launch": cycles}


PHASES = ["stage", "push+records", "crossers+scans", "index scatter", "cell sums", "reset", "flush"]


def phase_split(workload, steps, warmup):
    """Per-step split of the 3D push kernel's tile time by phase, from the
    phase-stamped build's cycle stamps (thread 0 of each workgroup, at the
    phases' closing barriers)."""
    import tmpg
    tmpg.load_library(os.path.join(os.path.dirname(tmpg.LIB_PATH), "libtmpg_phases.so"))
    w = WORKLOADS[workload]
    ctx, _ = make_context(w)
    ctx.step(warmup)
    ctx.sync()
    tmpg.phase_cycles(reset=True)
    rows = []
    for k in range(steps):
        ctx.step(1)
        ctx.sync()
        cyc, cnt = tmpg.phase_cycles(reset=True)
        tiles = max(int(cnt[0]), 1)
        tot = float(sum(int(c) for c in cyc[:7]))
        per = {PHASES[i]: int(cyc[i]) / tiles for i in range(7)}
        rows.append(per)
        print(f"step {warmup + k + 1}: " + ", ".join(f"{n} {per[n]:.0f} ({100 * int(cyc[i]) / tot:.1f}%)"
                                                     for i, n in enumerate(PHASES)), flush=True)
    ctx.close()
    print(json.dumps({"workload": workload, "cycles_per_tile": rows}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="weibel3d_ord3_512x256x256_ppc4", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-ratio", action="store_true", help="skip the denoised / plain step-time ratio")
    ap.add_argument("--no-counters", action="store_true", help="skip the profiled pass for VALU / LDS counters")
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="host cores of the CPU baseline (0: all, at most 16)")
    ap.add_argument("--sort-every", type=int, default=0,
                    help="steps between in-tile cell sorts of the particles (0: library default, <0: never)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (5 times): the known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--phase-split", type=int, default=0, metavar="STEPS",
                    help="only run STEPS steps of the workload with the phase-stamped library "
                         "(lib/libtmpg_phases.so) and print the push kernel's cycles per tile and phase")
    a = ap.parse_args()
    if a.phase_split:
        phase_split(a.workload, a.phase_split, a.warmup)
        return
    if a.copy_calibration:
        import torch
        n = a.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, device="cuda:0", dtype=torch.float32)
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (one periodic domain per GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    w = WORKLOADS[a.workload]
    import tmpg
    ctx, ncell = make_context(w, a.sort_every)
    npart = ctx.count(0) + ctx.count(1)
    ctx.step(a.warmup)
    ctx.sync()
    ctx.last_timings()          # discard warm-up timings
    t0 = time.perf_counter()
    ctx.step(a.steps)
    ctx.sync()
    t1 = time.perf_counter()
    ms = ctx.last_timings()
    dt = t1 - t0
    value = npart * a.steps / dt
    push_ms = ms[1] / max(ms[3], 1)
    kernel_ms = ms[5] / max(ms[3], 1)      # the tiled push+deposit kernel alone
    # read+write of the packed cell, the offsets (3 in 3D, 2 in 2D) and 3
    # momenta per particle; 6 field components read and 3 current
    # components read-modify-written per cell
    alg_bytes = (56.0 if w["dim"] == 3 else 48.0) * npart + 48.0 * ncell
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    traffic = read_traffic(a.workload)
    ctx.close()                 # free the device memory before the other runs
    print(f"bench: {a.workload} {1e3 * dt / a.steps:.3f} ms/step", file=sys.stderr, flush=True)
    headline = a.workload == "weibel3d_ord3_512x256x256_ppc4"
    ratio = None
    if headline and not a.no_ratio:
        try:
            ratio = denoise_ratio()
            print(f"bench: denoised / plain step time {ratio['ratio']:.3f}", file=sys.stderr, flush=True)
        except Exception as e:  # reported, never required
            print(f"denoise ratio failed: {e}", file=sys.stderr)
    counters = None
    if not a.no_counters:
        try:
            counters = push_counters(a.workload, "k_push_cells" if w["dim"] == 3 else "k_push_tiled", npart)
            print(f"bench: push counters {counters}", file=sys.stderr, flush=True)
        except Exception as e:  # reported, never required
            print(f"push counters failed: {e}", file=sys.stderr)
    cpu = None
    if not a.no_cpu:
        try:
            print("bench: CPU baseline", file=sys.stderr, flush=True)
            cpu = cpu_reference_baseline(w, max_cores=a.cpu_cores)
        except Exception as e:  # the baseline is reported, never required
            print(f"cpu baseline failed: {e}", file=sys.stderr)
    out = {
        "metric": f"particle-steps/s ({w['dim']}D ord-{w['order']} full step"
                  f"{', learned denoiser' if w.get('denoise') else ''}{', Vay pusher' if w.get('pusher') else ''})",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * dt / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (library loader: drifting Maxwell-Juttner beams, uniform positions)",
        "config": {"workload": a.workload, "grid": [w["nx"], w["ny"], w["nz"]], "order": w["order"],
                   "particles": npart, "ppc_per_species": w["ppc"], "filter_passes": w["ntimes"],
                   "phase_ms": {"fields": ms[0] / ms[3], "push_deposit": push_ms, "push_kernel": kernel_ms,
                                "filter": ms[2] / ms[3]}},
        "roofline": {"bound": "hbm", "kernel": "k_push_cells" if w["dim"] == 3 else "k_push_tiled",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": alg_bytes, "counters": counters},
        "cpu_baseline": cpu,
    }
    if ratio is not None:
        out["denoised_vs_plain"] = ratio
    print(json.dumps(out))


if __name__ == "__main__":
    main()
