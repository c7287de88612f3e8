"""Generate the golden fixtures of tests/golden/ from the compiled reference.

Runs the reference's own routines (oracle/_ref/tristan_ref_*, built by
`make -C oracle ref` from the reference's Fortran sources) on small seeded
states and stores inputs and outputs as compressed .npz files.  Only runs
where the reference sources are present; the fixtures themselves travel with
the repository.

Each fixture holds, for one (dim, order):
  params          : grid, physics inputs, normalisation
  f_in            : interior E,B (6, nz, ny, nx) float32 -- ghosts are filled
                    periodically by the first bc_b1/bc_e1 of the step
  pi_in, pe_in    : particles (7, n) float32, absolute 1-based coordinates
  step_{f32,f64}  : after one main-loop step: interior fields (9, nz, ny, nx)
                    (E, B, filtered J), particles (7, n) in input order
  dep_{f32,f64}   : currents after deposit + fold only, interior (3, nz, ny, nx)
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
import orc  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")

CASES = {
    2: dict(nx=24, ny=20, nz=1, npart=300),
    3: dict(nx=10, ny=8, nz=6, npart=250),
}
PHYS = dict(c=0.45, corr=1.025, ntimes=4, c_omp=3.0, gamma0=0.5, ppc0=2.0, me=1.0, mi=4.0)


def input_text(dim, nx, ny, nz, maxptl, highorder=0):
    p = PHYS
    return f"""<node_configuration>
sizex = 1
sizey = 1
<time>
last = 1
c = {p['c']}
<grid>
mx0 = {nx}
my0 = {ny}
mz0 = {nz if dim == 3 else 1}
<algorithm>
highorder = {highorder}
Corr = {p['corr']}
ntimes = {p['ntimes']}
<boundaries>
periodicx = 1
periodicy = 1
periodicz = 1
<fields>
btheta = 0
bphi = 90
<domain>
enlarge = 0
<particles>
sigma = 0.
maxptl0 = {maxptl}
ppc0 = {p['ppc0']}
delgam = 1.e-4
me = {p['me']}
mi = {p['mi']}
gamma0 = {p['gamma0']}
c_omp = {p['c_omp']}
<problem>
distr_dim = 3
temperature_ratio = 1
external_fields = 0
sigma_ext = 0.
user_part_bcs = 0
"""


def make_particles(rng, cfg, n):
    """absolute coordinates in [h+1, h+1+n) per axis; momenta up to |u|~3.
    A quarter of the particles sit on special offsets (0, 0.5, just below 1)
    or next to the periodic edges."""
    mx, my, mz, ng, ngz = orc.extent(cfg)
    h, hz = ng // 2, ngz // 2
    lo = np.array([h + 1, h + 1, hz + 1], dtype=np.float64)
    span = np.array([cfg.nx, cfg.ny, cfg.nz if cfg.dim == 3 else 1], dtype=np.float64)
    pos = lo[:, None] + rng.random((3, n)) * span[:, None]
    k = n // 16
    cells = np.floor(pos)
    pos[:, 0:k] = cells[:, 0:k]                        # offset exactly 0
    pos[:, k:2 * k] = cells[:, k:2 * k] + 0.5          # ord-2 branch point
    pos[:, 2 * k:3 * k] = cells[:, 2 * k:3 * k] + 0.99999
    pos[0, 3 * k:4 * k] = lo[0] + 0.02                 # next to the low x edge
    pos[1, 3 * k:4 * k] = lo[1] + span[1] - 0.03       # next to the high y edge
    if cfg.dim == 2:
        pos[2] = lo[2] + 0.5
    mom = rng.uniform(-1.0, 1.0, (3, n))
    mom[:, 4 * k:5 * k] *= 3.0
    mom[:, 5 * k:5 * k + 4] = 0.0
    ch = np.ones((1, n))
    return np.concatenate([pos, mom, ch]).astype(np.float32)


def write_state(path, f6_padded, pi, pe, dtype):
    with open(path, "wb") as fh:
        fh.write(np.ascontiguousarray(f6_padded, dtype=dtype).tobytes())
        fh.write(np.array([pi.shape[1], pe.shape[1]], dtype=np.int32).tobytes())
        for p in (pi, pe):
            for r in range(7):
                fh.write(np.ascontiguousarray(p[r], dtype=dtype).tobytes())


def read_state(path, cfg, dtype):
    mx, my, mz, _, _ = orc.extent(cfg)
    nc = mx * my * mz
    isz = np.dtype(dtype).itemsize
    raw = open(path, "rb").read()
    g = np.frombuffer(raw[: 9 * nc * isz], dtype=dtype).reshape(9, mz, my, mx).copy()
    off = 9 * nc * isz
    ni, ne = np.frombuffer(raw[off: off + 8], dtype=np.int32)
    off += 8
    out = []
    for n in (ni, ne):
        a = np.frombuffer(raw[off: off + 7 * n * isz], dtype=dtype).reshape(7, n).copy()
        off += 7 * n * isz
        ind = np.frombuffer(raw[off: off + 4 * n], dtype=np.int32).copy()
        off += 4 * n
        out.append((a, ind))
    return g, out


def run_ref(dim, order, prec, inp, state, mode, cfg, ni, ne, maxhlf, suffix="", exe=None):
    exe = exe or os.path.join(REF_DIR, f"tristan_ref_{dim}d_dd{order}_{prec}{suffix}")
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "input")
        open(fin, "w").write(inp)
        fout = os.path.join(td, "out.bin")
        subprocess.run([exe, fin, state, fout, "1", str(mode)], check=True, cwd=td,
                       stdout=subprocess.DEVNULL)
        dtype = np.float32 if prec == "f32" else np.float64
        g, ((a_i, ind_i), (a_e, ind_e)) = read_state(fout, cfg, dtype)
    # back to input order: ions carry ind 1..ni, electrons maxhlf+1..
    pi = np.empty_like(a_i)
    pi[:, ind_i - 1] = a_i
    pe = np.empty_like(a_e)
    pe[:, ind_e - maxhlf - 1] = a_e
    assert sorted(ind_i - 1) == list(range(ni)) and sorted(ind_e - maxhlf - 1) == list(range(ne))
    return g, pi, pe


def main():
    os.makedirs(HERE, exist_ok=True)
    for dim in (2, 3):
        for order in (1, 2, 3):
            case = CASES[dim]
            cfg = orc.make_cfg(dim, order, case["nx"], case["ny"], case["nz"], ntimes=PHYS["ntimes"])
            rng = np.random.default_rng(1000 * dim + order)
            nz = case["nz"] if dim == 3 else 1
            f_in = np.empty((6, nz, case["ny"], case["nx"]), dtype=np.float32)
            f_in[:3] = rng.uniform(-0.05, 0.05, f_in[:3].shape)
            f_in[3:] = rng.uniform(-0.1, 0.1, f_in[3:].shape)
            pi = make_particles(rng, cfg, case["npart"])
            pe = make_particles(rng, cfg, case["npart"] + 17)
            maxptl = 20000
            maxhlf = maxptl // 2
            inp = input_text(dim, case["nx"], case["ny"], case["nz"], maxptl)
            fix = dict(f_in=f_in, pi_in=pi, pe_in=pe)
            sl = orc.interior_slices(cfg)
            with tempfile.TemporaryDirectory() as td:
                for prec, dt in (("f32", np.float32), ("f64", np.float64)):
                    f6 = orc.fill_from_interior(cfg, f_in.astype(dt), dt)
                    st = os.path.join(td, f"state_{prec}.bin")
                    write_state(st, f6, pi, pe, dt)
                    g, qi_, qe_ = run_ref(dim, order, prec, inp, st, 0, cfg, pi.shape[1], pe.shape[1], maxhlf)
                    fix[f"step_{prec}"] = g[(slice(None),) + sl]
                    fix[f"step_pi_{prec}"] = qi_
                    fix[f"step_pe_{prec}"] = qe_
                    g, _, _ = run_ref(dim, order, prec, inp, st, 2, cfg, pi.shape[1], pe.shape[1], maxhlf)
                    fix[f"dep_{prec}"] = g[(slice(6, 9),) + sl]
            params = dict(dim=dim, order=order, nx=case["nx"], ny=case["ny"], nz=nz, **PHYS)
            fix["params"] = np.array(repr(params))
            out = os.path.join(HERE, f"ref_{dim}d_ord{order}.npz")
            np.savez_compressed(out, **fix)
            print(out, os.path.getsize(out))


def main_highorder():
    """ref_{dim}d_ord{order}_ho.npz: the same kind of state run with
    highorder = 1 (the 4th-order field solver, advance_b_halfstep_42 and
    advance_e_fullstep_42): one full main-loop step, and the Yee part alone
    (B half, B half, E; driver mode 4)"""
    for dim in (2, 3):
        for order in (1, 3):
            case = CASES[dim]
            cfg = orc.make_cfg(dim, order, case["nx"], case["ny"], case["nz"], ntimes=PHYS["ntimes"])
            rng = np.random.default_rng(5000 + 1000 * dim + order)
            nz = case["nz"] if dim == 3 else 1
            f_in = np.empty((6, nz, case["ny"], case["nx"]), dtype=np.float32)
            f_in[:3] = rng.uniform(-0.05, 0.05, f_in[:3].shape)
            f_in[3:] = rng.uniform(-0.1, 0.1, f_in[3:].shape)
            pi = make_particles(rng, cfg, case["npart"])
            pe = make_particles(rng, cfg, case["npart"] + 17)
            maxptl = 20000
            maxhlf = maxptl // 2
            inp = input_text(dim, case["nx"], case["ny"], case["nz"], maxptl, highorder=1)
            fix = dict(f_in=f_in, pi_in=pi, pe_in=pe)
            sl = orc.interior_slices(cfg)
            with tempfile.TemporaryDirectory() as td:
                for prec, dt in (("f32", np.float32), ("f64", np.float64)):
                    f6 = orc.fill_from_interior(cfg, f_in.astype(dt), dt)
                    st = os.path.join(td, f"state_{prec}.bin")
                    write_state(st, f6, pi, pe, dt)
                    g, qi_, qe_ = run_ref(dim, order, prec, inp, st, 0, cfg, pi.shape[1], pe.shape[1], maxhlf)
                    fix[f"step_{prec}"] = g[(slice(None),) + sl]
                    fix[f"step_pi_{prec}"] = qi_
                    fix[f"step_pe_{prec}"] = qe_
                    g, _, _ = run_ref(dim, order, prec, inp, st, 4, cfg, pi.shape[1], pe.shape[1], maxhlf)
                    fix[f"yee_{prec}"] = g[(slice(0, 6),) + sl]
            params = dict(dim=dim, order=order, nx=case["nx"], ny=case["ny"], nz=nz, highorder=1, **PHYS)
            fix["params"] = np.array(repr(params))
            out = os.path.join(HERE, f"ref_{dim}d_ord{order}_ho.npz")
            np.savez_compressed(out, **fix)
            print(out, os.path.getsize(out))


def main_vay():
    """ref_{dim}d_ord{order}_vay.npz: one main-loop step and the mover alone
    (driver mode 1) of the reference built with -Dvay (the Vay pusher);
    2D order 1 and 3D order 3"""
    for dim, order in ((2, 1), (3, 3)):
        case = CASES[dim]
        cfg = orc.make_cfg(dim, order, case["nx"], case["ny"], case["nz"], ntimes=PHYS["ntimes"])
        rng = np.random.default_rng(7000 + 1000 * dim + order)
        nz = case["nz"] if dim == 3 else 1
        # fields 10x those of the other fixtures, so that the Vay and Boris
        # pushes of the same state differ far beyond the parity bars
        f_in = np.empty((6, nz, case["ny"], case["nx"]), dtype=np.float32)
        f_in[:3] = rng.uniform(-0.5, 0.5, f_in[:3].shape)
        f_in[3:] = rng.uniform(-1.0, 1.0, f_in[3:].shape)
        pi = make_particles(rng, cfg, case["npart"])
        pe = make_particles(rng, cfg, case["npart"] + 17)
        maxptl = 20000
        maxhlf = maxptl // 2
        inp = input_text(dim, case["nx"], case["ny"], case["nz"], maxptl)
        fix = dict(f_in=f_in, pi_in=pi, pe_in=pe)
        sl = orc.interior_slices(cfg)
        with tempfile.TemporaryDirectory() as td:
            for prec, dt in (("f32", np.float32), ("f64", np.float64)):
                f6 = orc.fill_from_interior(cfg, f_in.astype(dt), dt)
                st = os.path.join(td, f"state_{prec}.bin")
                write_state(st, f6, pi, pe, dt)
                g, qi_, qe_ = run_ref(dim, order, prec, inp, st, 0, cfg, pi.shape[1], pe.shape[1], maxhlf, "_vay")
                fix[f"step_{prec}"] = g[(slice(None),) + sl]
                fix[f"step_pi_{prec}"] = qi_
                fix[f"step_pe_{prec}"] = qe_
                _, qi_, qe_ = run_ref(dim, order, prec, inp, st, 1, cfg, pi.shape[1], pe.shape[1], maxhlf, "_vay")
                fix[f"move_pi_{prec}"] = qi_
                fix[f"move_pe_{prec}"] = qe_
                _, bi_, _ = run_ref(dim, order, prec, inp, st, 1, cfg, pi.shape[1], pe.shape[1], maxhlf)
                d = np.linalg.norm(bi_[3:6] - qi_[3:6]) / np.linalg.norm(qi_[3:6])
                print(f"{dim}D ord{order} {prec}: Boris vs Vay momenta after one push {d:.2e}")
        params = dict(dim=dim, order=order, nx=case["nx"], ny=case["ny"], nz=nz, pusher="vay", **PHYS)
        fix["params"] = np.array(repr(params))
        out = os.path.join(HERE, f"ref_{dim}d_ord{order}_vay.npz")
        np.savez_compressed(out, **fix)
        print(out, os.path.getsize(out))


def main_filter2():
    """ref_{dim}d_ord{order}_filter2.npz: the reference built with -Dfilter2
    (apply_filter2_opt, code/optimized_filters.F90): the filter alone on
    random currents (driver mode 3) and one full main-loop step; 3D order 3
    and 2.  (In 2D the reference never creates the MPI datatypes filter2
    copies with: code/fields.F90:362-368 calls create_MPI_filter_datatypes
    only #ifndef twoD, so a 2D filter2 build stops in its first MPI call.)"""
    for dim, order in ((3, 3), (3, 2)):
        case = CASES[dim]
        cfg = orc.make_cfg(dim, order, case["nx"], case["ny"], case["nz"], ntimes=PHYS["ntimes"])
        rng = np.random.default_rng(9000 + 1000 * dim + order)
        nz = case["nz"] if dim == 3 else 1
        f_in = np.empty((6, nz, case["ny"], case["nx"]), dtype=np.float32)
        f_in[:3] = rng.uniform(-0.05, 0.05, f_in[:3].shape)
        f_in[3:] = rng.uniform(-0.1, 0.1, f_in[3:].shape)
        j_in = rng.uniform(-1.0, 1.0, (3, nz, case["ny"], case["nx"])).astype(np.float32)
        pi = make_particles(rng, cfg, case["npart"])
        pe = make_particles(rng, cfg, case["npart"] + 17)
        maxptl = 20000
        maxhlf = maxptl // 2
        inp = input_text(dim, case["nx"], case["ny"], case["nz"], maxptl)
        fix = dict(f_in=f_in, j_in=j_in, pi_in=pi, pe_in=pe)
        sl = orc.interior_slices(cfg)
        with tempfile.TemporaryDirectory() as td:
            for prec, dt in (("f32", np.float32), ("f64", np.float64)):
                f6 = orc.fill_from_interior(cfg, f_in.astype(dt), dt)
                st = os.path.join(td, f"state_{prec}.bin")
                write_state(st, f6, pi, pe, dt)
                g, qi_, qe_ = run_ref(dim, order, prec, inp, st, 0, cfg, pi.shape[1], pe.shape[1], maxhlf,
                                      "_filter2")
                fix[f"step_{prec}"] = g[(slice(None),) + sl]
                fix[f"step_pi_{prec}"] = qi_
                fix[f"step_pe_{prec}"] = qe_
                # filter alone: currents after the fields in the state file
                j3 = orc.fill_from_interior(cfg, j_in.astype(dt), dt)[:3]
                stj = os.path.join(td, f"statej_{prec}.bin")
                with open(stj, "wb") as fh:
                    fh.write(np.ascontiguousarray(f6, dtype=dt).tobytes())
                    fh.write(np.ascontiguousarray(j3, dtype=dt).tobytes())
                    fh.write(np.array([pi.shape[1], pe.shape[1]], dtype=np.int32).tobytes())
                    for pp in (pi, pe):
                        for r in range(7):
                            fh.write(np.ascontiguousarray(pp[r], dtype=dt).tobytes())
                g, _, _ = run_ref(dim, order, prec, inp, stj, 3, cfg, pi.shape[1], pe.shape[1], maxhlf, "_filter2")
                fix[f"filt_{prec}"] = g[(slice(6, 9),) + sl]
        params = dict(dim=dim, order=order, nx=case["nx"], ny=case["ny"], nz=nz, filter="filter2", **PHYS)
        fix["params"] = np.array(repr(params))
        out = os.path.join(HERE, f"ref_{dim}d_ord{order}_filter2.npz")
        np.savez_compressed(out, **fix)
        print(out, os.path.getsize(out))


def main_zzag():
    """ref_{dim}d_zzag.npz: the reference built with -Dzzag (the original
    trilinear mover, code/particles_movedeposit.F90:98-347, and the zigzag
    deposit, code/particles.F90:550-669): one main-loop step and the mover
    alone (driver mode 1), 2D and 3D; nghost = 5 as for order 1"""
    for dim in (2, 3):
        case = CASES[dim]
        cfg = orc.make_cfg(dim, 1, case["nx"], case["ny"], case["nz"], ntimes=PHYS["ntimes"])
        rng = np.random.default_rng(9100 + dim)
        nz = case["nz"] if dim == 3 else 1
        f_in = rng.uniform(-0.05, 0.05, (6, nz, case["ny"], case["nx"])).astype(np.float32)
        pi = make_particles(rng, cfg, case["npart"])
        pe = make_particles(rng, cfg, case["npart"] + 17)
        maxptl = 20000
        maxhlf = maxptl // 2
        inp = input_text(dim, case["nx"], case["ny"], case["nz"], maxptl)
        fix = dict(f_in=f_in, pi_in=pi, pe_in=pe)
        sl = orc.interior_slices(cfg)
        with tempfile.TemporaryDirectory() as td:
            for prec, dt in (("f32", np.float32), ("f64", np.float64)):
                exe = os.path.join(REF_DIR, f"tristan_ref_{dim}d_{prec}_zzag")
                f6 = orc.fill_from_interior(cfg, f_in.astype(dt), dt)
                st = os.path.join(td, f"state_{prec}.bin")
                write_state(st, f6, pi, pe, dt)
                g, qi_, qe_ = run_ref(dim, 1, prec, inp, st, 0, cfg, pi.shape[1], pe.shape[1], maxhlf, exe=exe)
                fix[f"step_{prec}"] = g[(slice(None),) + sl]
                fix[f"step_pi_{prec}"] = qi_
                fix[f"step_pe_{prec}"] = qe_
                _, qi_, qe_ = run_ref(dim, 1, prec, inp, st, 1, cfg, pi.shape[1], pe.shape[1], maxhlf, exe=exe)
                fix[f"move_pi_{prec}"] = qi_
                fix[f"move_pe_{prec}"] = qe_
        params = dict(dim=dim, order="zzag", nx=case["nx"], ny=case["ny"], nz=nz, **PHYS)
        fix["params"] = np.array(repr(params))
        out = os.path.join(HERE, f"ref_{dim}d_zzag.npz")
        np.savez_compressed(out, **fix)
        print(out, os.path.getsize(out))


if __name__ == "__main__":
    if sys.argv[1:] == ["filter2"]:
        main_filter2()
    elif sys.argv[1:] == ["highorder"]:
        main_highorder()
    elif sys.argv[1:] == ["vay"]:
        main_vay()
    elif sys.argv[1:] == ["zzag"]:
        main_zzag()
    else:
        main()
