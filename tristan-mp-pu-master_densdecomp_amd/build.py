"""Build libtmpg.so (HIP, gfx950) in-tree.

    python build.py            # compile csrc/*.hip -> lib/libtmpg.so
    python build.py --fortran  # also the Fortran host (bin/tristan-mp-gpu)

Each .hip file is compiled separately (in parallel), then linked.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(HERE, "bin")
LIB = os.path.join(LIBDIR, "libtmpg.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("TMPG_FC", "/opt/rocm/bin/amdflang")
ARCH = os.environ.get("TMPG_ARCH", "gfx950")

HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    # packed FP32 instructions stay disabled in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def compile_one(src, objdir=OBJ, defines=()):
    obj = os.path.join(objdir, src.replace(".hip", ".o"))
    cmd = [HIPCC] + HIPFLAGS + [f"-D{d}" for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_lib(jobs=8, phase_stamps=False):
    """libtmpg.so; with phase_stamps, the profiling build libtmpg_phases.so
    (per-phase cycle stamps in the 3D push, read by tmpg_phase_cycles)."""
    objdir = os.path.join(OBJ, "phases") if phase_stamps else OBJ
    lib = os.path.join(LIBDIR, "libtmpg_phases.so") if phase_stamps else LIB
    defines = ("TMPG_PHASE_STAMPS",) if phase_stamps else ()
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda f: compile_one(f, objdir, defines), sources()))
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    return lib


def build_fortran():
    """Fortran host: iso_c_binding module + main program, linked to libtmpg."""
    if not os.path.exists(FC):
        raise RuntimeError(f"Fortran compiler {FC} not found")
    os.makedirs(BINDIR, exist_ok=True)
    fobj = os.path.join(OBJ, "fortran")
    os.makedirs(fobj, exist_ok=True)
    srcs = ["m_tmpg.F90", "m_restart_io.F90", "m_host_setup.F90", "tristan_gpu.F90"]
    objs = []
    for s in srcs:
        o = os.path.join(fobj, s.replace(".F90", ".o"))
        cmd = [FC, "-cpp", "-O2", "-module-dir", fobj, "-I", fobj, "-c", os.path.join(FSRC, s), "-o", o]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        objs.append(o)
    exe = os.path.join(BINDIR, "tristan-mp-gpu")
    cmd = [FC, "-o", exe] + objs + ["-L" + LIBDIR, "-ltmpg", "-Wl,-rpath,$ORIGIN/../lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return exe


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fortran", action="store_true")
    ap.add_argument("--phase-stamps", action="store_true",
                    help="also build lib/libtmpg_phases.so, the push with per-phase cycle stamps")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build_lib(a.j))
    if a.phase_stamps:
        print(build_lib(a.j, phase_stamps=True))
    if a.fortran:
        print(build_fortran())


if __name__ == "__main__":
    sys.exit(main())
