"""Restart records beyond 2 GiB: a restprtl record of 60 M particles
(2.4e9 bytes, code/output.F90:2212-2223) through the Fortran host's record
layer (m_restart_io write_prtl_file / read_prtl_file) and the Python reader
and writer (python/restart_io.py).

* the Fortran writer splits the record into subrecords of at most 2**31 - 9
  bytes with gfortran's and Intel Fortran's marker signs (leading negative
  when continued, trailing negative when continuing), and the Python reader
  recovers every value;
* the Python writer's subrecords are read back by the Fortran reader;
* a record written by the compiler's own runtime (one record, markers modulo
  2**32) is read by both readers.
The values depend on the particle index, the quantity and the species, so a
subrecord seam in the wrong place, a dropped or doubled byte or a swapped
species shows as a mismatch.  Runs on the CPU (host code only)."""
import os
import subprocess

import numpy as np
import pytest

import restart_io as rio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tristan-mp-pu-master_densdecomp_amd")
FOBJ = os.path.join(PKG, "_build", "fortran")
LIBDIR = os.path.join(PKG, "lib")
FC = os.environ.get("TMPG_FC", "/opt/rocm/bin/amdflang")
IONS, LECS = 31_000_000, 29_000_000
MAXSUB = 2 ** 31 - 9

PROGRAM = r"""
program big
	use, intrinsic :: iso_c_binding
	use m_restart_io
	implicit none
	character(len=512) :: mode, path
	integer :: ions, lecs, n, k, q, s, u, ri, rl
	real(4), allocatable, target :: pr(:, :, :), rp(:, :)
	integer(4), allocatable, target :: ind(:, :), rind(:)
	integer(8) :: bad
	call get_command_argument(1, mode)
	call get_command_argument(2, path)
	ions = IONS_; lecs = LECS_
	n = max(ions, lecs)
	if (trim(mode) == "write" .or. trim(mode) == "native") then
		allocate(pr(n, 7, 2), ind(n, 2))
		do s = 1, 2
			do k = 1, 7
				do q = 1, n
					pr(q, k, s) = real(mod(7 * q + 131 * k + 1009 * s, 1048576), 4) * 0.5
				end do
			end do
			do q = 1, n
				ind(q, s) = q + 100000000 * s
			end do
		end do
		if (trim(mode) == "write") then
			call write_prtl_file(trim(path), ions, lecs, n, pr, ind)
		else
			allocate(rind(2 * (ions + lecs)))
			rind = 0
			open(newunit=u, file=trim(path), form="unformatted", access="sequential", status="replace")
			write(u) ions, lecs, 2 * n, n, ions + lecs, pr(1:ions, 1, 1), pr(1:lecs, 1, 2), &
				pr(1:ions, 2, 1), pr(1:lecs, 2, 2), pr(1:ions, 3, 1), pr(1:lecs, 3, 2), &
				pr(1:ions, 4, 1), pr(1:lecs, 4, 2), pr(1:ions, 5, 1), pr(1:lecs, 5, 2), &
				pr(1:ions, 6, 1), pr(1:lecs, 6, 2), pr(1:ions, 7, 1), pr(1:lecs, 7, 2), &
				ind(1:ions, 1), ind(1:lecs, 2), rind
			close(u)
		endif
		print *, "wrote"
	else
		call read_prtl_file(trim(path), ri, rl, rp, rind)
		bad = 0
		do k = 1, 7
			do q = 1, ri + rl
				s = merge(1, 2, q <= ri)
				n = merge(q, q - ri, q <= ri)
				if (rp(q, k) /= real(mod(7 * n + 131 * k + 1009 * s, 1048576), 4) * 0.5) bad = bad + 1
			end do
		end do
		do q = 1, ri + rl
			s = merge(1, 2, q <= ri)
			n = merge(q, q - ri, q <= ri)
			if (rind(q) /= n + 100000000 * s) bad = bad + 1
		end do
		print '(a,i0,a,i0,a,i0)', "read ions ", ri, " lecs ", rl, " bad ", bad
	endif
end program big
""".replace("IONS_", str(IONS)).replace("LECS_", str(LECS))


def _expected(k, s, n):
    q = np.arange(1, n + 1, dtype=np.int64)
    return (((7 * q + 131 * k + 1009 * s) % 1048576).astype(np.float32) * np.float32(0.5))


def _check_python_read(path):
    d = rio.read_particles(path)
    assert (d["ions"], d["lecs"]) == (IONS, LECS)
    assert d["totalpartnum"] == IONS + LECS
    for s, n in ((1, IONS), (2, LECS)):
        sp = d["species"][s - 1]
        for k, name in enumerate(rio.PART_REALS, start=1):
            assert np.array_equal(sp[name], _expected(k, s, n)), (name, s)
        assert np.array_equal(sp["ind"], np.arange(1, n + 1, dtype=np.int32) + 100000000 * s)
        assert not sp["proc"].any() and not sp["splitlev"].any()


def _markers(path):
    size = os.path.getsize(path)
    out, pos = [], 0
    with open(path, "rb") as f:
        while pos < size:
            f.seek(pos)
            lead = int(np.frombuffer(f.read(4), "<i4")[0])
            f.seek(pos + 4 + abs(lead))
            tail = int(np.frombuffer(f.read(4), "<i4")[0])
            out.append((lead, tail))
            pos += 8 + abs(lead)
    return out


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    if not (os.path.exists(FC) and os.path.exists(os.path.join(FOBJ, "m_restart_io.o"))):
        pytest.skip("Fortran host objects not built")
    d = tmp_path_factory.mktemp("bigrec")
    src = d / "big.F90"
    src.write_text(PROGRAM)
    exe = d / "big"
    r = subprocess.run([FC, "-O2", "-I", FOBJ, str(src), os.path.join(FOBJ, "m_restart_io.o"),
                        os.path.join(FOBJ, "m_tmpg.o"), "-L" + LIBDIR, "-ltmpg", "-Wl,-rpath," + LIBDIR,
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe, mode, path):
    r = subprocess.run([str(exe), mode, str(path)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_fortran_writes_subrecords_python_reads(prog, tmp_path):
    path = tmp_path / "restprtl.default.000.d"
    _run(prog, "write", path)
    payload = 20 + 40 * (IONS + LECS)
    assert payload > 2 ** 31
    rest = payload - MAXSUB
    assert _markers(path) == [(-MAXSUB, MAXSUB), (rest, -rest)]
    assert os.path.getsize(path) == payload + 16
    _check_python_read(path)
    os.remove(path)


def test_python_writes_subrecords_fortran_reads(prog, tmp_path):
    path = tmp_path / "restprtl.py.000.d"
    sp = []
    for s, n in ((1, IONS), (2, LECS)):
        d = {name: _expected(k, s, n) for k, name in enumerate(rio.PART_REALS, start=1)}
        d["ind"] = np.arange(1, n + 1, dtype=np.int32) + 100000000 * s
        sp.append(d)
    rio.write_particles(path, sp[0], sp[1])
    del sp
    rest = 20 + 40 * (IONS + LECS) - MAXSUB
    assert _markers(path) == [(-MAXSUB, MAXSUB), (rest, -rest)]
    out = _run(prog, "read", path)
    assert f"read ions {IONS} lecs {LECS} bad 0" in out
    os.remove(path)


def test_runtime_native_record_read_by_both(prog, tmp_path):
    path = tmp_path / "restprtl.native.000.d"
    _run(prog, "native", path)
    payload = 20 + 40 * (IONS + LECS)
    assert os.path.getsize(path) == payload + 8
    m = int(np.frombuffer(open(path, "rb").read(4), "<i4")[0])
    assert m % 2 ** 32 == payload % 2 ** 32
    out = _run(prog, "read", path)
    assert f"read ions {IONS} lecs {LECS} bad 0" in out
    _check_python_read(path)
    os.remove(path)
