"""The restart-file reader and writer (python/restart_io.py) against the
Fortran compiler's own unformatted records.

A small Fortran program writes the two records with the item lists and types
of the reference's write_restart (code/output.F90:2200-2201, 2212-2223) and
reads them back with the item lists of its restart() (code/restart.F90:
220-247, implied-do loops over the extents read from the same record).  The
Python reader must recover what the program wrote, and the program must
recover what the Python writer wrote.  The values are chosen so that a
swapped item, a wrong type width or a wrong species order changes them."""
import os
import subprocess

import numpy as np
import pytest

import restart_io as rio

FC = os.environ.get("TMPG_FC", "/opt/rocm/bin/amdflang")

PROGRAM = r"""
program rt
	implicit none
	character(len=512) :: mode, ffld, fprt
	integer :: mx, my, mz, i, j, k, n, lap, ions, lecs, maxptl, maxhlf, total, idum
	real(4), allocatable :: ex(:,:,:), ey(:,:,:), ez(:,:,:), bx(:,:,:), by(:,:,:), bz(:,:,:)
	real(4), allocatable :: px(:), py(:), pz(:), pu(:), pv(:), pw(:), pch(:)
	integer, allocatable :: pind(:), pproc(:), psplit(:)
	real(8) :: dseed, xinject, xinject2, xinject3, walloc
	real(4) :: leftwall
	call get_command_argument(1, mode)
	call get_command_argument(2, ffld)
	call get_command_argument(3, fprt)
	mx = 9; my = 7; mz = 5; ions = 11; lecs = 6; maxhlf = 20; maxptl = 40
	allocate(ex(mx,my,mz), ey(mx,my,mz), ez(mx,my,mz), bx(mx,my,mz), by(mx,my,mz), bz(mx,my,mz))
	! particles stored as the reference does: ions at 1..ions, electrons at maxhlf+1..
	allocate(px(maxptl), py(maxptl), pz(maxptl), pu(maxptl), pv(maxptl), pw(maxptl), pch(maxptl))
	allocate(pind(maxptl), pproc(maxptl), psplit(maxptl))
	if (trim(mode) == "write") then
		do k = 1, mz
			do j = 1, my
				do i = 1, mx
					ex(i,j,k) = i + 100*j + 10000*k
					ey(i,j,k) = -ex(i,j,k); ez(i,j,k) = 0.5*ex(i,j,k)
					bx(i,j,k) = 2*ex(i,j,k); by(i,j,k) = 3*ex(i,j,k); bz(i,j,k) = 4*ex(i,j,k)
				end do
			end do
		end do
		do n = 1, maxptl
			px(n) = n; py(n) = n + 0.25; pz(n) = n + 0.5; pu(n) = -n; pv(n) = 2*n; pw(n) = 3*n
			pch(n) = 1 + n/8.; pind(n) = 1000 + n; pproc(n) = 7; psplit(n) = n/3
		end do
		dseed = 123457.5d0; lap = 4321; xinject = 1.25d0; xinject2 = 2.5d0; xinject3 = 3.75d0
		leftwall = 15.5; walloc = 6.125d0
		open(unit=7, file=trim(ffld), form='unformatted')
		write(7) mx, my, mz, ex, ey, ez, bx, by, bz, dseed, lap, xinject, xinject2, xinject3, leftwall, walloc
		close(7)
		total = ions + lecs
		open(unit=8, file=trim(fprt), form='unformatted')
		write(8) ions, lecs, maxptl, maxhlf, total, &
			(px(n),n=1,ions), (px(n),n=maxhlf+1,maxhlf+lecs), (py(n),n=1,ions), (py(n),n=maxhlf+1,maxhlf+lecs), &
			(pz(n),n=1,ions), (pz(n),n=maxhlf+1,maxhlf+lecs), (pu(n),n=1,ions), (pu(n),n=maxhlf+1,maxhlf+lecs), &
			(pv(n),n=1,ions), (pv(n),n=maxhlf+1,maxhlf+lecs), (pw(n),n=1,ions), (pw(n),n=maxhlf+1,maxhlf+lecs), &
			(pch(n),n=1,ions), (pch(n),n=maxhlf+1,maxhlf+lecs), &
			(pind(n),n=1,ions), (pind(n),n=maxhlf+1,maxhlf+lecs), &
			(pproc(n),n=1,ions), (pproc(n),n=maxhlf+1,maxhlf+lecs), &
			(psplit(n),n=1,ions), (psplit(n),n=maxhlf+1,maxhlf+lecs)
		close(8)
	else
		ex = 0; ey = 0; ez = 0; bx = 0; by = 0; bz = 0
		open(unit=7, file=trim(ffld), form='unformatted')
		read(7) mx, my, mz, (((ex(i,j,k),i=1,mx),j=1,my),k=1,mz), (((ey(i,j,k),i=1,mx),j=1,my),k=1,mz), &
			(((ez(i,j,k),i=1,mx),j=1,my),k=1,mz), (((bx(i,j,k),i=1,mx),j=1,my),k=1,mz), &
			(((by(i,j,k),i=1,mx),j=1,my),k=1,mz), (((bz(i,j,k),i=1,mx),j=1,my),k=1,mz), &
			dseed, lap, xinject, xinject2, xinject3, leftwall, walloc
		close(7)
		open(unit=8, file=trim(fprt), form='unformatted')
		read(8) ions, lecs, idum, idum, total, &
			(px(n),n=1,ions), (px(n),n=maxhlf+1,maxhlf+lecs), (py(n),n=1,ions), (py(n),n=maxhlf+1,maxhlf+lecs), &
			(pz(n),n=1,ions), (pz(n),n=maxhlf+1,maxhlf+lecs), (pu(n),n=1,ions), (pu(n),n=maxhlf+1,maxhlf+lecs), &
			(pv(n),n=1,ions), (pv(n),n=maxhlf+1,maxhlf+lecs), (pw(n),n=1,ions), (pw(n),n=maxhlf+1,maxhlf+lecs), &
			(pch(n),n=1,ions), (pch(n),n=maxhlf+1,maxhlf+lecs), &
			(pind(n),n=1,ions), (pind(n),n=maxhlf+1,maxhlf+lecs), &
			(pproc(n),n=1,ions), (pproc(n),n=maxhlf+1,maxhlf+lecs), &
			(psplit(n),n=1,ions), (psplit(n),n=maxhlf+1,maxhlf+lecs)
		close(8)
		print '(3i6)', mx, my, mz
		print '(f14.3,i8,3f10.4,f8.3,f10.4)', dseed, lap, xinject, xinject2, xinject3, leftwall, walloc
		print '(6es16.8)', sum(real(ex,8)), sum(real(ey,8)), sum(real(ez,8)), sum(real(bx,8)), sum(real(by,8)), &
			sum(real(bz,8))
		print '(es16.8)', ex(2,3,4)
		print '(2i6,i8)', ions, lecs, total
		print '(es16.8,i8)', px(2), pind(maxhlf+3)
		print '(7es16.8)', sum(px(1:ions)), sum(pu(maxhlf+1:maxhlf+lecs)), sum(pch(1:ions)), &
			sum(pw(maxhlf+1:maxhlf+lecs)), sum(py(1:ions)), sum(pz(maxhlf+1:maxhlf+lecs)), sum(pv(1:ions))
		print '(3i10)', sum(pind(1:ions)), sum(pproc(maxhlf+1:maxhlf+lecs)), sum(psplit(1:ions))
	endif
end program rt
"""


@pytest.fixture(scope="module")
def fortran_rt(tmp_path_factory):
    if not os.path.exists(FC):
        pytest.skip(f"Fortran compiler {FC} not found")
    d = tmp_path_factory.mktemp("rt")
    src = d / "rt.F90"
    src.write_text(PROGRAM)
    exe = d / "rt"
    r = subprocess.run([FC, "-O1", str(src), "-o", str(exe)], capture_output=True, text=True, cwd=d)
    assert r.returncode == 0, r.stderr
    return exe


def _expected():
    mx, my, mz = 9, 7, 5
    k, j, i = np.meshgrid(np.arange(1, mz + 1), np.arange(1, my + 1), np.arange(1, mx + 1), indexing="ij")
    ex = (i + 100 * j + 10000 * k).astype(np.float32)
    f = dict(ex=ex, ey=-ex, ez=np.float32(0.5) * ex, bx=2 * ex, by=3 * ex, bz=4 * ex)
    n = np.arange(1, 41)
    p = dict(x=n.astype(np.float32), y=(n + 0.25).astype(np.float32), z=(n + 0.5).astype(np.float32),
             u=(-n).astype(np.float32), v=(2 * n).astype(np.float32), w=(3 * n).astype(np.float32),
             ch=(1 + n / 8.0).astype(np.float32), ind=(1000 + n).astype(np.int32), proc=np.full(40, 7, np.int32),
             splitlev=(n // 3).astype(np.int32))
    ions = {k: v[:11] for k, v in p.items()}
    lecs = {k: v[20:26] for k, v in p.items()}
    return f, ions, lecs


def test_reader_reads_fortran_records(fortran_rt, tmp_path):
    ffld, fprt = tmp_path / "restflds.default.000.d", tmp_path / "restprtl.default.000.d"
    r = subprocess.run([str(fortran_rt), "write", str(ffld), str(fprt)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    f = rio.read_fields(ffld)
    fe, ie, le = _expected()
    assert (f["mx"], f["my"], f["mz"]) == (9, 7, 5)
    for k in rio.FIELD_NAMES:
        assert np.array_equal(f[k], fe[k]), k
    assert (f["dseed"], f["lap"], f["xinject"], f["xinject2"], f["xinject3"], f["leftwall"], f["walloc"]) == \
        (123457.5, 4321, 1.25, 2.5, 3.75, 15.5, 6.125)
    p = rio.read_particles(fprt)
    assert (p["ions"], p["lecs"], p["maxptl"], p["maxhlf"], p["totalpartnum"]) == (11, 6, 40, 20, 17)
    for k in rio.PART_REALS + rio.PART_INTS:
        assert np.array_equal(p["species"][0][k], ie[k]), k
        assert np.array_equal(p["species"][1][k], le[k]), k


def test_fortran_reads_python_records(fortran_rt, tmp_path):
    fe, ie, le = _expected()
    ffld, fprt = tmp_path / "f.d", tmp_path / "p.d"
    rio.write_fields(ffld, fe, lap=4321, dseed=123457.5, xinject=1.25, xinject2=2.5, xinject3=3.75, leftwall=15.5,
                     walloc=6.125)
    rio.write_particles(fprt, ie, le, maxptl=40, maxhlf=20)
    r = subprocess.run([str(fortran_rt), "read", str(ffld), str(fprt)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    assert lines[0].split() == ["9", "7", "5"]
    v = lines[1].split()
    assert float(v[0]) == 123457.5 and int(v[1]) == 4321 and float(v[5]) == 15.5 and float(v[6]) == 6.125
    sums = [float(s) for s in lines[2].split()]
    np.testing.assert_allclose(sums, [fe[k].astype(np.float64).sum() for k in rio.FIELD_NAMES], rtol=1e-7)
    assert float(lines[3]) == fe["ex"][3, 2, 1]          # ex(2,3,4): x=2, y=3, z=4
    assert lines[4].split() == ["11", "6", "17"]
    a, b = lines[5].split()
    assert float(a) == ie["x"][1] and int(b) == le["ind"][2]
    got = [float(s) for s in lines[6].split()]
    want = [ie["x"].sum(), le["u"].sum(), ie["ch"].sum(), le["w"].sum(), ie["y"].sum(), le["z"].sum(), ie["v"].sum()]
    np.testing.assert_allclose(got, want, rtol=1e-6)
    assert [int(s) for s in lines[7].split()] == [int(ie["ind"].sum()), int(le["proc"].sum()),
                                                  int(ie["splitlev"].sum())]


def test_round_trip_and_errors(tmp_path):
    rng = np.random.default_rng(5)
    f = {k: rng.standard_normal((3, 4, 6)).astype(np.float32) for k in rio.FIELD_NAMES}
    rio.write_fields(tmp_path / "f", f, lap=12)
    g = rio.read_fields(tmp_path / "f")
    assert g["lap"] == 12 and g["dseed"] == 123457.0 and g["leftwall"] == 15.0
    for k in rio.FIELD_NAMES:
        assert np.array_equal(g[k], f[k])
    ions = {k: rng.random(5).astype(np.float32) for k in rio.PART_REALS}
    ions["ind"] = np.arange(5, dtype=np.int32)
    lecs = {k: np.zeros(0, np.float32) for k in rio.PART_REALS}       # an empty species
    rio.write_particles(tmp_path / "p", ions, lecs)
    p = rio.read_particles(tmp_path / "p")
    assert (p["ions"], p["lecs"], p["maxhlf"], p["maxptl"], p["totalpartnum"]) == (5, 0, 5, 10, 5)
    assert np.array_equal(p["species"][0]["ind"], ions["ind"]) and p["species"][1]["x"].size == 0
    assert np.array_equal(p["species"][0]["proc"], np.zeros(5, np.int32))
    # a truncated file and a record with bytes left over are errors
    raw = (tmp_path / "f").read_bytes()
    (tmp_path / "t").write_bytes(raw[:-8])
    with pytest.raises(ValueError):
        rio.read_fields(tmp_path / "t")
    with pytest.raises(ValueError):
        rio.read_particles(tmp_path / "f")


def test_subrecords_are_joined(tmp_path):
    # the continuation form of long records: negative markers on all but the last piece
    payload = np.arange(10, dtype="<i4").tobytes()
    pieces = [payload[:12], payload[12:28], payload[28:]]
    with open(tmp_path / "s", "wb") as fh:
        for i, pc in enumerate(pieces):
            m = np.int32(len(pc) if i == len(pieces) - 1 else -len(pc)).tobytes()
            fh.write(m + pc + m)
    assert rio.read_record(tmp_path / "s") == payload
