!-------------------------------------------------------------------------------
! m_restart_io -- checkpoint and resume of the Fortran host in the reference's
! restart format.
!
! Input keys (code/restart.F90:95-98): <restart> irestart, intrestart,
! laprestart, lnamedrestartint.  File names (code/restart.F90:140-170):
!   restart/restflds.<jobname>.000.d, restart/restprtl.<jobname>.000.d
! (rank 000, the only rank of a one-GPU run), jobname read from a file named
! "jobname", else "default".  Every lnamedrestartint laps the files are also
! written under restart/restflds.<jobname>.lapNNNNNN.000.d (and restprtl);
! laprestart /= 0 resumes from those named copies.
!
! Each file is one Fortran-unformatted sequential record, written and read
! here through stream access with gfortran's and Intel Fortran's record
! markers (subrecords beyond 2**31 - 9 bytes), so that checkpoints of any
! size are readable by a reference built with either:
!   fields     mx, my, mz, ex, ey, ez, bx, by, bz (mx*my*mz each, ghost cells
!              included), dseed, lap, xinject, xinject2, xinject3, leftwall,
!              walloc  (code/output.F90:2200-2201, read by
!              code/restart.F90:220-227)
!   particles  ions, lecs, maxptl, maxhlf, totalpartnum, then x, y, z, u, v,
!              w, ch, ind, proc, splitlev, each as all ions followed by all
!              electrons  (code/output.F90:2212-2223, read by
!              code/restart.F90:237-247)
! with the reference's types: default integers; real(4) field, particle and
! leftwall values; real(8) dseed, xinject*, walloc.  The two-region Weibel
! setup injects nothing, so dseed, xinject*, leftwall and walloc carry the
! reference's starting values (dseed = 123457 + rank, code/communications.F90:
! 224-231; leftwall = 15, code/particles.F90:298; the others 0) and are not
! used on resume.  maxhlf is written as the larger species count and maxptl
! as twice that (the reference's per-species capacity; its reader skips
! both).  ind holds the particle ids when the context tracks them, else 0;
! proc and splitlev are 0.
!
! Resume (code/restart.F90:209-254): fields and particles are read back into
! the library, and the main loop continues at lap + 1.  The particles are
! binned again by tmpg_set_particles, which takes the place of the
! reorder_particles call that follows the reference's read.
!-------------------------------------------------------------------------------
module m_restart_io
	use, intrinsic :: iso_c_binding
	use m_tmpg
	implicit none

	type restart_params
		integer :: irestart = 0, intrestart = 2000, laprestart = 0, namedrestartint = 1000000
		character(len=64) :: jobname = "default"
		logical :: track_ids = .false.
	end type restart_params

	! one record being written (rec_create/rec_put) or read (rec_open/rec_get)
	integer(8), parameter :: kMaxSub = 2147483639_8   ! 2**31 - 9
	type rec_file
		integer :: u = -1
		integer(8) :: total = 0, done = 0
		integer(8) :: sub_size = 0, sub_left = 0     ! writer: the current subrecord
		logical :: first = .true.
		integer(8), allocatable :: poff(:), plen(:)   ! reader: payload pieces (stream positions, lengths)
		integer :: piece = 1
		integer(8) :: piece_used = 0
	end type rec_file

contains

	subroutine read_restart_input(inp, rp)
		type(c_ptr), intent(in) :: inp
		type(restart_params), intent(inout) :: rp
		integer(c_int) :: iv
		logical :: exst
		integer :: u
		call tmpg_check(tmpg_input_get_int(inp, cstr("restart"), cstr("irestart"), 0_c_int, iv), "irestart", &
			c_null_ptr)
		rp%irestart = iv
		call tmpg_check(tmpg_input_get_int(inp, cstr("restart"), cstr("intrestart"), 2000_c_int, iv), &
			"intrestart", c_null_ptr)
		rp%intrestart = iv
		call tmpg_check(tmpg_input_get_int(inp, cstr("restart"), cstr("laprestart"), 0_c_int, iv), &
			"laprestart", c_null_ptr)
		rp%laprestart = iv
		call tmpg_check(tmpg_input_get_int(inp, cstr("restart"), cstr("lnamedrestartint"), 1000000_c_int, iv), &
			"lnamedrestartint", c_null_ptr)
		rp%namedrestartint = iv
		inquire(file="jobname", exist=exst)
		if (exst) then
			open(newunit=u, file="jobname", form="formatted", action="read")
			read(u, *) rp%jobname
			close(u)
		endif
	end subroutine read_restart_input

	function restart_name(rp, kind, lap) result(name)
		type(restart_params), intent(in) :: rp
		character(len=*), intent(in) :: kind        ! "flds" or "prtl"
		integer, intent(in) :: lap                  ! 0: the unnamed file
		character(len=256) :: name
		character(len=6) :: lapchar
		if (lap == 0) then
			name = "restart/rest"//kind//"."//trim(rp%jobname)//".000.d"
		else
			write(lapchar, "(i6.6)") lap
			name = "restart/rest"//kind//"."//trim(rp%jobname)//".lap"//lapchar//".000.d"
		endif
	end function restart_name

	! grid extent with ghosts: mx = nx + nghost, nghost 5 (order 1) or 7
	subroutine grid_extent(dim, order, nx, ny, nz, mx, my, mz)
		integer, intent(in) :: dim, order, nx, ny, nz
		integer, intent(out) :: mx, my, mz
		integer :: ng
		ng = merge(5, 7, order == 1)
		mx = nx + ng
		my = ny + ng
		mz = merge(nz + ng, 1, dim == 3)
	end subroutine grid_extent

	! code/output.F90:2085-2262: at lap mod intrestart == 0, or when a file
	! named "write_restart" exists (which is then deleted)
	subroutine maybe_write_restart(rp, ctx, lap, dim, order, nx, ny, nz)
		type(restart_params), intent(in) :: rp
		type(c_ptr), intent(in) :: ctx
		integer, intent(in) :: lap, dim, order, nx, ny, nz
		logical :: exst
		integer :: u
		! an interval of 0 or less turns the periodic checkpoints off
		inquire(file="write_restart", exist=exst)
		if (.not. (exst .or. (rp%intrestart > 0 .and. mod(lap, max(rp%intrestart, 1)) == 0))) return
		call write_restart_files(rp, ctx, lap, dim, order, nx, ny, nz, 0)
		if (rp%namedrestartint > 0 .and. mod(lap, max(rp%namedrestartint, 1)) == 0) &
			call write_restart_files(rp, ctx, lap, dim, order, nx, ny, nz, lap)
		if (exst) then
			open(newunit=u, file="write_restart", form="unformatted")
			close(u, status="delete")
		endif
	end subroutine maybe_write_restart

	subroutine write_restart_files(rp, ctx, lap, dim, order, nx, ny, nz, named)
		type(restart_params), intent(in) :: rp
		type(c_ptr), intent(in) :: ctx
		integer, intent(in) :: lap, dim, order, nx, ny, nz, named
		real(4), allocatable, target :: f(:, :)
		real(4), allocatable, target :: pr(:, :, :)
		integer(4), allocatable, target :: ind(:, :)
		integer(c_int64_t) :: cnt(2), got
		integer :: mx, my, mz, s, ions, lecs, maxhlf, ierr
		call execute_command_line("mkdir -p restart", exitstat=ierr)
		call grid_extent(dim, order, nx, ny, nz, mx, my, mz)
		allocate(f(int(mx, 8) * my * mz, 6))
		call tmpg_check(tmpg_get_fields(ctx, f(:, 1), f(:, 2), f(:, 3), f(:, 4), f(:, 5), f(:, 6)), &
			"write_restart: get fields", ctx)
		call write_flds_file(trim(restart_name(rp, "flds", named)), mx, my, mz, f, lap)
		deallocate(f)

		do s = 1, 2
			call tmpg_check(tmpg_count_particles(ctx, int(s - 1, c_int), cnt(s)), "write_restart: count", ctx)
		end do
		! the record's counts are default integers (code/output.F90:2212)
		if (max(cnt(1), cnt(2)) > huge(ions) / 2) then
			write(*, '(a,2(1x,i0))') "write_restart: particle counts beyond the format's integer range:", cnt
			stop 2
		endif
		ions = int(cnt(1))
		lecs = int(cnt(2))
		maxhlf = max(ions, lecs, 1)
		! pr(:, k, s): k = x, y, z, u, v, w, ch of species s
		allocate(pr(maxhlf, 7, 2), ind(maxhlf, 2))
		ind = 0
		do s = 1, 2
			if (cnt(s) == 0) cycle
			call tmpg_check(tmpg_get_particles(ctx, int(s - 1, c_int), cnt(s), got, c_loc(pr(1, 1, s)), &
				c_loc(pr(1, 2, s)), c_loc(pr(1, 3, s)), c_loc(pr(1, 4, s)), c_loc(pr(1, 5, s)), &
				c_loc(pr(1, 6, s)), c_loc(pr(1, 7, s)), merge(c_loc(ind(1, s)), c_null_ptr, rp%track_ids)), &
				"write_restart: get particles", ctx)
			if (got /= cnt(s)) then
				write(*, '(a,i0,a,i0)') "write_restart: got ", got, " particles of ", cnt(s)
				stop 2
			endif
		end do
		call write_prtl_file(trim(restart_name(rp, "prtl", named)), ions, lecs, maxhlf, pr, ind)
		write(*, '(a,i0,a,a)') "wrote restart at lap ", lap, ": ", trim(restart_name(rp, "prtl", named))
	end subroutine write_restart_files

	! code/restart.F90:209-254; returns the lap stored in the file
	subroutine read_restart(rp, ctx, dim, order, nx, ny, nz, lap)
		type(restart_params), intent(in) :: rp
		type(c_ptr), intent(in) :: ctx
		integer, intent(in) :: dim, order, nx, ny, nz
		integer, intent(out) :: lap
		real(4), allocatable, target :: f(:, :)
		real(4), allocatable, target :: pr(:, :)
		integer(4), allocatable, target :: ind(:)
		integer :: mx, my, mz, ions, lecs, s, n, off
		call grid_extent(dim, order, nx, ny, nz, mx, my, mz)
		call read_flds_file(trim(restart_name(rp, "flds", rp%laprestart)), mx, my, mz, f, lap)
		call tmpg_check(tmpg_set_fields(ctx, f(:, 1), f(:, 2), f(:, 3), f(:, 4), f(:, 5), f(:, 6)), &
			"restart: set fields", ctx)
		deallocate(f)
		call read_prtl_file(trim(restart_name(rp, "prtl", rp%laprestart)), ions, lecs, pr, ind)
		do s = 1, 2
			n = merge(ions, lecs, s == 1)
			off = merge(0, ions, s == 1)
			call tmpg_check(tmpg_set_particles(ctx, int(s - 1, c_int), int(n, c_int64_t), &
				c_loc(pr(off + 1, 1)), c_loc(pr(off + 1, 2)), c_loc(pr(off + 1, 3)), c_loc(pr(off + 1, 4)), &
				c_loc(pr(off + 1, 5)), c_loc(pr(off + 1, 6)), c_loc(pr(off + 1, 7)), &
				merge(c_loc(ind(off + 1)), c_null_ptr, rp%track_ids)), "restart: set particles", ctx)
		end do
		write(*, '(a,i0,a,i0,a,i0)') "read restart: lap ", lap, "  ions ", ions, "  electrons ", lecs
	end subroutine read_restart

	! ---- the two records on host arrays ---------------------------------

	! restflds: mx, my, mz, f(:, 1:6) (ex, ey, ez, bx, by, bz), dseed, lap,
	! xinject, xinject2, xinject3, leftwall, walloc with the reference's
	! starting values of the scalars the library does not use
	subroutine write_flds_file(path, mx, my, mz, f, lap)
		character(len=*), intent(in) :: path
		integer, intent(in) :: mx, my, mz, lap
		real(4), intent(in), target :: f(:, :)
		type(rec_file) :: w
		integer(4), target :: hdr(3), lap4
		real(8), target :: dseed, xin(3), walloc
		real(4), target :: leftwall
		integer(8) :: n
		integer :: k
		n = int(mx, 8) * my * mz
		hdr = [mx, my, mz]
		dseed = 123457d0
		lap4 = lap
		xin = 0d0
		leftwall = 15.
		walloc = 0d0
		call rec_create(w, path, 12_8 + 24_8 * n + 8 + 4 + 24 + 4 + 8)
		call rec_put(w, c_loc(hdr), 12_8)
		do k = 1, 6
			call rec_put(w, c_loc(f(1, k)), 4 * n)
		end do
		call rec_put(w, c_loc(dseed), 8_8)
		call rec_put(w, c_loc(lap4), 4_8)
		call rec_put(w, c_loc(xin), 24_8)
		call rec_put(w, c_loc(leftwall), 4_8)
		call rec_put(w, c_loc(walloc), 8_8)
		call rec_close(w)
	end subroutine write_flds_file

	subroutine read_flds_file(path, mx, my, mz, f, lap)
		character(len=*), intent(in) :: path
		integer, intent(in) :: mx, my, mz
		real(4), allocatable, target, intent(out) :: f(:, :)
		integer, intent(out) :: lap
		type(rec_file) :: r
		integer(4), target :: hdr(3), lap4
		real(8), target :: d8(5)
		real(4), target :: leftwall
		integer(8) :: n
		integer :: k
		call rec_open(r, path)
		call rec_get(r, c_loc(hdr), 12_8)
		if (hdr(1) /= mx .or. hdr(2) /= my .or. hdr(3) /= mz) then
			write(*, '(a,3i6,a,3i6)') "restart: grid in file ", hdr, " differs from the input's ", mx, my, mz
			stop 2
		endif
		n = int(mx, 8) * my * mz
		allocate(f(n, 6))
		do k = 1, 6
			call rec_get(r, c_loc(f(1, k)), 4 * n)
		end do
		call rec_get(r, c_loc(d8(1)), 8_8)      ! dseed
		call rec_get(r, c_loc(lap4), 4_8)
		call rec_get(r, c_loc(d8(2)), 24_8)     ! xinject, xinject2, xinject3
		call rec_get(r, c_loc(leftwall), 4_8)
		call rec_get(r, c_loc(d8(5)), 8_8)      ! walloc
		call rec_close(r)
		lap = lap4
	end subroutine read_flds_file

	! restprtl: ions, lecs, maxptl, maxhlf, totalpartnum, then x, y, z, u, v,
	! w, ch, ind, proc, splitlev, each the ions' then the electrons' values;
	! pr(:, k, s) holds quantity k of species s, ind(:, s) the ids; maxhlf is
	! written as n and maxptl as 2 n (the reference's reader skips both);
	! proc and splitlev are 0
	subroutine write_prtl_file(path, ions, lecs, n, pr, ind)
		character(len=*), intent(in) :: path
		integer, intent(in) :: ions, lecs, n
		real(4), intent(in), target :: pr(n, 7, 2)
		integer(4), intent(in), target :: ind(n, 2)
		type(rec_file) :: w
		integer(4), target :: hdr(5)
		integer(4), allocatable, target :: zeros(:)
		integer(8) :: tot, left, take
		integer :: k, q
		if (ions > n .or. lecs > n .or. ions < 0 .or. lecs < 0) call rec_fail("write_prtl_file: bad counts")
		tot = int(ions, 8) + lecs
		hdr = [ions, lecs, 2 * n, n, int(tot, 4)]
		call rec_create(w, path, 20_8 + 40_8 * tot)
		call rec_put(w, c_loc(hdr), 20_8)
		do k = 1, 7
			call rec_put(w, c_loc(pr(1, k, 1)), 4_8 * ions)
			call rec_put(w, c_loc(pr(1, k, 2)), 4_8 * lecs)
		end do
		call rec_put(w, c_loc(ind(1, 1)), 4_8 * ions)
		call rec_put(w, c_loc(ind(1, 2)), 4_8 * lecs)
		allocate(zeros(max(1_8, min(tot, 1048576_8))))
		zeros = 0
		do q = 1, 2
			left = tot
			do while (left > 0)
				take = min(left, int(size(zeros), 8))
				call rec_put(w, c_loc(zeros), 4 * take)
				left = left - take
			end do
		end do
		call rec_close(w)
	end subroutine write_prtl_file

	! pr(1:ions+lecs, k): quantity k, the ions then the electrons; ind the
	! ids; one spare row keeps c_loc(pr(ions + lecs + 1, k)) in bounds
	subroutine read_prtl_file(path, ions, lecs, pr, ind)
		character(len=*), intent(in) :: path
		integer, intent(out) :: ions, lecs
		real(4), allocatable, target, intent(out) :: pr(:, :)
		integer(4), allocatable, target, intent(out) :: ind(:)
		type(rec_file) :: r
		integer(4), target :: hdr(5)
		integer(8) :: tot
		integer :: k
		call rec_open(r, path)
		call rec_get(r, c_loc(hdr), 20_8)
		ions = hdr(1)
		lecs = hdr(2)
		if (ions < 0 .or. lecs < 0) call rec_fail("restart: negative particle count in "//path)
		tot = int(ions, 8) + lecs
		allocate(pr(tot + 1, 7), ind(tot + 1))
		do k = 1, 7
			call rec_get(r, c_loc(pr(1, k)), 4 * tot)
		end do
		call rec_get(r, c_loc(ind), 4 * tot)
		call rec_skip(r, 8 * tot)     ! proc, splitlev
		call rec_close(r)
	end subroutine read_prtl_file

	! ---- sequential unformatted records through stream access ------------
	!
	! The markers of gfortran and Intel Fortran: a record of more than
	! kMaxSub bytes is split into subrecords of at most kMaxSub bytes, each
	! framed by its 4-byte length, the leading length negative when another
	! subrecord follows, the trailing one negative when another precedes.  A
	! record of at most kMaxSub bytes is the plain form of every compiler.
	! The reader also takes the single record whose two lengths hold its byte
	! count modulo 2**32, the form this compiler's runtime gives a record
	! beyond 2 GiB (which gfortran and Intel Fortran do not read).

	subroutine rec_fail(msg)
		character(len=*), intent(in) :: msg
		write(*, '(a)') "restart record: "//msg
		stop 2
	end subroutine rec_fail

	subroutine rec_create(w, path, total)
		type(rec_file), intent(inout) :: w
		character(len=*), intent(in) :: path
		integer(8), intent(in) :: total
		integer :: ios
		open(newunit=w%u, file=path, access="stream", form="unformatted", status="replace", action="write", &
			iostat=ios)
		if (ios /= 0) call rec_fail("cannot create "//path)
		w%total = total
		w%done = 0
		w%first = .true.
		call rec_begin_sub(w)
	end subroutine rec_create

	subroutine rec_begin_sub(w)
		type(rec_file), intent(inout) :: w
		integer(4) :: m
		w%sub_size = min(kMaxSub, w%total - w%done)
		w%sub_left = w%sub_size
		m = int(w%sub_size, 4)
		if (w%done + w%sub_size < w%total) m = -m
		write(w%u) m
	end subroutine rec_begin_sub

	subroutine rec_end_sub(w)
		type(rec_file), intent(inout) :: w
		integer(4) :: m
		m = int(w%sub_size, 4)
		if (.not. w%first) m = -m
		write(w%u) m
		w%first = .false.
	end subroutine rec_end_sub

	subroutine rec_put(w, p, nbytes)
		type(rec_file), intent(inout) :: w
		type(c_ptr), intent(in) :: p
		integer(8), intent(in) :: nbytes
		integer(1), pointer :: b(:)
		integer(8) :: off, take
		if (nbytes <= 0) return
		if (w%done + nbytes > w%total) call rec_fail("write past the record's length")
		call c_f_pointer(p, b, [nbytes])
		off = 0
		do while (off < nbytes)
			if (w%sub_left == 0) call rec_begin_sub(w)
			take = min(nbytes - off, w%sub_left)
			write(w%u) b(off + 1:off + take)
			off = off + take
			w%sub_left = w%sub_left - take
			w%done = w%done + take
			if (w%sub_left == 0) call rec_end_sub(w)
		end do
	end subroutine rec_put

	subroutine rec_open(r, path)
		type(rec_file), intent(inout) :: r
		character(len=*), intent(in) :: path
		integer(8), parameter :: two32 = 4294967296_8
		integer(8) :: fsize, pos, sz
		integer(8), allocatable :: o(:), l(:), t8(:)
		integer(4) :: m, t
		integer :: n, ios
		logical :: ok
		open(newunit=r%u, file=path, access="stream", form="unformatted", status="old", action="read", &
			iostat=ios)
		if (ios /= 0) call rec_fail("cannot open "//path)
		inquire(unit=r%u, size=fsize)
		allocate(o(8), l(8))
		n = 0
		pos = 1
		ok = .false.
		do
			if (pos + 7 > fsize) exit
			read(r%u, pos=pos) m
			sz = abs(int(m, 8))
			if (pos + 7 + sz > fsize) exit
			read(r%u, pos=pos + 4 + sz) t
			if (abs(int(t, 8)) /= sz) exit
			if (n == size(o)) then
				allocate(t8(2 * n))
				t8(1:n) = o
				call move_alloc(t8, o)
				allocate(t8(2 * n))
				t8(1:n) = l
				call move_alloc(t8, l)
			endif
			n = n + 1
			o(n) = pos + 4
			l(n) = sz
			pos = pos + 8 + sz
			if (m >= 0) then
				ok = .true.
				exit
			endif
		end do
		if (.not. ok .and. fsize >= 8) then
			read(r%u, pos=1) m
			read(r%u, pos=fsize - 3) t
			if (m == t .and. modulo(fsize - 8, two32) == modulo(int(m, 8), two32)) then
				n = 1
				o(1) = 5
				l(1) = fsize - 8
				ok = .true.
			endif
		endif
		if (.not. ok) call rec_fail("not a sequential unformatted record: "//path)
		r%poff = o(1:n)
		r%plen = l(1:n)
		r%total = sum(l(1:n))
		r%done = 0
		r%piece = 1
		r%piece_used = 0
	end subroutine rec_open

	subroutine rec_read(r, p, nbytes, skip)
		type(rec_file), intent(inout) :: r
		type(c_ptr), intent(in) :: p
		integer(8), intent(in) :: nbytes
		logical, intent(in) :: skip
		integer(1), pointer :: b(:)
		integer(8) :: off, take, avail
		if (nbytes <= 0) return
		if (r%done + nbytes > r%total) call rec_fail("the record is shorter than its contents")
		if (.not. skip) call c_f_pointer(p, b, [nbytes])
		off = 0
		do while (off < nbytes)
			avail = r%plen(r%piece) - r%piece_used
			if (avail == 0) then
				r%piece = r%piece + 1
				r%piece_used = 0
				cycle
			endif
			take = min(avail, nbytes - off)
			if (.not. skip) read(r%u, pos=r%poff(r%piece) + r%piece_used) b(off + 1:off + take)
			off = off + take
			r%piece_used = r%piece_used + take
			r%done = r%done + take
		end do
	end subroutine rec_read

	subroutine rec_get(r, p, nbytes)
		type(rec_file), intent(inout) :: r
		type(c_ptr), intent(in) :: p
		integer(8), intent(in) :: nbytes
		call rec_read(r, p, nbytes, .false.)
	end subroutine rec_get

	subroutine rec_skip(r, nbytes)
		type(rec_file), intent(inout) :: r
		integer(8), intent(in) :: nbytes
		call rec_read(r, c_null_ptr, nbytes, .true.)
	end subroutine rec_skip

	! a writer checks that it wrote the declared length (and frames an empty
	! record); a reader that it consumed the whole record
	subroutine rec_close(r)
		type(rec_file), intent(inout) :: r
		if (r%done /= r%total) call rec_fail("record length and contents differ")
		if (.not. allocated(r%poff) .and. r%total == 0) call rec_end_sub(r)
		close(r%u)
		r%u = -1
	end subroutine rec_close

end module m_restart_io
