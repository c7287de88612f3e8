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
! Each file is one Fortran-unformatted sequential record:
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
		integer :: mx, my, mz, u, s, ions, lecs, maxhlf, ierr
		integer :: lap_out
		real(8) :: dseed, xinject, xinject2, xinject3, walloc
		real(4) :: leftwall
		call execute_command_line("mkdir -p restart", exitstat=ierr)
		call grid_extent(dim, order, nx, ny, nz, mx, my, mz)
		allocate(f(int(mx, 8) * my * mz, 6))
		call tmpg_check(tmpg_get_fields(ctx, f(:, 1), f(:, 2), f(:, 3), f(:, 4), f(:, 5), f(:, 6)), &
			"write_restart: get fields", ctx)
		dseed = 123457d0
		xinject = 0d0
		xinject2 = 0d0
		xinject3 = 0d0
		walloc = 0d0
		leftwall = 15.
		lap_out = lap
		open(newunit=u, file=trim(restart_name(rp, "flds", named)), form="unformatted", access="sequential", &
			status="replace", action="write")
		write(u) mx, my, mz, f(:, 1), f(:, 2), f(:, 3), f(:, 4), f(:, 5), f(:, 6), dseed, lap_out, xinject, &
			xinject2, xinject3, leftwall, walloc
		close(u)
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
		maxhlf = max(ions, lecs)
		! pr(:, k, s): k = x, y, z, u, v, w, ch of species s
		allocate(pr(max(maxhlf, 1), 7, 2), ind(max(maxhlf, 1), 2))
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
		open(newunit=u, file=trim(restart_name(rp, "prtl", named)), form="unformatted", access="sequential", &
			status="replace", action="write")
		write(u) ions, lecs, 2 * maxhlf, maxhlf, ions + lecs, &
			pr(1:ions, 1, 1), pr(1:lecs, 1, 2), pr(1:ions, 2, 1), pr(1:lecs, 2, 2), &
			pr(1:ions, 3, 1), pr(1:lecs, 3, 2), pr(1:ions, 4, 1), pr(1:lecs, 4, 2), &
			pr(1:ions, 5, 1), pr(1:lecs, 5, 2), pr(1:ions, 6, 1), pr(1:lecs, 6, 2), &
			pr(1:ions, 7, 1), pr(1:lecs, 7, 2), ind(1:ions, 1), ind(1:lecs, 2), &
			(0, s = 1, ions + lecs), (0, s = 1, ions + lecs)
		close(u)
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
		integer :: mx, my, mz, mxr, myr, mzr, u, ions, lecs, idum, tot, k, s, n, off
		real(8) :: dseed, xinject, xinject2, xinject3, walloc
		real(4) :: leftwall
		character(len=256) :: ffld, fprt
		ffld = restart_name(rp, "flds", rp%laprestart)
		fprt = restart_name(rp, "prtl", rp%laprestart)
		call grid_extent(dim, order, nx, ny, nz, mx, my, mz)
		open(newunit=u, file=trim(ffld), form="unformatted", access="sequential", status="old", action="read")
		read(u) mxr, myr, mzr
		if (mxr /= mx .or. myr /= my .or. mzr /= mz) then
			write(*, '(a,3i6,a,3i6)') "restart: grid in file ", mxr, myr, mzr, " differs from the input's ", &
				mx, my, mz
			stop 2
		endif
		rewind(u)
		allocate(f(int(mx, 8) * my * mz, 6))
		read(u) mxr, myr, mzr, f(:, 1), f(:, 2), f(:, 3), f(:, 4), f(:, 5), f(:, 6), dseed, lap, xinject, &
			xinject2, xinject3, leftwall, walloc
		close(u)
		call tmpg_check(tmpg_set_fields(ctx, f(:, 1), f(:, 2), f(:, 3), f(:, 4), f(:, 5), f(:, 6)), &
			"restart: set fields", ctx)
		deallocate(f)

		open(newunit=u, file=trim(fprt), form="unformatted", access="sequential", status="old", action="read")
		read(u) ions, lecs
		rewind(u)
		tot = ions + lecs
		! pr(:, k): k = x, y, z, u, v, w, ch over ions then electrons (one
		! spare row keeps c_loc(pr(off + 1, k)) in bounds for an empty species)
		allocate(pr(tot + 1, 7), ind(tot + 1))
		read(u) ions, lecs, idum, idum, idum, ((pr(n, k), n = 1, tot), k = 1, 7), (ind(n), n = 1, tot)
		close(u)
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

end module m_restart_io
