!-------------------------------------------------------------------------------
! tristan-mp-gpu -- Fortran host of the MI355X per-step PIC loop.
!
!   tristan-mp-gpu [-i input] [-o energy-file]
!
! The program keeps the reference's shape (code/tristan.F90:33-37,
! code/tristanmainloop.F90:90-346): initialize, then one main-loop iteration
! per lap, with diagnostics every <output> interval laps.  All grid and
! particle work runs in libtmpg on the GPU through m_tmpg.  With <gpu>
! fused = 0 the lap issues the reference's call sequence one routine at a
! time; fused = 1 (default) issues tmpg_step, which runs the same sequence
! with move+deposit and E+add_current fused.
!-------------------------------------------------------------------------------
program tristan_gpu
	use, intrinsic :: iso_c_binding
	use m_tmpg
	use m_host_setup
	use m_restart_io
	implicit none
	type(host_params) :: hp
	type(tmpg_config) :: cfg
	type(c_ptr) :: ctx
	character(len=512) :: infile, outfile, arg
	integer :: lap, k, nargs, u, lapst
	real(c_double) :: d(8), ms(6), gs(4)
	integer(c_int64_t) :: ni, ne
	real(8) :: t0, t1, tacc
	integer(8) :: c0, c1, crate

	infile = "input"
	outfile = ""
	nargs = command_argument_count()
	k = 1
	do while (k <= nargs)
		call get_command_argument(k, arg)
		select case (trim(arg))
		case ("-i", "--input")
			k = k + 1
			call get_command_argument(k, infile)
		case ("-o", "--output")
			k = k + 1
			call get_command_argument(k, outfile)
		end select
		k = k + 1
	end do

	! initialize() (code/initialize.F90:113-175)
	call read_host_input(trim(infile), hp)
	call make_config(hp, cfg)
	call tmpg_check(tmpg_create(cfg, ctx), "tmpg_create", c_null_ptr)
	lapst = 1
	if (hp%rst%irestart == 1) then
		! code/initialize.F90:155,165: a resume reads the state instead of loading it
		call read_restart(hp%rst, ctx, hp%dim, hp%order, hp%nx, hp%ny, hp%nz, lap)
		lapst = lap + 1
	else
		call init_particle_distribution_user(hp, ctx)  ! fields start at zero (user_weibel init_EMfields)
	endif
	call tmpg_check(tmpg_count_particles(ctx, 0_c_int, ni), "count", ctx)
	call tmpg_check(tmpg_count_particles(ctx, 1_c_int, ne), "count", ctx)
	write(*, '(a,i0,a,i0,a,i0,a,i0,a,i0,a,i0)') "grid ", hp%nx, " x ", hp%ny, " x ", hp%nz, &
		"  order ", hp%order, "  ions ", ni, "  electrons ", ne
	write(*, '(a,a,a,a)') "field solver ", trim(merge("4th order (highorder)", "2nd order            ", &
		hp%highorder /= 0)), "  pusher ", trim(merge("Vay  ", "Boris", hp%pusher == 1))
	u = 0
	if (len_trim(outfile) > 0) then
		open(newunit=u, file=trim(outfile), status="replace", action="write")
		write(u, '(a)') "# lap  E_energy  B_energy  KE_ions  KE_electrons"
	endif

	! mainloop (code/tristanmainloop.F90:107-344)
	call system_clock(c0, crate)
	tacc = 0d0
	! Diagnostics() runs at lap 0 and at the top of every lap, before its
	! step (code/tristanmainloop.F90:102-109); a resumed run skips lap 0
	if (lapst == 1) call maybe_write_flds_tot(hp, ctx, 0)
	do lap = lapst, hp%last
		call maybe_write_flds_tot(hp, ctx, lap)
		if (hp%fused /= 0) then
			call tmpg_check(tmpg_step(ctx, 1_c_int), "tmpg_step", ctx)
		else
			call tmpg_check(tmpg_advance_b_half(ctx), "advance_Bhalfstep", ctx)
			call tmpg_check(tmpg_move(ctx), "move_particles", ctx)
			call tmpg_check(tmpg_advance_b_half(ctx), "advance_Bhalfstep", ctx)
			call tmpg_check(tmpg_advance_e(ctx), "advance_Efield", ctx)
			call tmpg_check(tmpg_reset_currents(ctx), "reset_currents", ctx)
			call tmpg_check(tmpg_deposit(ctx), "deposit_particles", ctx)
			call tmpg_check(tmpg_fold_currents(ctx), "exchange_current", ctx)
			call tmpg_check(tmpg_filter(ctx, int(hp%ntimes, c_int)), "apply_filter1_opt", ctx)
			call tmpg_check(tmpg_add_current(ctx), "add_current", ctx)
		endif
		if (mod(lap, hp%interval) == 0 .or. lap == hp%last) then
			call tmpg_check(tmpg_diag(ctx, d), "tmpg_diag", ctx)
			call tmpg_check(tmpg_gauss_residual(ctx, gs), "tmpg_gauss_residual", ctx)
			write(*, '(a,i7,5(a,es14.6))') "lap ", lap, "  E ", d(1), "  B ", d(2), "  KEi ", d(3), "  KEe ", d(4), &
				"  max|divE-rho| ", gs(1)
			if (u /= 0) write(u, '(i7,4es24.15)') lap, d(1), d(2), d(3), d(4)
		endif
		call maybe_write_restart(hp%rst, ctx, lap, hp%dim, hp%order, hp%nx, hp%ny, hp%nz)
	end do
	call tmpg_check(tmpg_sync(ctx), "tmpg_sync", ctx)
	call system_clock(c1)
	t1 = real(c1 - c0, 8) / real(crate, 8)
	call tmpg_check(tmpg_last_timings(ctx, ms), "timings", ctx)
	! print_timers (code/communications.F90:242-326)
	if (ms(4) > 0d0) write(*, '(a,f10.3,a,f10.3,a,f10.3,a,f10.3)') "ms/lap: fields ", ms(1) / ms(4), &
		"  push+deposit ", ms(2) / ms(4), "  filter ", ms(3) / ms(4), "  total ", ms(5) / ms(4)
	write(*, '(a,es12.4,a)') "particle-steps/s ", real(ni + ne, 8) * max(hp%last - lapst + 1, 0) / t1, &
		" (host clock, incl. diagnostics)"
	if (u /= 0) close(u)
	call tmpg_check(tmpg_destroy(ctx), "tmpg_destroy", c_null_ptr)
end program tristan_gpu
