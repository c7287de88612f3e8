!-------------------------------------------------------------------------------
! m_host_setup -- input reading, normalisation and the problem setup hooks of
! the Fortran host.
!
! Input keys, defaults and the normalisation follow the reference:
!   <time> c, last                       code/initialize.F90:267-276
!   <grid> mx0, my0, mz0                 code/fields.F90:154-192
!   <algorithm> highorder, Corr, ntimes  code/initialize.F90:287-307
!   <particles> ppc0, c_omp, gamma0, delgam, me, mi
!                                        code/particles.F90:189-256
!   <problem> temperature_ratio          user/user_weibel.F90:82-120
!   <restart> irestart, intrestart, laprestart, lnamedrestartint
!                                        code/restart.F90:95-98
! The problem (<gpu> problem, the user module the reference links):
!   0 Weibel, user/user_weibel.F90:252-292: two plasma regions over the
!     whole box drifting -x and +x, each holding both species;
!   1 two-stream, user/user_twostream.F90:236-305: the same two beams with
!     the in-plane Maxwellian (pcosthmult = 0), then the ions removed;
!   2 plasma oscillation, user/user_plasmaosc.F90:236-305: one plasma at
!     rest, the electrons kicked by beta sin(2 pi x / (mx0 - 5)), the ions
!     removed (an immobile background).
! <gpu> loader = 1 (default) draws the plasma with the reference's own
! random stream (tmpg_inject_plasma_region, seed 123457 + <gpu> seed, like
! 123457 + rank); 0 uses the library's device loader (tmpg_load_uniform:
! the same distributions from a counter-based generator, Weibel only).
!-------------------------------------------------------------------------------
module m_host_setup
	use, intrinsic :: iso_c_binding
	use m_tmpg
	use m_restart_io
	implicit none

	type host_params
		integer :: dim = 2, order = 3
		integer :: nx = 100, ny = 100, nz = 1
		integer :: last = 100, ntimes = 4, highorder = 1, interval = 10
		integer :: fused = 1, seed = 0, pusher = 0, track_ids = 0, problem = 0, loader = 1, scheme = 0
		! flds.tot output (code/output.F90:143-149, 2546-2556): written when
		! flds_tot /= 0, every interval laps from pltstart, sampled every istep
		integer :: istep = 2, pltstart = 0, flds_tot = 0, pcosthmult = 1
		real(8) :: c = 0.45d0, corr = 1.025d0
		real(8) :: ppc0 = 1d0, c_omp = 0d0, gamma0 = 1d0, delgam = 1d-8, me = 1d0, mi = 20d0
		real(8) :: temperature_ratio = 1d0, sigma = 0d0
		! normalised
		real(8) :: qe, qi, qme, qmi, mass_e, mass_i, beta, g0
		type(restart_params) :: rst
	end type host_params

contains

	subroutine read_host_input(path, hp)
		character(len=*), intent(in) :: path
		type(host_params), intent(inout) :: hp
		type(c_ptr) :: inp
		integer(c_int) :: iv
		real(c_double) :: dv
		call tmpg_check(tmpg_input_open(cstr(path), inp), "tmpg_input_open "//trim(path), c_null_ptr)
		call gi("time", "last", 100, hp%last)
		call gd("time", "c", 0.45d0, hp%c)
		call gi("grid", "mx0", 100, hp%nx)
		call gi("grid", "my0", 100, hp%ny)
		call gi("grid", "mz0", 1, hp%nz)
		call gi("algorithm", "highorder", 1, hp%highorder)
		call gd("algorithm", "Corr", 1.025d0, hp%corr)
		call gi("algorithm", "ntimes", 4, hp%ntimes)
		call gd("particles", "ppc0", 1d0, hp%ppc0)
		call gd("particles", "c_omp", 0d0, hp%c_omp)
		call gd("particles", "gamma0", 1d0, hp%gamma0)
		call gd("particles", "delgam", 1d-8, hp%delgam)
		call gd("particles", "me", 1d0, hp%me)
		call gd("particles", "mi", 20d0, hp%mi)
		call gd("problem", "temperature_ratio", 1d0, hp%temperature_ratio)
		call gd("particles", "sigma", 0d0, hp%sigma)
		call gi("output", "interval", 10, hp%interval)
		call gi("output", "istep", 2, hp%istep)
		call gi("output", "pltstart", 0, hp%pltstart)
		! keys of this host (the shape order and dimensionality are compile-time
		! cpp flags of the reference, code/particles_movedeposit.F90:69-84)
		call gi("gpu", "dim", merge(3, 2, hp%nz > 1), hp%dim)
		call gi("gpu", "order", 3, hp%order)
		call gi("gpu", "fused", 1, hp%fused)
		call gi("gpu", "seed", 0, hp%seed)
		call gi("gpu", "pusher", 0, hp%pusher)
		call gi("gpu", "track_ids", 0, hp%track_ids)
		call gi("gpu", "problem", 0, hp%problem)
		call gi("gpu", "loader", 1, hp%loader)
		call gi("gpu", "scheme", 0, hp%scheme)
		call gi("gpu", "flds_tot", 0, hp%flds_tot)
		! the user files' pcosthmult: 1 in user_weibel, 0 in user_twostream and
		! user_plasmaosc (it selects the 2D flds.tot dataset list)
		hp%pcosthmult = merge(1, 0, hp%problem == 0)
		call read_restart_input(inp, hp%rst)
		hp%rst%track_ids = hp%track_ids /= 0
		call tmpg_check(tmpg_input_close(inp), "tmpg_input_close", c_null_ptr)
		if (hp%dim == 2) hp%nz = 1
		call normalise(hp)
	contains
		subroutine gi(b, k, d, v)
			character(len=*), intent(in) :: b, k
			integer, intent(in) :: d
			integer, intent(out) :: v
			call tmpg_check(tmpg_input_get_int(inp, cstr(b), cstr(k), int(d, c_int), iv), "input "//k, c_null_ptr)
			v = iv
		end subroutine gi
		subroutine gd(b, k, d, v)
			character(len=*), intent(in) :: b, k
			real(8), intent(in) :: d
			real(8), intent(out) :: v
			call tmpg_check(tmpg_input_get_double(inp, cstr(b), cstr(k), real(d, c_double), dv), "input "//k, &
				c_null_ptr)
			v = dv
		end subroutine gd
	end subroutine read_host_input

	! output_tot (code/output.F90:2494-3875) for this lap: output/flds.tot.NNN
	! with NNN = (lap - pltstart) / interval
	subroutine maybe_write_flds_tot(hp, ctx, lap)
		type(host_params), intent(in) :: hp
		type(c_ptr), intent(in) :: ctx
		integer, intent(in) :: lap
		character(len=64) :: path
		integer(c_int) :: rc
		logical :: in_plane
		interface
			integer(c_int) function c_mkdir(p, mode) bind(C, name="mkdir")
				import :: c_int, c_char
				character(kind=c_char), intent(in) :: p(*)
				integer(c_int), value :: mode
			end function
		end interface
		if (hp%flds_tot == 0 .or. lap < hp%pltstart) return
		if (mod(lap - hp%pltstart, hp%interval) /= 0) return
		rc = c_mkdir(cstr("output"), int(o'755', c_int))   ! an existing directory is fine
		write(path, '(a,i3.3)') "output/flds.tot.", (lap - hp%pltstart) / hp%interval
		in_plane = hp%dim == 2 .and. hp%sigma == 0d0 .and. hp%pcosthmult == 0
		call tmpg_check(tmpg_write_flds_tot(ctx, cstr(path), int(hp%istep, c_int), merge(1_c_int, 0_c_int, in_plane)), &
			"tmpg_write_flds_tot "//trim(path), ctx)
	end subroutine maybe_write_flds_tot

	! code/particles.F90:206-235
	subroutine normalise(hp)
		type(host_params), intent(inout) :: hp
		real(8) :: omp
		hp%g0 = hp%gamma0
		if (hp%g0 < 1d0) hp%g0 = sqrt(1d0 / (1d0 - hp%g0**2))
		omp = hp%c / hp%c_omp
		hp%beta = sqrt(1d0 - 1d0 / hp%g0**2)
		hp%qe = -(omp**2 * hp%g0) / ((hp%ppc0 * 0.5d0) * (1d0 + hp%me / hp%mi))
		hp%qi = -hp%qe
		hp%mass_e = hp%me * abs(hp%qi)
		hp%mass_i = hp%mi * abs(hp%qi)
		hp%qme = hp%qe / hp%mass_e
		hp%qmi = hp%qi / hp%mass_i
	end subroutine normalise

	subroutine make_config(hp, cfg)
		type(host_params), intent(in) :: hp
		type(tmpg_config), intent(out) :: cfg
		cfg%dim = hp%dim
		cfg%order = hp%order
		cfg%nx = hp%nx
		cfg%ny = hp%ny
		cfg%nz = hp%nz
		cfg%ntimes = hp%ntimes
		cfg%compat = TMPG_COMPAT_DEFAULT
		cfg%track_ids = merge(1, 0, hp%track_ids /= 0)
		cfg%c = real(hp%c, c_float)
		cfg%corr = real(hp%corr, c_float)
		cfg%qm = [real(hp%qmi, c_float), real(hp%qme, c_float)]
		cfg%q = [real(hp%qi, c_float), real(hp%qe, c_float)]
		cfg%mass = [real(hp%mass_i, c_float), real(hp%mass_e, c_float)]
		cfg%capacity = 0
		cfg%tile = 0
		cfg%sort_every = 0
		cfg%highorder = merge(1, 0, hp%highorder /= 0)
		cfg%pusher = hp%pusher
		cfg%scheme = hp%scheme
		cfg%reserved = 0
	end subroutine make_config

	! init_particle_distribution_user of the problem's user module
	subroutine init_particle_distribution_user(hp, ctx)
		type(host_params), intent(in) :: hp
		type(c_ptr), intent(in) :: ctx
		real(c_double) :: dseed
		real(c_float) :: g0, beta, delgam_i, delgam_e, ppc, x2, y2, z2
		integer :: ng, ngz, mx0, my0, mz0
		if (hp%problem < 0 .or. hp%problem > 2) then
			write(*, '(a,i0)') "tristan-mp-gpu: unknown <gpu> problem ", hp%problem
			stop 1
		endif
		if (hp%loader == 0) then
			if (hp%problem /= 0) then
				write(*, '(a)') "tristan-mp-gpu: the device loader (<gpu> loader = 0) sets up the Weibel problem only"
				stop 1
			endif
			call weibel_device_loader(hp, ctx)
			return
		endif
		! the reference's grid sizes include the ghost cells (code/fields.F90:186)
		ng = merge(5, 7, hp%order == 1)
		ngz = merge(ng, 5, hp%dim == 3)
		mx0 = hp%nx + ng
		my0 = hp%ny + ng
		mz0 = merge(hp%nz + ngz, 1, hp%dim == 3)
		! code/particles.F90:189-235 in the reference's precision
		g0 = real(hp%gamma0, c_float)
		if (g0 < 1.) g0 = sqrt(1. / (1. - g0**2))
		beta = sqrt(1. - 1. / g0**2)
		delgam_i = real(hp%delgam, c_float)
		delgam_e = delgam_i * real(hp%mi, c_float) / real(hp%me, c_float) * real(hp%temperature_ratio, c_float)
		dseed = 123457d0 + hp%seed   ! code/communications.F90:224-231
		select case (hp%problem)
		case (0)
			! user/user_weibel.F90:252-292
			x2 = mx0 - 1. * (ng / 2)
			y2 = my0 - 1. * (ng / 2)
			z2 = mz0 - 2.
			ppc = real(hp%ppc0, c_float) / 2.
			call region(1. * (ng / 2 + 1), x2, 1. * (ng / 2 + 1), y2, 3., z2, ppc, -g0, 1, .true., .false.)
			call region(1. * (ng / 2 + 1), x2, 1. * (ng / 2 + 1), y2, 3., z2, ppc, g0, 1, .true., .false.)
		case (1)
			! user/user_twostream.F90:236-305
			ppc = real(hp%ppc0, c_float) / 2.
			call region(3., mx0 - 2., 3., my0 - 2., 3., mz0 - 2., ppc, -g0, 0, .false., .false.)
			call region(3., mx0 - 2., 3., my0 - 2., 3., mz0 - 2., ppc, g0, 0, .false., .false.)
		case (2)
			! user/user_plasmaosc.F90:236-305
			call region(3., mx0 - 2., 3., my0 - 2., 3., mz0 - 2., real(hp%ppc0, c_float), 0., 0, .false., .true.)
		end select
	contains
		subroutine region(x1, x2r, y1, y2r, z1, z2r, ppcr, gd, pcm, keep_ions, kick)
			real(c_float), intent(in) :: x1, x2r, y1, y2r, z1, z2r, ppcr, gd
			integer, intent(in) :: pcm
			logical, intent(in) :: keep_ions, kick
			type(tmpg_plasma_region) :: r
			real(c_float), allocatable, target :: ions(:, :), lecs(:, :)
			real(8), parameter :: pi = 3.1415927   ! code/particles.F90:292
			integer(c_int64_t) :: n
			integer(c_int) :: rc
			integer :: k
			r%dim = hp%dim
			r%order = hp%order
			r%nx = hp%nx
			r%ny = hp%ny
			r%nz = hp%nz
			r%c = real(hp%c, c_float)
			r%x1 = x1; r%x2 = x2r; r%y1 = y1; r%y2 = y2r; r%z1 = z1; r%z2 = z2r
			r%ppc = ppcr
			r%gamma_drift = gd
			r%delgam_i = delgam_i
			r%delgam_e = delgam_e
			r%weight = 1.
			r%direction = 1
			r%pcosthmult = pcm
			r%sigma = real(hp%sigma, c_float)
			r%upsamp_e = 1
			rc = tmpg_inject_plasma_region(r, dseed, 0_c_int64_t, n, c_null_ptr, c_null_ptr)
			if (rc /= TMPG_OK .and. rc /= TMPG_EOVERFLOW) call tmpg_check(rc, "tmpg_inject_plasma_region", ctx)
			allocate(ions(max(n, 1_c_int64_t), 7), lecs(max(n, 1_c_int64_t), 7))
			call tmpg_check(tmpg_inject_plasma_region(r, dseed, max(n, 1_c_int64_t), n, c_loc(ions), c_loc(lecs)), &
				"tmpg_inject_plasma_region", ctx)
			if (kick) then
				! the electrons' sinusoidal kick, with the reference's promotions
				do k = 1, int(n)
					lecs(k, 4) = lecs(k, 4) + beta * sin(2 * pi / (mx0 - 5.) * lecs(k, 1))
				enddo
			endif
			if (keep_ions) call add(0, ions, n)
			call add(1, lecs, n)
			deallocate(ions, lecs)
		end subroutine region
		subroutine add(species, a, n)
			integer, intent(in) :: species
			real(c_float), target, intent(in) :: a(:, :)
			integer(c_int64_t), intent(in) :: n
			if (n <= 0) return
			call tmpg_check(tmpg_add_particles(ctx, int(species, c_int), n, c_loc(a(1, 1)), c_loc(a(1, 2)), &
				c_loc(a(1, 3)), c_loc(a(1, 4)), c_loc(a(1, 5)), c_loc(a(1, 6)), c_loc(a(1, 7)), c_null_ptr), &
				"tmpg_add_particles", ctx)
		end subroutine add
	end subroutine init_particle_distribution_user

	! the Weibel setup on the device loader (ppc0/4 per species and region)
	subroutine weibel_device_loader(hp, ctx)
		type(host_params), intent(in) :: hp
		type(c_ptr), intent(in) :: ctx
		integer(c_int) :: ppc_half
		real(c_float) :: gb, th_i, th_e
		ppc_half = int(hp%ppc0 / 4d0, c_int)
		gb = real(hp%g0 * hp%beta, c_float)
		th_i = real(hp%delgam, c_float)
		th_e = real(hp%delgam * hp%mi / hp%me * hp%temperature_ratio, c_float)
		call tmpg_check(tmpg_load_uniform(ctx, 0_c_int, ppc_half, -gb, th_i, int(4 * hp%seed + 1, c_int64_t)), &
			"load ions -x", ctx)
		call tmpg_check(tmpg_load_uniform(ctx, 1_c_int, ppc_half, -gb, th_e, int(4 * hp%seed + 2, c_int64_t)), &
			"load electrons -x", ctx)
		call tmpg_check(tmpg_load_uniform(ctx, 0_c_int, ppc_half, gb, th_i, int(4 * hp%seed + 3, c_int64_t)), &
			"load ions +x", ctx)
		call tmpg_check(tmpg_load_uniform(ctx, 1_c_int, ppc_half, gb, th_e, int(4 * hp%seed + 4, c_int64_t)), &
			"load electrons +x", ctx)
	end subroutine weibel_device_loader

end module m_host_setup
