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
! The setup is the counter-streaming Weibel problem of
! user/user_weibel.F90:252-292: two plasma regions drifting -x and +x, each
! holding both species.
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
		integer :: fused = 1, seed = 0, pusher = 0, track_ids = 0
		real(8) :: c = 0.45d0, corr = 1.025d0
		real(8) :: ppc0 = 1d0, c_omp = 0d0, gamma0 = 1d0, delgam = 1d-8, me = 1d0, mi = 20d0
		real(8) :: temperature_ratio = 1d0
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
		call gi("output", "interval", 10, hp%interval)
		! keys of this host (the shape order and dimensionality are compile-time
		! cpp flags of the reference, code/particles_movedeposit.F90:69-84)
		call gi("gpu", "dim", merge(3, 2, hp%nz > 1), hp%dim)
		call gi("gpu", "order", 3, hp%order)
		call gi("gpu", "fused", 1, hp%fused)
		call gi("gpu", "seed", 0, hp%seed)
		call gi("gpu", "pusher", 0, hp%pusher)
		call gi("gpu", "track_ids", 0, hp%track_ids)
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
		cfg%scheme = 0
		cfg%reserved = 0
	end subroutine make_config

	! user/user_weibel.F90:252-292: two regions of ppc0/2 particles per cell,
	! drifting -x and +x, each with ions and electrons (ppc0/4 per species);
	! electron temperature delgam*mi/me*temperature_ratio
	subroutine init_particle_distribution_user(hp, ctx)
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
	end subroutine init_particle_distribution_user

end module m_host_setup
