!-------------------------------------------------------------------------------
! m_tmpg -- iso_c_binding interface to libtmpg (include/tmpg.h).
!
! Each interface binds one C entry point; the comments name the reference
! routine the entry point replaces (code/tristanmainloop.F90:117-242).
!-------------------------------------------------------------------------------
module m_tmpg
	use, intrinsic :: iso_c_binding
	implicit none

	integer(c_int), parameter :: TMPG_OK = 0
	integer(c_int), parameter :: TMPG_Q_GATHER2 = 1, TMPG_Q_BXY_Z = 2
	integer(c_int), parameter :: TMPG_COMPAT_DEFAULT = 3

	! struct tmpg_config
	type, bind(C) :: tmpg_config
		integer(c_int32_t) :: dim, order, nx, ny, nz, ntimes, compat, track_ids
		real(c_float) :: c, corr
		real(c_float) :: qm(2), q(2), mass(2)
		integer(c_int64_t) :: capacity(2)
		integer(c_int32_t) :: tile(3)
		integer(c_int32_t) :: sort_every
		integer(c_int32_t) :: highorder
		integer(c_int32_t) :: pusher
		integer(c_int32_t) :: scheme
		integer(c_int32_t) :: reserved(1)
	end type tmpg_config

	! tmpg_plasma_region: the arguments of inject_plasma_region
	! (code/particles.F90:2549-2938) and the domain they refer to
	type, bind(C) :: tmpg_plasma_region
		integer(c_int32_t) :: dim, order, nx, ny, nz
		real(c_float) :: c
		real(c_float) :: x1, x2, y1, y2, z1, z2
		real(c_float) :: ppc, gamma_drift, delgam_i, delgam_e, weight
		integer(c_int32_t) :: direction, pcosthmult
		real(c_float) :: sigma
		integer(c_int32_t) :: upsamp_e
	end type tmpg_plasma_region
	integer(c_int), parameter :: TMPG_EOVERFLOW = -4
	integer(c_int), parameter :: TMPG_SCHEME_ESIRKEPOV = 0, TMPG_SCHEME_ZIGZAG = 1

	interface
		integer(c_int) function tmpg_create(cfg, ctx) bind(C, name="tmpg_create")
			import :: c_int, c_ptr, tmpg_config
			type(tmpg_config), intent(in) :: cfg
			type(c_ptr), intent(out) :: ctx
		end function
		integer(c_int) function tmpg_destroy(ctx) bind(C, name="tmpg_destroy")
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		type(c_ptr) function tmpg_last_error(ctx) bind(C, name="tmpg_last_error")
			import :: c_ptr
			type(c_ptr), value :: ctx
		end function
		integer(c_int) function tmpg_sync(ctx) bind(C, name="tmpg_sync")
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		! ex..bz / curx..curz module arrays (Fortran (mx,my,mz) with ghosts)
		integer(c_int) function tmpg_set_fields(ctx, ex, ey, ez, bx, by, bz) bind(C, name="tmpg_set_fields")
			import :: c_int, c_ptr, c_float
			type(c_ptr), value :: ctx
			real(c_float), intent(in) :: ex(*), ey(*), ez(*), bx(*), by(*), bz(*)
		end function
		integer(c_int) function tmpg_get_fields(ctx, ex, ey, ez, bx, by, bz) bind(C, name="tmpg_get_fields")
			import :: c_int, c_ptr, c_float
			type(c_ptr), value :: ctx
			real(c_float), intent(out) :: ex(*), ey(*), ez(*), bx(*), by(*), bz(*)
		end function
		integer(c_int) function tmpg_get_currents(ctx, jx, jy, jz) bind(C, name="tmpg_get_currents")
			import :: c_int, c_ptr, c_float
			type(c_ptr), value :: ctx
			real(c_float), intent(out) :: jx(*), jy(*), jz(*)
		end function
		integer(c_int) function tmpg_count_particles(ctx, species, n) bind(C, name="tmpg_count_particles")
			import :: c_int, c_ptr, c_int64_t
			type(c_ptr), value :: ctx
			integer(c_int), value :: species
			integer(c_int64_t), intent(out) :: n
		end function
		! p(:) of one species as SoA; ch and id may be c_null_ptr
		integer(c_int) function tmpg_set_particles(ctx, species, n, x, y, z, u, v, w, ch, id) &
				bind(C, name="tmpg_set_particles")
			import :: c_int, c_ptr, c_int64_t
			type(c_ptr), value :: ctx
			integer(c_int), value :: species
			integer(c_int64_t), value :: n
			type(c_ptr), value :: x, y, z, u, v, w, ch, id
		end function
		integer(c_int) function tmpg_add_particles(ctx, species, n, x, y, z, u, v, w, ch, id) &
				bind(C, name="tmpg_add_particles")
			import :: c_int, c_ptr, c_int64_t
			type(c_ptr), value :: ctx
			integer(c_int), value :: species
			integer(c_int64_t), value :: n
			type(c_ptr), value :: x, y, z, u, v, w, ch, id
		end function
		integer(c_int) function tmpg_inject_plasma_region(r, dseed, cap, n, ions, lecs) &
				bind(C, name="tmpg_inject_plasma_region")
			import :: c_int, c_ptr, c_int64_t, c_double, tmpg_plasma_region
			type(tmpg_plasma_region), intent(in) :: r
			real(c_double), intent(inout) :: dseed
			integer(c_int64_t), value :: cap
			integer(c_int64_t), intent(out) :: n
			type(c_ptr), value :: ions, lecs
		end function
		integer(c_int) function tmpg_get_particles(ctx, species, n_capacity, n, x, y, z, u, v, w, ch, id) &
				bind(C, name="tmpg_get_particles")
			import :: c_int, c_ptr, c_int64_t
			type(c_ptr), value :: ctx
			integer(c_int), value :: species
			integer(c_int64_t), value :: n_capacity
			integer(c_int64_t), intent(out) :: n
			type(c_ptr), value :: x, y, z, u, v, w, ch, id
		end function
		integer(c_int) function tmpg_sort_particles(ctx) bind(C, name="tmpg_sort_particles")
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		! per-step kernels
		integer(c_int) function tmpg_advance_b_half(ctx) bind(C, name="tmpg_advance_b_half")  ! advance_Bhalfstep
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		integer(c_int) function tmpg_advance_e(ctx) bind(C, name="tmpg_advance_e")            ! advance_Efield
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		integer(c_int) function tmpg_reset_currents(ctx) bind(C, name="tmpg_reset_currents")  ! reset_currents
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		integer(c_int) function tmpg_move(ctx) bind(C, name="tmpg_move")                      ! move_particles
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		integer(c_int) function tmpg_deposit(ctx) bind(C, name="tmpg_deposit")                ! deposit_particles
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		integer(c_int) function tmpg_push_deposit(ctx) bind(C, name="tmpg_push_deposit")
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		integer(c_int) function tmpg_fold_currents(ctx) bind(C, name="tmpg_fold_currents")    ! exchange_current
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		integer(c_int) function tmpg_filter(ctx, npasses) bind(C, name="tmpg_filter")         ! apply_filter1_opt
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
			integer(c_int), value :: npasses
		end function
		integer(c_int) function tmpg_add_current(ctx) bind(C, name="tmpg_add_current")        ! add_current
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
		end function
		integer(c_int) function tmpg_step(ctx, nsteps) bind(C, name="tmpg_step")
			import :: c_int, c_ptr
			type(c_ptr), value :: ctx
			integer(c_int), value :: nsteps
		end function
		integer(c_int) function tmpg_diag(ctx, out) bind(C, name="tmpg_diag")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), intent(out) :: out(8)
		end function
		integer(c_int) function tmpg_gauss_residual(ctx, out) bind(C, name="tmpg_gauss_residual")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), intent(out) :: out(4)
		end function
		integer(c_int) function tmpg_write_flds_tot(ctx, path, istep, in_plane) bind(C, name="tmpg_write_flds_tot")
			import :: c_int, c_ptr, c_char
			type(c_ptr), value :: ctx
			character(kind=c_char), intent(in) :: path(*)
			integer(c_int), value :: istep, in_plane
		end function
		integer(c_int) function tmpg_load_uniform(ctx, species, ppc, drift_gb, theta, seed) &
				bind(C, name="tmpg_load_uniform")
			import :: c_int, c_ptr, c_float, c_int64_t
			type(c_ptr), value :: ctx
			integer(c_int), value :: species, ppc
			real(c_float), value :: drift_gb, theta
			integer(c_int64_t), value :: seed
		end function
		integer(c_int) function tmpg_last_timings(ctx, ms) bind(C, name="tmpg_last_timings")
			import :: c_int, c_ptr, c_double
			type(c_ptr), value :: ctx
			real(c_double), intent(out) :: ms(6)
		end function
		! input file in the reference's <block> key = value format
		integer(c_int) function tmpg_input_open(path, inp) bind(C, name="tmpg_input_open")
			import :: c_int, c_ptr, c_char
			character(kind=c_char), intent(in) :: path(*)
			type(c_ptr), intent(out) :: inp
		end function
		integer(c_int) function tmpg_input_close(inp) bind(C, name="tmpg_input_close")
			import :: c_int, c_ptr
			type(c_ptr), value :: inp
		end function
		integer(c_int) function tmpg_input_get_int(inp, block, key, def, val) bind(C, name="tmpg_input_get_int")
			import :: c_int, c_ptr, c_char
			type(c_ptr), value :: inp
			character(kind=c_char), intent(in) :: block(*), key(*)
			integer(c_int), value :: def
			integer(c_int), intent(out) :: val
		end function
		integer(c_int) function tmpg_input_get_double(inp, block, key, def, val) bind(C, name="tmpg_input_get_double")
			import :: c_int, c_ptr, c_char, c_double
			type(c_ptr), value :: inp
			character(kind=c_char), intent(in) :: block(*), key(*)
			real(c_double), value :: def
			real(c_double), intent(out) :: val
		end function
		integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
			import :: c_size_t, c_ptr
			type(c_ptr), value :: s
		end function
	end interface

contains

	! stop with the library's message when a call fails (the reference stops
	! on fatal conditions too, e.g. code/particles.F90:362-370)
	subroutine tmpg_check(rc, what, ctx)
		integer(c_int), intent(in) :: rc
		character(len=*), intent(in) :: what
		type(c_ptr), intent(in) :: ctx
		character(kind=c_char), pointer :: msg(:)
		integer :: n
		if (rc == TMPG_OK) return
		n = int(c_strlen(tmpg_last_error(ctx)))
		call c_f_pointer(tmpg_last_error(ctx), msg, [n])
		write(*, '(a,a,i0,a,*(a))') trim(what), ": error ", rc, ": ", msg(1:n)
		stop 2
	end subroutine tmpg_check

	function cstr(s) result(r)
		character(len=*), intent(in) :: s
		character(kind=c_char, len=len_trim(s)+1) :: r
		r = trim(s)//c_null_char
	end function cstr

end module m_tmpg
