!-------------------------------------------------------------------------------
! ref_driver.F90 -- test harness around the reference's own hot-path modules.
!
! This program is part of the oracle (test infrastructure).  It is linked
! against objects compiled, unmodified, from the reference sources
! (code/fields.F90, code/fieldboundaries.F90 + code/filter.F90,
! code/particles.F90, code/particles_movedeposit.F90, ...) and drives them in
! the order of the reference main loop (code/tristanmainloop.F90:107-274).
!
! Usage:  tristan_ref <input-file> <state-in> <state-out> <nsteps> <mode>
!   input-file : reference-format input (read with the reference parser)
!   state-in   : stream file: ex,ey,ez,bx,by,bz (mx,my,mz) then
!                int32 ions, lecs; then x,y,z,u,v,w,ch arrays for ions, then
!                for electrons (all real(sprec))
!   state-out  : same layout + curx,cury,curz appended after the fields
!   mode       : 0 full step (main-loop order), 1 mover only,
!                2 deposit only (reset, deposit, fold), 3 filter only
!                (state-in currents are read after the fields;
!                apply_filter2_opt in a -Dfilter2 build, as the main loop
!                at code/tristanmainloop.F90:217-219),
!                4 Yee only (B half, B half, E), 5 deposit + fold + filter,
!                6 initial plasma: the particles of state-in are dropped, the
!                random seed is initialised (code/communications.F90:224-231)
!                and the user module's init_particle_distribution_user runs
!                (code/overload.F90:148); no steps
!   optional   : <history-file> <every>: in mode 0, every <every> steps one
!                line "step E_energy B_energy KE_ions KE_electrons" in FP64:
!                0.5*sum(E^2), 0.5*sum(B^2) over the interior cells and
!                sum((gamma-1) c^2 m) per species, the quantities of the
!                reference's energy() (code/tristanmainloop.F90:384-494)
!                without its index-range and normalisation choices
!
! The main loop itself (code/tristanmainloop.F90) cannot be compiled here:
! it depends on m_output, whose save_param calls HDF5 outside #ifdef HDF5
! (code/output.F90:2308-2345) and the image has no flang-compatible HDF5.
! This driver therefore reissues the main-loop call sequence itself.
!-------------------------------------------------------------------------------
#ifdef twoD
#else
#endif

program tristan_ref
#ifdef twoD
	use m_system
	use m_aux
	use m_communications
	use m_inputparser
	use m_fields
	use m_fieldboundaries
	use m_particles
	use m_particles_movedeposit
	use m_user
#else
	use m_system_3d
	use m_aux_3d
	use m_communications_3d
	use m_inputparser_3d
	use m_fields_3d
	use m_fieldboundaries_3d
	use m_particles_3d
	use m_particles_movedeposit_3d
	use m_user_3d
#endif
	implicit none
	character(len=512) :: inpfile, fin, fout, arg
	integer :: nsteps, mode, step, lhighorder, ierr2, n, ni, ne, ptype
	integer(kind=4) :: ni4, ne4
	real(sprec), allocatable :: buf(:)
	real(sprec) :: lc
	character(len=512) :: histfile
	integer :: every

	call get_command_argument(1, inpfile)
	call get_command_argument(2, fin)
	call get_command_argument(3, fout)
	call get_command_argument(4, arg)
	read(arg,*) nsteps
	call get_command_argument(5, arg)
	read(arg,*) mode
	histfile = ""
	every = 0
	if (command_argument_count() .ge. 7) then
		call get_command_argument(6, histfile)
		call get_command_argument(7, arg)
		read(arg,*) every
	endif

	call inputpar_open(trim(inpfile))
	call read_input_communications()
	call init_communications()
	! the reference registers MPI_REAL for its field messages and a 7-real
	! particle record (code/communications.F90:124-150); in a build with
	! -fdefault-real-8 both must describe 8-byte reals instead
	! (the particle record is registered after allocate_particles below).
	if (sprec .eq. 8) mpi_read = mpi_dbl
	! read_input_time / read_input_algorithm live in m_initialize
	! (code/initialize.F90:267-307), which cannot be compiled here; the
	! same keys and defaults are read directly.
	call inputpar_getd_def("time", "c", 0.45_sprec, lc)
	c = lc
	call inputpar_geti_def("algorithm", "highorder", 1, lhighorder)
	highorder = (lhighorder .eq. 1)
	call inputpar_getd_def("algorithm", "Corr", 1.025_sprec, corr)
	call inputpar_geti_def("algorithm", "ntimes", 4, ntimes)
	irestart = 0
	call read_input_grid()
	call read_input_boundaries()
	call read_input_fields()
	call read_input_particles()
	call read_input_user()
	call allocate_fields()
	call allocate_particles()
	if (sprec .eq. 8) then
		! one record per particle, padding included (a sequence type of
		! 7 real(8) + 3 integer(4) is laid out with an 8-byte stride)
		call ref_record_type(storage_size(p(1))/8, ptype)
		particletype = ptype
	endif

	open(unit=21, file=trim(fin), access='stream', form='unformatted', status='old')
	read(21) ex, ey, ez, bx, by, bz
	if (mode .eq. 3) read(21) curx, cury, curz
	read(21) ni4, ne4
	ni = ni4
	ne = ne4
	ions = ni
	lecs = ne
	allocate(buf(max(ni, ne, 1)))
	call read_species(1, ni)
	call read_species(maxhlf + 1, ne)
	close(21)

	if (mode .eq. 6) then
		ions = 0
		lecs = 0
		call initialize_random_seed()
		call init_particle_distribution_user()
		nsteps = 0
	endif
	if (every .gt. 0) open(unit=23, file=trim(histfile), status='replace', action='write')
	do step = 1, nsteps
		select case (mode)
		case (0)
			! code/tristanmainloop.F90:107-274 (periodic box, no user hooks)
			call bc_b1()
			call bc_e1()
			call advance_Bhalfstep()
			call bc_b1()
			call move_particles()
			call advance_Bhalfstep()
			call bc_b1()
			call bc_b2()
			call advance_Efield()
			call bc_e2()
			call reset_currents()
			call bc_e1()
			call bc_b1()
			call deposit_particles()
			call exchange_particles()
			call exchange_current()
#ifdef filter2
			call apply_filter2_opt()
#else
			call apply_filter1_opt()
#endif
			call add_current()
			call inject_others()
			call exchange_particles()
			call inject_others()
		case (1)
			call bc_b1()
			call bc_e1()
			call move_particles()
		case (2)
			call reset_currents()
			call deposit_particles()
			call exchange_particles()
			call exchange_current()
			call inject_others()
			call exchange_particles()
			call inject_others()
		case (3)
#ifdef filter2
			call apply_filter2_opt()
#else
			call apply_filter1_opt()
#endif
		case (4)
			call bc_b1()
			call bc_e1()
			call advance_Bhalfstep()
			call bc_b1()
			call advance_Bhalfstep()
			call bc_b1()
			call advance_Efield()
			call bc_e1()
		case (5)
			call reset_currents()
			call deposit_particles()
			call exchange_particles()
			call exchange_current()
#ifdef filter2
			call apply_filter2_opt()
#else
			call apply_filter1_opt()
#endif
			call inject_others()
			call exchange_particles()
			call inject_others()
		end select
		if (every .gt. 0) then
			if (mod(step, every) .eq. 0) call write_energies(step)
		endif
	enddo
	if (every .gt. 0) close(23)

	open(unit=22, file=trim(fout), access='stream', form='unformatted', status='replace')
	write(22) ex, ey, ez, bx, by, bz, curx, cury, curz
	ni4 = ions
	ne4 = lecs
	write(22) ni4, ne4
	deallocate(buf)
	allocate(buf(max(ions, lecs, 1)))
	call write_species(1, ions)
	call write_species(maxhlf + 1, lecs)
	close(22)
	call MPI_Finalize(ierr2)

contains

	subroutine write_energies(st)
		integer, intent(in) :: st
		integer :: i, j, k, i1, i2, j1, j2, k1, k2
		real(8) :: ee, be, ki, ke
		i1 = nghost / 2 + 1
		i2 = mx0 - (nghost - nghost / 2)
		j1 = i1
		j2 = my0 - (nghost - nghost / 2)
#ifdef twoD
		k1 = 1
		k2 = 1
#else
		k1 = nghostz / 2 + 1
		k2 = mz0 - (nghostz - nghostz / 2)
#endif
		ee = 0d0
		be = 0d0
		do k = k1, k2
			do j = j1, j2
				do i = i1, i2
					ee = ee + real(ex(i,j,k), 8)**2 + real(ey(i,j,k), 8)**2 + real(ez(i,j,k), 8)**2
					be = be + real(bx(i,j,k), 8)**2 + real(by(i,j,k), 8)**2 + real(bz(i,j,k), 8)**2
				enddo
			enddo
		enddo
		ki = 0d0
		do n = 1, ions
			ki = ki + (sqrt(1d0 + real(p(n)%u, 8)**2 + real(p(n)%v, 8)**2 + real(p(n)%w, 8)**2) - 1d0) &
				* real(p(n)%ch, 8)
		enddo
		ke = 0d0
		do n = maxhlf + 1, maxhlf + lecs
			ke = ke + (sqrt(1d0 + real(p(n)%u, 8)**2 + real(p(n)%v, 8)**2 + real(p(n)%w, 8)**2) - 1d0) &
				* real(p(n)%ch, 8)
		enddo
		write(23, '(i8, 4es25.16)') st, 0.5d0 * ee, 0.5d0 * be, ki * real(c, 8)**2 * real(mi, 8), &
			ke * real(c, 8)**2 * real(me, 8)
	end subroutine write_energies

	subroutine read_species(i0, cnt)
		integer, intent(in) :: i0, cnt
		if (cnt .le. 0) return
		read(21) buf(1:cnt); p(i0:i0+cnt-1)%x = buf(1:cnt)
		read(21) buf(1:cnt); p(i0:i0+cnt-1)%y = buf(1:cnt)
		read(21) buf(1:cnt); p(i0:i0+cnt-1)%z = buf(1:cnt)
		read(21) buf(1:cnt); p(i0:i0+cnt-1)%u = buf(1:cnt)
		read(21) buf(1:cnt); p(i0:i0+cnt-1)%v = buf(1:cnt)
		read(21) buf(1:cnt); p(i0:i0+cnt-1)%w = buf(1:cnt)
		read(21) buf(1:cnt); p(i0:i0+cnt-1)%ch = buf(1:cnt)
		do n = i0, i0 + cnt - 1
			p(n)%ind = n
			p(n)%proc = 0
			p(n)%splitlev = 1
		enddo
	end subroutine read_species

	subroutine write_species(i0, cnt)
		integer, intent(in) :: i0, cnt
		integer(kind=4), allocatable :: ib(:)
		if (cnt .le. 0) return
		buf(1:cnt) = p(i0:i0+cnt-1)%x; write(22) buf(1:cnt)
		buf(1:cnt) = p(i0:i0+cnt-1)%y; write(22) buf(1:cnt)
		buf(1:cnt) = p(i0:i0+cnt-1)%z; write(22) buf(1:cnt)
		buf(1:cnt) = p(i0:i0+cnt-1)%u; write(22) buf(1:cnt)
		buf(1:cnt) = p(i0:i0+cnt-1)%v; write(22) buf(1:cnt)
		buf(1:cnt) = p(i0:i0+cnt-1)%w; write(22) buf(1:cnt)
		buf(1:cnt) = p(i0:i0+cnt-1)%ch; write(22) buf(1:cnt)
		allocate(ib(cnt))
		ib = p(i0:i0+cnt-1)%ind
		write(22) ib
		deallocate(ib)
	end subroutine write_species

end program tristan_ref

! A contiguous byte record of nbytes, for the particle exchange of a build
! with 8-byte reals (kept out of the program unit: m_communications
! re-exports part of mpif.h and the two cannot be mixed in one scope).
subroutine ref_record_type(nbytes, t)
	implicit none
	include "mpif.h"
	integer, intent(in) :: nbytes
	integer, intent(out) :: t
	integer :: ierr
	call MPI_Type_contiguous(nbytes, MPI_BYTE, t, ierr)
	call MPI_Type_commit(t, ierr)
end subroutine ref_record_type
