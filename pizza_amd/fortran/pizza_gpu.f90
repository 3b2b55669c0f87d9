! pizza_gpu — driver of the GPU IMEX step, the counterpart of pizza.f90's
! main program (pizza.f90:63-199) and time_loop (step_time.f90:96-405) for
! the path the library implements.  It reads the reference's input
! namelist, creates the run, applies the namelist's initial conditions
! (init_fields.f90:260-509) or reads its start_file checkpoint
! (read_checkpoint, checkpoints.f90:185-707), steps through libpizza_amd and
! writes the time series of get_time_series in the reference's files and
! formats (outputs.f90:79-121, 672-717, 775-777, 824-826: e_kin, e_kin_3D,
! power, power_3D, reynolds or rossby, reynolds_3D or rossby_3D,
! length_scales, heat, composition; no 3D file in a non-rotating run) on
! every log step, the frame files
! frame_<field>_<n>.TAG on the frame steps (outputs.f90:293-312) and the
! checkpoint_end.TAG file at the end of the run (step_time.f90:190-193,
! checkpoints.f90:33-183), as time_loop does (l_correct_step,
! step_time.f90:145-160).
!
! usage: pizza_gpu input.nml [-n n_time_steps] [--print-config] [--stages]
!
! --stages steps through the reference's stage-loop procedures of module
! step_stages (radial_loop, dt_courant, finish_explicit_assembly, mloop)
! one step at a time instead of the captured pz_step.
program pizza_gpu
   use iso_c_binding
   use pizza_amd_gpu
   use step_stages
   use namelists
   implicit none

   character(len=512) :: fname, arg
   type(pz_config) :: cfg
   type(c_ptr) :: h
   integer :: nargs, i, n_steps, n_done, chunk, u, n_frame_out
   integer :: uf(8)
   type(pz_series) :: ts
   logical :: print_only, by_stages
   real(c_double) :: time, dt, e(4)

   nargs = command_argument_count()
   if (nargs < 1) then
      write(*, '(a)') 'usage: pizza_gpu input.nml [-n n_time_steps] [--print-config] [--stages]'
      error stop 1
   end if
   call get_command_argument(1, fname)
   n_steps = -1
   print_only = .false.
   by_stages = .false.
   i = 2
   do while (i <= nargs)
      call get_command_argument(i, arg)
      if (trim(arg) == '-n') then
         i = i + 1
         call get_command_argument(i, arg)
         read(arg, *) n_steps
      else if (trim(arg) == '--print-config') then
         print_only = .true.
      else if (trim(arg) == '--stages') then
         by_stages = .true.
      else
         write(*, '(a)') 'pizza_gpu: unknown argument '//trim(arg)
         error stop 1
      end if
      i = i + 1
   end do

   call read_namelists(trim(fname))
   call fill_config(cfg)
   if (n_steps < 0) n_steps = n_time_steps
   if (print_only) then
      call print_config(cfg)
   else
      call check_supported()
      call run()
   end if

contains

   subroutine run()

      call check(pz_create(cfg, h), 'pz_create')
      ! get_start_fields (init_fields.f90:58-112): the checkpoint first, then
      ! the namelist's perturbations on top of it
      if (l_start_file) call check(pz_read_checkpoint(h, cstr(start_file)), 'pz_read_checkpoint')
      call check(pz_init_fields(h, int(init_t, c_int), real(amp_t, c_double), int(init_xi, c_int), &
                                real(amp_xi, c_double), int(init_u, c_int), real(amp_u, c_double)), 'pz_init_fields')
      call check(pz_start(h), 'pz_start')

      ! a non-rotating run writes no 3D series (outputs.f90:102-121)
      u = 0
      if (.not. l_non_rot) open(newunit=u, file='e_kin_3D.'//trim(tag), status='replace', action='write')
      call open_series()
      n_done = 0
      n_frame_out = 0
      do
         if (l_correct_step(n_done, n_steps, n_log_step, 0) .or. n_done == 0) then
            call check(pz_time(h, time, dt), 'pz_time')
            if (u /= 0) then
               call check(pz_ekin_3d(h, e), 'pz_ekin_3d')
               write(u, '(1P, es20.12, 4es16.8)') time, e(1), e(2), e(3), e(4)
            end if
            call write_series()
         end if
         if (l_correct_step(n_done, n_steps, n_frame_step, n_frames)) call write_frames()
         if (n_done >= n_steps) exit
         chunk = next_output(n_done) - n_done
         if (by_stages) then
            do i = 1, chunk
               call time_loop_step(h)
            end do
         else
            call check(pz_step(h, int(chunk, c_int)), 'pz_step')
         end if
         n_done = n_done + chunk
      end do
      if (u /= 0) close(u)
      do i = 1, size(uf)
         if (uf(i) /= 0) close(uf(i))
      end do
      call check(pz_write_checkpoint(h, cstr('checkpoint_end.'//trim(tag))), 'pz_write_checkpoint')
      call check(pz_destroy(h), 'pz_destroy')

   end subroutine run

   ! the time-series files under the names of outputs.f90:79-121: rossby
   ! instead of reynolds with the rotational time scale, heat and composition
   ! only with the corresponding field, the 3D files only when rotating
   subroutine open_series()
      character(len=16) :: names(8)
      integer :: k
      names = [character(len=16) :: 'e_kin', 'power', 'power_3D', 'reynolds', 'reynolds_3D', &
               'length_scales', 'heat', 'composition']
      if (has(time_scale, 'ROT')) then
         names(4) = 'rossby'
         names(5) = 'rossby_3D'
      end if
      uf = 0
      do k = 1, size(names)
         if (k == 7 .and. ra == 0.0_c_double) cycle
         if (k == 8 .and. raxi == 0.0_c_double) cycle
         if ((k == 3 .or. k == 5) .and. l_non_rot) cycle
         open(newunit=uf(k), file=trim(names(k))//'.'//trim(tag), status='replace', action='write')
      end do
   end subroutine open_series

   ! one row of each (outputs.f90:672-717, 775-777, 824-826)
   subroutine write_series()
      call check(pz_time_series(h, ts), 'pz_time_series')
      write(uf(1), '(1P, es20.12, 3es16.8)') ts%time, ts%ekin_2d
      write(uf(2), '(1P, es20.12, 5es16.8)') ts%time, ts%power_2d
      if (uf(3) /= 0) write(uf(3), '(1P, es20.12, 5es16.8)') ts%time, ts%power_3d
      write(uf(4), '(1P, es20.12, 3es16.8)') ts%time, ts%reynolds_2d
      if (uf(5) /= 0) write(uf(5), '(1P, es20.12, 3es16.8)') ts%time, ts%reynolds_3d
      write(uf(6), '(1P, es20.12, 5es16.8)') ts%time, ts%lengths
      if (uf(7) /= 0) write(uf(7), '(1P, ES20.12, 7ES16.8)') ts%time, ts%heat
      if (uf(8) /= 0) write(uf(8), '(1P, ES20.12, 6ES16.8)') ts%time, ts%chem
   end subroutine write_series

   ! the next step count after n at which time_loop writes an output
   integer function next_output(n)
      integer, intent(in) :: n
      integer :: k
      do k = n + 1, n_steps
         if (l_correct_step(k, n_steps, n_log_step, 0) .or. l_correct_step(k, n_steps, n_frame_step, n_frames)) exit
      end do
      next_output = min(k, n_steps)
   end function next_output

   ! frame files of outputs.f90:293-312
   subroutine write_frames()
      character(len=32) :: num
      n_frame_out = n_frame_out + 1
      write(num, '(i0)') n_frame_out
      if (ra /= 0.0_c_double) call frame(PZ_TEMP, 'temp', num)
      if (raxi /= 0.0_c_double) call frame(PZ_XI, 'xi', num)
      call frame(PZ_US, 'us', num)
      call frame(PZ_UP, 'up', num)
      call frame(PZ_OM, 'om', num)
   end subroutine write_frames

   subroutine frame(field, name, num)
      integer(c_int), intent(in) :: field
      character(len=*), intent(in) :: name, num
      call check(pz_write_frame(h, field, cstr('frame_'//name//'_'//trim(num)//'.'//trim(tag))), 'pz_write_frame')
   end subroutine frame

   ! l_correct_step (useful.f90:181-222) with the step offset of time_loop:
   ! true after n of n_max steps when the action of n_step steps (or of
   ! n_intervals evenly spaced ones) falls there, and at the last step
   logical function l_correct_step(n, n_max, n_step, n_intervals)
      integer, intent(in) :: n, n_max, n_step, n_intervals
      integer :: n_offset, n_every
      l_correct_step = .false.
      if (n_intervals /= 0) then
         n_every = max(n_max / n_intervals, 1)
         n_offset = n_max - n_every * n_intervals
         if (n > n_offset .and. mod(n - n_offset, n_every) == 0) l_correct_step = .true.
      else if (n_step /= 0) then
         if (n == n_max .or. mod(n, n_step) == 0) l_correct_step = .true.
      end if
   end function l_correct_step

   ! a Fortran string as a NUL-terminated C string
   function cstr(f) result(c)
      character(len=*), intent(in) :: f
      character(kind=c_char, len=:), allocatable :: c
      c = trim(f)//c_null_char
   end function cstr

   subroutine check(ierr, what)
      integer(c_int), intent(in) :: ierr
      character(len=*), intent(in) :: what
      if (ierr /= 0) then
         write(*, '(a, i0, a)') 'pizza_gpu: '//what//' failed (', ierr, '): '//pz_last_error(h)
         error stop 3
      end if
   end subroutine check

   subroutine print_config(c)
      type(pz_config), intent(in) :: c
      write(*, '(a, i0)') 'n_r_max ', c%n_r_max
      write(*, '(a, i0)') 'n_cheb_max ', c%n_cheb_max
      write(*, '(a, i0)') 'm_max ', c%m_max
      write(*, '(a, i0)') 'minc ', c%minc
      write(*, '(a, i0)') 'time_scheme ', c%time_scheme
      write(*, '(a, i0)') 'l_cheb_coll ', c%l_cheb_coll
      write(*, '(a, i0)') 'l_direct_solve ', c%l_direct_solve
      write(*, '(a, i0)') 'l_coriolis_imp ', c%l_coriolis_imp
      write(*, '(a, i0)') 'l_buo_imp ', c%l_buo_imp
      write(*, '(a, i0)') 'l_non_rot ', c%l_non_rot
      write(*, '(a, i0)') 'time_scale_rot ', c%time_scale_rot
      write(*, '(a, i0)') 'l_rerror_fix ', c%l_rerror_fix
      write(*, '(a, es24.16)') 'rerror_fac ', c%rerror_fac
      write(*, '(a, es24.16)') 'alpha ', c%alpha
      write(*, '(a, es24.16)') 'dtMax ', c%dtMax
      write(*, '(a, es24.16)') 'courfac ', c%courfac
      write(*, '(a, es24.16)') 'dt_fac ', c%dt_fac
      write(*, '(a, es24.16)') 'ra ', c%ra
      write(*, '(a, es24.16)') 'ek ', c%ek
      write(*, '(a, es24.16)') 'pr ', c%pr
      write(*, '(a, es24.16)') 'raxi ', c%raxi
      write(*, '(a, es24.16)') 'sc ', c%sc
      write(*, '(a, es24.16)') 'radratio ', c%radratio
      write(*, '(a, es24.16)') 'g0 ', c%g0
      write(*, '(a, es24.16)') 'g1 ', c%g1
      write(*, '(a, es24.16)') 'g2 ', c%g2
      write(*, '(a, i0)') 'ktopt ', c%ktopt
      write(*, '(a, i0)') 'kbott ', c%kbott
      write(*, '(a, i0)') 'ktopv ', c%ktopv
      write(*, '(a, i0)') 'kbotv ', c%kbotv
      write(*, '(a, i0)') 'ktopxi ', c%ktopxi
      write(*, '(a, i0)') 'kbotxi ', c%kbotxi
      write(*, '(a, i0)') 'l_ek_pump ', c%l_ek_pump
      write(*, '(a, i0)') 'l_temp_advz ', c%l_temp_advz
      write(*, '(a, es24.16)') 'tcond_fac ', c%tcond_fac
      write(*, '(a, es24.16)') 'xicond_fac ', c%xicond_fac
      write(*, '(a, es24.16)') 'fd_stretch ', c%fd_stretch
      write(*, '(a, es24.16)') 'fd_ratio ', c%fd_ratio
      write(*, '(a, i0)') 'l_finite_diff ', c%l_finite_diff
      write(*, '(a, i0)') 'l_fd_bc_pattern ', c%l_fd_bc_pattern
      write(*, '(a, i0)') 'init_t ', init_t
      write(*, '(a, es24.16)') 'amp_t ', amp_t
      write(*, '(a, i0)') 'init_xi ', init_xi
      write(*, '(a, es24.16)') 'amp_xi ', amp_xi
      write(*, '(a, i0)') 'init_u ', init_u
      write(*, '(a, es24.16)') 'amp_u ', amp_u
      write(*, '(a, i0)') 'n_log_step ', n_log_step
      write(*, '(a, i0)') 'n_time_steps ', n_steps
      write(*, '(a, a)') 'tag ', trim(tag)
   end subroutine print_config

end program pizza_gpu
