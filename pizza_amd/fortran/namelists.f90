! Namelist input of the GPU driver: the reference's groups and names
! (namelists.f90:162-189) with its defaults (namelists.f90:530-672), so an
! input file of the reference reads unchanged.  Options outside the GPU path
! are accepted; the unsupported ones stop the run in check_supported.
module namelists
   use iso_c_binding
   use pizza_amd_gpu
   implicit none
   private

   integer, parameter :: cp = c_double
   integer, parameter :: n_t_bounds = PZ_N_T_BOUNDS

   ! &grid
   integer, public :: n_r_max, n_cheb_max, m_max, minc
   real(cp), public :: fd_ratio, fd_stretch, rcut_m
   logical, public :: l_var_m
   ! &control
   character(len=72), public :: tag, map_function, time_scheme, cheb_method, time_scale
   character(len=72), public :: matrix_solve, corio_term, buo_term, bc_method, mpi_transp, radial_scheme
   integer, public :: n_time_steps, runHours, runMinutes, runSeconds, n_fft_optim_lev
   real(cp), public :: alpha, alph1, alph2, dtMax, courfac, tEND, dt_fac, rerror_fac, c_sgs
   logical, public :: l_newmap, l_non_rot, l_rerror_fix, l_packed_transp
   ! &hdif
   real(cp), public :: hdif_temp, hdif_vel, hdif_comp
   integer, public :: hdif_exp, hdif_m
   ! &phys_param
   real(cp), public :: ra, ek, pr, raxi, sc, stef, radratio, g0, g1, g2, tcond_fac, beta_shift
   real(cp), public :: xicond_fac, h_temp, h_xi, beta_fac, amp_forcing, radius_forcing, dx_forcing
   real(cp), public :: epsPhase, phaseDiffFac, penaltyFac, tmelt
   integer, public :: ktopt, kbott, ktopv, kbotv, n_rings, ktopxi, kbotxi, ktopphi, kbotphi
   logical, public :: l_ek_pump, l_temp_3D, l_temp_advz, l_xi_3D
   real(cp), public :: t_bot(3*n_t_bounds), t_top(3*n_t_bounds), xi_bot(3*n_t_bounds), xi_top(3*n_t_bounds)
   character(len=72), public :: container, forcing_type
   ! &start_field
   logical, public :: l_start_file, l_reset_t
   character(len=72), public :: start_file
   real(cp), public :: scale_t, amp_t, scale_u, amp_u, amp_xi, scale_xi
   integer, public :: init_t, init_u, init_xi, init_phi
   ! &output_control
   integer, public :: n_log_step, n_checkpoints, n_checkpoint_step, n_frames, n_frame_step
   integer, public :: n_specs, n_spec_step, m_max_transfer
   logical, public :: l_vphi_balance, l_vort_balance, l_2D_spectra, l_2D_SD, l_corr, l_energy_trans
   real(cp), public :: bl_cut

   namelist/grid/n_r_max,n_cheb_max,m_max,minc,fd_ratio,fd_stretch, &
   &             l_var_m,rcut_m
   namelist/control/tag,n_time_steps,alpha,l_newmap,map_function,&
   &                alph1,alph2,dtMax,courfac,tEND,runHours,     &
   &                runMinutes,runSeconds,l_non_rot,dt_fac,      &
   &                n_fft_optim_lev,time_scheme,cheb_method,     &
   &                l_rerror_fix, rerror_fac, time_scale,        &
   &                matrix_solve,corio_term,buo_term,bc_method,  &
   &                mpi_transp,l_packed_transp,radial_scheme,    &
   &                c_sgs
   namelist/hdif/hdif_temp,hdif_vel,hdif_exp,hdif_m,hdif_comp
   namelist/phys_param/ra,ek,pr,raxi,sc,stef,radratio,g0,g1,g2,  &
   &                   ktopt,kbott,ktopv,kbotv,l_ek_pump,        &
   &                   l_temp_3D,tcond_fac,l_temp_advz,n_rings,  &
   &                   beta_shift,ktopxi,kbotxi,t_bot,t_top,     &
   &                   xi_bot,xi_top, l_xi_3D, xicond_fac,       &
   &                   h_temp, h_xi, container, beta_fac,        &
   &                   amp_forcing, radius_forcing, forcing_type,&
   &                   dx_forcing, ktopphi, kbotphi, epsPhase,   &
   &                   phaseDiffFac, penaltyFac, tmelt
   namelist/start_field/l_start_file,start_file,scale_t,init_t,amp_t, &
   &                    scale_u,init_u,amp_u,l_reset_t,amp_xi,init_xi,&
   &                    scale_xi, init_phi
   namelist/output_control/n_log_step,n_checkpoints, n_checkpoint_step, &
   &                       n_frames, n_frame_step, n_specs, n_spec_step,&
   &                       l_vphi_balance, l_vort_balance, bl_cut,      &
   &                       l_2D_spectra, l_2D_SD, l_corr,               &
   &                       l_energy_trans, m_max_transfer

   public :: read_namelists, fill_config, check_supported, has

contains

   ! reference defaults (namelists.f90:530-672)
   subroutine default_namelists()
      n_r_max = 32; n_cheb_max = 32; m_max = 32; minc = 1
      fd_stretch = 0.3_cp; fd_ratio = 0.1_cp; l_var_m = .false.; rcut_m = 0.1_cp
      n_time_steps = 100; tag = 'test'; alpha = 0.5_cp; l_newmap = .false.; c_sgs = 0.0_cp
      alph1 = 0.8_cp; alph2 = 0.0_cp; map_function = 'arcsin'
      dtMax = 1.0e-5_cp; courfac = 1.0e3_cp; dt_fac = 2.0_cp; tEND = 0.0_cp
      runHours = 0; runMinutes = 10; runSeconds = 0; n_fft_optim_lev = 1
      time_scheme = 'CNAB2'; radial_scheme = 'cheb'; cheb_method = 'colloc'; bc_method = 'tau-lanczos'
      matrix_solve = 'DIRECT'; l_rerror_fix = .false.; rerror_fac = 500.0_cp
      corio_term = 'IMPLICIT'; buo_term = 'IMPLICIT'; time_scale = 'VISC'; mpi_transp = 'AUTO'
      l_packed_transp = .false.
      hdif_vel = 0.0_cp; hdif_temp = 0.0_cp; hdif_comp = 0.0_cp; hdif_m = 0; hdif_exp = 0
      l_non_rot = .false.; l_ek_pump = .false.; l_temp_advz = .false.; l_temp_3D = .false.; l_xi_3D = .false.
      ra = 1.0e5_cp; pr = 1.0_cp; ek = 1.0e-3_cp; raxi = 0.0_cp; sc = 0.0_cp; stef = 0.0_cp
      tmelt = 0.5_cp; phaseDiffFac = 1.0_cp; epsPhase = 1.0e-2_cp; penaltyFac = 1.0_cp
      radratio = 0.35_cp; tcond_fac = 1.0_cp; xicond_fac = 1.0_cp; beta_shift = 0.0_cp; beta_fac = 0.4_cp
      g0 = 0.0_cp; g1 = 1.0_cp; g2 = 0.0_cp; h_temp = 0.0_cp; h_xi = 0.0_cp
      ktopt = 1; kbott = 1; ktopxi = 1; kbotxi = 1; ktopv = 2; kbotv = 2; ktopphi = 1; kbotphi = 1
      t_bot = 0.0_cp; t_top = 0.0_cp; xi_bot = 0.0_cp; xi_top = 0.0_cp
      container = 'sphere'; n_rings = 0; radius_forcing = 1.0e-2_cp; amp_forcing = 0.0_cp
      dx_forcing = 0.0_cp; forcing_type = 'polar'
      l_reset_t = .false.; l_start_file = .false.; start_file = 'no_start_file'
      init_t = 0; scale_t = 1.0_cp; amp_t = 0.0_cp; init_xi = 0; scale_xi = 1.0_cp; amp_xi = 0.0_cp
      init_u = 0; scale_u = 1.0_cp; amp_u = 0.0_cp; init_phi = 0
      n_log_step = 50; n_frames = 0; n_frame_step = 0; n_checkpoints = 1; n_checkpoint_step = 0
      n_specs = 0; n_spec_step = 0; l_vphi_balance = .false.; l_vort_balance = .false.
      l_2D_spectra = .false.; l_2D_SD = .false.; l_corr = .false.; bl_cut = 1.0e-3_cp
      l_energy_trans = .false.; m_max_transfer = -1
   end subroutine default_namelists

   ! readNamelists (namelists.f90:191-230): every group is optional
   subroutine read_namelists(fname)
      character(len=*), intent(in) :: fname
      integer :: u, ios
      call default_namelists()
      open(newunit=u, file=fname, status='old', action='read', iostat=ios)
      if (ios /= 0) then
         write(*, '(a)') 'pizza_gpu: cannot open input file '//trim(fname)
         error stop 1
      end if
      read(u, nml=grid, iostat=ios); call group_ok('grid')
      rewind(u); read(u, nml=control, iostat=ios); call group_ok('control')
      rewind(u); read(u, nml=hdif, iostat=ios); call group_ok('hdif')
      rewind(u); read(u, nml=phys_param, iostat=ios); call group_ok('phys_param')
      rewind(u); read(u, nml=start_field, iostat=ios); call group_ok('start_field')
      rewind(u); read(u, nml=output_control, iostat=ios); call group_ok('output_control')
      close(u)
   contains
      subroutine group_ok(name)
         character(len=*), intent(in) :: name
         if (ios > 0) then
            write(*, '(a)') 'pizza_gpu: error reading namelist &'//name
            error stop 1
         end if
         rewind(u)
      end subroutine group_ok
   end subroutine read_namelists

   logical function has(str, key)
      character(len=*), intent(in) :: str, key
      has = index(upper(str), key) /= 0
   end function has

   function upper(s) result(u)
      character(len=*), intent(in) :: s
      character(len=len(s)) :: u
      integer :: i, c
      u = s
      do i = 1, len(s)
         c = iachar(s(i:i))
         if (c >= iachar('a') .and. c <= iachar('z')) u(i:i) = achar(c - 32)
      end do
   end function upper

   integer(c_int) function flag(l)
      logical, intent(in) :: l
      flag = merge(1_c_int, 0_c_int, l)
   end function flag

   ! The scheme a time_scheme string selects: the family test of
   ! select_time_scheme.f90:15-49, then the first name that initialize()
   ! matches (dirk_schemes.f90:55-230, multistep_schemes.f90:56-104).
   integer(c_int) function scheme_id(name)
      character(len=*), intent(in) :: name
      character(len=10), parameter :: dirk(22) = [character(len=10) :: 'ARS222', 'SSP222', 'LSDIRK222', &
         'ARS232', 'ARS233', 'SSP322', 'SSP332', 'SSP433', 'DP1-242', 'DP2-242', 'LZ232', 'CK232', &
         'ARS443', 'LZ453', 'BPR353', 'PC2', 'ARS343', 'MARS343', 'KC343', 'KC564', 'KC785', 'DBM453']
      integer(c_int), parameter :: dirk_id(22) = [PZ_ARS222, PZ_SSP222, PZ_LSDIRK222, PZ_ARS232, &
         PZ_ARS233, PZ_SSP322, PZ_SSP332, PZ_SSP433, PZ_DP1_242, PZ_DP2_242, PZ_LZ232, PZ_CK232, &
         PZ_ARS443, PZ_LZ453, PZ_BPR353, PZ_PC2, PZ_ARS343, PZ_MARS343, PZ_KC343, PZ_KC564, PZ_KC785, &
         PZ_DBM453]
      character(len=10), parameter :: ms(10) = [character(len=10) :: 'CNAB2', 'MODCNAB', 'CNLF', &
         'BDF2AB2', 'SBDF2', 'BDF3AB3', 'SBDF3', 'TVB33', 'BDF4AB4', 'SBDF4']
      integer(c_int), parameter :: ms_id(10) = [PZ_CNAB2, PZ_MODCNAB, PZ_CNLF, PZ_SBDF2, PZ_SBDF2, &
         PZ_SBDF3, PZ_SBDF3, PZ_TVB33, PZ_SBDF4, PZ_SBDF4]
      integer :: i
      logical :: l_dirk
      l_dirk = .false.
      do i = 1, size(dirk)
         if (has(name, trim(dirk(i)))) l_dirk = .true.
      end do
      scheme_id = 0
      if (l_dirk) then
         do i = 1, size(dirk)
            if (has(name, trim(dirk(i)))) then
               scheme_id = dirk_id(i)
               exit
            end if
         end do
      else
         do i = 1, size(ms)
            if (has(name, trim(ms(i)))) then
               scheme_id = ms_id(i)
               exit
            end if
         end do
      end if
      if (scheme_id <= 0) then
         write(*, '(a)') 'pizza_gpu: time scheme '//trim(name)//' is not supported'
         error stop 1
      end if
   end function scheme_id

   ! namelist values -> pz_config (the mapping of namelists.f90:233-300)
   subroutine fill_config(cfg)
      type(pz_config), intent(out) :: cfg
      call pz_config_default(cfg)
      cfg%n_r_max = n_r_max; cfg%n_cheb_max = n_cheb_max; cfg%m_max = m_max; cfg%minc = minc
      cfg%time_scheme = scheme_id(time_scheme)
      cfg%l_cheb_coll = flag(has(cheb_method, 'COLL'))
      cfg%l_finite_diff = flag(.not. has(radial_scheme, 'CHEB'))
      cfg%fd_stretch = fd_stretch; cfg%fd_ratio = fd_ratio
      cfg%l_direct_solve = flag(has(matrix_solve, 'DIRECT'))
      cfg%l_coriolis_imp = flag(has(corio_term, 'IMP'))
      cfg%l_buo_imp = flag(has(buo_term, 'IMP'))
      cfg%time_scale_rot = flag(has(time_scale, 'ROT'))
      cfg%l_non_rot = flag(l_non_rot); cfg%l_rerror_fix = flag(l_rerror_fix)
      cfg%rerror_fac = rerror_fac; cfg%alpha = alpha; cfg%dtMax = dtMax; cfg%courfac = courfac
      cfg%dt_fac = dt_fac
      cfg%ra = ra; cfg%ek = ek; cfg%pr = pr; cfg%raxi = raxi; cfg%sc = sc; cfg%radratio = radratio
      cfg%g0 = g0; cfg%g1 = g1; cfg%g2 = g2
      cfg%ktopt = ktopt; cfg%kbott = kbott; cfg%ktopv = ktopv; cfg%kbotv = kbotv
      cfg%ktopxi = ktopxi; cfg%kbotxi = kbotxi
      cfg%l_ek_pump = flag(l_ek_pump); cfg%l_temp_advz = flag(l_temp_advz)
      cfg%tcond_fac = tcond_fac; cfg%xicond_fac = xicond_fac
      cfg%t_top = t_top; cfg%t_bot = t_bot; cfg%xi_top = xi_top; cfg%xi_bot = xi_bot
   end subroutine fill_config

   ! options the GPU path does not implement stop the run here, as the
   ! reference's abortRun would for an invalid combination
   subroutine check_supported()
      if (hdif_temp /= 0.0_cp .or. hdif_vel /= 0.0_cp .or. hdif_comp /= 0.0_cp) &
         call unsupported('hyperdiffusion (&hdif)')
      if (l_temp_3D .or. l_xi_3D) call unsupported('l_temp_3D / l_xi_3D')
      if (h_temp /= 0.0_cp .or. h_xi /= 0.0_cp) call unsupported('volumetric heating h_temp / h_xi')
      if (l_var_m) call unsupported('l_var_m')
      if (l_newmap) call unsupported('l_newmap')
      if (has(bc_method, 'GALERKIN')) call unsupported('bc_method galerkin')
      if (n_rings /= 0 .or. amp_forcing /= 0.0_cp) call unsupported('forcing')
      if (init_phi /= 0 .or. stef /= 0.0_cp) call unsupported('phase field')
   contains
      subroutine unsupported(what)
         character(len=*), intent(in) :: what
         write(*, '(a)') 'pizza_gpu: unsupported option: '//what
         error stop 2
      end subroutine unsupported
   end subroutine check_supported

end module namelists
