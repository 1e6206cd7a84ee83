! Driver of the GPU eff_pot path.  Same namelist (/input/, DFT4He3d.f90:104-134
! subset plus the Morse/relaxation keys), the same density-file format
! (readenc mode 3, readenc.f90:151-168) and the same stdout energy labels
! (DFT4He3d.f90:460-463, formats 6050/6060) as the Fortran reference, so that
! get_impurity_energy.sh and loop.sh keep working; the field operators run on
! the GPU through effpot_c.
program effpot_driver
  use iso_c_binding
  use effpot_c
  implicit none

  character(len=40) :: title = 'Helium4 - 3dim.  '
  character(len=60) :: filedenin = 'he4.input.density'
  character(len=3) :: core4 = 'OTC'
  character(len=80) :: selec_pi = 'Morse'
  character(len=80) :: selec_plus = 'Ba_plus_2D2'   ! X+-He potential of uplus (V_impur.f90:14-16)
  ! P/D internal state (selec_pi other than 'Morse'/'null'): V_impur.f90:5-13, modules.f90:299-338
  character(len=80) :: selec_sigma = 'Ba_plus_2D0', selec_delta = 'Ba_plus_2D2'
  character(len=1) :: lstate = 'P'
  integer :: instate = 0
  integer :: n4 = 300, mode = 0, nx = 64, ny = 64, nz = 64, npd = 13, icon = 13
  integer :: niter = 0, pener = 50, nthread = 1, precision = 0, ninvar = 10
  integer(c_int64_t) :: seed = 20261203_c_int64_t
  real(8) :: xmax = 24.d0, ymax = 24.d0, zmax = 24.d0, xc = 0.d0, yc = 0.d0, zc = 0.d0
  real(8) :: eps4 = 10.22d0, sigma4 = 2.556d0, h_core = 2.377d0, l = 1.d0
  real(8) :: cp4 = -2.41186d4, cpp4 = 1.85850d6, den4c = 0.062d0, alphas = 54.31d0, den0s = 0.04d0
  real(8) :: h2o2m4 = 6.05969638298d0, deltat = 0.d0, deltatps = 0.d0
  real(8) :: ximp = 0.d0, yimp = 0.d0, zimp = 0.d0
  real(8) :: morse_d = 30.d0, morse_re = 4.d0, morse_a = 1.d0
  real(8) :: r_cutoff_pi = 0.d0, umax_pi = 0.d0, afermi = 0.6d0, rho0 = 0.0218d0, noise = 1.d-6
  real(8) :: r_cutoff_plus = 1.05d0, umax_plus = 6.260015d4
  real(8) :: r_cutoff_sigma = 1.05d0, umax_sigma = 9.674836d4, r_cutoff_delta = 1.05d0, umax_delta = 6.260015d4
  real(8) :: als_p = 0.d0, als_d = 0.d0
  ! instates' choice of invar (modules.f90:284-318)
  logical :: ldiag_jz = .false., lexcite_state_fix = .false., laverage_p_value = .false., lmixture = .false.
  logical :: lexciplex_state_fix = .false., lfilter_exciplex_force = .false.
  character(len=4) :: ljz = ''
  character(len=6) :: exciplex = 'Ring'
  character(len=1) :: lexciplex_axis = 'Z'
  real(8) :: a_pi_3o2 = 0.94d0, a_sig_1o2 = 0.34117444d0, r_exc = 3.9d0
  real(8) :: z_exciplex_exclusion = 0.d0   ! energy.f90:155 (modules.f90 default 0)
  real(8) :: tol = 1.d-9
  logical :: fast_io = .true.   ! mode 3: the library's reader instead of list-directed READ
  ! The other keys of the reference's /input/ (DFT4He3d.f90:104-134; types and
  ! defaults of modules.f90, DFT4He3d.f90:26-99, V_impur.f90:2-4,51-53): read
  ! and not used by this driver (real-time, surface, constraint and output
  ! options of the TDDFT code), so that an input.dat written for the reference,
  ! as prep.sh edits it, reads unchanged.
  character(len=15) :: fftwplan = 'FFTW_PATIENT'
  character(len=60) :: fileout = 'DFT4He3d.res', filerimp = 'rimp.out', filevimp = 'vimp.out', &
                       fileaimp = 'aimp.out', filedenout = 'he4.output.density', fileimpin = 'X.input.wf', &
                       fileimpout = 'X.output.wf'
  character(len=80) :: selec_gs = 'Ba_plus_gs'
  integer :: nsfiles = 1, pchem = 50, irespar = 0, pdenpar = 50, ppot = 1, iron = 1, ironx = 1, norder = 0, &
             nv_iter = 0, iter0 = 0, pcurr = 0, icurr = 0, ndmax = 2
  real(8) :: denmin = 1.d-60, psimin = 1.d-30, rimpur = 5.5d0, mimpur = 0.d0, xlamda = 0.d0, xlamdax = 0.d0, &
             xlamday = 0.d0, xlamdaz = 0.d0, xlamdax_x = 0.d0, xlamdax_y = 0.d0, xlamdax_z = 0.d0, &
             time0 = -1.d0, zsurf = -25.d0, lambdah = 0.d0, tzmean = 0.d0, tzsurf = 0.d0, vximp = 0.d0, &
             vyimp = 0.d0, vzimp = 0.d0, r_cutoff_gs = 1.05d0, umax_gs = 6.836177d3, ev0 = 0.d0, &
             quita_c4_ba_plus_gs_fix_c4 = 0.d0, quita_c4_ba_plus_pi_fix_c4 = 0.d0, &
             quita_c4_ba_plus_sigma_fix_c4 = 0.d0, z_exciplex_position = -3.5d0
  logical :: printpot = .false., lrkpc = .true., lrk = .false., lsurf = .false., lsurf3d = .false., &
             lexcite_state = .false., lexcite_state_external = .false., lfix_lambda = .false., &
             l_v_aditional = .false.
  real(8), parameter :: ktops = 7.63822291d0
  namelist /input/ title, filedenin, core4, selec_pi, n4, mode, nx, ny, nz, npd, icon, niter, pener, &
                   nthread, precision, seed, xmax, ymax, zmax, xc, yc, zc, eps4, sigma4, h_core, l, &
                   cp4, cpp4, den4c, alphas, den0s, h2o2m4, deltat, deltatps, ximp, yimp, zimp, &
                   morse_d, morse_re, morse_a, r_cutoff_pi, umax_pi, afermi, rho0, noise, tol, ninvar, &
                   fast_io, selec_plus, r_cutoff_plus, umax_plus, selec_sigma, selec_delta, lstate, &
                   instate, r_cutoff_sigma, umax_sigma, r_cutoff_delta, umax_delta, als_p, als_d, &
                   ldiag_jz, ljz, lexcite_state_fix, laverage_p_value, lmixture, a_pi_3o2, a_sig_1o2, &
                   lexciplex_state_fix, exciplex, lexciplex_axis, r_exc, z_exciplex_exclusion, &
                   lfilter_exciplex_force, &
                   fftwplan, nsfiles, fileout, filerimp, filevimp, fileaimp, filedenout, fileimpin, &
                   fileimpout, denmin, psimin, printpot, pchem, irespar, pdenpar, ppot, iron, ironx, &
                   rimpur, mimpur, norder, nv_iter, lrkpc, lrk, xlamda, xlamdax, xlamday, xlamdaz, &
                   xlamdax_x, xlamdax_y, xlamdax_z, iter0, time0, zsurf, pcurr, icurr, lsurf, lsurf3d, &
                   lambdah, tzmean, tzsurf, vximp, vyimp, vzimp, ndmax, selec_gs, r_cutoff_gs, umax_gs, &
                   ev0, lexcite_state, quita_c4_ba_plus_gs_fix_c4, quita_c4_ba_plus_pi_fix_c4, &
                   quita_c4_ba_plus_sigma_fix_c4, lexcite_state_external, lfix_lambda, l_v_aditional, &
                   z_exciplex_position

  type(effpot_params) :: p
  type(effpot_impurity) :: imp
  type(effpot_plus) :: plus
  type(effpot_state_imp) :: simp
  type(effpot_energies) :: e
  type(effpot_solve_opts) :: opt
  type(effpot_result) :: res
  type(c_ptr) :: ctx
  complex(8), allocatable, target :: psi(:,:,:)
  type(effpot_snapshot) :: snap
  real(8) :: xmaxp, ymaxp, zmaxp, hxp, hyp, hzp, rimp(3), vimp(3), hx, hy, hz
  integer :: nxp, nyp, nzp, nninvar0, isalto, islash, iaxis
  complex(8), allocatable :: invar0(:)
  real(8) :: dist

  read(5, nml=input)
  open(10, file="DFT4He3d.namelist.read")
  write(10, nml=input)
  close(10)
  if (tstgrid(nx) /= 0) stop 'SEVERE ERROR: NX is not correct'
  if (tstgrid(ny) /= 0) stop 'SEVERE ERROR: NY is not correct'
  if (tstgrid(nz) /= 0) stop 'SEVERE ERROR: NZ is not correct'
  if (deltat == 0.d0) then
    if (deltatps == 0.d0) then
      if (niter > 0) stop 'You must specify either Deltat (in kelvin) or Deltatps (in picosecond)'
    else
      deltat = deltatps / ktops
    end if
  else if (deltatps /= 0.d0 .and. deltatps /= deltat * ktops) then
    stop 'Inconsistent deltat - deltatps'
  end if

  p%nx = nx; p%ny = ny; p%nz = nz; p%npd = npd; p%icon = icon; p%precision = precision; p%batch = 1
  p%psi_complex = 1; p%reserved_ = 0   ! snapshots (mode 3) may carry a phase
  select case (core4)
  case ('OT ', 'OT')
    p%core = EFFPOT_CORE_OT
  case ('OTC')
    p%core = EFFPOT_CORE_OTC
  case default
    stop ' ... The program stops due to a severe error.'
  end select
  p%xmax = xmax; p%ymax = ymax; p%zmax = zmax; p%n4 = dble(n4); p%h_core = h_core
  p%eps4 = eps4; p%sigma4 = sigma4; p%cp4 = cp4; p%cpp4 = cpp4; p%alphas = alphas; p%l = l
  p%den0s = den0s; p%h2o2m4 = h2o2m4; p%cpen = 3.1577504d4; p%beta = 40.d0; p%den_m = 0.37d0
  call check(c_null_ptr, effpot_ctx_create(p, 0_c_int, ctx), 'effpot_ctx_create')

  hx = 2.d0 * xmax / nx; hy = 2.d0 * ymax / ny; hz = 2.d0 * zmax / nz
  rimp = (/ximp, yimp, zimp/)
  select case (mode)
  case (0)   ! Fermi droplet (the start readenc.f90:15-17 describes)
    call check(ctx, effpot_init_fermi(ctx, 0_c_int, rho0, 2.22d0 * dble(n4)**(1.d0/3.d0), afermi, noise, &
                                      seed, (/xc, yc, zc/)), 'effpot_init_fermi')
  case (3)   ! wavefunction snapshot (readenc.f90:151-168)
    allocate(psi(nx, ny, nz))
    if (fast_io) then   ! one pass over the file in C: formatted or unformatted
      call check(c_null_ptr, effpot_snapshot_read(trim(filedenin) // c_null_char, snap, c_null_ptr), &
                 'effpot_snapshot_read')
      if (snap%nx /= nx .or. snap%ny /= ny .or. snap%nz /= nz) stop 'readenc: grid of the file differs from the namelist'
      call check(c_null_ptr, effpot_snapshot_read(trim(filedenin) // c_null_char, snap, c_loc(psi)), &
                 'effpot_snapshot_read')
      rimp = snap%rimp
      call check(ctx, effpot_upload_psi(ctx, 0_c_int, psi), 'effpot_upload_psi')
    else
    open(unit=1, file=filedenin, status='old')
    call titols(1, '#', isalto)
    read(1, *) xmaxp, ymaxp, zmaxp, hxp, hyp, hzp, nxp, nyp, nzp
    if (nxp /= nx .or. nyp /= ny .or. nzp /= nz) stop 'readenc: grid of the file differs from the namelist'
    read(1, *) rimp
    read(1, *) vimp
    read(1, *) nninvar0
    allocate(invar0(nninvar0))
    read(1, *) invar0
    read(1, *) psi
    close(1)
    call check(ctx, effpot_upload_psi(ctx, 0_c_int, psi), 'effpot_upload_psi')
    end if
  case default
    stop 'ERROR READEN: This mode is still not programmed'
  end select

  imp%pad_ = 0
  if (trim(selec_pi) == 'Morse' .or. trim(selec_pi) == 'null') then   ! isotropic impurity
    imp%kind = merge(EFFPOT_POT_MORSE, EFFPOT_POT_NULL, trim(selec_pi) == 'Morse')
    imp%d = morse_d; imp%re = morse_re; imp%a = morse_a
    imp%r_cutoff = r_cutoff_pi; imp%umax = umax_pi; imp%cap = 1000.d0; imp%rimp = rimp
    call check(ctx, effpot_set_impurity(ctx, 0_c_int, imp), 'effpot_set_impurity')
  else   ! P or D internal state: potenimpini, instates (readenc.f90:176-221), potenimp
    select case (lstate)
    case ('P')
      simp%state = EFFPOT_STATE_P; simp%als = als_p
    case ('D')
      simp%state = EFFPOT_STATE_D; simp%als = als_d
    case default
      stop 'Lstate must be P or D'
    end select
    if (instate < 0) stop 'instate less than 0, still not programmed'
    simp%instate = instate; simp%rimp = rimp
    call preset(selec_pi, r_cutoff_pi, umax_pi, simp%v_pi)
    call preset(selec_sigma, r_cutoff_sigma, umax_sigma, simp%v_sigma)
    call preset(selec_delta, r_cutoff_delta, umax_delta, simp%v_delta)
    ! the branches of instates in their order of precedence (instates.f90:228-364, :613-726)
    simp%select%mode = EFFPOT_SELECT_PAIR; simp%select%jz2 = 0; simp%select%fixed_moments = 0
    simp%select%pad_ = 0; simp%select%a_pi_3o2 = a_pi_3o2; simp%select%a_sig_1o2 = a_sig_1o2
    if (ldiag_jz) then
      simp%select%mode = EFFPOT_SELECT_JZ
      if (len_trim(ljz) > 0) then
        islash = index(ljz, '/2')
        if (islash < 2) then
          write(6, '("Wrong Ljz value...:",A)') ljz
          stop 'Wrong Ljz value..'
        end if
        read(ljz(1:islash-1), *) simp%select%jz2
      end if
    else if (lstate == 'P' .and. lexcite_state_fix) then
      simp%select%mode = EFFPOT_SELECT_FIX
    else if (lstate == 'P' .and. laverage_p_value) then
      simp%select%mode = EFFPOT_SELECT_AVERAGE
    else if (lstate == 'P' .and. lmixture) then
      simp%select%mode = EFFPOT_SELECT_MIXTURE
    end if
    simp%moments = 0.d0
    if (lstate == 'P' .and. lexciplex_state_fix) then   ! instates.f90:54-113
      iaxis = index('XYZ', lexciplex_axis) - 1
      if (iaxis < 0) stop 'Instates ---> Iaxis = 0'
      if (trim(exciplex) /= 'Ring' .and. trim(exciplex) /= 'Linear') then
        write(6, '("From Instates: Exciplex variable is not correct...",A)') exciplex
        stop 'From instates  001'
      end if
      call check(ctx, effpot_exciplex_moments(simp%v_pi, simp%v_sigma, merge(6_c_int, 2_c_int, &
                 trim(exciplex) == 'Ring'), int(iaxis, c_int), r_exc, simp%moments(1:7)), 'effpot_exciplex_moments')
      simp%select%fixed_moments = 1
    end if
    call check(ctx, effpot_set_impurity_state(ctx, 0_c_int, simp), 'effpot_set_impurity_state')
  end if
  call preset(selec_plus, r_cutoff_plus, umax_plus, plus)
  call check(ctx, effpot_set_plus(ctx, 0_c_int, plus), 'effpot_set_plus')
  call check(ctx, effpot_set_exclusion(ctx, 0_c_int, z_exciplex_exclusion), 'effpot_set_exclusion')
  call check(ctx, effpot_set_force_filter(ctx, 0_c_int, merge(1_c_int, 0_c_int, lfilter_exciplex_force)), &
             'effpot_set_force_filter')

  if (niter > 0) then
    opt%dt = deltat; opt%tol = tol; opt%n_rk = 3; opt%pener = pener; opt%max_steps = niter
    opt%polish_steps = niter; opt%tol_fp32 = 2.d-8; opt%check_finite = 1; opt%graph_steps = 24
    call check(ctx, effpot_solve(ctx, opt, res), 'effpot_solve')
    e = res%e
    write(6, '(" Relaxation: steps",i8,"  converged",i2,"  |dE/E| per step",1p,e12.4)') res%steps, &
      res%converged, res%rel_de
  else
    call check(ctx, effpot_poten(ctx), 'effpot_poten')
    call check(ctx, effpot_energy(ctx, e), 'effpot_energy')
  end if

  write(6, '("Number of He4 atoms",1P,E15.6)') e%natoms
  write(6, 6050) e%natoms, e%etot4, e%etot4 / e%natoms, e%ekin4, e%elj4, e%ealphas, e%esolid, e%ecor4
  write(6, 6060) e%eimpu, e%ekinx, e%ehex, e%uplus, e%eso, e%etot
  dist = sqrt((rimp(1) - e%xcm)**2 + (rimp(2) - e%ycm)**2 + (rimp(3) - e%zcm)**2)
  write(6, *) "X-HE CoM distance: ", dist
  call check(ctx, effpot_ctx_destroy(ctx), 'effpot_ctx_destroy')

6050 format(//,T5,'FIRST ENERGY BALANCE: ',                    //     &
              ,T5,'Number of particles ..........: ',F18.6,'  ',/,    &
               T5,'TOTAL   energy (He) ..........: ',F18.6,' K',/,    &
               T5,'Energy per particle (He) .....: ',F18.6,' K',/,    &
               T5,'Kinetic energy (He) ..........: ',F18.6,' K',/,    &
               T5,'Lennard-Jones energy (He) ....: ',F18.6,' K',/,    &
               T5,'Alpha_s term  energy (He) ....: ',F18.6,' K',/,    &
               T5,'Solid energy (He)  ...........: ',F18.6,' K',/,    &
               T5,'Correlation energy   (He) ....: ',F18.6,' K')
6060 format(1x,T5,'Impurity energy (X) ..........: ',F18.6,' K',/,    &
               T5,'Kinetic energy (X) ...........: ',F18.6,' K',/,    &
               T5,'Interaction energy (X*-He) ...: ',F18.6,' K',/,    &
               T5,'Interaction energy (X+-He) ...: ',F18.6,' K',/,    &
               T5,'Spin-Orbit energy (X) ........: ',F18.6,' K',/,    &
               T5,'TOTAL ENERGY (He+X) ..........: ',F18.6,' K',/)

contains

  subroutine preset(selec, r_cutoff, umax, v)   ! Select_Pot by name, or its stop
    character(len=*), intent(in) :: selec
    real(8), intent(in) :: r_cutoff, umax
    type(effpot_plus), intent(out) :: v
    if (effpot_plus_preset(trim(selec) // c_null_char, r_cutoff, umax, v) /= EFFPOT_OK) then
      write(6, '(A,": aquest potencial no est definit, de part de: Selec_Pot")') trim(selec)
      stop 'From V_impur: 0001'
    end if
  end subroutine preset

  integer function tstgrid(n)   ! n = 2^a 3^b 5^c <= 256 (tstgrid.f90:8-12)
    integer, intent(in) :: n
    integer :: m
    tstgrid = 1
    if (n < 16 .or. n > 256) return
    m = n
    do while (mod(m, 2) == 0); m = m / 2; end do
    do while (mod(m, 3) == 0); m = m / 3; end do
    do while (mod(m, 5) == 0); m = m / 5; end do
    if (m == 1) tstgrid = 0
  end function tstgrid

  subroutine titols(ulog, cchar, isalto)   ! skip the leading comment lines (titols.f90)
    integer, intent(in) :: ulog
    character(len=1), intent(in) :: cchar
    integer, intent(out) :: isalto
    character(len=1) :: pchar
    integer :: i, ios
    isalto = 0
    do
      read(ulog, '(A1)', iostat=ios) pchar
      if (ios /= 0) stop 'STOP due to severe errors in routine TITOLS'
      if (pchar /= cchar) exit
      isalto = isalto + 1
    end do
    rewind(ulog)
    do i = 1, isalto
      read(ulog, *)
    end do
  end subroutine titols
end program effpot_driver
