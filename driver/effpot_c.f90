! iso_c_binding interface of the gfx950 solver library (include/effpot.h).
! The derived types mirror the C structs field for field; every function
! returns the C status (0 = EFFPOT_OK).
module effpot_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: EFFPOT_OK = 0
  integer(c_int), parameter :: EFFPOT_CORE_OT = 0, EFFPOT_CORE_OTC = 1
  integer(c_int), parameter :: EFFPOT_PREC_FP64 = 0, EFFPOT_PREC_MIXED = 1
  integer(c_int), parameter :: EFFPOT_POT_NULL = 0, EFFPOT_POT_MORSE = 1

  type, bind(c) :: effpot_params
    integer(c_int32_t) :: nx, ny, nz, core, npd, icon, precision, batch, psi_complex, reserved_
    real(c_double) :: xmax, ymax, zmax, n4, h_core, eps4, sigma4, cp4, cpp4
    real(c_double) :: alphas, l, den0s, h2o2m4, cpen, beta, den_m
  end type

  type, bind(c) :: effpot_impurity
    integer(c_int32_t) :: kind, pad_
    real(c_double) :: d, re, a, r_cutoff, umax, cap
    real(c_double) :: rimp(3)
  end type

  type, bind(c) :: effpot_energies
    real(c_double) :: natoms, etot4, ekin4, elj4, ealphas, esolid, ecor4, eimpu
    real(c_double) :: ekinx, ehex, uplus, eso, etot, mu, xcm, ycm, zcm, resid
  end type

  type, bind(c) :: effpot_solve_opts
    real(c_double) :: dt, tol
    integer(c_int32_t) :: n_rk, pener, max_steps, polish_steps
    real(c_double) :: tol_fp32
    integer(c_int32_t) :: check_finite, graph_steps
  end type

  type, bind(c) :: effpot_snapshot
    integer(c_int32_t) :: nx, ny, nz, ninvar
    real(c_double) :: xmax, ymax, zmax, hx, hy, hz
    real(c_double) :: rimp(3), vimp(3)
    real(c_double) :: invar(32)
    integer(c_int32_t) :: binary, reserved_
  end type

  type, bind(c) :: effpot_plus
    integer(c_int32_t) :: kind, nterms, k0, pad_
    real(c_double) :: r_cutoff, umax, rcutoff, fcutoff, r0, aa, bb, cc
    real(c_double) :: pg(16), pg_lo(16)
  end type

  integer(c_int), parameter :: EFFPOT_STATE_P = 1, EFFPOT_STATE_D = 2
  integer(c_int), parameter :: EFFPOT_SELECT_PAIR = 0, EFFPOT_SELECT_JZ = 1, EFFPOT_SELECT_FIX = 2, &
                               EFFPOT_SELECT_AVERAGE = 3, EFFPOT_SELECT_MIXTURE = 4
  type, bind(c) :: effpot_select
    integer(c_int32_t) :: mode, jz2, fixed_moments, pad_
    real(c_double) :: a_pi_3o2, a_sig_1o2
  end type
  type, bind(c) :: effpot_state_imp
    integer(c_int32_t) :: state, instate
    real(c_double) :: als, rimp(3)
    type(effpot_plus) :: v_pi, v_sigma, v_delta
    real(c_double) :: invar(20), eigen(10), moments(31), eso
    type(effpot_select) :: select
  end type

  type, bind(c) :: effpot_result
    type(effpot_energies) :: e
    integer(c_int32_t) :: steps, converged
    real(c_double) :: rel_de, err_he
  end type

  interface
    integer(c_int) function effpot_ctx_create(p, device, ctx) bind(c, name="effpot_ctx_create")
      import :: c_int, c_ptr, effpot_params
      type(effpot_params), intent(in) :: p
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function effpot_ctx_destroy(ctx) bind(c, name="effpot_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function effpot_last_error(ctx) bind(c, name="effpot_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function effpot_upload_psi(ctx, slot, psi) bind(c, name="effpot_upload_psi")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      complex(c_double_complex), intent(in) :: psi(*)
    end function
    integer(c_int) function effpot_download_psi(ctx, slot, psi) bind(c, name="effpot_download_psi")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      complex(c_double_complex), intent(out) :: psi(*)
    end function
    integer(c_int) function effpot_init_fermi(ctx, slot, rho0, r, afermi, noise, seed, centre) &
        bind(c, name="effpot_init_fermi")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      real(c_double), value :: rho0, r, afermi, noise
      integer(c_int64_t), value :: seed
      real(c_double), intent(in) :: centre(3)
    end function
    integer(c_int) function effpot_set_impurity(ctx, slot, imp) bind(c, name="effpot_set_impurity")
      import :: c_int, c_ptr, effpot_impurity
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      type(effpot_impurity), intent(in) :: imp
    end function
    integer(c_int) function effpot_poten(ctx) bind(c, name="effpot_poten")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function effpot_energy(ctx, e) bind(c, name="effpot_energy")
      import :: c_int, c_ptr, effpot_energies
      type(c_ptr), value :: ctx
      type(effpot_energies), intent(out) :: e
    end function
    integer(c_int) function effpot_solve(ctx, opt, res) bind(c, name="effpot_solve")
      import :: c_int, c_ptr, effpot_solve_opts, effpot_result
      type(c_ptr), value :: ctx
      type(effpot_solve_opts), intent(in) :: opt
      type(effpot_result), intent(out) :: res
    end function
    ! per-step replacements of steprk(deltat) / steppc(deltat, errHe, ...)
    integer(c_int) function effpot_step_rk(ctx, dt) bind(c, name="effpot_step_rk")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function effpot_step_pc(ctx, dt, err_he, err_imp, err_vimp, err_iv) bind(c, name="effpot_step_pc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      real(c_double), intent(out) :: err_he, err_imp, err_vimp, err_iv
    end function
    integer(c_int) function effpot_renormalize(ctx) bind(c, name="effpot_renormalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function effpot_set_natoms(ctx, slot, n4) bind(c, name="effpot_set_natoms")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      real(c_double), value :: n4
    end function
    integer(c_int) function effpot_set_dt(ctx, slot, dt) bind(c, name="effpot_set_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      real(c_double), value :: dt
    end function
    integer(c_int) function effpot_download_den(ctx, slot, den) bind(c, name="effpot_download_den")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      real(c_double), intent(out) :: den(*)
    end function
    ! Select_Pot(selec_plus, ...) of the uplus term: preset by name, then per slot
    integer(c_int) function effpot_plus_preset(selec, r_cutoff, umax, plus) bind(c, name="effpot_plus_preset")
      import :: c_int, c_char, c_double, effpot_plus
      character(kind=c_char), intent(in) :: selec(*)
      real(c_double), value :: r_cutoff, umax
      type(effpot_plus), intent(out) :: plus
    end function
    integer(c_int) function effpot_set_plus(ctx, slot, plus) bind(c, name="effpot_set_plus")
      import :: c_int, c_ptr, effpot_plus
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      type(effpot_plus), intent(in) :: plus
    end function
    ! z_exciplex_exclusion of energy(): points nearer than r leave eHeX and uplus
    integer(c_int) function effpot_set_exclusion(ctx, slot, r) bind(c, name="effpot_set_exclusion")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      real(c_double), value :: r
    end function
    ! Lfilter_exciplex_force: the z force of a moving impurity beyond the exclusion only
    integer(c_int) function effpot_set_force_filter(ctx, slot, on) bind(c, name="effpot_set_force_filter")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot, on
    end function
    ! P/D internal-state impurity: potenimpini + instates + potenimp
    integer(c_int) function effpot_set_impurity_state(ctx, slot, imp) bind(c, name="effpot_set_impurity_state")
      import :: c_int, c_ptr, effpot_state_imp
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      type(effpot_state_imp), intent(inout) :: imp
    end function
    ! Lexciplex_state_fix: P moments of a Ring (6 He) or Linear (2 He) exciplex
    integer(c_int) function effpot_exciplex_moments(v_pi, v_sigma, n_he, axis, r_exc, moments) &
        bind(c, name="effpot_exciplex_moments")
      import :: c_int, c_double, effpot_plus
      type(effpot_plus), intent(in) :: v_pi, v_sigma
      integer(c_int), value :: n_he, axis
      real(c_double), value :: r_exc
      real(c_double), intent(out) :: moments(7)
    end function
    ! readenc mode 3 (formatted or unformatted file); psi = c_null_ptr reads the header only
    integer(c_int) function effpot_snapshot_read(path, hdr, psi) bind(c, name="effpot_snapshot_read")
      import :: c_int, c_char, c_ptr, effpot_snapshot
      character(kind=c_char), intent(in) :: path(*)
      type(effpot_snapshot), intent(out) :: hdr
      type(c_ptr), value :: psi
    end function
    ! derivnd(id, nn, h, iv, f, df, icon) on a real (nx,ny,nz) array, grid and icon of the context
    integer(c_int) function effpot_derivnd(ctx, id, iv, f, df) bind(c, name="effpot_derivnd")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: id, iv
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: df(*)
    end function
  end interface

contains

  ! Turns a non-zero status into `error stop` with the library's message
  ! (the reference STOPs at the same points).
  subroutine check(ctx, rc, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == EFFPOT_OK) return
    text = ''
    ! a null context gives the thread's last snapshot (file) error
    call c_f_pointer(effpot_last_error(ctx), msg, [512])
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(0, '(a,": status ",i0,": ",a)') where, rc, trim(text)
    error stop 'effpot: severe error'
  end subroutine check
end module effpot_c
