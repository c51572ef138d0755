! snac_amd.f90 -- Fortran binding of the C-ABI in snac_amd.h.
!
! The bind(C) interface idiom of the reference's fftw.f90:1-29, for the
! entry points a Fortran host needs in place of the hot-path routines of
! main.f90:893-1069 (rk_mom, bounduvw, fillps, the Poisson solve, correc,
! updt_pressure, chkdiv, chkdt, ...) and the set-up of main.f90:219-768.
! Every function returns 0 on success; snac_check turns a failure into the
! reference's way of reporting one: the message on stderr, then error stop.
! Strings cross as NUL-terminated character(kind=c_char) arrays
! ('dns.nml'//c_null_char).  Fields are real(c_double) arrays in the
! reference's own layout, (lo-1:hi+1) along each axis, column-major.
module mod_snac_amd
  use, intrinsic :: iso_c_binding
  use, intrinsic :: iso_fortran_env, only: error_unit
  implicit none
  private :: error_unit
  integer(c_int), parameter :: SNAC_U = 0, SNAC_V = 1, SNAC_W = 2, SNAC_P = 3, SNAC_PP = 4, SNAC_PO = 5, &
                               SNAC_DUDTRKO = 6, SNAC_DVDTRKO = 7, SNAC_DWDTRKO = 8, SNAC_SCAL0 = 16
  ! scalar iscal (1-based): SNAC_SCAL0+2*(iscal-1) = scalars(iscal)%val, +1 = %dsdtrko
  integer(c_int), parameter :: SNAC_SOLVER_AUTO = 0, SNAC_SOLVER_TENSOR = 1, SNAC_SOLVER_MG = 2
  integer(c_int), parameter :: SNAC_PREC_FP64 = 0, SNAC_PREC_MIXED = 1, SNAC_PREC_LEARNED = 2
  interface
    type(c_ptr) function snac_last_error() bind(C, name='snac_last_error')
      import
    end function
    integer(c_size_t) function snac_c_strlen(s) bind(C, name='strlen')
      import; type(c_ptr), value :: s
    end function
    ! set-up: read_input + initmpi + initgrid + init_bc_rhs + init_fft_reduction +
    ! init_n_2d_matrices + create/setup_n_solvers (main.f90:219-768); device -1: host-only
    integer(c_int) function snac_ctx_create_from_nml(ctx, dns, blocks, fft_axis, device) bind(C)
      import; type(c_ptr) :: ctx; character(kind=c_char), dimension(*) :: dns, blocks
      integer(c_int), value :: fft_axis, device
    end function
    integer(c_int) function snac_ctx_destroy(ctx) bind(C)           ! finalize_n_matrices / fftend
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function snac_ctx_num_blocks(ctx) bind(C)
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function snac_ctx_block_info(ctx, b, n3, lo_g3, nb6, periods3) bind(C)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: b
      integer(c_int), dimension(3) :: n3, lo_g3, periods3; integer(c_int), dimension(6) :: nb6
    end function
    integer(c_int) function snac_ctx_get_metric(ctx, b, which, out) bind(C)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: b, which; real(c_double), dimension(*) :: out
    end function
    integer(c_int) function snac_ctx_get_params(ctx, dparams, iparams) bind(C)
      import; type(c_ptr), value :: ctx; real(c_double), dimension(11) :: dparams; integer(c_int), dimension(8) :: iparams
    end function
    integer(c_int) function snac_ctx_set_solver(ctx, solver) bind(C)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: solver
    end function
    integer(c_int) function snac_set_precision(ctx, mode) bind(C)   ! no reference counterpart
      import; type(c_ptr), value :: ctx; integer(c_int), value :: mode
    end function
    integer(c_int) function snac_solver_stats(ctx, iters, relres) bind(C) ! HYPRE iteration count / residual
      import; type(c_ptr), value :: ctx; integer(c_int) :: iters; real(c_double) :: relres
    end function
    integer(c_int) function snac_upload_field(ctx, b, fid, a) bind(C)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: b, fid; real(c_double), dimension(*) :: a
    end function
    integer(c_int) function snac_download_field(ctx, b, fid, a) bind(C)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: b, fid; real(c_double), dimension(*) :: a
    end function
    integer(c_int) function snac_load_one(ctx, io, filename, b, fid, time, istep) bind(C) ! load.f90:84-187
      import; type(c_ptr), value :: ctx; character(kind=c_char), value :: io
      character(kind=c_char), dimension(*) :: filename; integer(c_int), value :: b, fid
      real(c_double) :: time; integer(c_int) :: istep
    end function
    integer(c_int) function snac_restart_io(ctx, io, datadir, time, istep) bind(C) ! main.f90:543-560,1090-1134
      import; type(c_ptr), value :: ctx; character(kind=c_char), value :: io
      character(kind=c_char), dimension(*) :: datadir; real(c_double) :: time; integer(c_int) :: istep
    end function
    integer(c_int) function snac_write_visu_3d(ctx, datadir, fname_bin, fname_log, varname, b, fid, &
                                               nmin, nmax, nskip, time, istep) bind(C) ! output.f90:289-303
      import; type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*) :: datadir, fname_bin, fname_log, varname
      integer(c_int), value :: b, fid; integer(c_int), dimension(3) :: nmin, nmax, nskip
      real(c_double), value :: time; integer(c_int), value :: istep
    end function
    integer(c_int) function snac_save_grid(ctx, datadir) bind(C)    ! initgrid.f90:128-143, main.f90:429-437
      import; type(c_ptr), value :: ctx; character(kind=c_char), dimension(*) :: datadir
    end function
    integer(c_int) function snac_initflow(ctx) bind(C)              ! initflow.f90:10-262 (no noise)
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function snac_rk_mom(ctx, irk, dt) bind(C)         ! rk.f90:7-128
      import; type(c_ptr), value :: ctx; integer(c_int), value :: irk; real(c_double), value :: dt
    end function
    integer(c_int) function snac_set_buoyancy(ctx, enable) bind(C)   ! the BOUSSINESQ_BUOYANCY=1 build
      import; type(c_ptr), value :: ctx; integer(c_int), value :: enable
    end function
    integer(c_int) function snac_rk_scal(ctx, irk, dt) bind(C)        ! rk.f90:130-204 + bulk_forcing_s
      import; type(c_ptr), value :: ctx; integer(c_int), value :: irk; real(c_double), value :: dt
    end function
    integer(c_int) function snac_scalar_forcing(ctx, f) bind(C)      ! s%f of main.f90:911-913
      import; type(c_ptr), value :: ctx; real(c_double), dimension(*) :: f
    end function
    integer(c_int) function snac_bounduvw(ctx, is_correc) bind(C)    ! bound.f90:8-66
      import; type(c_ptr), value :: ctx; integer(c_int), value :: is_correc
    end function
    integer(c_int) function snac_inflow(ctx, is_correc) bind(C)      ! bound.f90:438-477
      import; type(c_ptr), value :: ctx; integer(c_int), value :: is_correc
    end function
    integer(c_int) function snac_boundp(ctx, fid) bind(C)            ! bound.f90:68-106
      import; type(c_ptr), value :: ctx; integer(c_int), value :: fid
    end function
    integer(c_int) function snac_fillps(ctx, dtrk) bind(C)           ! fillps.f90 + updt_rhs
      import; type(c_ptr), value :: ctx; real(c_double), value :: dtrk
    end function
    integer(c_int) function snac_solve_pressure(ctx) bind(C)        ! main.f90:1004-1022
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function snac_correc(ctx, dtrk) bind(C)           ! correc.f90
      import; type(c_ptr), value :: ctx; real(c_double), value :: dtrk
    end function
    integer(c_int) function snac_updt_pressure(ctx) bind(C)         ! updt_pressure.f90
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function snac_set_impdiff(ctx, enable) bind(C)    ! the IMPDIFF=1 build (run-time switch)
      import; type(c_ptr), value :: ctx; integer(c_int), value :: enable
    end function
    integer(c_int) function snac_solve_impdiff(ctx, fid, alpha) bind(C) ! solve_helmholtz.f90:30-136
      import; type(c_ptr), value :: ctx; integer(c_int), value :: fid; real(c_double), value :: alpha
    end function
    integer(c_int) function snac_impdiff_stats(ctx, fid, iters, relres) bind(C) ! HYPRE counts of a Helmholtz solve
      import; type(c_ptr), value :: ctx; integer(c_int), value :: fid; integer(c_int) :: iters
      real(c_double) :: relres
    end function
    integer(c_int) function snac_updt_pressure_impdiff(ctx, alpha) bind(C) ! updt_pressure.f90 (_IMPDIFF)
      import; type(c_ptr), value :: ctx; real(c_double), value :: alpha
    end function
    integer(c_int) function snac_chkdiv(ctx, divtot, divmax) bind(C)  ! chkdiv.f90
      import; type(c_ptr), value :: ctx; real(c_double) :: divtot, divmax
    end function
    integer(c_int) function snac_chkdt(ctx, dtmax) bind(C)           ! chkdt.f90
      import; type(c_ptr), value :: ctx; real(c_double) :: dtmax
    end function
    integer(c_int) function snac_init_bounds(ctx) bind(C)           ! main.f90:562-639
      import; type(c_ptr), value :: ctx
    end function
    integer(c_int) function snac_step(ctx, dt) bind(C)               ! main.f90:906-1033
      import; type(c_ptr), value :: ctx; real(c_double), value :: dt
    end function
    integer(c_int) function snac_steps(ctx, dt, nsteps) bind(C)      ! nsteps of main.f90:906-1033, dt fixed
      import; type(c_ptr), value :: ctx; real(c_double), value :: dt; integer(c_int), value :: nsteps
    end function
    integer(c_int) function snac_sync(ctx) bind(C)
      import; type(c_ptr), value :: ctx
    end function
  end interface
contains
  ! the message of the last failure on this thread
  function snac_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = snac_last_error()
    n = 0
    if (c_associated(p)) n = int(snac_c_strlen(p))
    allocate(character(len=n) :: msg)
    if (n == 0) return
    call c_f_pointer(p, chars, [n])
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function snac_message

  ! the reference reports failures with a message and error stop
  subroutine snac_check(ierr)
    integer(c_int), intent(in) :: ierr
    if (ierr /= 0) then
      write(error_unit, '(a)') 'ERROR: '//snac_message()
      error stop
    end if
  end subroutine snac_check
end module mod_snac_amd
