! The reference's time loop (main.f90:893-1069, no scalars, explicit
! diffusion) written against the Fortran binding, one C-ABI call per
! reference routine in the reference's order, on the case in the working
! directory (dns.nml, blocks.nml): FFT_AXIS from the first argument, the
! step count from the second (default: nstep of dns.nml).  dt follows
! chkdt every icheck steps, merge(dt_f, min(cfl*dt_cfl, dtmax), dt_f > 0),
! and the run stops on a NaN or growing divergence as the reference does.
! At the end each block's v (ghosted, the reference's array layout) is
! written to v_b_NNN.bin (stream access, FP64) and divtot/divmax printed.
program ldc_steps
  use mod_snac_amd
  implicit none
  real(c_double), parameter :: rkc(2, 3) = reshape([32._c_double/60, 0._c_double, &
                                                    25._c_double/60, -17._c_double/60, &
                                                    45._c_double/60, -25._c_double/60], [2, 3])
  type(c_ptr) :: ctx
  real(c_double) :: dpar(11), dt, dt_cfl, dtrk, divtot, divmax, cfl, dtmax, dt_f
  integer(c_int) :: ipar(8), nb, b, n3(3), lo3(3), nb6(6), per3(3), istep, irk, nstep, icheck, fft_axis
  integer :: u
  character(len=64) :: arg
  character(len=3) :: cb
  real(c_double), allocatable :: v(:, :, :)
  call get_command_argument(1, arg)
  read(arg, *) fft_axis
  call snac_check(snac_ctx_create_from_nml(ctx, 'dns.nml'//c_null_char, 'blocks.nml'//c_null_char, fft_axis, 0_c_int))
  call snac_check(snac_ctx_get_params(ctx, dpar, ipar))
  cfl = dpar(2); dtmax = dpar(3); dt_f = dpar(4)
  nstep = ipar(1); icheck = ipar(2)
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    read(arg, *) nstep
  end if
  call snac_check(snac_initflow(ctx))
  call snac_check(snac_init_bounds(ctx))
  call snac_check(snac_chkdt(ctx, dt_cfl))
  dt = merge(dt_f, min(cfl*dt_cfl, dtmax), dt_f > 0)
  do istep = 1, nstep
    do irk = 1, 3
      dtrk = sum(rkc(:, irk))*dt
      call snac_check(snac_rk_mom(ctx, irk, dt))
      call snac_check(snac_bounduvw(ctx, 0_c_int))
      call snac_check(snac_inflow(ctx, 0_c_int))
      call snac_check(snac_fillps(ctx, dtrk))
      call snac_check(snac_solve_pressure(ctx))
      call snac_check(snac_boundp(ctx, SNAC_PP))
      call snac_check(snac_correc(ctx, dtrk))
      call snac_check(snac_bounduvw(ctx, 1_c_int))
      call snac_check(snac_inflow(ctx, 1_c_int))
      call snac_check(snac_updt_pressure(ctx))
      call snac_check(snac_boundp(ctx, SNAC_P))
    end do
    if (mod(istep, icheck) == 0) then
      call snac_check(snac_chkdt(ctx, dt_cfl))
      dt = merge(dt_f, min(cfl*dt_cfl, dtmax), dt_f > 0)
      call snac_check(snac_chkdiv(ctx, divtot, divmax))
      if (divmax /= divmax .or. divmax > 1.e-6_c_double) then
        write(*, *) 'ERROR: divergence ', divmax, ' at step ', istep
        error stop
      end if
    end if
  end do
  call snac_check(snac_chkdiv(ctx, divtot, divmax))
  nb = snac_ctx_num_blocks(ctx)
  do b = 0, nb - 1
    call snac_check(snac_ctx_block_info(ctx, b, n3, lo3, nb6, per3))
    allocate(v(0:n3(1)+1, 0:n3(2)+1, 0:n3(3)+1))
    call snac_check(snac_download_field(ctx, b, SNAC_V, v))
    write(cb, '(i3.3)') b + 1
    open(newunit=u, file='v_b_'//cb//'.bin', access='stream', form='unformatted', status='replace')
    write(u) v
    close(u)
    deallocate(v)
  end do
  write(*, '(a,i0,a,es12.4,a,es12.4)') 'steps ', nstep, ' divtot ', divtot, ' divmax ', divmax
  call snac_check(snac_ctx_destroy(ctx))
end program ldc_steps
