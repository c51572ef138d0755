! Host-only use of the Fortran binding (no GPU): a context of the case in
! the working directory (dns.nml, blocks.nml) created without a device,
! its blocks' placement and neighbours printed one line per block, then a
! failing call (a missing namelist file) whose message comes back through
! snac_message.
program ldc_info
  use mod_snac_amd
  implicit none
  type(c_ptr) :: ctx, bad
  integer(c_int) :: nb, b, n3(3), lo3(3), nb6(6), per3(3), ierr
  real(c_double) :: dpar(11)
  integer(c_int) :: ipar(8)
  call snac_check(snac_ctx_create_from_nml(ctx, 'dns.nml'//c_null_char, 'blocks.nml'//c_null_char, 1_c_int, &
                                           -1_c_int))
  nb = snac_ctx_num_blocks(ctx)
  call snac_check(snac_ctx_get_params(ctx, dpar, ipar))
  write(*, '(a,i0,a,i0,a,i0)') 'blocks ', nb, ' nstep ', ipar(1), ' icheck ', ipar(2)
  do b = 0, nb - 1
    call snac_check(snac_ctx_block_info(ctx, b, n3, lo3, nb6, per3))
    write(*, '(a,i0,15(1x,i0))') 'block ', b, n3, lo3, nb6, per3
  end do
  call snac_check(snac_ctx_destroy(ctx))
  ierr = snac_ctx_create_from_nml(bad, 'no_such_dns.nml'//c_null_char, 'blocks.nml'//c_null_char, 1_c_int, -1_c_int)
  write(*, '(a,i0,a,a)') 'missing file: status ', ierr, ' message: ', snac_message()
end program ldc_info
