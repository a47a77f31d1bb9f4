!> Fortran caller of the bind(C) module: generalized eigenproblem of a small symmetric pair and
!> the density matrix, checked by residuals (prints PASS/FAIL; exit status 1 on failure).
program test_dftbhip
  use dftbhip
  implicit none
  integer, parameter :: n = 40
  real(8) :: h(n,n), s(n,n), h0(n,n), s0(n,n), w(n), f(n), dm(n,n), r
  integer :: info, i, j, ngpu
  call dftbhip_gpus_available(ngpu)
  print '(a,i0)', "gpus: ", ngpu
  if (ngpu < 1) then
    print '(a)', "SKIP: no GPU"
    stop
  end if
  do j = 1, n
    do i = 1, n
      h0(i,j) = 1.0d0 / real(i + j, 8) + merge(real(i, 8), 0.0d0, i == j)
      s0(i,j) = merge(1.0d0, 0.1d0 / real(abs(i - j) + 1, 8)**2, i == j)
    end do
  end do
  h = h0
  s = s0
  call dftbhip_hegvd_dreal(h, s, w, "V", info)
  if (info /= 0) then
    print '(a,i0,1x,a)', "FAIL: info = ", info, trim(dftbhip_error_message())
    error stop 1
  end if
  r = maxval(abs(matmul(h0, h) - matmul(s0, h) * spread(w, 1, n)))
  print '(a,es10.3)', "residual ", r
  f = 0.0d0
  f(1:n/2) = 2.0d0
  call dftbhip_dm_dreal(dm, h, f, info)
  if (r > 1.0d-10 .or. info /= 0) then
    print '(a)', "FAIL"
    error stop 1
  end if
  print '(a)', "PASS"
end program test_dftbhip
