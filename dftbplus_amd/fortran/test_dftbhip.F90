!> Fortran caller of the bind(C) module: generalized eigenproblems of a small symmetric and a
!> small Hermitian pair checked by residuals, their density matrices checked against the explicit
!> sums over the filled levels, and the neighbour list of a periodic cell checked by its
!> invariants (prints PASS/FAIL; exit status 1 on failure).
program test_dftbhip
  use dftbhip
  implicit none
  integer, parameter :: n = 40
  real(8) :: h(n,n), s(n,n), h0(n,n), s0(n,n), w(n), f(n), dm(n,n), r, ref
  complex(8) :: hc(n,n), sc(n,n), hc0(n,n), sc0(n,n), dmc(n,n), refc
  integer :: info, i, j, k, ngpu, nAll
  real(8) :: x0(3,2), lat(3,3), cut
  real(8), allocatable :: xall(:,:), d2(:,:)
  integer, allocatable :: i2c(:), inb(:,:), nnb(:)
  ! host pieces (no device work): Fermi filling of four levels with two electrons per spin
  ! channel at kT = 0.01 Ha, and two Broyden steps on a three-element vector
  block
    use iso_c_binding, only : c_ptr
    real(8) :: e4(4), f4(4), kw(1), ef, ebs, ts, e0, q3(3), d3(3)
    type(c_ptr) :: mixer
    e4 = [-0.5d0, -0.3d0, 0.2d0, 0.4d0]
    kw = 1.0d0
    info = dftbhip_efilling(4, 1, e4, kw, 2.0d0, 0.01d0, f4, ef, ebs, ts, e0)
    if (info /= 0 .or. abs(sum(f4) - 2.0d0) > 1.0d-12 .or. abs(ef + 0.05d0) > 1.0d-12) then
      print '(a)', "FAIL: efilling"
      error stop 1
    end if
    info = dftbhip_mixer_create(mixer, 10, 0.2d0, 0.01d0, 1.0d0, 1.0d5, 1.0d-2)
    info = info + dftbhip_mixer_reset(mixer, 3)
    q3 = [6.0d0, 1.0d0, 1.0d0]
    d3 = [0.8d0, -0.4d0, -0.4d0]
    info = info + dftbhip_mixer_mix(mixer, q3, d3)
    if (info /= 0 .or. abs(q3(1) - 6.16d0) > 1.0d-14) then
      print '(a)', "FAIL: mixer"
      error stop 1
    end if
    info = dftbhip_mixer_destroy(mixer)
    print '(a)', "host pieces ok"
  end block
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
  ! lower triangle of sum_k f_k c_k c_k^T
  r = 0.0d0
  do j = 1, n
    do i = j, n
      ref = sum(f * h(i, :) * h(j, :))
      r = max(r, abs(dm(i, j) - ref))
    end do
  end do
  print '(a,es10.3)', "density matrix ", r
  if (r > 1.0d-12) then
    print '(a)', "FAIL: density matrix"
    error stop 1
  end if
  ! complex Hermitian pair: H with an imaginary part antisymmetric in (i, j), S Hermitian
  ! positive definite
  do j = 1, n
    do i = 1, n
      hc0(i,j) = cmplx(h0(i,j), 0.05d0 * real(i - j, 8) / real(i + j, 8), 8)
      sc0(i,j) = cmplx(s0(i,j), merge(0.0d0, 0.02d0 * sign(1.0d0, real(i - j, 8)) /&
          & real(abs(i - j) + 1, 8)**2, i == j), 8)
    end do
  end do
  hc = hc0
  sc = sc0
  call dftbhip_hegvd_dcomplex(hc, sc, w, "V", info)
  if (info /= 0) then
    print '(a,i0,1x,a)', "FAIL: complex info = ", info, trim(dftbhip_error_message())
    error stop 1
  end if
  r = maxval(abs(matmul(hc0, hc) - matmul(sc0, hc) * spread(w, 1, n)))
  print '(a,es10.3)', "complex residual ", r
  call dftbhip_dm_dcomplex(dmc, hc, f, info)
  if (r > 1.0d-10 .or. info /= 0) then
    print '(a)', "FAIL: complex eigenproblem"
    error stop 1
  end if
  r = 0.0d0
  do j = 1, n
    do i = j, n
      refc = sum(f * hc(i, :) * conjg(hc(j, :)))
      r = max(r, abs(dmc(i, j) - refc))
    end do
  end do
  print '(a,es10.3)', "complex density matrix ", r
  if (r > 1.0d-12) then
    print '(a)', "FAIL: complex density matrix"
    error stop 1
  end if
  ! two atoms in a cubic cell of 3 bohr, cutoff 4.5: every atom lists itself first, then
  ! images of atoms with an index not below its own, by non-decreasing distance within the cutoff
  x0 = reshape([0.0d0, 0.0d0, 0.0d0, 1.5d0, 1.5d0, 1.5d0], [3, 2])
  lat = 0.0d0
  lat(1,1) = 3.0d0
  lat(2,2) = 3.0d0
  lat(3,3) = 3.0d0
  cut = 4.5d0
  call dftbhip_update_neighbour_list(x0, cut, xall, i2c, inb, nnb, d2, nAll, info, lat)
  if (info /= 0) then
    print '(a,i0,1x,a)', "FAIL: neighbour list info = ", info, trim(dftbhip_error_message())
    error stop 1
  end if
  print '(a,i0,a,2(1x,i0))', "images ", nAll, ", neighbours", nnb
  do i = 1, 2
    if (inb(0, i) /= i .or. d2(0, i) /= 0.0d0 .or. nnb(i) < 1) error stop 1
    do k = 1, nnb(i)
      j = inb(k, i)
      if (j < 1 .or. j > nAll) error stop 1
      if (i2c(j) < i .or. d2(k, i) > cut**2 .or. d2(k, i) < d2(k - 1, i)) error stop 1
      if (abs(sum((xall(:, j) - x0(:, i))**2) - d2(k, i)) > 1.0d-12) error stop 1
    end do
  end do
  print '(a)', "PASS"
end program test_dftbhip
