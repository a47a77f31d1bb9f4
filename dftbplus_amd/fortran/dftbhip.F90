!> ISO_C_BINDING interface to libdftbhip, the MI355X backend of the SCC-DFTB hot path.
!>
!> Mirrors the pattern of DFTB+'s MAGMA binding (src/dftbp/extlibs/magma.F90:30-60): thin
!> bind(C) interfaces plus Fortran wrappers with the argument lists of the routines they
!> replace (magmaHegvd in src/dftbp/math/eigensolver.F90:1143-1245, makeDensityMtxRealGPU /
!> makeEnergyDensityMtxRealGPU in src/dftbp/dftb/densitymatrix.F90:1105-1395).
module dftbhip
  use iso_c_binding, only : c_int, c_double, c_double_complex, c_char, c_ptr, c_null_char,&
      & c_null_ptr, c_loc
  implicit none
  private
  public :: dftbhip_gpus_available, dftbhip_hegvd_dreal, dftbhip_dm_dreal, dftbhip_edm_dreal
  public :: dftbhip_hegvd_dcomplex, dftbhip_dm_dcomplex, dftbhip_edm_dcomplex
  public :: dftbhip_efilling, dftbhip_mixer_create, dftbhip_mixer_reset, dftbhip_mixer_mix
  public :: dftbhip_mixer_destroy
  public :: dftbhip_error_message, dftbhip_update_neighbour_list

  interface
    function c_gpus(ngpus) result(rc) bind(C, name="dftbhip_get_gpus_available")
      import :: c_int
      integer(c_int), intent(out) :: ngpus
      integer(c_int) :: rc
    end function c_gpus

    function c_sygvd(itype, jobz, uplo, n, a, lda, b, ldb, w) result(rc)&
        & bind(C, name="dftbhip_sygvd")
      import :: c_int, c_double, c_char
      integer(c_int), value :: itype, n, lda, ldb
      character(kind=c_char), value :: jobz, uplo
      real(c_double), intent(inout) :: a(lda, *), b(ldb, *)
      real(c_double), intent(out) :: w(*)
      integer(c_int) :: rc
    end function c_sygvd

    function c_dm(n, nlevels, evec, lde, filling, dm, lddm) result(rc)&
        & bind(C, name="dftbhip_dm_real")
      import :: c_int, c_double
      integer(c_int), value :: n, nlevels, lde, lddm
      real(c_double), intent(in) :: evec(lde, *), filling(*)
      real(c_double), intent(inout) :: dm(lddm, *)
      integer(c_int) :: rc
    end function c_dm

    function c_edm(n, nlevels, evec, lde, filling, eigen, edm, lddm) result(rc)&
        & bind(C, name="dftbhip_edm_real")
      import :: c_int, c_double
      integer(c_int), value :: n, nlevels, lde, lddm
      real(c_double), intent(in) :: evec(lde, *), filling(*), eigen(*)
      real(c_double), intent(inout) :: edm(lddm, *)
      integer(c_int) :: rc
    end function c_edm

    function c_hegvd(itype, jobz, uplo, n, a, lda, b, ldb, w) result(rc)&
        & bind(C, name="dftbhip_hegvd")
      import :: c_int, c_double, c_double_complex, c_char
      integer(c_int), value :: itype, n, lda, ldb
      character(kind=c_char), value :: jobz, uplo
      complex(c_double_complex), intent(inout) :: a(lda, *), b(ldb, *)
      real(c_double), intent(out) :: w(*)
      integer(c_int) :: rc
    end function c_hegvd

    function c_dm_cmplx(n, nlevels, evec, lde, filling, dm, lddm) result(rc)&
        & bind(C, name="dftbhip_dm_cmplx")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: n, nlevels, lde, lddm
      complex(c_double_complex), intent(in) :: evec(lde, *)
      real(c_double), intent(in) :: filling(*)
      complex(c_double_complex), intent(inout) :: dm(lddm, *)
      integer(c_int) :: rc
    end function c_dm_cmplx

    function c_edm_cmplx(n, nlevels, evec, lde, filling, eigen, edm, lddm) result(rc)&
        & bind(C, name="dftbhip_edm_cmplx")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: n, nlevels, lde, lddm
      complex(c_double_complex), intent(in) :: evec(lde, *)
      real(c_double), intent(in) :: filling(*), eigen(*)
      complex(c_double_complex), intent(inout) :: edm(lddm, *)
      integer(c_int) :: rc
    end function c_edm_cmplx

    function c_neigh(nAtom, coords, latVecs, cutoff, nAll, mNeighbour, coordsAll, img2CentCell,&
        & iNeighbour, nNeighbour, neighDist2) result(rc) bind(C, name="dftbhip_neighbour_list")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nAtom
      real(c_double), intent(in) :: coords(3, *)
      type(c_ptr), value :: latVecs
      real(c_double), value :: cutoff
      integer(c_int), intent(inout) :: nAll, mNeighbour
      type(c_ptr), value :: coordsAll, img2CentCell, iNeighbour, nNeighbour, neighDist2
      integer(c_int) :: rc
    end function c_neigh

    !> Efilling (src/dftbp/dftb/etemp.F90:59-194), Fermi distribution, one spin channel
    function dftbhip_efilling(nLevels, nKPoints, eigenvals, kWeight, nElectrons, kT, filling,&
        & Ef, Ebs, TS, E0) result(rc) bind(C, name="dftbhip_efilling")
      import :: c_int, c_double
      integer(c_int), value :: nLevels, nKPoints
      real(c_double), intent(in) :: eigenvals(nLevels, *), kWeight(*)
      real(c_double), value :: nElectrons, kT
      real(c_double), intent(out) :: filling(nLevels, *), Ef, Ebs, TS, E0
      integer(c_int) :: rc
    end function dftbhip_efilling

    !> TBroydenMixerReal (src/dftbp/mixer/broydenmixer.F90:112-342): init, reset, mix
    function dftbhip_mixer_create(mixer, maxSccIter, mixParam, omega0, minWeight, maxWeight,&
        & weightFac) result(rc) bind(C, name="dftbhip_mixer_create")
      import :: c_int, c_double, c_ptr
      type(c_ptr), intent(out) :: mixer
      integer(c_int), value :: maxSccIter
      real(c_double), value :: mixParam, omega0, minWeight, maxWeight, weightFac
      integer(c_int) :: rc
    end function dftbhip_mixer_create

    function dftbhip_mixer_reset(mixer, nElem) result(rc) bind(C, name="dftbhip_mixer_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: mixer
      integer(c_int), value :: nElem
      integer(c_int) :: rc
    end function dftbhip_mixer_reset

    function dftbhip_mixer_mix(mixer, qInpResult, qDiff) result(rc)&
        & bind(C, name="dftbhip_mixer_mix")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: mixer
      real(c_double), intent(inout) :: qInpResult(*)
      real(c_double), intent(in) :: qDiff(*)
      integer(c_int) :: rc
    end function dftbhip_mixer_mix

    function dftbhip_mixer_destroy(mixer) result(rc) bind(C, name="dftbhip_mixer_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: mixer
      integer(c_int) :: rc
    end function dftbhip_mixer_destroy

    function c_last_error(buf, blen) result(rc) bind(C, name="dftbhip_last_error")
      import :: c_int, c_char
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int), value :: blen
      integer(c_int) :: rc
    end function c_last_error
  end interface

contains

  !> Number of visible GPUs (replaces getGpusAvailable).
  subroutine dftbhip_gpus_available(ngpus)
    integer, intent(out) :: ngpus
    integer(c_int) :: n, rc
    rc = c_gpus(n)
    ngpus = n
  end subroutine dftbhip_gpus_available

  !> Generalized symmetric-definite eigenproblem, itype 1, lower storage (replaces
  !> magmaHegvd for real matrices). info follows LAPACK: > n means a non-positive-definite
  !> leading minor (info - n) of b; < 0 is an argument or device error.
  subroutine dftbhip_hegvd_dreal(a, b, w, jobz, info)
    real(c_double), intent(inout) :: a(:,:), b(:,:)
    real(c_double), intent(out) :: w(:)
    character, intent(in) :: jobz
    integer, intent(out) :: info
    integer(c_int) :: n
    n = size(a, dim=1)
    info = c_sygvd(1_c_int, jobz, "L", n, a, n, b, size(b, dim=1), w)
  end subroutine dftbhip_hegvd_dreal

  !> Generalized Hermitian eigenproblem, itype 1, lower storage (replaces magmaHegvd for
  !> complex matrices); info as dftbhip_hegvd_dreal.
  subroutine dftbhip_hegvd_dcomplex(a, b, w, jobz, info)
    complex(c_double_complex), intent(inout) :: a(:,:), b(:,:)
    real(c_double), intent(out) :: w(:)
    character, intent(in) :: jobz
    integer, intent(out) :: info
    integer(c_int) :: n
    n = size(a, dim=1)
    info = c_hegvd(1_c_int, jobz, "L", n, a, n, b, size(b, dim=1), w)
  end subroutine dftbhip_hegvd_dcomplex

  !> Complex density matrix from eigenvectors and fillings (lower triangle; replaces
  !> makeDensityMtxCmplxGPU).
  subroutine dftbhip_dm_dcomplex(dm, eigenvecs, filling, info)
    complex(c_double_complex), intent(inout) :: dm(:,:)
    complex(c_double_complex), intent(in) :: eigenvecs(:,:)
    real(c_double), intent(in) :: filling(:)
    integer, intent(out) :: info
    integer :: nLevels
    nLevels = size(filling)
    do while (nLevels > 0)
      if (filling(nLevels) > epsilon(1.0_c_double)) exit
      nLevels = nLevels - 1
    end do
    info = c_dm_cmplx(int(size(eigenvecs, dim=1), c_int), int(nLevels, c_int), eigenvecs,&
        & int(size(eigenvecs, dim=1), c_int), filling, dm, int(size(dm, dim=1), c_int))
  end subroutine dftbhip_dm_dcomplex

  !> Complex energy-weighted density matrix (lower triangle; replaces
  !> makeEnergyDensityMtxCmplxGPU).
  subroutine dftbhip_edm_dcomplex(edm, eigenvecs, filling, eigen, info)
    complex(c_double_complex), intent(inout) :: edm(:,:)
    complex(c_double_complex), intent(in) :: eigenvecs(:,:)
    real(c_double), intent(in) :: filling(:), eigen(:)
    integer, intent(out) :: info
    integer :: nLevels
    nLevels = size(filling)
    do while (nLevels > 0)
      if (filling(nLevels) > epsilon(1.0_c_double)) exit
      nLevels = nLevels - 1
    end do
    info = c_edm_cmplx(int(size(eigenvecs, dim=1), c_int), int(nLevels, c_int), eigenvecs,&
        & int(size(eigenvecs, dim=1), c_int), filling, eigen, edm, int(size(edm, dim=1), c_int))
  end subroutine dftbhip_edm_dcomplex

  !> Density matrix from eigenvectors and fillings (lower triangle).
  subroutine dftbhip_dm_dreal(dm, eigenvecs, filling, info)
    real(c_double), intent(inout) :: dm(:,:)
    real(c_double), intent(in) :: eigenvecs(:,:), filling(:)
    integer, intent(out) :: info
    integer :: nLevels
    nLevels = size(filling)
    do while (nLevels > 0)
      if (filling(nLevels) > epsilon(1.0_c_double)) exit
      nLevels = nLevels - 1
    end do
    info = c_dm(int(size(eigenvecs, dim=1), c_int), int(nLevels, c_int), eigenvecs,&
        & int(size(eigenvecs, dim=1), c_int), filling, dm, int(size(dm, dim=1), c_int))
  end subroutine dftbhip_dm_dreal

  !> Energy-weighted density matrix (lower triangle).
  subroutine dftbhip_edm_dreal(edm, eigenvecs, filling, eigen, info)
    real(c_double), intent(inout) :: edm(:,:)
    real(c_double), intent(in) :: eigenvecs(:,:), filling(:), eigen(:)
    integer, intent(out) :: info
    integer :: nLevels
    nLevels = size(filling)
    do while (nLevels > 0)
      if (filling(nLevels) > epsilon(1.0_c_double)) exit
      nLevels = nLevels - 1
    end do
    info = c_edm(int(size(eigenvecs, dim=1), c_int), int(nLevels, c_int), eigenvecs,&
        & int(size(eigenvecs, dim=1), c_int), filling, eigen, edm, int(size(edm, dim=1), c_int))
  end subroutine dftbhip_edm_dreal

  !> Half neighbour list on the device (the arrays updateNeighbourList fills, src/dftbp/dftb/
  !> periodic.F90:446-696): coordinates of central atoms and images, image -> central atom,
  !> neighbours iNeighbour(0:, iAt) (0: the atom itself) by increasing distance, their counts
  !> and squared distances. latVecs (columns: lattice vectors) absent for a cluster.
  subroutine dftbhip_update_neighbour_list(coord0, cutoff, coord, img2CentCell, iNeighbour,&
      & nNeighbour, neighDist2, nAllAtom, info, latVecs)
    real(c_double), intent(in) :: coord0(:,:)
    real(c_double), intent(in) :: cutoff
    real(c_double), allocatable, target, intent(out) :: coord(:,:), neighDist2(:,:)
    integer(c_int), allocatable, target, intent(out) :: img2CentCell(:), iNeighbour(:,:)
    integer(c_int), allocatable, target, intent(out) :: nNeighbour(:)
    integer, intent(out) :: nAllAtom, info
    real(c_double), target, intent(in), optional :: latVecs(3, 3)
    type(c_ptr) :: lat
    integer(c_int) :: nAtom, nAll, mNb
    nAtom = size(coord0, dim=2)
    lat = c_null_ptr
    if (present(latVecs)) lat = c_loc(latVecs)
    nAll = 0
    mNb = 0
    info = c_neigh(nAtom, coord0, lat, cutoff, nAll, mNb, c_null_ptr, c_null_ptr, c_null_ptr,&
        & c_null_ptr, c_null_ptr)
    nAllAtom = nAll
    if (info /= 0) return
    allocate(coord(3, nAll), img2CentCell(nAll), iNeighbour(0:mNb, nAtom), nNeighbour(nAtom))
    allocate(neighDist2(0:mNb, nAtom))
    info = c_neigh(nAtom, coord0, lat, cutoff, nAll, mNb, c_loc(coord), c_loc(img2CentCell),&
        & c_loc(iNeighbour), c_loc(nNeighbour), c_loc(neighDist2))
    nAllAtom = nAll
  end subroutine dftbhip_update_neighbour_list

  !> Message of the last failing call.
  function dftbhip_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char) :: buf(512)
    integer(c_int) :: rc
    integer :: ii
    rc = c_last_error(buf, 512_c_int)
    msg = ""
    do ii = 1, 512
      if (buf(ii) == c_null_char) exit
      msg(ii:ii) = buf(ii)
    end do
  end function dftbhip_error_message

end module dftbhip
