! fmm3d_mi355x.f90 -- ISO_C_BINDING drop-in for fmmlib3d's particle FMM.
!
! Exports the reference's F77 entry points with their exact calling
! sequences (src/lfmm3dpart.f:36-141, :1391-1394, src/hfmm3dpart.f:39-144,
! :1268-1271): lfmm3dpart, lfmm3dpartself, lfmm3dparttarg, l3dpartdirect,
! hfmm3dpart, hfmm3dpartself, hfmm3dparttarg and h3dpartdirect; and the
! triangle FMMs (src/lfmm3dtria.f:39-125, :1385, :1566; src/hfmm3dtria.f:
! 51-137, :1456, :1668): lfmm3dtria, lfmm3dtriaself, lfmm3dtriatarg,
! lfmm3dtriampftarg, l3dtriadirect and their Helmholtz twins hfmm3dtria,
! hfmm3dtriaself, hfmm3dtriatarg, hfmm3dtriampftarg, h3dtriadirect.  Each one forwards its
! arguments to the C ABI of libfmm3d_mi355x.so (include/fmm3d_mi355x.h), so a
! Fortran program links against this library instead of the reference
! objects and runs the FMM on the GPU.  Arrays keep the reference layouts and
! are passed through without copies; ier keeps the reference's meaning.
module fmm3d_mi355x_c
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function c_lfmm3dparttarg(iprec, nsource, source, ifcharge, charge, &
        ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, &
        pottarg, iffldtarg, fldtarg) bind(C, name="fmm3d_lfmm3dparttarg")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: iprec, nsource, ifcharge, ifdipole, ifpot, iffld, ntarget
      integer(c_int), value :: ifpottarg, iffldtarg
      real(c_double) :: source(*), dipvec(*), target(*)
      complex(c_double_complex) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
    end function c_lfmm3dparttarg

    subroutine c_l3dpartdirect(nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, &
        ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg) &
        bind(C, name="fmm3d_l3dpartdirect")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nsource, ifcharge, ifdipole, ifpot, iffld, ntarget
      integer(c_int), value :: ifpottarg, iffldtarg
      real(c_double) :: source(*), dipvec(*), target(*)
      complex(c_double_complex) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
    end subroutine c_l3dpartdirect

    integer(c_int) function c_hfmm3dparttarg(iprec, zk, nsource, source, ifcharge, charge, &
        ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, &
        pottarg, iffldtarg, fldtarg) bind(C, name="fmm3d_hfmm3dparttarg")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: iprec, nsource, ifcharge, ifdipole, ifpot, iffld, ntarget
      integer(c_int), value :: ifpottarg, iffldtarg
      complex(c_double_complex) :: zk
      real(c_double) :: source(*), dipvec(*), target(*)
      complex(c_double_complex) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
    end function c_hfmm3dparttarg

    subroutine c_h3dpartdirect(zk, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, &
        ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg) &
        bind(C, name="fmm3d_h3dpartdirect")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nsource, ifcharge, ifdipole, ifpot, iffld, ntarget
      integer(c_int), value :: ifpottarg, iffldtarg
      complex(c_double_complex) :: zk
      real(c_double) :: source(*), dipvec(*), target(*)
      complex(c_double_complex) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
    end subroutine c_h3dpartdirect
    integer(c_int) function c_lfmm3dtriatarg(iprec, nsource, triaflat, trianorm, source, &
        ifcharge, charge, ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, &
        ifpottarg, pottarg, iffldtarg, fldtarg) bind(C, name="fmm3d_lfmm3dtriatarg")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: iprec, nsource, ifcharge, ifdipole, ifpot, iffld, ntarget
      integer(c_int), value :: ifpottarg, iffldtarg
      real(c_double) :: triaflat(*), trianorm(*), source(*), dipvec(*), target(*)
      complex(c_double_complex) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
    end function c_lfmm3dtriatarg

    subroutine c_l3dtriadirect(nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, &
        dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, &
        fldtarg) bind(C, name="fmm3d_l3dtriadirect")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nsource, ifcharge, ifdipole, ifpot, iffld, ntarget
      integer(c_int), value :: ifpottarg, iffldtarg
      real(c_double) :: triaflat(*), trianorm(*), source(*), dipvec(*), target(*)
      complex(c_double_complex) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
    end subroutine c_l3dtriadirect

    integer(c_int) function c_hfmm3dtriatarg(iprec, zk, nsource, triaflat, trianorm, source, &
        ifcharge, charge, ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, &
        ifpottarg, pottarg, iffldtarg, fldtarg) bind(C, name="fmm3d_hfmm3dtriatarg")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: iprec, nsource, ifcharge, ifdipole, ifpot, iffld, ntarget
      integer(c_int), value :: ifpottarg, iffldtarg
      complex(c_double_complex) :: zk
      real(c_double) :: triaflat(*), trianorm(*), source(*), dipvec(*), target(*)
      complex(c_double_complex) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
    end function c_hfmm3dtriatarg

    subroutine c_h3dtriadirect(nqtri, zk, nsource, triaflat, trianorm, source, ifcharge, charge, &
        ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, &
        iffldtarg, fldtarg) bind(C, name="fmm3d_h3dtriadirect")
      import :: c_int, c_double, c_double_complex
      integer(c_int), value :: nqtri, nsource, ifcharge, ifdipole, ifpot, iffld, ntarget
      integer(c_int), value :: ifpottarg, iffldtarg
      complex(c_double_complex) :: zk
      real(c_double) :: triaflat(*), trianorm(*), source(*), dipvec(*), target(*)
      complex(c_double_complex) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
    end subroutine c_h3dtriadirect
  end interface
end module fmm3d_mi355x_c

subroutine lfmm3dparttarg(ier, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, &
    dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
  use fmm3d_mi355x_c
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld, ntarget, ifpottarg, iffldtarg
  real(8) :: source(*), dipvec(*), target(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
  ier = c_lfmm3dparttarg(iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, &
      ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
end subroutine lfmm3dparttarg

subroutine lfmm3dpartself(ier, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, &
    dipvec, ifpot, pot, iffld, fld)
  use fmm3d_mi355x_c
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld
  real(8) :: source(*), dipvec(*), dummy(3)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), cdummy(3)
  ier = c_lfmm3dparttarg(iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, &
      ifpot, pot, iffld, fld, 0, dummy, 0, cdummy, 0, cdummy)
end subroutine lfmm3dpartself

subroutine lfmm3dpart(ier, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, &
    dipvec, ifpot, pot, iffld, fld)
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld
  real(8) :: source(*), dipvec(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*)
  call lfmm3dpartself(ier, iprec, nsource, source, ifcharge, charge, ifdipole, dipstr, &
      dipvec, ifpot, pot, iffld, fld)
end subroutine lfmm3dpart

subroutine l3dpartdirect(nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, &
    ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
  use fmm3d_mi355x_c
  implicit none
  integer :: nsource, ifcharge, ifdipole, ifpot, iffld, ntarget, ifpottarg, iffldtarg
  real(8) :: source(*), dipvec(*), target(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
  call c_l3dpartdirect(nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, &
      ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
end subroutine l3dpartdirect

subroutine hfmm3dparttarg(ier, iprec, zk, nsource, source, ifcharge, charge, ifdipole, dipstr, &
    dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
  use fmm3d_mi355x_c
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld, ntarget, ifpottarg, iffldtarg
  complex(8) :: zk
  real(8) :: source(*), dipvec(*), target(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
  ier = c_hfmm3dparttarg(iprec, zk, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, &
      ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
end subroutine hfmm3dparttarg

subroutine hfmm3dpartself(ier, iprec, zk, nsource, source, ifcharge, charge, ifdipole, dipstr, &
    dipvec, ifpot, pot, iffld, fld)
  use fmm3d_mi355x_c
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld
  complex(8) :: zk
  real(8) :: source(*), dipvec(*), dummy(3)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), cdummy(3)
  ier = c_hfmm3dparttarg(iprec, zk, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, &
      ifpot, pot, iffld, fld, 0, dummy, 0, cdummy, 0, cdummy)
end subroutine hfmm3dpartself

subroutine hfmm3dpart(ier, iprec, zk, nsource, source, ifcharge, charge, ifdipole, dipstr, &
    dipvec, ifpot, pot, iffld, fld)
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld
  complex(8) :: zk
  real(8) :: source(*), dipvec(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*)
  call hfmm3dpartself(ier, iprec, zk, nsource, source, ifcharge, charge, ifdipole, dipstr, &
      dipvec, ifpot, pot, iffld, fld)
end subroutine hfmm3dpart

subroutine h3dpartdirect(zk, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, &
    ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
  use fmm3d_mi355x_c
  implicit none
  integer :: nsource, ifcharge, ifdipole, ifpot, iffld, ntarget, ifpottarg, iffldtarg
  complex(8) :: zk
  real(8) :: source(*), dipvec(*), target(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
  call c_h3dpartdirect(zk, nsource, source, ifcharge, charge, ifdipole, dipstr, dipvec, &
      ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
end subroutine h3dpartdirect

subroutine lfmm3dtriatarg(ier, iprec, nsource, triaflat, trianorm, source, ifcharge, charge, &
    ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, &
    iffldtarg, fldtarg)
  use fmm3d_mi355x_c
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld, ntarget, ifpottarg, iffldtarg
  real(8) :: triaflat(*), trianorm(*), source(*), dipvec(*), target(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
  ier = c_lfmm3dtriatarg(iprec, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, &
      dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
end subroutine lfmm3dtriatarg

subroutine lfmm3dtriaself(ier, iprec, nsource, triaflat, trianorm, source, ifcharge, charge, &
    ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld)
  use fmm3d_mi355x_c
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld
  real(8) :: triaflat(*), trianorm(*), source(*), dipvec(*), dummy(3)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), cdummy(3)
  ier = c_lfmm3dtriatarg(iprec, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, &
      dipstr, dipvec, ifpot, pot, iffld, fld, 0, dummy, 0, cdummy, 0, cdummy)
end subroutine lfmm3dtriaself

subroutine lfmm3dtria(ier, iprec, nsource, triaflat, trianorm, source, ifcharge, charge, &
    ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld)
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld
  real(8) :: triaflat(*), trianorm(*), source(*), dipvec(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*)
  call lfmm3dtriaself(ier, iprec, nsource, triaflat, trianorm, source, ifcharge, charge, &
      ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld)
end subroutine lfmm3dtria

subroutine lfmm3dtriampftarg(ier, iprec, nsource, triaflat, trianorm, source, ifcharge, charge, &
    ifdipole, dipstr, dipvec, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
  use fmm3d_mi355x_c
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ntarget, ifpottarg, iffldtarg
  real(8) :: triaflat(*), trianorm(*), source(*), dipvec(*), target(*)
  complex(8) :: charge(*), dipstr(*), pottarg(*), fldtarg(*), cdummy(3)
  ier = c_lfmm3dtriatarg(iprec, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, &
      dipstr, dipvec, 0, cdummy, 0, cdummy, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
end subroutine lfmm3dtriampftarg

subroutine l3dtriadirect(nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, &
    dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
  use fmm3d_mi355x_c
  implicit none
  integer :: nsource, ifcharge, ifdipole, ifpot, iffld, ntarget, ifpottarg, iffldtarg
  real(8) :: triaflat(*), trianorm(*), source(*), dipvec(*), target(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
  call c_l3dtriadirect(nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, dipstr, &
      dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
end subroutine l3dtriadirect

subroutine hfmm3dtriatarg(ier, iprec, zk, nsource, triaflat, trianorm, source, ifcharge, charge, &
    ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, &
    iffldtarg, fldtarg)
  use fmm3d_mi355x_c
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld, ntarget, ifpottarg, iffldtarg
  complex(8) :: zk
  real(8) :: triaflat(*), trianorm(*), source(*), dipvec(*), target(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
  ier = c_hfmm3dtriatarg(iprec, zk, nsource, triaflat, trianorm, source, ifcharge, charge, &
      ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, &
      iffldtarg, fldtarg)
end subroutine hfmm3dtriatarg

subroutine hfmm3dtriaself(ier, iprec, zk, nsource, triaflat, trianorm, source, ifcharge, charge, &
    ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld)
  use fmm3d_mi355x_c
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld
  complex(8) :: zk
  real(8) :: triaflat(*), trianorm(*), source(*), dipvec(*), dummy(3)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), cdummy(3)
  ier = c_hfmm3dtriatarg(iprec, zk, nsource, triaflat, trianorm, source, ifcharge, charge, &
      ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, 0, dummy, 0, cdummy, 0, cdummy)
end subroutine hfmm3dtriaself

subroutine hfmm3dtria(ier, iprec, zk, nsource, triaflat, trianorm, source, ifcharge, charge, &
    ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld)
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ifpot, iffld
  complex(8) :: zk
  real(8) :: triaflat(*), trianorm(*), source(*), dipvec(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*)
  call hfmm3dtriaself(ier, iprec, zk, nsource, triaflat, trianorm, source, ifcharge, charge, &
      ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld)
end subroutine hfmm3dtria

subroutine hfmm3dtriampftarg(ier, iprec, zk, nsource, triaflat, trianorm, source, ifcharge, &
    charge, ifdipole, dipstr, dipvec, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
  use fmm3d_mi355x_c
  implicit none
  integer :: ier, iprec, nsource, ifcharge, ifdipole, ntarget, ifpottarg, iffldtarg
  complex(8) :: zk
  real(8) :: triaflat(*), trianorm(*), source(*), dipvec(*), target(*)
  complex(8) :: charge(*), dipstr(*), pottarg(*), fldtarg(*), cdummy(3)
  ier = c_hfmm3dtriatarg(iprec, zk, nsource, triaflat, trianorm, source, ifcharge, charge, &
      ifdipole, dipstr, dipvec, 0, cdummy, 0, cdummy, ntarget, target, ifpottarg, pottarg, &
      iffldtarg, fldtarg)
end subroutine hfmm3dtriampftarg

subroutine h3dtriadirect(nqtri, zk, nsource, triaflat, trianorm, source, ifcharge, charge, &
    ifdipole, dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, &
    iffldtarg, fldtarg)
  use fmm3d_mi355x_c
  implicit none
  integer :: nqtri, nsource, ifcharge, ifdipole, ifpot, iffld, ntarget, ifpottarg, iffldtarg
  complex(8) :: zk
  real(8) :: triaflat(*), trianorm(*), source(*), dipvec(*), target(*)
  complex(8) :: charge(*), dipstr(*), pot(*), fld(*), pottarg(*), fldtarg(*)
  call c_h3dtriadirect(nqtri, zk, nsource, triaflat, trianorm, source, ifcharge, charge, ifdipole, &
      dipstr, dipvec, ifpot, pot, iffld, fld, ntarget, target, ifpottarg, pottarg, iffldtarg, fldtarg)
end subroutine h3dtriadirect
