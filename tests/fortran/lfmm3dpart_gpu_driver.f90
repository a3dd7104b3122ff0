! A Fortran caller of the drop-in: sources on a sphere of radius 1/2 and
! targets on a second sphere shifted by 2 in x (the geometry of the
! reference's sample driver), charges and dipoles, iprec = 1.  It calls
! lfmm3dparttarg and l3dpartdirect exactly as a reference user would (the
! names resolve to libfmm3d_mi355x_f.so) and prints the relative L2 errors
! of the FMM against the direct sum at the first 100 sources and targets;
! then the same with hfmm3dparttarg / h3dpartdirect at zk = 6 + 0.1i
! (targets only).
program lfmm3dpart_gpu_driver
  implicit none
  integer, parameter :: n = 16000, nt = 16000, m = 100
  real(8) :: source(3, n), target(3, nt), dipvec(3, n), u, v, w, r, pi
  complex(8) :: charge(n), dipstr(n), pot(n), fld(3, n), pottarg(nt), fldtarg(3, nt)
  complex(8) :: pot2(m), fld2(3, m), pt2(m), ft2(3, m), cdum(3)
  real(8) :: rdum(3)
  integer :: ier, i, seed(64), ns
  real(8) :: e1, e2, e3, e4
  complex(8) :: zk
  pi = 4 * atan(1.0d0)
  call random_seed(size=ns)
  seed = 20120418
  call random_seed(put=seed(1:ns))
  do i = 1, n
    call random_number(u)
    call random_number(v)
    w = 2 * pi * v
    u = 2 * u - 1
    r = sqrt(1 - u * u)
    source(1, i) = 0.5d0 * r * cos(w)
    source(2, i) = 0.5d0 * r * sin(w)
    source(3, i) = 0.5d0 * u
    call random_number(u)
    call random_number(v)
    charge(i) = cmplx(u - 0.5d0, v - 0.5d0, kind=8)
    call random_number(u)
    call random_number(v)
    dipstr(i) = cmplx(u - 0.5d0, v - 0.5d0, kind=8)
    call random_number(dipvec(:, i))
  end do
  do i = 1, nt
    target(:, i) = source(:, i)
    target(1, i) = target(1, i) + 2.0d0
  end do
  call lfmm3dparttarg(ier, 1, n, source, 1, charge, 1, dipstr, dipvec, 1, pot, 1, fld, &
      nt, target, 1, pottarg, 1, fldtarg)
  if (ier /= 0) then
    print *, 'ier=', ier
    stop 1
  end if
  ! direct sum at the first m sources (all sources acting) and first m targets
  call l3dpartdirect(n, source, 1, charge, 1, dipstr, dipvec, 0, cdum, 0, cdum, &
      m, target, 1, pt2, 1, ft2)
  call direct_at_sources(n, source, charge, dipstr, dipvec, m, pot2, fld2)
  e1 = sqrt(sum(abs(pot(1:m) - pot2)**2) / sum(abs(pot2)**2))
  e2 = sqrt(sum(abs(fld(:, 1:m) - fld2)**2) / sum(abs(fld2)**2))
  e3 = sqrt(sum(abs(pottarg(1:m) - pt2)**2) / sum(abs(pt2)**2))
  e4 = sqrt(sum(abs(fldtarg(:, 1:m) - ft2)**2) / sum(abs(ft2)**2))
  write (*, '(a,4es12.4)') 'relerr pot fld pottarg fldtarg ', e1, e2, e3, e4
  zk = cmplx(6.0d0, 0.1d0, kind=8)
  call hfmm3dparttarg(ier, 1, zk, n, source, 1, charge, 1, dipstr, dipvec, 1, pot, 1, fld, &
      nt, target, 1, pottarg, 1, fldtarg)
  if (ier /= 0) then
    print *, 'hfmm ier=', ier
    stop 1
  end if
  call h3dpartdirect(zk, n, source, 1, charge, 1, dipstr, dipvec, 0, cdum, 0, cdum, &
      m, target, 1, pt2, 1, ft2)
  e3 = sqrt(sum(abs(pottarg(1:m) - pt2)**2) / sum(abs(pt2)**2))
  e4 = sqrt(sum(abs(fldtarg(:, 1:m) - ft2)**2) / sum(abs(ft2)**2))
  write (*, '(a,2es12.4)') 'hrelerr pottarg fldtarg ', e3, e4
contains
  subroutine direct_at_sources(n, src, q, ds, dv, m, p, f)
    integer :: n, m, i, j
    real(8) :: src(3, n), dv(3, n), d(3), r2, ri, ri3, dot
    complex(8) :: q(n), ds(n), p(m), f(3, m)
    p = 0
    f = 0
    do j = 1, m
      do i = 1, n
        if (i == j) cycle
        d = src(:, j) - src(:, i)
        r2 = sum(d * d)
        ri = 1 / sqrt(r2)
        ri3 = ri * ri * ri
        dot = sum(dv(:, i) * d)
        p(j) = p(j) + q(i) * ri + ds(i) * dot * ri3
        f(:, j) = f(:, j) + q(i) * d * ri3 + ds(i) * (3 * dot * d * ri3 * ri * ri - dv(:, i) * ri3)
      end do
    end do
  end subroutine direct_at_sources
end program lfmm3dpart_gpu_driver
