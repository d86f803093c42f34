!> Calls the drop-in JCG through the ISO_C_BINDING interface exactly the way
!> src/gwce.F:2002-2005 calls ITPACK (1-D Laplacian + mass system, ADCIRC's
!> IPARM/RPARM settings) and prints the iteration count and the residual.
program jcg_dropin_check
  use iso_c_binding
  use adcgpu_iface
  implicit none
  integer, parameter :: n = 1000, mnei = 4
  integer(c_int32_t) :: neitab(n, mnei), iparm(12)
  real(c_double) :: coef(n, mnei), rhs(n), etas(n), rparm(12), x(n), r, b(n)
  integer :: i, ier
  neitab = 0; coef = 0.d0
  do i = 1, n
     neitab(i, 1) = i; coef(i, 1) = 2.1d0
     if (i > 1) then
        neitab(i, 2) = i - 1; coef(i, 2) = -1.d0
     end if
     if (i < n) then
        neitab(i, 3) = i + 1; coef(i, 3) = -1.d0
     end if
     x(i) = sin(0.01d0 * i)
  end do
  do i = 1, n
     b(i) = 2.1d0 * x(i)
     if (i > 1) b(i) = b(i) - x(i - 1)
     if (i < n) b(i) = b(i) - x(i + 1)
  end do
  rhs = b
  etas = 0.d0
  ! dfault (src/itpackv.F:1135-1160) with ADCIRC's overrides (src/gwce.F:146-150):
  ! IPARM(1)=ITMAX, IPARM(2)=ISLDIA, IPARM(4)=log unit, RPARM(1)=CONVCR
  iparm = (/ 100, 0, 0, 16, 0, 1, 1, 0, -2, 0, 0, 3 /)
  rparm = 0.d0
  rparm(1) = 1.d-10; rparm(4) = 0.75d0; rparm(5) = 1.d0; rparm(7) = 0.25d0
  rparm(8) = 128.d0 * epsilon(1.d0) / 2.d0
  ier = adcgpu_jcg(n, n, mnei, neitab, coef, rhs, etas, iparm, rparm)
  r = maxval(abs(etas - x))
  print '(a,i0,a,i0,a,es10.3)', 'jcg ier=', ier, ' NUMITR=', iparm(1), ' max|x-x*|=', r
  if (ier < 0) print '(2a)', 'adcgpu_last_error: ', adcgpu_error_message()
  if (ier /= 0 .or. r > 1.d-6) stop 1
end program
