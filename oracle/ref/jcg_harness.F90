! Test harness: runs the reference's own ITPACKV JCG (src/itpackv.F:63) on a
! sequence of ELLPACK systems read from a binary file, in one process so that
! ITPACK's SAVEd state carries from one solve to the next exactly as it does
! across ADCIRC timesteps.  NOT part of the product.
!
! Input stream (little-endian, no record markers), repeated until n = 0:
!   int64 n, maxnz, itmax ; real64 zeta ;
!   int32 jcoef(n,maxnz) ; real64 coef(n,maxnz), rhs(n), u(n)
! Output stream per system:
!   int64 ier, iparm(1), iparm(8) ; real64 rparm(1:3), rparm(11), rparm(12) ;
!   real64 u(n), rhs(n), coef(n,maxnz) ; int32 jcoef(n,maxnz)
program jcg_ref
  use itpackv, only : jcg, dfault
  implicit none
  integer(8) :: n8, m8, it8
  real(8) :: zeta
  integer :: n, maxnz, ier, uin, uout, nw
  integer, allocatable :: jcoef(:,:), iwksp(:)
  real(8), allocatable :: coef(:,:), rhs(:), u(:), wksp(:)
  integer :: iparm(12)
  real(8) :: rparm(12)
  character(len=512) :: fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=uin, file=trim(fin), form='unformatted', access='stream', status='old')
  open(newunit=uout, file=trim(fout), form='unformatted', access='stream', status='replace')
  do
    read(uin) n8
    if (n8 == 0) exit
    read(uin) m8, it8, zeta
    n = int(n8); maxnz = int(m8)
    allocate(jcoef(n, maxnz), coef(n, maxnz), rhs(n), u(n), iwksp(3 * n))
    nw = 4 * n + 4 * int(it8)
    allocate(wksp(nw))
    read(uin) jcoef, coef, rhs, u
    ! src/gwce.F:141-150, 2002
    call dfault(iparm, rparm)
    iparm(1) = int(it8)
    iparm(2) = 0
    iparm(4) = 16
    rparm(1) = zeta
    call jcg(n, n, maxnz, jcoef, coef, rhs, u, iwksp, nw, wksp, iparm, rparm, ier)
    write(uout) int(ier, 8), int(iparm(1), 8), int(iparm(8), 8)
    write(uout) rparm(1), rparm(2), rparm(3), rparm(11), rparm(12)
    write(uout) u, rhs, coef, jcoef
    deallocate(jcoef, coef, rhs, u, iwksp, wksp)
  end do
  close(uin); close(uout)
end program jcg_ref
