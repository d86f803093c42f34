!> Drives the per-step C ABI from Fortran the way ADCIRC's TIMESTEP would
!> (src/timestep.F:1069, 1088, 1103): the mesh and parameters go to
!> adcgpu_create through the bind(C) types adcgpu_mesh / adcgpu_params, with
!> c_loc pointers to Fortran arrays in the reference's layout; each step calls
!> adcgpu_elevation_bc, adcgpu_begin_step, adcgpu_friction,
!> adcgpu_set_forcing, adcgpu_solve_gwce, adcgpu_wetdry and adcgpu_momentum.
!>
!> usage: resident_check IN OUT
!>   IN   stream file: int32 header(18), float64 header(10), then the arrays
!>        (written by tests/test_gpu_fortran.py)
!>   OUT  stream file: int32 numitr, ier; float64 ETA2, UU2, VV2, ETAMAX (NP
!>        each); int32 NODECODE (NP), NOFF (NE)
program resident_check
  use iso_c_binding
  use adcgpu_iface
  implicit none
  type(adcgpu_mesh) :: m
  type(adcgpu_params) :: p
  integer(c_int32_t) :: hi(18)
  real(c_double) :: hd(10)
  integer :: np, ne, mnei, neta, nvel, nvelme, nbfr, nsteps, it0, nws, s, rc
  integer(c_int) :: numitr, ier
  integer(c_int32_t), allocatable, target :: nm(:), nneigh(:), neitab(:), csr(:), nbd(:), &
       nbv(:), lbc(:), me2gw(:), nodecode(:), noff(:)
  real(c_double), allocatable, target :: x(:), y(:), dp(:), fdxe(:), fdye(:), areas(:), &
       csii(:), siii(:), tau0(:), evm(:), fric(:), corif(:), eta2(:), amig(:), ff(:), &
       face(:), emo(:), efa(:), th(:), ramp(:), bc(:), wsx(:), wsy(:), pr(:)
  real(c_double), allocatable, target :: outd(:)
  integer(c_int32_t), allocatable, target :: outi(:)
  character(len=4096) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(10) hi
  read(10) hd
  np = hi(1); ne = hi(2); mnei = hi(3); neta = hi(4); nvel = hi(5); nvelme = hi(6)
  nws = hi(12); nbfr = hi(15); nsteps = hi(16); it0 = hi(17)
  allocate(nm(3*ne), x(np), y(np), dp(np), fdxe(3*ne), fdye(3*ne), areas(ne), nneigh(np), &
       neitab(np*mnei), csr(6*ne), nbd(max(neta, 1)), nbv(max(nvel, 1)), lbc(max(nvel, 1)), &
       csii(max(nvel, 1)), siii(max(nvel, 1)), me2gw(max(nvelme, 1)), tau0(np), evm(np), &
       fric(np), corif(np), eta2(np), nodecode(np), noff(ne), amig(max(nbfr, 1)), &
       ff(max(nbfr, 1)), face(max(nbfr, 1)), emo(max(neta*nbfr, 1)), efa(max(neta*nbfr, 1)), &
       th(nsteps), ramp(nsteps), bc(max(neta, 1)), wsx(np), wsy(np), pr(np))
  read(10) nm, x, y, dp, fdxe, fdye, areas, nneigh, neitab, csr
  read(10) nbd
  if (nvel > 0) read(10) nbv(1:nvel), lbc(1:nvel), csii(1:nvel), siii(1:nvel)
  if (nvelme > 0) read(10) me2gw(1:nvelme)
  read(10) tau0, evm, fric, corif, eta2, nodecode, noff
  if (nbfr > 0) read(10) amig(1:nbfr), ff(1:nbfr), face(1:nbfr)
  if (neta*nbfr > 0) read(10) emo(1:neta*nbfr), efa(1:neta*nbfr)
  read(10) th, ramp

  m%np = np; m%ne = ne; m%mnei = mnei; m%neta = neta; m%nvel = nvel; m%nvelme = nvelme
  m%nm = c_loc(nm); m%x = c_loc(x); m%y = c_loc(y); m%dp = c_loc(dp)
  m%fdxe = c_loc(fdxe); m%fdye = c_loc(fdye); m%areas = c_loc(areas)
  m%nneigh = c_loc(nneigh); m%neitab = c_loc(neitab); m%csr_position = c_loc(csr)
  m%nbd = c_loc(nbd); m%nbv = c_loc(nbv); m%lbcodei = c_loc(lbc); m%csii = c_loc(csii)
  m%siii = c_loc(siii); m%me2gw = c_loc(me2gw)
  p%dtdp = hd(1); p%g = hd(2); p%h0 = hd(3); p%velmin = hd(4); p%a00 = hd(5); p%b00 = hd(6)
  p%c00 = hd(7); p%convcr = hd(8); p%prbckgrnd_mh2o = hd(9); p%errorelev = hd(10)
  p%nolifa = hi(7); p%nolica = hi(8); p%nolicat = hi(9); p%nolibf = hi(10); p%itmax = hi(11)
  p%nws = nws; p%ilump = hi(13); p%warm_start = hi(14); p%precision = hi(18)
  p%tau0var = c_loc(tau0); p%evm = c_loc(evm); p%fric = c_loc(fric); p%corif = c_loc(corif)

  call check(adcgpu_create(m, p, 0_c_int, adcgpu_ctx), 'adcgpu_create')
  call check(adcgpu_cold_start(adcgpu_ctx, eta2, nodecode, noff), 'adcgpu_cold_start')
  if (nws /= 0) then
     ! the cold-start forcing: the first step's WSX1/WSY1/PR1
     read(10) wsx, wsy, pr
     call check(adcgpu_set_field(adcgpu_ctx, F_WSX2, c_loc(wsx), int(np, c_int64_t)), 'wsx2')
     call check(adcgpu_set_field(adcgpu_ctx, F_WSY2, c_loc(wsy), int(np, c_int64_t)), 'wsy2')
     call check(adcgpu_set_field(adcgpu_ctx, F_PR2, c_loc(pr), int(np, c_int64_t)), 'pr2')
  end if
  do s = 1, nsteps
     call check(adcgpu_elevation_bc(neta, nbfr, amig, ff, face, emo, efa, th(s), ramp(s), bc), &
          'adcgpu_elevation_bc')
     call check(adcgpu_begin_step(adcgpu_ctx, it0 + s - 1), 'adcgpu_begin_step')
     call check(adcgpu_friction(adcgpu_ctx), 'adcgpu_friction')
     if (nws /= 0) then
        read(10) wsx, wsy, pr
        call check(adcgpu_set_forcing(adcgpu_ctx, wsx, wsy, pr), 'adcgpu_set_forcing')
     end if
     call check(adcgpu_solve_gwce(adcgpu_ctx, it0 + s - 1, bc), 'adcgpu_solve_gwce')
     call check(adcgpu_wetdry(adcgpu_ctx, it0 + s - 1), 'adcgpu_wetdry')
     call check(adcgpu_momentum(adcgpu_ctx), 'adcgpu_momentum')
  end do
  close(10)
  call check(adcgpu_last_numitr(adcgpu_ctx, numitr, ier), 'adcgpu_last_numitr')

  allocate(outd(np), outi(max(np, ne)))
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(11) int(numitr, c_int32_t), int(ier, c_int32_t)
  call field_d(F_ETA2); call field_d(F_UU2); call field_d(F_VV2); call field_d(F_ETAMAX)
  call check(adcgpu_get_field(adcgpu_ctx, F_NODECODE, c_loc(outi), int(np, c_int64_t)), 'nodecode')
  write(11) outi(1:np)
  call check(adcgpu_get_field(adcgpu_ctx, F_NOFF, c_loc(outi), int(ne, c_int64_t)), 'noff')
  write(11) outi(1:ne)
  close(11)
  call adcgpu_destroy(adcgpu_ctx)
  print '(a,i0,a,i0,a,i0)', 'resident_check: steps=', nsteps, ' numitr=', numitr, ' ier=', ier

contains

  subroutine field_d(f)
    integer(c_int), intent(in) :: f
    call check(adcgpu_get_field(adcgpu_ctx, f, c_loc(outd), int(np, c_int64_t)), 'get_field')
    write(11) outd
  end subroutine

  subroutine check(rc_in, what)
    integer(c_int), intent(in) :: rc_in
    character(len=*), intent(in) :: what
    if (rc_in /= 0) then
       print '(a,a,a,i0,a,a)', 'resident_check: ', what, ' failed (', rc_in, '): ', &
            adcgpu_error_message()
       stop 1
    end if
  end subroutine
end program resident_check
