!> ISO_C_BINDING interface to libadcgpu.so (include/adcgpu.h) for ADCIRC's
!> Fortran timestep driver.  The module holds the device context and mirrors
!> the call sites it replaces:
!>
!>   CALL JCG(NP,MNP,MNEI,NEITAB,COEF,GWCE_LV,ETAS,IWKSP,NW,WKSP,IPARM,RPARM,IER)
!>        (src/gwce.F:2003-2004)  ->  IER = adcgpu_jcg(...)
!>   CALL solveGWCE(IT,ITIME_BGN,TimeLoc,TimeH)        (src/timestep.F:1069)
!>   CALL computeWettingAndDrying(IT)                  (src/timestep.F:1088)
!>   CALL solveMomentumEq()                            (src/timestep.F:1103)
!>        ->  adcgpu_begin_step / adcgpu_friction / adcgpu_set_forcing /
!>            adcgpu_solve_gwce / adcgpu_wetdry / adcgpu_momentum
module adcgpu_iface
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: ADCGPU_FP64 = 0, ADCGPU_MIXED = 1
  integer(c_int), parameter, public :: F_ETA1 = 0, F_ETA2 = 1, F_ETAS = 2, F_UU1 = 3, &
       F_VV1 = 4, F_UU2 = 5, F_VV2 = 6, F_QX1 = 7, F_QY1 = 8, F_QX2 = 9, F_QY2 = 10, &
       F_H1 = 11, F_H2 = 12, F_TK = 13, F_ETAMAX = 14, F_TOTALAREA = 15, F_WSX1 = 16, &
       F_WSY1 = 17, F_WSX2 = 18, F_WSY2 = 19, F_PR1 = 20, F_PR2 = 21, F_UMAX = 23, &
       F_NODECODE = 100, F_NOFF = 101, F_MJU = 102, F_ETAMAX_IT = 103, F_UMAX_IT = 104, F_COEF = 200

  type, bind(C), public :: adcgpu_mesh
     integer(c_int32_t) :: np, ne, mnei, neta, nvel, nvelme
     type(c_ptr) :: nm, x, y, dp, fdxe, fdye, areas, nneigh, neitab, csr_position
     type(c_ptr) :: nbd, nbv, lbcodei, csii, siii, me2gw
  end type

  type, bind(C), public :: adcgpu_params
     real(c_double) :: dtdp, g, h0, velmin, a00, b00, c00, convcr, prbckgrnd_mh2o
     integer(c_int32_t) :: nolifa, nolica, nolicat, nolibf, itmax, nws, precision
     type(c_ptr) :: tau0var, evm, fric, corif
     integer(c_int32_t) :: ilump
     integer(c_int32_t) :: warm_start
     real(c_double) :: errorelev
  end type

  !> the context of this process (one GPU per process)
  type(c_ptr), public, save :: adcgpu_ctx = c_null_ptr

  public :: adcgpu_create, adcgpu_destroy, adcgpu_jcg, adcgpu_cold_start, adcgpu_begin_step, &
       adcgpu_friction, adcgpu_set_forcing, adcgpu_solve_gwce, adcgpu_wetdry, &
       adcgpu_momentum, adcgpu_get_field, adcgpu_set_field, adcgpu_last_numitr, &
       adcgpu_elevation_bc, adcgpu_last_error, adcgpu_error_message

  interface
     !> message of the last failed call on this thread (NUL-terminated C string)
     type(c_ptr) function adcgpu_last_error() bind(C, name="adcgpu_last_error")
       import :: c_ptr
     end function
     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
     end function
     integer(c_int) function adcgpu_create(mesh, prm, device, out) bind(C, name="adcgpu_create")
       import :: c_int, c_ptr, adcgpu_mesh, adcgpu_params
       type(adcgpu_mesh), intent(in) :: mesh
       type(adcgpu_params), intent(in) :: prm
       integer(c_int), value :: device
       type(c_ptr), intent(out) :: out
     end function
     subroutine adcgpu_destroy(ctx) bind(C, name="adcgpu_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine
     !> drop-in for ITPACKV's JCG (src/itpackv.F:63): same arrays, same in-place
     !> semantics; returns ier
     integer(c_int) function adcgpu_jcg(n, ndim, maxnz, jcoef, coef, rhs, u, iparm, rparm) &
          bind(C, name="adcgpu_jcg")
       import :: c_int, c_int32_t, c_double
       integer(c_int), value :: n, ndim, maxnz
       integer(c_int32_t), intent(inout) :: jcoef(ndim, maxnz)
       real(c_double), intent(inout) :: coef(ndim, maxnz), rhs(n), u(n)
       integer(c_int32_t), intent(inout) :: iparm(12)
       real(c_double), intent(inout) :: rparm(12)
     end function
     integer(c_int) function adcgpu_cold_start(ctx, eta2, nodecode, noff) &
          bind(C, name="adcgpu_cold_start")
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: eta2(*)
       integer(c_int32_t), intent(in) :: nodecode(*), noff(*)
     end function
     integer(c_int) function adcgpu_begin_step(ctx, it) bind(C, name="adcgpu_begin_step")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: it
     end function
     integer(c_int) function adcgpu_friction(ctx) bind(C, name="adcgpu_friction")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function adcgpu_set_forcing(ctx, wsx2, wsy2, pr2) &
          bind(C, name="adcgpu_set_forcing")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: wsx2(*), wsy2(*), pr2(*)
     end function
     integer(c_int) function adcgpu_solve_gwce(ctx, it, eta_bc) bind(C, name="adcgpu_solve_gwce")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: it
       real(c_double), intent(in) :: eta_bc(*)
     end function
     integer(c_int) function adcgpu_wetdry(ctx, it) bind(C, name="adcgpu_wetdry")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: it
     end function
     integer(c_int) function adcgpu_momentum(ctx) bind(C, name="adcgpu_momentum")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function adcgpu_get_field(ctx, field, host, count) bind(C, name="adcgpu_get_field")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       type(c_ptr), value :: host
       integer(c_int64_t), value :: count
     end function
     integer(c_int) function adcgpu_set_field(ctx, field, host, count) bind(C, name="adcgpu_set_field")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       type(c_ptr), value :: host
       integer(c_int64_t), value :: count
     end function
     integer(c_int) function adcgpu_last_numitr(ctx, numitr, ier) bind(C, name="adcgpu_last_numitr")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: numitr, ier
     end function
     integer(c_int) function adcgpu_elevation_bc(neta, nbfr, amig, ff, face, emo, efa, timeh, &
          rampelev, out) bind(C, name="adcgpu_elevation_bc")
       import :: c_int, c_double
       integer(c_int), value :: neta, nbfr
       real(c_double), intent(in) :: amig(*), ff(*), face(*), emo(*), efa(*)
       real(c_double), value :: timeh, rampelev
       real(c_double), intent(out) :: out(*)
     end function
     ! hybrid bottom friction (NOLIBF = 2): fort.15's HBREAK, FTHETA, FGAMMA
     integer(c_int) function adcgpu_set_hybrid_friction(ctx, hbreak, ftheta, fgamma) &
          bind(C, name="adcgpu_set_hybrid_friction")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: hbreak, ftheta, fgamma
     end function
     ! specified normal flux boundaries: BNDLEN2O3 once, then each step's QN2
     ! (src/timestep.F:861-903) after adcgpu_begin_step
     integer(c_int) function adcgpu_set_flux_boundary(ctx, bndlen2o3) &
          bind(C, name="adcgpu_set_flux_boundary")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: bndlen2o3(*)
     end function
     integer(c_int) function adcgpu_set_flux(ctx, qn2) bind(C, name="adcgpu_set_flux")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: qn2(*)
     end function
     integer(c_int) function adcgpu_flux_bc(nvel, lbcodei, nffr, famig, fff, fface, qnam, &
          qnph, timeh, rampextflux, out) bind(C, name="adcgpu_flux_bc")
       import :: c_int, c_double
       integer(c_int), value :: nvel, nffr
       integer(c_int), intent(in) :: lbcodei(*)
       real(c_double), intent(in) :: famig(*), fff(*), fface(*), qnam(*), qnph(*)
       real(c_double), value :: timeh, rampextflux
       real(c_double), intent(out) :: out(*)
     end function
  end interface

contains

  !> adcgpu_last_error as a Fortran string
  function adcgpu_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: k, n
    p = adcgpu_last_error()
    if (.not. c_associated(p)) then
       msg = ''
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do k = 1, n
       msg(k:k) = chars(k)
    end do
  end function
end module adcgpu_iface
