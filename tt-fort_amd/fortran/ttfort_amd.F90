! Fortran bindings of the tt-fort MI355X library (C ABI in include/ttfort_amd.h).
!
! Build with -DTTF_FINT64 for a default-integer-8 program (the reference's
! i8 preset) and link libttfort_amd_i8.so; otherwise link libttfort_amd.so.
!
!  * dbfun3: the reference's bind(c) symbol (tt_linalg.f90:127-159) with the
!    same by-reference argument list, so existing callers (GMRES, PRIMME and
!    Expokit adapters) reach the GPU by linking this library in place of the
!    reference definition.
!  * tt_eigb_gpu, tt_amen_solve_gpu, ztt_ksl_gpu: the argument lists of
!    tt_eigb (tt_eigb.f90:110-120), tt_amen_solve (tt_adapt_als.f90:513-525)
!    and ztt_ksl (tt_ksl.f90:431-445) with the same optional arguments and
!    defaults; results land in the module allocatables result_core /
!    zresult_core, as in the reference modules.
module ttfort_amd
  use iso_c_binding
  implicit none

#ifdef TTF_FINT64
  integer, parameter :: fint = c_int64_t
#else
  integer, parameter :: fint = c_int32_t
#endif

  real(c_double), allocatable :: result_core(:)
  complex(c_double_complex), allocatable :: zresult_core(:)

  type, bind(c) :: ttf_eigb_stats
    integer(c_int) :: sweeps
    integer(c_long) :: total_matvecs
    real(c_double) :: last_ermax
    integer(c_int) :: max_rank
    real(c_double) :: seconds, solve_seconds, trunc_seconds, phi_seconds
    integer(c_int) :: nrec
    real(c_double) :: sweep_seconds(64), sweep_flops(64)
    integer(c_long) :: sweep_matvecs(64)
    real(c_double) :: sweep_ermax(64), sweep_lambda_sum(64)
    integer(c_int) :: converged
  end type ttf_eigb_stats

  type, bind(c) :: ttf_amen_stats
    integer(c_int) :: sweeps
    integer(c_long) :: total_matvecs
    real(c_double) :: last_res_max, last_err_max
    integer(c_int) :: max_rank
    real(c_double) :: seconds, solve_seconds, trunc_seconds, phi_seconds, flops
    integer(c_long) :: learned_solves, learned_fallbacks
  end type ttf_amen_stats

  type, bind(c) :: ttf_ksl_stats
    integer(c_long) :: nsolves, matvecs, krylov_steps
    real(c_double) :: seconds, flops
  end type ttf_ksl_stats

  interface
    subroutine dbfun3(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi1, A, phi2, x, y, res1, res2) bind(c, name='dbfun3')
      import :: fint, c_double
      integer(fint), intent(in) :: rx1, m, rx2, ry1, n, ry2, ra1, ra2
      real(c_double), intent(in) :: phi1(*), A(*), phi2(*), x(*)
      real(c_double), intent(inout) :: y(*), res1(*), res2(*)
    end subroutine dbfun3

    integer(c_int) function ttf_init(device) bind(c, name='ttf_init')
      import :: c_int
      integer(c_int), value :: device
    end function ttf_init

    type(c_ptr) function ttf_last_error() bind(c, name='ttf_last_error')
      import :: c_ptr
    end function ttf_last_error

    integer(c_int) function c_tt_eigb(d, n, m, ra, crA, crY0, ry, eps, rmax, lambda, B, kickrank, nswp, &
                                      max_full_size, verb, result, result_cap, result_size, stats) &
        bind(c, name='ttf_tt_eigb')
      import :: c_int, c_double, c_int64_t, ttf_eigb_stats
      integer(c_int), value :: d, rmax, B, kickrank, nswp, max_full_size, verb
      integer(c_int), intent(in) :: n(*), m(*), ra(*)
      integer(c_int), intent(inout) :: ry(*)
      real(c_double), intent(in) :: crA(*), crY0(*)
      real(c_double), value :: eps
      real(c_double), intent(out) :: lambda(*), result(*)
      integer(c_int64_t), value :: result_cap
      integer(c_int64_t), intent(out) :: result_size
      type(ttf_eigb_stats), intent(out) :: stats
    end function c_tt_eigb

    integer(c_int) function c_tt_amen_solve(d, n, m, ry, ra, crA, crY, crX0, rx, eps, kickrank, nswp, verb, &
                                            prec, nrestart, niters, rmax, result, result_cap, result_size, stats) &
        bind(c, name='ttf_tt_amen_solve')
      import :: c_int, c_double, c_int64_t, c_char, ttf_amen_stats
      integer(c_int), value :: d, kickrank, nswp, verb, nrestart, niters, rmax
      integer(c_int), intent(in) :: n(*), m(*), ry(*), ra(*)
      integer(c_int), intent(inout) :: rx(*)
      real(c_double), intent(in) :: crA(*), crY(*), crX0(*)
      real(c_double), value :: eps
      character(kind=c_char), value :: prec
      real(c_double), intent(out) :: result(*)
      integer(c_int64_t), value :: result_cap
      integer(c_int64_t), intent(out) :: result_size
      type(ttf_amen_stats), intent(out) :: stats
    end function c_tt_amen_solve

    integer(c_int) function c_ztt_ksl(d, n, m, ra, crA, crY0, ry, tau, rmax, kickrank, nswp, verb, typ, order, &
                                      usenrm, result, result_cap, result_size, stats) bind(c, name='ttf_ztt_ksl')
      import :: c_int, c_double, c_double_complex, c_int64_t, ttf_ksl_stats
      integer(c_int), value :: d, rmax, kickrank, nswp, verb, typ, order, usenrm
      integer(c_int), intent(in) :: n(*), m(*), ra(*)
      integer(c_int), intent(inout) :: ry(*)
      complex(c_double_complex), intent(in) :: crA(*), crY0(*)
      real(c_double), value :: tau
      complex(c_double_complex), intent(out) :: result(*)
      integer(c_int64_t), value :: result_cap
      integer(c_int64_t), intent(out) :: result_size
      type(ttf_ksl_stats), intent(out) :: stats
    end function c_ztt_ksl
  end interface

contains

  subroutine check_status(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    if (rc == 0) return
    call c_f_pointer(ttf_last_error(), msg, [1024])
    write(*, '(a)', advance='no') where // ': '
    do i = 1, 1024
      if (msg(i) == c_null_char) exit
      write(*, '(a)', advance='no') msg(i)
    end do
    write(*, *)
    error stop 1
  end subroutine check_status

  subroutine tt_eigb_gpu(d, n, m, ra, crA, crY0, ry, eps, rmax, lambda, B, kickrank, nswp, max_full_size, verb)
    integer, intent(in) :: d, n(d), m(d), ra(d + 1), rmax, B
    integer, intent(inout) :: ry(d + 1)
    real(c_double), intent(in) :: crA(*), crY0(*), eps
    real(c_double), intent(out) :: lambda(B)
    integer, intent(in), optional :: kickrank, nswp, max_full_size, verb
    integer(c_int) :: n_(d), m_(d), ra_(d + 1), ry_(d + 1), kick0, nswp0, mfs0, verb0
    integer(c_int64_t) :: cap, sz
    type(ttf_eigb_stats) :: st
    kick0 = 5; nswp0 = 20; mfs0 = 2000; verb0 = 1     ! tt_eigb.f90:165-190
    if (present(kickrank)) kick0 = kickrank
    if (present(nswp)) nswp0 = nswp
    if (present(max_full_size)) mfs0 = max_full_size
    if (present(verb)) verb0 = verb
    n_ = n; m_ = m; ra_ = ra; ry_ = ry
    cap = int(sum(int(max(rmax, maxval(ry)), 8)**2 * n) * B, c_int64_t)
    if (allocated(result_core)) deallocate(result_core)
    allocate(result_core(cap))
    call check_status(c_tt_eigb(int(d, c_int), n_, m_, ra_, crA, crY0, ry_, eps, int(rmax, c_int), lambda, &
                                int(B, c_int), kick0, nswp0, mfs0, verb0, result_core, cap, sz, st), 'tt_eigb')
    ry = ry_
  end subroutine tt_eigb_gpu

  subroutine tt_amen_solve_gpu(d, n, m, ry, ra, crA, crY, crX0, rx, eps, kickrank, nswp, verb, prec, nrestart, &
                               niters)
    integer, intent(in) :: d, n(d), m(d), ry(d + 1), ra(d + 1)
    integer, intent(inout) :: rx(d + 1)
    real(c_double), intent(in) :: crA(*), crY(*), crX0(*), eps
    integer, intent(in), optional :: kickrank, nswp, verb, nrestart, niters
    character, intent(in), optional :: prec
    integer(c_int) :: n_(d), m_(d), ry_(d + 1), ra_(d + 1), rx_(d + 1), kick0, nswp0, verb0, nrs0, nit0
    character(kind=c_char) :: prec0
    integer(c_int64_t) :: cap, sz
    type(ttf_amen_stats) :: st
    kick0 = 4; nswp0 = 20; verb0 = 1; prec0 = 'n'; nrs0 = 40; nit0 = 2   ! tt_adapt_als.f90:549-573
    if (present(kickrank)) kick0 = kickrank
    if (present(nswp)) nswp0 = nswp
    if (present(verb)) verb0 = verb
    if (present(prec)) prec0 = prec
    if (present(nrestart)) nrs0 = nrestart
    if (present(niters)) nit0 = niters
    n_ = n; m_ = m; ry_ = ry; ra_ = ra; rx_ = rx
    cap = int(256_8 * 256_8 * sum(int(n, 8)), c_int64_t)
    if (allocated(result_core)) deallocate(result_core)
    allocate(result_core(cap))
    call check_status(c_tt_amen_solve(int(d, c_int), n_, m_, ry_, ra_, crA, crY, crX0, rx_, eps, kick0, nswp0, &
                                      verb0, prec0, nrs0, nit0, 0_c_int, result_core, cap, sz, st), 'tt_amen_solve')
    rx = rx_
  end subroutine tt_amen_solve_gpu

  subroutine ztt_ksl_gpu(d, n, m, ra, crA, crY0, ry, tau, rmax, kickrank, nswp, verb, typ0, order0, usenrm0)
    integer, intent(in) :: d, n(d), m(d), ra(d + 1), rmax
    integer, intent(inout) :: ry(d + 1)
    complex(c_double_complex), intent(in) :: crA(*), crY0(*)
    real(c_double), intent(in) :: tau
    integer, intent(in), optional :: kickrank, nswp, verb, typ0, order0, usenrm0
    integer(c_int) :: n_(d), m_(d), ra_(d + 1), ry_(d + 1), kick0, nswp0, verb0, typ, order, usenrm
    integer(c_int64_t) :: cap, sz
    type(ttf_ksl_stats) :: st
    kick0 = 5; nswp0 = 20; verb0 = 1; typ = 2; order = 8; usenrm = 1   ! tt_ksl.f90:467-500
    if (present(kickrank)) kick0 = kickrank
    if (present(nswp)) nswp0 = nswp
    if (present(verb)) verb0 = verb
    if (present(typ0)) typ = typ0
    if (present(order0)) order = order0
    if (present(usenrm0)) usenrm = usenrm0
    n_ = n; m_ = m; ra_ = ra; ry_ = ry
    cap = int(sum(int(maxval(ry), 8)**2 * n), c_int64_t)
    if (allocated(zresult_core)) deallocate(zresult_core)
    allocate(zresult_core(cap))
    call check_status(c_ztt_ksl(int(d, c_int), n_, m_, ra_, crA, crY0, ry_, tau, int(rmax, c_int), kick0, nswp0, &
                                verb0, typ, order, usenrm, zresult_core, cap, sz, st), 'ztt_ksl')
    ry = ry_
  end subroutine ztt_ksl_gpu

end module ttfort_amd
