! Fortran bindings of the tt-fort MI355X library (C ABI in include/ttfort_amd.h).
!
! Build with -DTTF_FINT64 for a default-integer-8 program (the reference's
! i8 preset) and link libttfort_amd_i8.so; otherwise link libttfort_amd.so.
!
!  * dbfun3, dqr, drow_add, drow_cut, dtransp, dperm1324, djac_gen,
!    djac_apply, duchol_fort, dgmresr_hh_fort: the reference's bind(c)
!    symbols (tt_linalg.f90:30-845) with the same by-reference argument
!    lists, so existing callers reach the GPU by linking this library in
!    place of the reference definitions.
!  * init_bfun_sizes, init_bfun_main, primme_matvec, zmatvec,
!    zmatvec_transp: the callback binding of one local step and the PRIMME /
!    Expokit callback shapes (tt_eigb.f90:15-49, tt_ksl.f90:20-90).
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
    integer(c_long) :: total_matvecs32, refreshes
    integer(c_long) :: sweep_matvecs32(64)
    integer(c_long) :: block_prec_solves, block_prec_fallbacks
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
    integer(c_long) :: matvecs32
  end type ttf_ksl_stats

  interface
    subroutine dbfun3(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi1, A, phi2, x, y, res1, res2) bind(c, name='dbfun3')
      import :: fint, c_double
      integer(fint), intent(in) :: rx1, m, rx2, ry1, n, ry2, ra1, ra2
      real(c_double), intent(in) :: phi1(*), A(*), phi2(*), x(*)
      real(c_double), intent(inout) :: y(*), res1(*), res2(*)
    end subroutine dbfun3

    subroutine dqr(n, m, A, R, work, lwork, tau) bind(c, name='dqr')
      import :: fint, c_double
      integer(fint), intent(in) :: n, m, lwork
      real(c_double), intent(inout) :: A(*), R(*), work(*), tau(*)
    end subroutine dqr

    subroutine drow_add(m, n, k, A, B) bind(c, name='drow_add')
      import :: fint, c_double
      integer(fint), intent(in) :: m, n, k
      real(c_double), intent(inout) :: A(*)
      real(c_double), intent(in) :: B(*)
    end subroutine drow_add

    subroutine drow_cut(m, n, k, A) bind(c, name='drow_cut')
      import :: fint, c_double
      integer(fint), intent(in) :: m, n, k
      real(c_double), intent(inout) :: A(*)
    end subroutine drow_cut

    subroutine dtransp(n, m, A, B) bind(c, name='dtransp')
      import :: fint, c_double
      integer(fint), intent(in) :: n, m
      real(c_double), intent(in) :: A(*)
      real(c_double), intent(inout) :: B(*)
    end subroutine dtransp

    subroutine dperm1324(n1, n2, n3, n4, A, B) bind(c, name='dperm1324')
      import :: fint, c_double
      integer(fint), intent(in) :: n1, n2, n3, n4
      real(c_double), intent(in) :: A(*)
      real(c_double), intent(inout) :: B(*)
    end subroutine dperm1324

    subroutine djac_gen(ptype, rx1, n, rx2, ra1, ra2, Phi1, A, Phi2, jacs) bind(c, name='djac_gen')
      import :: fint, c_double, c_char
      character(kind=c_char), intent(in) :: ptype
      integer(fint), intent(in) :: rx1, n, rx2, ra1, ra2
      real(c_double), intent(in) :: Phi1(*), A(*), Phi2(*)
      real(c_double), intent(inout) :: jacs(*)
    end subroutine djac_gen

    subroutine djac_apply(ptype, rx1, n, rx2, jacs, x, y, work1) bind(c, name='djac_apply')
      import :: fint, c_double, c_char
      character(kind=c_char), intent(in) :: ptype
      integer(fint), intent(in) :: rx1, n, rx2
      real(c_double), intent(in) :: jacs(*), x(*)
      real(c_double), intent(inout) :: y(*), work1(*)
    end subroutine djac_apply

    subroutine duchol_fort(trans, m, n, dA, numvec, du, lam, real_numvec) bind(c, name='duchol_fort')
      import :: fint, c_double, c_char
      character(kind=c_char), intent(in) :: trans
      integer(fint), intent(in) :: m, n, numvec
      real(c_double), intent(in) :: dA(*)
      real(c_double), intent(inout) :: du(*), lam(*)
      integer(fint), intent(inout) :: real_numvec
    end subroutine duchol_fort

    subroutine dgmresr_hh_fort(Phi1, A, Phi2, rhs, rx1, n, rx2, ra1, ra2, nrestart, tol, niters, ptype, jacs, &
                               sol, verb) bind(c, name='dgmresr_hh_fort')
      import :: fint, c_double, c_char
      real(c_double), intent(in) :: Phi1(*), A(*), Phi2(*), rhs(*), jacs(*)
      integer(fint), intent(in) :: rx1, n, rx2, ra1, ra2, nrestart, niters, verb
      real(c_double), intent(in) :: tol
      character(kind=c_char), intent(in) :: ptype
      real(c_double), intent(inout) :: sol(*)
    end subroutine dgmresr_hh_fort

    ! the callback binding of one local step (tt_eigb.f90:29-49) and the
    ! PRIMME / Expokit callback shapes
    subroutine init_bfun_sizes(rx1, m, rx2, ry1, n, ry2, ra1, ra2, xsize, ysize) bind(c, name='ttf_init_bfun_sizes')
      import :: fint
      integer(fint), intent(in) :: rx1, m, rx2, ry1, n, ry2, ra1, ra2, xsize, ysize
    end subroutine init_bfun_sizes

    subroutine init_bfun_main(phi1, A, phi2) bind(c, name='ttf_init_bfun_main')
      import :: c_double
      real(c_double), intent(in) :: phi1(*), A(*), phi2(*)
    end subroutine init_bfun_main

    subroutine zinit_bfun_main(phi1, A, phi2) bind(c, name='ttf_zinit_bfun_main')
      import :: c_double_complex
      complex(c_double_complex), intent(in) :: phi1(*), A(*), phi2(*)
    end subroutine zinit_bfun_main

    subroutine primme_matvec(x, y, k, primme) bind(c, name='ttf_primme_matvec')
      import :: fint, c_double, c_ptr
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
      integer(fint), intent(in) :: k
      type(c_ptr), value :: primme
    end subroutine primme_matvec

    subroutine zmatvec(x, y) bind(c, name='ttf_zmatvec')
      import :: c_double_complex
      complex(c_double_complex), intent(in) :: x(*)
      complex(c_double_complex), intent(inout) :: y(*)
    end subroutine zmatvec

    subroutine zmatvec_transp(x, y) bind(c, name='ttf_zmatvec_transp')
      import :: c_double_complex
      complex(c_double_complex), intent(in) :: x(*)
      complex(c_double_complex), intent(inout) :: y(*)
    end subroutine zmatvec_transp

    integer(c_int) function ttf_my_chop3(n, s, eps) bind(c, name='ttf_my_chop3')
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double), intent(in) :: s(*)
      real(c_double), value :: eps
    end function ttf_my_chop3

    type(c_ptr) function ttf_version() bind(c, name='ttf_version')
      import :: c_ptr
    end function ttf_version

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
