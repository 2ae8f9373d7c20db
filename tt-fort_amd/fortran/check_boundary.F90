! A Fortran caller of the library through the ttfort_amd module, built the
! way a tt-fort program would be (default integers of the preset's kind,
! plain arrays passed to the reference's routine names).
!
!   check_boundary host   -- calls that need no GPU (library version, the
!                            chop rule on exact singular values)
!   check_boundary        -- additionally dbfun3 against a direct loop over
!                            its index definition (ttals.f90:18-29), and dqr
!                            (tt_linalg.f90:30-51) checked by Q^T Q = I and
!                            Q R = A
! Exit status 0 when every check holds.
program check_boundary
  use iso_c_binding
  use ttfort_amd
  implicit none
  integer :: rx1, m, rx2, ry1, n, ry2, ra1, ra2
  integer :: i, j, a1, a2, b1, b2, c1, c2, k, nq, mq, lwork
  real(8), allocatable :: phi1(:, :, :), A(:, :, :, :), phi2(:, :, :), x(:, :, :), y(:, :, :), yref(:, :, :)
  real(8), allocatable :: res1(:), res2(:), Aq(:, :), Aq0(:, :), Rq(:, :), work(:), tau(:)
  real(8) :: s(5), err, nrm
  character(len=16) :: arg
  character(kind=c_char), pointer :: ver(:)
  logical :: gpu

  call get_command_argument(1, arg)
  gpu = trim(arg) /= 'host'
  call c_f_pointer(ttf_version(), ver, [64])
  i = 1
  do while (ver(i) /= c_null_char .and. i < 64)
    i = i + 1
  end do
  write(*, '(a)', advance='no') 'library: '
  write(*, '(64a)') ver(1:i - 1)
  ! the chop rule on singular values whose sums are exact: ||s|| = 5,
  ! eps = 0.4, bound 4.0, tail 4.0 kept by the strict comparison -> rank 5
  s = [3d0, 2d0, 2d0, 2d0, 2d0]
  if (ttf_my_chop3(5_c_int, s, 0.4d0) /= 5) then
    print *, 'my_chop3 mismatch'
    stop 1
  end if
  print *, 'my_chop3 ok'
  if (.not. gpu) stop 0

  rx1 = 5; m = 2; rx2 = 4; ry1 = 3; n = 3; ry2 = 6; ra1 = 2; ra2 = 3
  allocate(phi1(ry1, rx1, ra1), A(ra1, n, m, ra2), phi2(rx2, ra2, ry2), x(rx1, m, rx2), y(ry1, n, ry2), &
           yref(ry1, n, ry2), res1(1000), res2(1000))
  call fill3(phi1, 1d0); call fill3(phi2, 2d0); call fill3(x, 3d0)
  do a2 = 1, ra2
    do j = 1, m
      do i = 1, n
        do a1 = 1, ra1
          A(a1, i, j, a2) = sin(0.7d0 * a1 + 1.3d0 * i + 0.4d0 * j + 2.1d0 * a2)
        end do
      end do
    end do
  end do
  call dbfun3(rx1, m, rx2, ry1, n, ry2, ra1, ra2, phi1, A, phi2, x, y, res1, res2)
  yref = 0d0
  do c2 = 1, ry2
    do i = 1, n
      do c1 = 1, ry1
        do a2 = 1, ra2
          do b2 = 1, rx2
            do j = 1, m
              do a1 = 1, ra1
                do b1 = 1, rx1
                  yref(c1, i, c2) = yref(c1, i, c2) + phi1(c1, b1, a1) * A(a1, i, j, a2) * phi2(b2, a2, c2) * &
                                    x(b1, j, b2)
                end do
              end do
            end do
          end do
        end do
      end do
    end do
  end do
  err = maxval(abs(y - yref)) / maxval(abs(yref))
  print *, 'dbfun3 relative error', err
  if (err > 1d-12) stop 2

  nq = 40; mq = 12; k = min(nq, mq); lwork = 64
  allocate(Aq(nq, mq), Aq0(nq, mq), Rq(k, nq), work(lwork), tau(mq))
  do j = 1, mq
    do i = 1, nq
      Aq(i, j) = cos(0.37d0 * i * j + 0.11d0 * i)
    end do
  end do
  Aq0 = Aq
  call dqr(nq, mq, Aq, Rq, work, lwork, tau)
  nrm = maxval(abs(matmul(transpose(Aq(:, 1:k)), Aq(:, 1:k)) - eye(k)))
  err = maxval(abs(matmul(Aq(:, 1:k), Rq(:, 1:mq)) - Aq0)) / maxval(abs(Aq0))
  print *, 'dqr orthonormality', nrm, ' reconstruction', err
  if (nrm > 1d-12 .or. err > 1d-12) stop 3
  print *, 'boundary checks passed'

contains

  subroutine fill3(t, seed)
    real(8), intent(inout) :: t(:, :, :)
    real(8), intent(in) :: seed
    integer :: p, q, r
    do r = 1, size(t, 3)
      do q = 1, size(t, 2)
        do p = 1, size(t, 1)
          t(p, q, r) = sin(seed + 0.9d0 * p + 1.7d0 * q + 0.3d0 * r)
        end do
      end do
    end do
  end subroutine fill3

  function eye(nn) result(e)
    integer, intent(in) :: nn
    real(8) :: e(nn, nn)
    integer :: p
    e = 0d0
    do p = 1, nn
      e(p, p) = 1d0
    end do
  end function eye
end program check_boundary
