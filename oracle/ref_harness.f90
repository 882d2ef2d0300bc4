! Test harness over the reference's own line-operator modules.
!
! TEST INFRASTRUCTURE ONLY.  Compiled together with the unmodified reference
! sources src/{constdef,commtype,commvar,commfunc,singleton,derivative,
! filter}.F90 (taken from where they lie, never copied) into
! oracle/_ref/librefops.so, which tests/golden/make_golden.py loads to write
! golden vectors and which tests use to pin the C restatement.  It exposes
! C-callable wrappers; every wrapper only calls the reference routine named
! in its comment.
module ref_harness

  use iso_c_binding
  use commtype,   only : compact_scheme
  use commvar,    only : hm, alfa_filter
  use commfunc,   only : tridiagonal_thomas_proprocess, tridiagonal_thomas_solver
  use derivative, only : fd_scheme_initiate, fds, compact_central, diff6ec
  use filter,     only : compact_filter_initiate, filter_coefficient_cal,   &
                         compact_filter, coef2i, coef4i, coef6i, coef8i,     &
                         coef10i, coefb, coefh

  implicit none

contains

  ! src/derivative.F90:63-198  (fd_scheme_initiate + df_compact)
  subroutine ref_df_compact(ntype, dim, f, df) bind(c, name='ref_df_compact')
    integer(c_int), value :: ntype, dim
    real(c_double), intent(in) :: f(-hm:dim+hm)
    real(c_double), intent(out) :: df(0:dim)
    type(compact_scheme) :: s
    if (.not. allocated(fds)) allocate(compact_central :: fds)
    call fd_scheme_initiate(asolver=s, scheme='643c', ntype=ntype, dim=dim, dir=1)
    df = fds%central(s, f, dim)
  end subroutine ref_df_compact

  ! src/derivative.F90:63-158: the pre-factored LHS of the derivative.
  subroutine ref_fd_ac(ntype, dim, n, ac) bind(c, name='ref_fd_ac')
    integer(c_int), value :: ntype, dim
    integer(c_int), intent(out) :: n
    real(c_double), intent(out) :: ac(3, dim+3)
    type(compact_scheme) :: s
    call fd_scheme_initiate(asolver=s, scheme='643c', ntype=ntype, dim=dim, dir=1)
    n = s%last_node - s%first_node + 1
    ac(:, 1:n) = s%ac(:, s%first_node:s%last_node)
    ac(2:3, 1) = (/1.d0, 0.d0/)
  end subroutine ref_fd_ac

  ! src/derivative.F90:350-413
  subroutine ref_diff6ec(ntype, dim, f, df) bind(c, name='ref_diff6ec')
    integer(c_int), value :: ntype, dim
    real(c_double), intent(in) :: f(-hm:dim+hm)
    real(c_double), intent(out) :: df(0:dim)
    df = diff6ec(f, dim, ntype)
  end subroutine ref_diff6ec

  subroutine set_filter_coef(alfa, beta_halo, beta_bound)
    real(c_double), intent(in) :: alfa, beta_halo, beta_bound
    if (allocated(coef2i)) deallocate(coef2i, coef4i, coef6i, coef8i, coef10i, coefb, coefh)
    alfa_filter = alfa
    call filter_coefficient_cal(alfa, beta_halo, beta_bound)
  end subroutine set_filter_coef

  ! src/filter.F90:31-144, 299-432
  subroutine ref_compact_filter(ntype, dim, alfa, beta_halo, beta_bound, f, ff) &
      bind(c, name='ref_compact_filter')
    integer(c_int), value :: ntype, dim
    real(c_double), value :: alfa, beta_halo, beta_bound
    real(c_double), intent(in) :: f(-hm:dim+hm)
    real(c_double), intent(out) :: ff(0:dim)
    type(compact_scheme) :: s
    call set_filter_coef(alfa, beta_halo, beta_bound)
    call compact_filter_initiate(s, ntype, dim)
    ff = compact_filter(s, f, dim)
  end subroutine ref_compact_filter

  ! src/filter.F90:31-100: the pre-factored LHS of the filter.
  subroutine ref_filter_ac(ntype, dim, alfa, n, ac) bind(c, name='ref_filter_ac')
    integer(c_int), value :: ntype, dim
    real(c_double), value :: alfa
    integer(c_int), intent(out) :: n
    real(c_double), intent(out) :: ac(3, dim+7)
    type(compact_scheme) :: s
    alfa_filter = alfa
    call compact_filter_initiate(s, ntype, dim)
    n = s%last_node - s%first_node + 1
    ac(:, 1:n) = s%ac(:, s%first_node:s%last_node)
    ac(2:3, 1) = (/1.d0, 0.d0/)
  end subroutine ref_filter_ac

  ! src/filter.F90:299-432: coef10i(0:5), coefb(0:4,0:8), coefh(0:4,0:10),
  ! coef6i(0:3), coef8i(0:4) in the layout of orc_filter_coef.
  subroutine ref_filter_coef(alfa, beta_halo, beta_bound, cf) bind(c, name='ref_filter_coef')
    real(c_double), value :: alfa, beta_halo, beta_bound
    real(c_double), intent(out) :: cf(0:114)
    integer :: k
    call set_filter_coef(alfa, beta_halo, beta_bound)
    cf = 0.d0
    cf(0:5) = coef10i(0:5)
    do k = 0, 4
      cf(6 + k*9 : 6 + k*9 + 8) = coefb(k, 0:8)
      cf(51 + k*11 : 51 + k*11 + 10) = coefh(k, 0:10)
    enddo
    ! rows 3 and 4 of coefb/coefh are never set by the reference
    cf(6 + 4*9 : 6 + 4*9 + 8) = 0.d0
    cf(51 + 3*11 : 51 + 4*11 + 10) = 0.d0
    cf(106:109) = coef6i(0:3)
    cf(110:114) = coef8i(0:4)
  end subroutine ref_filter_coef

  ! src/commfunc.F90:752-813
  subroutine ref_thomas(n, a, c, d, x) bind(c, name='ref_thomas')
    integer(c_int), value :: n
    real(c_double), intent(in) :: a(n), c(n)
    real(c_double), intent(inout) :: d(n)
    real(c_double), intent(out) :: x(n)
    real(8), allocatable :: ac(:,:)
    ac = tridiagonal_thomas_proprocess(a, c)
    x = tridiagonal_thomas_solver(ac, d)
  end subroutine ref_thomas

end module ref_harness
