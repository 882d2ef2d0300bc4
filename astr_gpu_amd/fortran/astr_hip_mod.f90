! iso_c_binding interface of the HIP solver (include/astr_hip.h).
!
! The Fortran host keeps its module state (commarray, commvar) and replaces
! the calls of time_integration_rk (src/mainloop.F90:392-476) by the entry
! points below.  Host arrays are passed as they are declared in the
! reference, e.g. q(-hm:im+hm,-hm:jm+hm,-hm:km+hm,1:numq).  A nonzero
! return code is turned into a stop by astr_check, which is what mpistop
! does in the reference.
module astr_hip

  use iso_c_binding

  implicit none

  ! struct astr_params
  type, bind(c) :: astr_params
    integer(c_int32_t) :: im, jm, km, hm, numq
    integer(c_int32_t) :: npdci, npdcj, npdck
    integer(c_int32_t) :: variant, reserved
    real(c_double) :: alfa_filter, reynolds, mach, prandtl, gamma
    real(c_double) :: ref_tem, sutherland, beta_halo, beta_bound
    integer(c_int32_t) :: convection, deriv_explicit, recon_schem, lchardecomp
    integer(c_int32_t) :: lfilter, diffterm
    real(c_double) :: shkcrt, bfacmpld
  end type astr_params

  ! astr_set_bc: udf_bc of the 2-D Riemann example
  integer(c_int32_t), parameter :: ASTR_UDF_RIEMANN2D = 1

  interface

    subroutine astr_default_params(p, n) bind(c, name='astr_default_params')
      import :: astr_params, c_int32_t
      type(astr_params), intent(out) :: p
      integer(c_int32_t), value :: n
    end subroutine astr_default_params

    integer(c_int) function astr_create(p, device, ctx) bind(c, name='astr_create')
      import :: astr_params, c_int, c_int32_t, c_ptr
      type(astr_params), intent(in) :: p
      integer(c_int32_t), value :: device
      type(c_ptr), intent(out) :: ctx
    end function astr_create

    integer(c_int) function astr_destroy(ctx) bind(c, name='astr_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function astr_destroy

    integer(c_int) function astr_upload_geom(ctx, jacob, dxi) bind(c, name='astr_upload_geom')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: jacob(*), dxi(*)
    end function astr_upload_geom

    integer(c_int) function astr_upload_state(ctx, q, rho, vel, prs, tmp) &
        bind(c, name='astr_upload_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: q(*), rho(*), vel(*), prs(*), tmp(*)
    end function astr_upload_state

    integer(c_int) function astr_download_state(ctx, q, rho, vel, prs, tmp) &
        bind(c, name='astr_download_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: q(*), rho(*), vel(*), prs(*), tmp(*)
    end function astr_download_state

    ! call filterq (src/mainloop.F90:395)
    integer(c_int) function astr_filterq(ctx) bind(c, name='astr_filterq')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function astr_filterq

    ! call qswap (src/mainloop.F90:415)
    integer(c_int) function astr_qswap(ctx) bind(c, name='astr_qswap')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function astr_qswap

    ! qrhs=0; call rhscal (src/mainloop.F90:404,417)
    integer(c_int) function astr_rhscal(ctx) bind(c, name='astr_rhscal')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function astr_rhscal

    ! RK3 combine and updatefvar (src/mainloop.F90:423-476)
    integer(c_int) function astr_rk3_update(ctx, stage, dt) bind(c, name='astr_rk3_update')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: stage
      real(c_double), value :: dt
    end function astr_rk3_update

    ! kenergycal, enstophycal, diss_rate_cal (src/statistic.F90:870-1042)
    integer(c_int) function astr_stats(ctx, out3) bind(c, name='astr_stats')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out3(3)
    end function astr_stats

    ! the statistics into device memory (out3_dev), queued, allocating nothing
    integer(c_int) function astr_stats_device(ctx, out3_dev) bind(c, name='astr_stats_device')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, out3_dev
    end function astr_stats_device

    ! call time_integration_rk (src/mainloop.F90:104); stats as rkfirst
    integer(c_int) function astr_rk3_step(ctx, dt, stats) bind(c, name='astr_rk3_step')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      real(c_double), intent(out) :: stats(3)
    end function astr_rk3_step

    integer(c_int) function astr_rk3_steps(ctx, nsteps, dt, stats) bind(c, name='astr_rk3_steps')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
      real(c_double), value :: dt
      real(c_double), intent(out) :: stats(*)
    end function astr_rk3_steps

    integer(c_int) function astr_download_rhs(ctx, qrhs, dvel, dtmp) &
        bind(c, name='astr_download_rhs')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: qrhs(*), dvel(*), dtmp(*)
    end function astr_download_rhs

    ! lshock(0:im,0:jm,0:km) of the last ducrossensor, as bytes
    integer(c_int) function astr_download_shock(ctx, lshock) bind(c, name='astr_download_shock')
      import :: c_int, c_ptr, c_int8_t
      type(c_ptr), value :: ctx
      integer(c_int8_t), intent(out) :: lshock(*)
    end function astr_download_shock

    ! fds%central / compact_filter of a whole haloed field along dir
    integer(c_int) function astr_deriv_field(ctx, dir, f, df) bind(c, name='astr_deriv_field')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: dir
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: df(*)
    end function astr_deriv_field

    integer(c_int) function astr_filter_field(ctx, dir, f, ff) bind(c, name='astr_filter_field')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: dir
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: ff(*)
    end function astr_filter_field

    ! upwind interface values of one split-flux field (flux_compact or
    ! recons_exp, by the context's convection), wind +1 / -1
    integer(c_int) function astr_flux_field(ctx, dir, wind, shock, f, fh) &
        bind(c, name='astr_flux_field')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: dir, wind, shock
      real(c_double), intent(in) :: f(*)
      real(c_double), intent(out) :: fh(*)
    end function astr_flux_field

    ! transpose of the derivative (op 0) or filter (op 1) along dir
    integer(c_int) function astr_line_adjoint(ctx, op, dir, wrap, gbar, fbar) &
        bind(c, name='astr_line_adjoint')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: op, dir, wrap
      real(c_double), intent(in) :: gbar(*)
      real(c_double), intent(out) :: fbar(*)
    end function astr_line_adjoint

    ! batched ensemble: ctxs(1:n) advance together (one HIP graph per step)
    integer(c_int) function astr_batch_create(ctxs, n, batch) bind(c, name='astr_batch_create')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), intent(in) :: ctxs(*)
      integer(c_int32_t), value :: n
      type(c_ptr), intent(out) :: batch
    end function astr_batch_create

    integer(c_int) function astr_batch_destroy(batch) bind(c, name='astr_batch_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function astr_batch_destroy

    integer(c_int) function astr_batch_rk3_steps(batch, nsteps, dt) bind(c, name='astr_batch_rk3_steps')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: nsteps
      real(c_double), value :: dt
    end function astr_batch_rk3_steps

    integer(c_int) function astr_batch_set_mode(batch, mode) bind(c, name='astr_batch_set_mode')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: mode
    end function astr_batch_set_mode

    ! line-queue engine (mode 2): trial steps per launch configuration,
    ! the choices kept, and recorded choices applied
    integer(c_int) function astr_batch_autotune(batch, dt) bind(c, name='astr_batch_autotune')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      real(c_double), value :: dt
    end function astr_batch_autotune

    integer(c_int) function astr_batch_get_tuning(batch, rows, cap, nrows) &
        bind(c, name='astr_batch_get_tuning')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: batch
      real(c_double), intent(out) :: rows(*)
      integer(c_int32_t), value :: cap
      integer(c_int32_t), intent(out) :: nrows
    end function astr_batch_get_tuning

    integer(c_int) function astr_batch_set_tuning(batch, choice) bind(c, name='astr_batch_set_tuning')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: batch
      integer(c_int32_t), intent(in) :: choice(*)
    end function astr_batch_set_tuning

    ! uniform spacing of the mini-app variant (variant 1)
    integer(c_int) function astr_set_spacing(ctx, dx, dy, dz) bind(c, name='astr_set_spacing')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dx, dy, dz
    end function astr_set_spacing

    integer(c_int) function astr_set_bc(ctx, bctype, udf) bind(c, name='astr_set_bc')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: bctype(6)
      integer(c_int32_t), value :: udf
    end function astr_set_bc

    ! gridcube's box lengths, for the boundary conditions that read node
    ! coordinates (gcnscbc)
    integer(c_int) function astr_set_lengths(ctx, lx, ly, lz) bind(c, name='astr_set_lengths')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: lx, ly, lz
    end function astr_set_lengths

    ! wall temperature of each face for bctype 41 (twall of src/bc.F90:29)
    integer(c_int) function astr_set_twall(ctx, twall) bind(c, name='astr_set_twall')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: twall(6)
    end function astr_set_twall

    ! slip-nonslip point of bctype 411 (xslip of src/bc.F90:29)
    integer(c_int) function astr_set_xslip(ctx, xslip) bind(c, name='astr_set_xslip')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: xslip
    end function astr_set_xslip

    integer(c_int) function astr_boucon(ctx) bind(c, name='astr_boucon')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function astr_boucon

    integer(c_int) function astr_set_time(ctx, time) bind(c, name='astr_set_time')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time
    end function astr_set_time

    integer(c_int) function astr_get_time(ctx, time) bind(c, name='astr_get_time')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: time
    end function astr_get_time

    ! maxmincal's max_q: the largest q(m) over the nodes 0..n
    integer(c_int) function astr_qmax(ctx, qmax) bind(c, name='astr_qmax')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: qmax(5)
    end function astr_qmax

    ! the discrete adjoint of nsteps RK3 steps, J = enstrophy of the final
    ! state; qbar(-hm:..,1:5), pbar(-hm:..,1:6) of the initial state
    integer(c_int) function astr_adjoint_record(ctx, nsteps, dt, ckpt, j) &
        bind(c, name='astr_adjoint_record')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps, ckpt
      real(c_double), value :: dt
      real(c_double), intent(out) :: j
    end function astr_adjoint_record

    integer(c_int) function astr_adjoint_reverse(ctx, qbar, pbar, ms) &
        bind(c, name='astr_adjoint_reverse')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: qbar(*), pbar(*), ms
    end function astr_adjoint_reverse

    integer(c_int) function astr_adjoint_enstrophy(ctx, nsteps, dt, ckpt, j, qbar, pbar, ms2) &
        bind(c, name='astr_adjoint_enstrophy')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps, ckpt
      real(c_double), value :: dt
      real(c_double), intent(out) :: j, qbar(*), pbar(*), ms2(2)
    end function astr_adjoint_enstrophy

    integer(c_int) function astr_synchronize(ctx) bind(c, name='astr_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function astr_synchronize

    integer(c_int) function astr_set_timing(ctx, on) bind(c, name='astr_set_timing')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function astr_set_timing

    integer(c_int) function astr_get_timing(ctx, ms9) bind(c, name='astr_get_timing')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms9(9)
    end function astr_get_timing

    ! measurement utility: field-sized copy, ms per copy and bytes moved
    integer(c_int) function astr_bench_copy(ctx, reps, ms, nbytes) bind(c, name='astr_bench_copy')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: reps
      real(c_double), intent(out) :: ms
      integer(c_int64_t), intent(out) :: nbytes
    end function astr_bench_copy

    type(c_ptr) function astr_last_error() bind(c, name='astr_last_error')
      import :: c_ptr
    end function astr_last_error

  end interface

contains

  ! Stop with the library's message on a nonzero code (the reference stops
  ! through mpistop on any error).
  subroutine astr_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(astr_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    enddo
    print '(a,a,i0,a,1024a)', what, ' failed, code ', rc, ': ', msg(1:n)
    stop 1
  end subroutine astr_check

end module astr_hip
