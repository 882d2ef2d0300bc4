! Fortran host driving the HIP solver through iso_c_binding: the Taylor-Green
! case of examples/Taylor_Green_Vortex set up as the reference does
! (gridcube, src/gridgeneration.F90:233-267; tgvini + updateq,
! src/initialisation.F90:1079-1149; uniform-box metrics) and advanced with
! astr_rk3_step in place of time_integration_rk (src/mainloop.F90:104).
!
!   tgv_hip N NSTEPS DT      prints "nstep ke enstrophy dissipation" per step
program tgv_hip

  use iso_c_binding
  use astr_hip

  implicit none

  integer, parameter :: hm = 5
  integer :: im, nsteps, i, j, k, n
  real(8) :: dt, pi, dx, x, y, z, u, v, w, p, t, pinf, const1, const2
  real(8) :: st(3)
  real(8), allocatable :: q(:,:,:,:), rho(:,:,:), vel(:,:,:,:), prs(:,:,:), tmp(:,:,:)
  real(8), allocatable :: jacob(:,:,:), dxi(:,:,:,:,:)
  character(len=32) :: arg
  type(astr_params) :: prm
  type(c_ptr) :: ctx

  call get_command_argument(1, arg); read(arg, *) im
  call get_command_argument(2, arg); read(arg, *) nsteps
  call get_command_argument(3, arg); read(arg, *) dt

  call astr_default_params(prm, int(im, c_int32_t))
  call astr_check(astr_create(prm, 0_c_int32_t, ctx), 'astr_create')

  allocate(q(-hm:im+hm,-hm:im+hm,-hm:im+hm,1:5), rho(-hm:im+hm,-hm:im+hm,-hm:im+hm), &
           vel(-hm:im+hm,-hm:im+hm,-hm:im+hm,1:3), prs(-hm:im+hm,-hm:im+hm,-hm:im+hm), &
           tmp(-hm:im+hm,-hm:im+hm,-hm:im+hm), jacob(-hm:im+hm,-hm:im+hm,-hm:im+hm), &
           dxi(-hm:im+hm,-hm:im+hm,-hm:im+hm,1:3,1:3))

  pi = 4.d0*atan(1.d0)
  dx = 2.d0*pi/real(im, 8)
  jacob = dx*dx*dx
  dxi = 0.d0
  do n = 1, 3
    dxi(:,:,:,n,n) = 1.d0/dx
  enddo
  call astr_check(astr_upload_geom(ctx, jacob, dxi), 'astr_upload_geom')

  const1 = 1.d0/(prm%gamma*(prm%gamma-1.d0)*prm%mach**2)
  const2 = prm%gamma*prm%mach**2
  pinf = 1.d0/const2
  q = 0.d0; rho = 0.d0; vel = 0.d0; prs = 0.d0; tmp = 0.d0
  do k = 0, im
  do j = 0, im
  do i = 0, im
    x = 2.d0*pi/real(im, 8)*real(i, 8)
    y = 2.d0*pi/real(im, 8)*real(j, 8)
    z = 2.d0*pi/real(im, 8)*real(k, 8)
    u =  sin(x)*cos(y)*cos(z)
    v = -cos(x)*sin(y)*cos(z)
    w = 0.d0
    p = pinf + 1.d0/16.d0*(cos(2.d0*x)+cos(2.d0*y))*(cos(2.d0*z)+2.d0)
    t = p/1.d0*const2
    rho(i,j,k) = 1.d0
    vel(i,j,k,:) = (/u, v, w/)
    prs(i,j,k) = p
    tmp(i,j,k) = t
    q(i,j,k,1) = 1.d0
    q(i,j,k,2) = u
    q(i,j,k,3) = v
    q(i,j,k,4) = w
    q(i,j,k,5) = 1.d0*(t*const1 + 0.5d0*(u**2 + v**2 + w**2))
  enddo
  enddo
  enddo
  call astr_check(astr_upload_state(ctx, q, rho, vel, prs, tmp), 'astr_upload_state')

  do n = 1, nsteps
    call astr_check(astr_rk3_step(ctx, dt, st), 'astr_rk3_step')
    write(*, '(i6,3es24.16)') n - 1, st
  enddo

  call astr_check(astr_download_state(ctx, q, rho, vel, prs, tmp), 'astr_download_state')
  write(*, '(a,es24.16)') ' max|u| ', maxval(abs(vel(0:im,0:im,0:im,1)))
  call astr_check(astr_destroy(ctx), 'astr_destroy')

end program tgv_hip
