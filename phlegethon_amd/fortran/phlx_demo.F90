! phlx_demo.F90 -- a Fortran caller of libphlx through phlx_iface, in the
! shape of the reference's time_loop: create a context, upload the initial
! state (SoA arrays, interior cells), run the step-0 initialisation, advance
! with the library's device-resident loop, download and check.
!
! Problem: 3D periodic ideal-gas density wave, rho = 1 + 0.1 sin(2 pi x),
! v = (0.5, 0.25, 0.125), p = 1, LHLLC + LIM2ND, RK2.  Checks: the step
! counter, conservation of mass (1e-13 relative) and finite output.
! Prints PASS on success and stops with a non-zero code otherwise.
program phlx_demo
  use, intrinsic :: iso_c_binding
  use phlx_iface
  implicit none
  integer, parameter :: n = 32, nv = 5, nsteps = 20
  type(phlx_config) :: pc
  type(c_ptr) :: ctx
  integer(c_int) :: ierr
  real(c_double), allocatable, target :: prim(:, :, :, :), out(:, :, :, :)
  real(c_double) :: t, dt, smax, x, mass0, mass1, pi
  integer(c_int64_t) :: step
  integer :: i

  pi = 4.0d0 * atan(1.0d0)
  pc = phlx_config(sdims=3, nx1=n, nx2=n, nx3=n, nas=0, &
                   x1l=0.0d0, x1u=1.0d0, x2l=0.0d0, x2u=1.0d0, x3l=0.0d0, x3u=1.0d0, &
                   gamma=5.0d0 / 3.0d0, mu=1.0d0, recon=PHLX_LIM2ND_REC, &
                   flux=PHLX_LHLLC_FLUX, all_mach=0, rk_stages=2, cfl=0.8d0, &
                   mach_cutoff=1.0d-4, mach_ct=1.0d-6, tmax=1.0d30, bc=PHLX_BC_PERIODIC, &
                   use_wb=0, use_gravity=0, evolve_etot=0, use_edot=0, &
                   eos=PHLX_EOS_IDEAL, fasteos=0, precision=PHLX_PREC_FP64, device=0, &
                   reserved=0)
  ! SoA: prim(i, j, k, v), x1 fastest, as the C side's [v][k][j][i]
  allocate(prim(n, n, n, nv), out(n, n, n, nv))
  do i = 1, n
    x = (real(i, c_double) - 0.5d0) / real(n, c_double)
    prim(i, :, :, 1) = 1.0d0 + 0.1d0 * sin(2.0d0 * pi * x)
  end do
  prim(:, :, :, 2) = 0.5d0
  prim(:, :, :, 3) = 0.25d0
  prim(:, :, :, 4) = 0.125d0
  prim(:, :, :, 5) = 1.0d0
  mass0 = sum(prim(:, :, :, 1))

  ierr = phlx_create(pc, ctx)
  call check(ierr, "phlx_create")
  call check(phlx_upload_state(ctx, c_loc(prim), c_null_ptr, c_null_ptr, c_null_ptr, &
                               c_null_ptr, c_null_ptr), "phlx_upload_state")
  call check(phlx_init(ctx), "phlx_init")
  call check(phlx_run(ctx, int(nsteps, c_int32_t)), "phlx_run")
  call check(phlx_get_scalars(ctx, t, dt, smax, step), "phlx_get_scalars")
  call check(phlx_download_state(ctx, c_loc(out), c_null_ptr, c_null_ptr, c_null_ptr, &
                                 c_null_ptr, c_null_ptr, c_null_ptr), "phlx_download_state")
  call check(phlx_destroy(ctx), "phlx_destroy")

  mass1 = sum(out(:, :, :, 1))
  write(*, '(a, i0, a, es12.5, a, es10.3)') 'steps ', step, ' time ', t, &
      ' mass drift ', abs(mass1 / mass0 - 1.0d0)
  if (step /= nsteps) stop 2
  if (.not. (abs(mass1 / mass0 - 1.0d0) < 1.0d-13)) stop 3
  if (.not. (maxval(abs(out)) < 1.0d10)) stop 4
  write(*, '(a)') 'PASS'

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(*), intent(in) :: what
    if (code /= 0) then
      write(*, '(a, a, i0)') what, ' failed: ', code
      stop 1
    end if
  end subroutine check

end program phlx_demo
