! phlx_iface.F90 -- ISO_C_BINDING interface to libphlx (include/phlx.h).
!
! A Phlegethon build adds this module and replaces, inside time_loop, the
! calls to hydro_step / compute_hyperbolic_dt by phlx_hydro_step /
! phlx_compute_dt on a context created once after the initial conditions
! are set.  Arrays are passed as C pointers (c_loc) with their sizes implied
! by the configuration; scalars by value.
module phlx_iface
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: PHLX_LIM2ND_REC = 0, PHLX_PPH_REC = 1, PHLX_LIM5TH_REC = 2
  integer(c_int), parameter :: PHLX_HLLC_FLUX = 0, PHLX_LHLLC_FLUX = 1, &
                               PHLX_HLLD_FLUX = 2, PHLX_LHLLD_FLUX = 3
  integer(c_int), parameter :: PHLX_BC_PERIODIC = 0, PHLX_BC_REFLECTIVE = 1, &
                               PHLX_BC_OUTFLOW = 2, PHLX_BC_DIODE = 3

  type, bind(c) :: phlx_config
    integer(c_int32_t) :: sdims, nx1, nx2, nx3, nas
    real(c_double) :: x1l, x1u, x2l, x2u, x3l, x3u
    real(c_double) :: gamma, mu
    integer(c_int32_t) :: recon, flux, all_mach, rk_stages
    real(c_double) :: cfl, mach_cutoff, mach_ct, tmax
    integer(c_int32_t) :: bc(6)
    integer(c_int32_t) :: use_wb, use_gravity, evolve_etot, use_edot
    integer(c_int32_t) :: eos, fasteos, precision, device
    integer(c_int32_t) :: reserved(8)
  end type phlx_config

  interface
    integer(c_int) function phlx_create(cfg, ctx) bind(c, name="phlx_create")
      import :: c_int, c_ptr, phlx_config
      type(phlx_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function phlx_destroy(ctx) bind(c, name="phlx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! prim(1:nvars, lx1-ngc:ux1+ngc, lx2-ngc:ux2+ngc, lx3-ngc:ux3+ngc) as declared
    ! in type(locgrid) (source/source.F90:1219-1227)
    integer(c_int) function phlx_upload_prim_aos(ctx, prim, ngc) bind(c, name="phlx_upload_prim_aos")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, prim
      integer(c_int32_t), value :: ngc
    end function
    integer(c_int) function phlx_download_prim_aos(ctx, prim, ngc) bind(c, name="phlx_download_prim_aos")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, prim
      integer(c_int32_t), value :: ngc
    end function
    integer(c_int) function phlx_upload_state(ctx, prim, eint, temp, b1, b2, b3) &
        bind(c, name="phlx_upload_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, prim, eint, temp, b1, b2, b3
    end function
    integer(c_int) function phlx_set_profiles(ctx, eq_rho_cc, eq_p_cc, eq_rho_f, eq_p_f, &
        grav_cc, phi_cc, phi_f, edot_cc) bind(c, name="phlx_set_profiles")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, eq_rho_cc, eq_p_cc, eq_rho_f, eq_p_f, grav_cc, phi_cc, &
                            phi_f, edot_cc
    end function
    integer(c_int) function phlx_init(ctx) bind(c, name="phlx_init")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function phlx_hydro_step(ctx) bind(c, name="phlx_hydro_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function phlx_compute_dt(ctx, dt, smax) bind(c, name="phlx_compute_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt, smax
    end function
    integer(c_int) function phlx_run(ctx, nsteps) bind(c, name="phlx_run")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function phlx_get_scalars(ctx, time, dt, smax, step) bind(c, name="phlx_get_scalars")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: time, dt, smax
      integer(c_int64_t), intent(out) :: step
    end function
    integer(c_int) function phlx_set_scalars(ctx, time, dt, smax, step) bind(c, name="phlx_set_scalars")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), value :: time, dt, smax
      integer(c_int64_t), value :: step
    end function
    integer(c_int) function phlx_divb_max(ctx, divb) bind(c, name="phlx_divb_max")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: divb
    end function
  end interface
end module phlx_iface
