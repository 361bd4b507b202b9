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
  integer(c_int), parameter :: PHLX_EOS_IDEAL = 0, PHLX_EOS_HELMHOLTZ = 1
  integer(c_int), parameter :: PHLX_PREC_FP64 = 0, PHLX_PREC_MIXED = 1
  integer(c_int), parameter :: PHLX_GUESS_WARM = 0, PHLX_GUESS_MLP = 1, PHLX_GUESS_BRACKET = 2

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
    integer(c_int) function phlx_download_state(ctx, prim, eint, temp, b1, b2, b3, bcc) &
        bind(c, name="phlx_download_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, prim, eint, temp, b1, b2, b3, bcc
    end function
    ! restart: upload the restart file's prim, temp, faces, then resume with
    ! its time, dt, smax, step (read_restart + time_loop's step /= 0 branch)
    integer(c_int) function phlx_resume(ctx, time, dt, smax, step) bind(c, name="phlx_resume")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), value :: time, dt, smax
      integer(c_int64_t), value :: step
    end function
    ! ---- Helmholtz EoS (load_helm_table and the per-zone / batched calls)
    ! f9 = helm_table_var(9, nrho, nT), c4 = dpdrho_table_var(4, nrho, nT)
    integer(c_int) function phlx_load_helm_table(ctx, f9, c4, nrho, nt, ldlo, ldhi, ltlo, lthi) &
        bind(c, name="phlx_load_helm_table")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx, f9, c4
      integer(c_int32_t), value :: nrho, nt
      real(c_double), value :: ldlo, ldhi, ltlo, lthi
    end function
    integer(c_int) function phlx_load_pig_table(ctx, f9, c4, nrho, nt, ldlo, ldhi, ltlo, lthi) &
        bind(c, name="phlx_load_pig_table")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx, f9, c4
      integer(c_int32_t), value :: nrho, nt
      real(c_double), value :: ldlo, ldhi, ltlo, lthi
    end function
    integer(c_int) function phlx_set_eos_profiles(ctx, eq_gammae_cc, eq_gammae_f) &
        bind(c, name="phlx_set_eos_profiles")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, eq_gammae_cc, eq_gammae_f
    end function
    integer(c_int) function phlx_download_gammaf(ctx, gammaf) bind(c, name="phlx_download_gammaf")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, gammaf
    end function
    ! batched helm_rhoT_given_full / helm_rhoe_given / helm_rhoP_given on
    ! device arrays of n points
    integer(c_int) function phlx_helm_rhoT_full_batch(ctx, n, rho, t, abar, zbar, p, e, sound, cv) &
        bind(c, name="phlx_helm_rhoT_full_batch")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, rho, t, abar, zbar, p, e, sound, cv
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function phlx_helm_rhoe_given_batch(ctx, n, rho, e, abar, zbar, t, p, sound, &
        return_sound, iters) bind(c, name="phlx_helm_rhoe_given_batch")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx, rho, e, abar, zbar, t, p, sound, iters
      integer(c_int64_t), value :: n
      integer(c_int32_t), value :: return_sound
    end function
    integer(c_int) function phlx_helm_rhoP_given_batch(ctx, n, rho, p, abar, zbar, t, e, sound, &
        return_sound, iters) bind(c, name="phlx_helm_rhoP_given_batch")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx, rho, p, abar, zbar, t, e, sound, iters
      integer(c_int64_t), value :: n
      integer(c_int32_t), value :: return_sound
    end function
    integer(c_int) function phlx_helm_rhoe_given_guess_batch(ctx, n, rho, e, abar, zbar, t, p, &
        sound, return_sound, guess_mode, iters, fallback) &
        bind(c, name="phlx_helm_rhoe_given_guess_batch")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx, rho, e, abar, zbar, t, p, sound, iters, fallback
      integer(c_int64_t), value :: n
      integer(c_int32_t), value :: return_sound, guess_mode
    end function
    ! ---- T(rho, e) start inside the update, surrogate, mixed-precision EoS step
    integer(c_int) function phlx_set_eos_guess(ctx, mode, resid_tol, count_iterations) &
        bind(c, name="phlx_set_eos_guess")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode, count_iterations
      real(c_double), value :: resid_tol
    end function
    integer(c_int) function phlx_set_eos_lin_tol(ctx, lin_tol) bind(c, name="phlx_set_eos_lin_tol")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: lin_tol
    end function
    integer(c_int) function phlx_eos_stats(ctx, newton_iters, solves, fallbacks, reset) &
        bind(c, name="phlx_eos_stats")
      import :: c_int, c_ptr, c_int64_t, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: newton_iters, solves, fallbacks
      integer(c_int32_t), value :: reset
    end function
    ! THERMAL_DIFFUSION_STS; fix_t(6) wall temperatures (NaN: zero gradient)
    integer(c_int) function phlx_set_thermal_diffusion(ctx, enable, evaluate_dt, fix_t) &
        bind(c, name="phlx_set_thermal_diffusion")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable, evaluate_dt
      real(c_double), intent(in) :: fix_t(6)
    end function
    ! kappa(nx1+2, nx2+2, nx3+2): the opacity with one ghost layer
    integer(c_int) function phlx_set_kappa(ctx, kappa) bind(c, name="phlx_set_kappa")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: kappa(*)
    end function
    integer(c_int) function phlx_parabolic_dt(ctx, dt_par) bind(c, name="phlx_parabolic_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt_par
    end function
    integer(c_int) function phlx_diffusion_step(ctx, strang) bind(c, name="phlx_diffusion_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: strang
    end function
    integer(c_int) function phlx_diffusion_stages(ctx, stages) &
        bind(c, name="phlx_diffusion_stages")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: stages
    end function
    ! USE_SHOCK_FLATTENING with eps_sf_make (FP64 contexts)
    integer(c_int) function phlx_set_shock_flattening(ctx, enable, eps_sf) &
        bind(c, name="phlx_set_shock_flattening")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
      real(c_double), value :: eps_sf
    end function
    ! hydro_step's fused one-pass sweep: 1 wherever admitted, 0 never, -1 default
    integer(c_int) function phlx_set_fused_sweep(ctx, mode) &
        bind(c, name="phlx_set_fused_sweep")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode
    end function
    ! phlx_run replays a captured HIP graph per step: 1 on, 0 off (default)
    integer(c_int) function phlx_set_step_graph(ctx, on) bind(c, name="phlx_set_step_graph")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    ! box(8): log10(rho*Ye) lo, hi, log10 T lo, hi, abar lo, hi, Ye lo, hi
    integer(c_int) function phlx_mlp_set_train_box(ctx, box) bind(c, name="phlx_mlp_set_train_box")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, box
    end function
    integer(c_int) function phlx_mlp_train(ctx, nsamples, epochs, batch, lr, seed, train_loss, &
        val_rms_logt, val_rms_loge) bind(c, name="phlx_mlp_train")
      import :: c_int, c_ptr, c_int64_t, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nsamples, seed
      integer(c_int32_t), value :: epochs, batch
      real(c_double), value :: lr
      real(c_double), intent(out) :: train_loss, val_rms_logt, val_rms_loge
    end function
    ! compute_hyperbolic_dt on caller-owned device arrays, allocating nothing
    ! (callers under graph capture or that forbid hipMalloc); dt_smax(2) on
    ! the device; stream a hipStream_t (c_null_ptr: the null stream)
    integer(c_int) function phlx_compute_dt_noalloc(cfg, prim, bcc, dt_smax, stream) &
        bind(c, name="phlx_compute_dt_noalloc")
      import :: c_int, c_ptr, phlx_config
      type(phlx_config), intent(in) :: cfg
      type(c_ptr), value :: prim, bcc, dt_smax, stream
    end function
  end interface
end module phlx_iface
