!----------------------------------------------------------------------
! sdpd_c.f90 -- iso_c_binding interface to the MI355X SDPD step
! (include/sdpd.h).  A Fortran host that keeps MCF's marching structure
! calls these in place of the Particles / Technique procedures it used:
!
!   technique_build_list            -> sdpd_build_cells
!   particles_compute_density       -> sdpd_density  (EOS fused)
!   particles_compute_pressure      -> sdpd_eos
!   particles_compute_interaction   -> sdpd_forces   (body force, act fused)
!   particles_integrate_velocity    -> sdpd_kick
!   particles_integrate_position +
!   particles_adjust_particles      -> sdpd_drift_wrap
!   particles_integrate_ct          -> sdpd_conformation
!   statistic_compute_statistic     -> sdpd_stats
!
! Every function returns 0 on success and a negative code otherwise,
! like the reference's stat_info.
!----------------------------------------------------------------------
MODULE sdpd_c
  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE

  INTEGER(c_int), PARAMETER :: SDPD_OK = 0
  INTEGER(c_int), PARAMETER :: SDPD_ERR_ARG = -1
  INTEGER(c_int), PARAMETER :: SDPD_ERR_STATE = -2
  INTEGER(c_int), PARAMETER :: SDPD_ERR_DEVICE = -3
  INTEGER(c_int), PARAMETER :: SDPD_ERR_PHYSICS = -4
  INTEGER(c_int32_t), PARAMETER :: SDPD_PRECISION_FP64 = 0
  INTEGER(c_int32_t), PARAMETER :: SDPD_PRECISION_MIXED = 1

  ! Layout of struct sdpd_params: ten int32 fields, then doubles, then
  ! the 64-bit noise key.
  TYPE, BIND(C) :: sdpd_params
     INTEGER(c_int32_t) :: num_dim
     INTEGER(c_int32_t) :: symmetry
     INTEGER(c_int32_t) :: kernel_type
     INTEGER(c_int32_t) :: rhs_density_type
     INTEGER(c_int32_t) :: rhs_force_type
     INTEGER(c_int32_t) :: state_equation_type
     INTEGER(c_int32_t) :: newtonian
     INTEGER(c_int32_t) :: brownian
     INTEGER(c_int32_t) :: body_force_type
     INTEGER(c_int32_t) :: precision
     REAL(c_double)     :: min_phys(3)
     REAL(c_double)     :: max_phys(3)
     REAL(c_double)     :: cut_off
     REAL(c_double)     :: c, rho, rho_ref, gamma
     REAL(c_double)     :: eta, eta_coef, kt, dt
     REAL(c_double)     :: tau, n_p, kt_p
     REAL(c_double)     :: body_force(3)
     INTEGER(c_int64_t) :: seed
  END TYPE sdpd_params

  ! struct sdpd_tuning (launch shapes of the pair kernels)
  TYPE, BIND(C) :: sdpd_tuning
     INTEGER(c_int32_t) :: density_shape
     INTEGER(c_int32_t) :: chunk_order
  END TYPE sdpd_tuning

  ! struct sdpd_colloid_params (coll_* keys of physics_config.mcf)
  TYPE, BIND(C) :: sdpd_colloid_params
     INTEGER(c_int32_t) :: noslip
     INTEGER(c_int32_t) :: rho_type
     INTEGER(c_int32_t) :: translate
     INTEGER(c_int32_t) :: rotate
     INTEGER(c_int32_t) :: integrate_ab
     INTEGER(c_int32_t) :: sub_steps
     INTEGER(c_int32_t) :: body_force_type
     INTEGER(c_int32_t) :: pad
     REAL(c_double)     :: body_force(3)
     REAL(c_double)     :: dout
  END TYPE sdpd_colloid_params

  INTERFACE
     INTEGER(c_int) FUNCTION sdpd_create(params, ctx) BIND(C, name="sdpd_create")
       IMPORT :: c_int, c_ptr, sdpd_params
       TYPE(sdpd_params), INTENT(IN) :: params
       TYPE(c_ptr), INTENT(OUT)      :: ctx
     END FUNCTION sdpd_create

     INTEGER(c_int) FUNCTION sdpd_upload(ctx, x, v, m, id, n) BIND(C, name="sdpd_upload")
       IMPORT :: c_int, c_ptr, c_double, c_int32_t
       TYPE(c_ptr), VALUE               :: ctx
       REAL(c_double), INTENT(IN)       :: x(*), v(*), m(*)
       INTEGER(c_int32_t), INTENT(IN)   :: id(*)
       INTEGER(c_int32_t), VALUE        :: n
     END FUNCTION sdpd_upload

     INTEGER(c_int) FUNCTION sdpd_set_replicas(ctx, nrep, box_lo, box_hi) &
          BIND(C, name="sdpd_set_replicas")
       IMPORT :: c_int, c_ptr, c_double, c_int32_t
       TYPE(c_ptr), VALUE         :: ctx
       INTEGER(c_int32_t), VALUE  :: nrep
       REAL(c_double), INTENT(IN) :: box_lo(3, *), box_hi(3, *)
     END FUNCTION sdpd_set_replicas

     INTEGER(c_int) FUNCTION sdpd_upload_replicas(ctx, x, v, m, id, rep, n) &
          BIND(C, name="sdpd_upload_replicas")
       IMPORT :: c_int, c_ptr, c_double, c_int32_t
       TYPE(c_ptr), VALUE               :: ctx
       REAL(c_double), INTENT(IN)       :: x(*), v(*), m(*)
       INTEGER(c_int32_t), INTENT(IN)   :: id(*), rep(*)
       INTEGER(c_int32_t), VALUE        :: n
     END FUNCTION sdpd_upload_replicas

     INTEGER(c_int) FUNCTION sdpd_stats_replicas(ctx, mom, ekin) &
          BIND(C, name="sdpd_stats_replicas")
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE          :: ctx
       REAL(c_double), INTENT(OUT) :: mom(3, *)
       REAL(c_double), INTENT(OUT) :: ekin(*)
     END FUNCTION sdpd_stats_replicas

     INTEGER(c_int) FUNCTION sdpd_set_lees_edwards(ctx, le_dim, v_lower, v_upper) &
          BIND(C, name="sdpd_set_lees_edwards")
       IMPORT :: c_int, c_ptr, c_double, c_int32_t
       TYPE(c_ptr), VALUE         :: ctx
       INTEGER(c_int32_t), VALUE  :: le_dim
       REAL(c_double), INTENT(IN) :: v_lower(3), v_upper(3)
     END FUNCTION sdpd_set_lees_edwards

     INTEGER(c_int) FUNCTION sdpd_set_colloids(ctx, cprm, num, radius, m, mmi, x, v, omega) &
          BIND(C, name="sdpd_set_colloids")
       IMPORT :: c_int, c_ptr, c_double, c_int32_t, sdpd_colloid_params
       TYPE(c_ptr), VALUE                    :: ctx
       TYPE(sdpd_colloid_params), INTENT(IN) :: cprm
       INTEGER(c_int32_t), VALUE             :: num
       REAL(c_double), INTENT(IN)            :: radius(*), m(*)
       REAL(c_double), INTENT(IN)            :: mmi(3, *), x(3, *), v(3, *), omega(3, *)
     END FUNCTION sdpd_set_colloids

     INTEGER(c_int) FUNCTION sdpd_get_colloids(ctx, x, v, omega, drag, torque) &
          BIND(C, name="sdpd_get_colloids")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
       TYPE(c_ptr), VALUE :: x, v, omega, drag, torque
     END FUNCTION sdpd_get_colloids

     ! device-path tuning (no counterpart in the reference; results do not
     ! depend on it): stored neighbour entries per particle before upload,
     ! launch shapes, the autotuner, graph replay of sdpd_step
     INTEGER(c_int) FUNCTION sdpd_set_list_capacity(ctx, cap) &
          BIND(C, name="sdpd_set_list_capacity")
       IMPORT :: c_int, c_ptr, c_int32_t
       TYPE(c_ptr), VALUE        :: ctx
       INTEGER(c_int32_t), VALUE :: cap
     END FUNCTION sdpd_set_list_capacity

     INTEGER(c_int) FUNCTION sdpd_set_tuning(ctx, t) BIND(C, name="sdpd_set_tuning")
       IMPORT :: c_int, c_ptr, sdpd_tuning
       TYPE(c_ptr), VALUE            :: ctx
       TYPE(sdpd_tuning), INTENT(IN) :: t
     END FUNCTION sdpd_set_tuning

     INTEGER(c_int) FUNCTION sdpd_autotune(ctx, reps, best, ms_best) &
          BIND(C, name="sdpd_autotune")
       IMPORT :: c_int, c_ptr, c_int32_t, c_double, sdpd_tuning
       TYPE(c_ptr), VALUE             :: ctx
       INTEGER(c_int32_t), VALUE      :: reps
       TYPE(sdpd_tuning), INTENT(OUT) :: best
       REAL(c_double), INTENT(OUT)    :: ms_best
     END FUNCTION sdpd_autotune

     INTEGER(c_int) FUNCTION sdpd_set_graphs(ctx, enable) BIND(C, name="sdpd_set_graphs")
       IMPORT :: c_int, c_ptr, c_int32_t
       TYPE(c_ptr), VALUE        :: ctx
       INTEGER(c_int32_t), VALUE :: enable
     END FUNCTION sdpd_set_graphs

     INTEGER(c_int) FUNCTION sdpd_upload_ct(ctx, ct) BIND(C, name="sdpd_upload_ct")
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE         :: ctx
       REAL(c_double), INTENT(IN) :: ct(*)
     END FUNCTION sdpd_upload_ct

     INTEGER(c_int) FUNCTION sdpd_build_cells(ctx) BIND(C, name="sdpd_build_cells")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION sdpd_build_cells

     INTEGER(c_int) FUNCTION sdpd_density(ctx) BIND(C, name="sdpd_density")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION sdpd_density

     INTEGER(c_int) FUNCTION sdpd_eos(ctx) BIND(C, name="sdpd_eos")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION sdpd_eos

     INTEGER(c_int) FUNCTION sdpd_forces(ctx) BIND(C, name="sdpd_forces")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION sdpd_forces

     INTEGER(c_int) FUNCTION sdpd_kick(ctx, lambda) BIND(C, name="sdpd_kick")
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE    :: ctx
       REAL(c_double), VALUE :: lambda
     END FUNCTION sdpd_kick

     INTEGER(c_int) FUNCTION sdpd_drift_wrap(ctx) BIND(C, name="sdpd_drift_wrap")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION sdpd_drift_wrap

     INTEGER(c_int) FUNCTION sdpd_conformation(ctx, lambda) BIND(C, name="sdpd_conformation")
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE    :: ctx
       REAL(c_double), VALUE :: lambda
     END FUNCTION sdpd_conformation

     INTEGER(c_int) FUNCTION sdpd_stats(ctx, mom, ekin) BIND(C, name="sdpd_stats")
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE          :: ctx
       REAL(c_double), INTENT(OUT) :: mom(3)
       REAL(c_double), INTENT(OUT) :: ekin
     END FUNCTION sdpd_stats

     INTEGER(c_int) FUNCTION sdpd_download(ctx, x, v, rho, p, f) BIND(C, name="sdpd_download")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
       TYPE(c_ptr), VALUE :: x, v, rho, p, f
     END FUNCTION sdpd_download

     INTEGER(c_int) FUNCTION sdpd_download_tensors(ctx, ct, vgt) BIND(C, name="sdpd_download_tensors")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
       TYPE(c_ptr), VALUE :: ct, vgt
     END FUNCTION sdpd_download_tensors

     INTEGER(c_int) FUNCTION sdpd_step(ctx, nsteps) BIND(C, name="sdpd_step")
       IMPORT :: c_int, c_ptr, c_int32_t
       TYPE(c_ptr), VALUE        :: ctx
       INTEGER(c_int32_t), VALUE :: nsteps
     END FUNCTION sdpd_step

     INTEGER(c_int) FUNCTION sdpd_set_step(ctx, step) BIND(C, name="sdpd_set_step")
       IMPORT :: c_int, c_ptr, c_int64_t
       TYPE(c_ptr), VALUE        :: ctx
       INTEGER(c_int64_t), VALUE :: step
     END FUNCTION sdpd_set_step

     INTEGER(c_int) FUNCTION sdpd_get_step(ctx, step) BIND(C, name="sdpd_get_step")
       IMPORT :: c_int, c_ptr, c_int64_t
       TYPE(c_ptr), VALUE              :: ctx
       INTEGER(c_int64_t), INTENT(OUT) :: step
     END FUNCTION sdpd_get_step

     INTEGER(c_int) FUNCTION sdpd_sync(ctx) BIND(C, name="sdpd_sync")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION sdpd_sync

     TYPE(c_ptr) FUNCTION sdpd_last_error_c(ctx) BIND(C, name="sdpd_last_error")
       IMPORT :: c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION sdpd_last_error_c

     INTEGER(c_int) FUNCTION sdpd_destroy(ctx) BIND(C, name="sdpd_destroy")
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION sdpd_destroy

     FUNCTION c_strlen(s) BIND(C, name="strlen")
       IMPORT :: c_ptr, c_size_t
       TYPE(c_ptr), VALUE :: s
       INTEGER(c_size_t)  :: c_strlen
     END FUNCTION c_strlen
  END INTERFACE

CONTAINS

  ! Message of the last failed call on ctx as a Fortran string.
  FUNCTION sdpd_last_error(ctx) RESULT(msg)
    TYPE(c_ptr), INTENT(IN) :: ctx
    CHARACTER(len=:), ALLOCATABLE :: msg
    TYPE(c_ptr) :: p
    CHARACTER(kind=c_char), POINTER :: chars(:)
    INTEGER :: k, n
    p = sdpd_last_error_c(ctx)
    n = INT(c_strlen(p))
    CALL c_f_pointer(p, chars, [n])
    ALLOCATE(CHARACTER(len=n) :: msg)
    DO k = 1, n
       msg(k:k) = chars(k)
    END DO
  END FUNCTION sdpd_last_error

END MODULE sdpd_c
