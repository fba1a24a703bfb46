!----------------------------------------------------------------------
! mcf_sdpd.f90 -- Fortran host of the MI355X SDPD step.
!
! Usage:  mcf_sdpd [ctrl_file]          (default ctrl.mcf)
!
! Reads MCF's control file and the physics / io files it names
! (key = value lines, '#' comments; io_read_ctrl.F90:157-357,
! io_read_physics.F90), sets up the particles (a square / cubic lattice
! at (i - 1/2) dx as particles_init_global_inter_square.F90:133-154, or
! the read_external file of io_read_particles.F90:162-204) and marches
! velocity-Verlet steps with the call order of marching_integrate_VV
! (marching_integrate_VV.F90:303-1619), every particle operation going
! through the C-ABI of include/sdpd.h:
!
!   prologue (marching_marching.F90:528-1123):
!     build cells, density + EOS, forces (+ body force, act)
!   each step:
!     kick 1/2, drift + wrap, [ct 1/2], build cells, density + EOS,
!     forces (+ body force, act), [ct 1/2], kick 1/2, statistics
!
! Keys beyond the reference's: precision = FP64 | mixed, fused_step = T
! (whole steps through sdpd_step between statistics), lattice_jitter
! (uniform jitter of the lattice in units of dx, seeded by random_seed).
!----------------------------------------------------------------------
MODULE mcf_sdpd_config
  IMPLICIT NONE
  INTEGER, PARAMETER :: MAX_KEYS = 512, KEY_LEN = 64, VAL_LEN = 256
  TYPE config
     INTEGER :: n = 0
     CHARACTER(len=KEY_LEN) :: key(MAX_KEYS)
     CHARACTER(len=VAL_LEN) :: val(MAX_KEYS)
  END TYPE config

CONTAINS

  FUNCTION lower(s) RESULT(t)
    CHARACTER(len=*), INTENT(IN) :: s
    CHARACTER(len=LEN(s)) :: t
    INTEGER :: k, c
    t = s
    DO k = 1, LEN(s)
       c = IACHAR(s(k:k))
       IF (c >= 65 .AND. c <= 90) t(k:k) = ACHAR(c + 32)
    END DO
  END FUNCTION lower

  ! Appends the key = value lines of a file; returns stat /= 0 when the
  ! file cannot be opened (io_read_ctrl.F90 aborts the same way).
  SUBROUTINE config_read(cfg, path, stat)
    TYPE(config), INTENT(INOUT) :: cfg
    CHARACTER(len=*), INTENT(IN) :: path
    INTEGER, INTENT(OUT) :: stat
    CHARACTER(len=1024) :: line
    INTEGER :: u, ios, eq, hash
    OPEN(newunit=u, file=TRIM(path), status='old', action='read', iostat=stat)
    IF (stat /= 0) RETURN
    DO
       READ(u, '(A)', iostat=ios) line
       IF (ios /= 0) EXIT
       hash = INDEX(line, '#')
       IF (hash > 0) line(hash:) = ' '
       eq = INDEX(line, '=')
       IF (eq <= 1) CYCLE
       IF (cfg%n >= MAX_KEYS) EXIT
       cfg%n = cfg%n + 1
       cfg%key(cfg%n) = lower(ADJUSTL(line(1:eq-1)))
       cfg%val(cfg%n) = ADJUSTL(line(eq+1:))
    END DO
    CLOSE(u)
    stat = 0
  END SUBROUTINE config_read

  ! Index of a key (the last occurrence wins), 0 when absent.
  INTEGER FUNCTION config_find(cfg, key)
    TYPE(config), INTENT(IN) :: cfg
    CHARACTER(len=*), INTENT(IN) :: key
    INTEGER :: k
    config_find = 0
    DO k = 1, cfg%n
       IF (TRIM(cfg%key(k)) == lower(key)) config_find = k
    END DO
  END FUNCTION config_find

  ! Index of the n-th occurrence of a key (keys the reference reads once
  ! per colloid, io_read_physics.F90:1242-1350), 0 when absent.
  INTEGER FUNCTION config_find_nth(cfg, key, nth)
    TYPE(config), INTENT(IN) :: cfg
    CHARACTER(len=*), INTENT(IN) :: key
    INTEGER, INTENT(IN) :: nth
    INTEGER :: k, seen
    config_find_nth = 0
    seen = 0
    DO k = 1, cfg%n
       IF (TRIM(cfg%key(k)) == lower(key)) THEN
          seen = seen + 1
          IF (seen == nth) THEN
             config_find_nth = k
             RETURN
          END IF
       END IF
    END DO
  END FUNCTION config_find_nth

  SUBROUTINE get_reals_nth(cfg, key, nth, n, vals)
    TYPE(config), INTENT(IN) :: cfg
    CHARACTER(len=*), INTENT(IN) :: key
    INTEGER, INTENT(IN) :: nth, n
    DOUBLE PRECISION, INTENT(INOUT) :: vals(n)
    DOUBLE PRECISION :: tmp(n)
    INTEGER :: k, ios
    k = config_find_nth(cfg, key, nth)
    IF (k > 0) THEN
       tmp = vals
       READ(cfg%val(k), *, iostat=ios) tmp
       IF (ios == 0) vals = tmp
    END IF
  END SUBROUTINE get_reals_nth

  INTEGER FUNCTION get_int(cfg, key, default)
    TYPE(config), INTENT(IN) :: cfg
    CHARACTER(len=*), INTENT(IN) :: key
    INTEGER, INTENT(IN) :: default
    INTEGER :: k, ios
    get_int = default
    k = config_find(cfg, key)
    IF (k > 0) THEN
       READ(cfg%val(k), *, iostat=ios) get_int
       IF (ios /= 0) get_int = default
    END IF
  END FUNCTION get_int

  DOUBLE PRECISION FUNCTION get_real(cfg, key, default)
    TYPE(config), INTENT(IN) :: cfg
    CHARACTER(len=*), INTENT(IN) :: key
    DOUBLE PRECISION, INTENT(IN) :: default
    INTEGER :: k, ios
    get_real = default
    k = config_find(cfg, key)
    IF (k > 0) THEN
       READ(cfg%val(k), *, iostat=ios) get_real
       IF (ios /= 0) get_real = default
    END IF
  END FUNCTION get_real

  SUBROUTINE get_reals(cfg, key, n, vals)
    TYPE(config), INTENT(IN) :: cfg
    CHARACTER(len=*), INTENT(IN) :: key
    INTEGER, INTENT(IN) :: n
    DOUBLE PRECISION, INTENT(INOUT) :: vals(n)
    INTEGER :: k, ios
    DOUBLE PRECISION :: tmp(n)
    k = config_find(cfg, key)
    IF (k > 0) THEN
       tmp = vals
       READ(cfg%val(k), *, iostat=ios) tmp
       IF (ios == 0) vals = tmp
    END IF
  END SUBROUTINE get_reals

  SUBROUTINE get_ints(cfg, key, n, vals)
    TYPE(config), INTENT(IN) :: cfg
    CHARACTER(len=*), INTENT(IN) :: key
    INTEGER, INTENT(IN) :: n
    INTEGER, INTENT(INOUT) :: vals(n)
    INTEGER :: k, ios
    INTEGER :: tmp(n)
    k = config_find(cfg, key)
    IF (k > 0) THEN
       tmp = vals
       READ(cfg%val(k), *, iostat=ios) tmp
       IF (ios == 0) vals = tmp
    END IF
  END SUBROUTINE get_ints

  SUBROUTINE get_ints_nth(cfg, key, nth, n, vals)
    TYPE(config), INTENT(IN) :: cfg
    CHARACTER(len=*), INTENT(IN) :: key
    INTEGER, INTENT(IN) :: nth, n
    INTEGER, INTENT(INOUT) :: vals(n)
    INTEGER :: tmp(n)
    INTEGER :: k, ios
    k = config_find_nth(cfg, key, nth)
    IF (k > 0) THEN
       tmp = vals
       READ(cfg%val(k), *, iostat=ios) tmp
       IF (ios == 0) vals = tmp
    END IF
  END SUBROUTINE get_ints_nth

  ! key = value lines of a restart physics file, each followed by a blank
  ! line: reals E16.8 and integers Iw, vectors joined by commas (the
  ! (n(A,E16.8)) records of io_write_restart_physics.F90), logicals L2
  SUBROUTINE put_reals(u, key, vals)
    INTEGER, INTENT(IN) :: u
    CHARACTER(len=*), INTENT(IN) :: key
    DOUBLE PRECISION, INTENT(IN) :: vals(:)
    CHARACTER(len=1024) :: buf
    INTEGER :: k, p
    buf = key//' = '
    p = LEN(key) + 4
    DO k = 1, SIZE(vals)
       IF (k > 1) THEN
          buf(p:p) = ','
          p = p + 1
       END IF
       WRITE(buf(p:p + 15), '(E16.8)') vals(k)
       p = p + 16
    END DO
    WRITE(u, '(A)') buf(1:p - 1)
    WRITE(u, '(A)') ' '
  END SUBROUTINE put_reals

  SUBROUTINE put_real(u, key, val)
    INTEGER, INTENT(IN) :: u
    CHARACTER(len=*), INTENT(IN) :: key
    DOUBLE PRECISION, INTENT(IN) :: val
    CALL put_reals(u, key, [val])
  END SUBROUTINE put_real

  SUBROUTINE put_ints(u, key, vals, w)
    INTEGER, INTENT(IN) :: u, w
    CHARACTER(len=*), INTENT(IN) :: key
    INTEGER, INTENT(IN) :: vals(:)
    CHARACTER(len=1024) :: buf
    CHARACTER(len=16) :: form
    INTEGER :: k, p
    WRITE(form, '(A,I0,A)') '(I', w, ')'
    buf = key//' = '
    p = LEN(key) + 4
    DO k = 1, SIZE(vals)
       IF (k > 1) THEN
          buf(p:p) = ','
          p = p + 1
       END IF
       WRITE(buf(p:p + w - 1), form) vals(k)
       p = p + w
    END DO
    WRITE(u, '(A)') buf(1:p - 1)
    WRITE(u, '(A)') ' '
  END SUBROUTINE put_ints

  SUBROUTINE put_int(u, key, val, w)
    INTEGER, INTENT(IN) :: u, val, w
    CHARACTER(len=*), INTENT(IN) :: key
    CALL put_ints(u, key, [val], w)
  END SUBROUTINE put_int

  SUBROUTINE put_logical(u, key, val)
    INTEGER, INTENT(IN) :: u
    CHARACTER(len=*), INTENT(IN) :: key
    LOGICAL, INTENT(IN) :: val
    WRITE(u, '(A,L2)') key//' = ', val
    WRITE(u, '(A)') ' '
  END SUBROUTINE put_logical

  LOGICAL FUNCTION get_logical(cfg, key, default)
    TYPE(config), INTENT(IN) :: cfg
    CHARACTER(len=*), INTENT(IN) :: key
    LOGICAL, INTENT(IN) :: default
    CHARACTER(len=VAL_LEN) :: v
    INTEGER :: k
    get_logical = default
    k = config_find(cfg, key)
    IF (k > 0) THEN
       v = lower(ADJUSTL(cfg%val(k)))
       IF (v(1:1) == '.') v = v(2:)
       IF (v(1:1) == 't') get_logical = .TRUE.
       IF (v(1:1) == 'f') get_logical = .FALSE.
    END IF
  END FUNCTION get_logical

  FUNCTION get_string(cfg, key, default) RESULT(s)
    TYPE(config), INTENT(IN) :: cfg
    CHARACTER(len=*), INTENT(IN) :: key, default
    CHARACTER(len=VAL_LEN) :: s
    INTEGER :: k
    s = default
    k = config_find(cfg, key)
    IF (k > 0) s = cfg%val(k)
  END FUNCTION get_string

END MODULE mcf_sdpd_config


PROGRAM mcf_sdpd
  USE, INTRINSIC :: iso_c_binding
  USE sdpd_c
  USE mcf_sdpd_config
  IMPLICIT NONE

  TYPE(config) :: cfg
  TYPE(sdpd_params), TARGET :: prm
  TYPE(c_ptr) :: ctx
  CHARACTER(len=VAL_LEN) :: ctrl_file, dir, physics_file, io_file, part_file, stat_file, out_file
  INTEGER :: stat, dim, n, k, d, nsteps, step, step_start, stat_freq, out_freq, su, rc
  INTEGER :: num_part(3), idx(3)
  INTEGER(c_int32_t) :: rc32
  LOGICAL :: read_external, fused, newtonian
  ! output files (io_config.mcf keys): statistic, particles, conformation,
  ! restart; FORMATTED (E16.8 records) or UNFORMATTED (binary records)
  CHARACTER(len=VAL_LEN) :: restart_file, conf_out_file, restart_conf_file, read_conf_file
  CHARACTER(len=VAL_LEN) :: restart_phys_file
  INTEGER :: restart_freq, conf_freq, next_out
  LOGICAL :: out_unf, restart_unf, conf_unf
  DOUBLE PRECISION :: time_cur
  DOUBLE PRECISION, ALLOCATABLE, TARGET :: ct(:,:)
  DOUBLE PRECISION :: dx(3), jitter, mom(3), ekin, t0, t1, r(3)
  DOUBLE PRECISION, ALLOCATABLE, TARGET :: x(:,:), v(:,:), m(:), rho(:)
  INTEGER(c_int32_t), ALLOCATABLE :: id(:,:)
  INTEGER :: seed_size
  INTEGER, ALLOCATABLE :: seed_arr(:)
  INTEGER(c_int64_t) :: cur_step
  INTEGER :: bcdef(6), le_dim, ncoll
  TYPE(sdpd_colloid_params) :: cprm
  DOUBLE PRECISION, ALLOCATABLE :: c_rad(:), c_m(:), c_mmi(:,:), c_x(:,:), c_v(:,:), c_w(:,:)
  DOUBLE PRECISION :: rr(3), dxc, cpi
  DOUBLE PRECISION :: shear_in(12)
  REAL(c_double) :: v_lo(3), v_hi(3)

  ctrl_file = 'ctrl.mcf'
  IF (COMMAND_ARGUMENT_COUNT() >= 1) CALL GET_COMMAND_ARGUMENT(1, ctrl_file)
  dir = ''
  k = INDEX(ctrl_file, '/', back=.TRUE.)
  IF (k > 0) dir = ctrl_file(1:k)

  CALL config_read(cfg, ctrl_file, stat)
  IF (stat /= 0) CALL die('cannot open control file '//TRIM(ctrl_file))
  physics_file = get_string(cfg, 'mcf_physics_config_file', 'physics_config.mcf')
  io_file = get_string(cfg, 'mcf_io_config_file', 'io_config.mcf')
  CALL config_read(cfg, TRIM(dir)//TRIM(physics_file), stat)
  IF (stat /= 0) CALL die('cannot open physics file '//TRIM(physics_file))
  CALL config_read(cfg, TRIM(dir)//TRIM(io_file), stat)
  IF (stat /= 0) CALL die('cannot open io file '//TRIM(io_file))

  IF (get_int(cfg, 'integrate_type', 2) /= 2) &
       CALL die('only integrate_type = 2 (velocity Verlet) is supported')

  ! ---- parameters (physics_adjust_parameters.F90 conventions) ----
  dim = get_int(cfg, 'num_dim', 3)
  prm%num_dim = dim
  prm%symmetry = MERGE(1, 0, get_logical(cfg, 'symmetry', .TRUE.))
  prm%kernel_type = get_int(cfg, 'kernel_type', 1)
  prm%rhs_density_type = get_int(cfg, 'rhs_density_type', 2)
  prm%rhs_force_type = get_int(cfg, 'rhs_force_type', 3)
  prm%state_equation_type = get_int(cfg, 'stateequation_type', 2)
  newtonian = get_logical(cfg, 'newtonian', .TRUE.)
  prm%newtonian = MERGE(1, 0, newtonian)
  prm%brownian = MERGE(1, 0, get_logical(cfg, 'brownian', .FALSE.))
  prm%body_force_type = get_int(cfg, 'body_force_type', 0)
  prm%precision = SDPD_PRECISION_FP64
  IF (lower(ADJUSTL(get_string(cfg, 'precision', 'fp64'))) == 'mixed') &
       prm%precision = SDPD_PRECISION_MIXED
  prm%min_phys = 0.0d0
  prm%max_phys = 1.0d0
  CALL get_reals(cfg, 'min_phys', dim, prm%min_phys(1:dim))
  CALL get_reals(cfg, 'max_phys', dim, prm%max_phys(1:dim))
  prm%cut_off = get_real(cfg, 'cut_off', 0.0d0)
  prm%c = get_real(cfg, 'c', 10.0d0)
  prm%rho = get_real(cfg, 'rho', 1.0d0)
  prm%rho_ref = get_real(cfg, 'rho_ref', 1.0d0)
  prm%gamma = get_real(cfg, 'gamma', 7.0d0)
  prm%eta = get_real(cfg, 'eta', 0.1d0)
  prm%eta_coef = get_real(cfg, 'eta_coef', 0.0d0)
  prm%kt = get_real(cfg, 'kt', 0.0d0)
  IF (prm%brownian == 0) prm%kt = 0.0d0   ! physics_adjust_parameters.F90:415-417
  prm%dt = get_real(cfg, 'dt', -1.0d0)
  prm%tau = get_real(cfg, 'tau', 0.0d0)
  prm%n_p = get_real(cfg, 'n_p', 0.0d0)
  prm%kt_p = get_real(cfg, 'kt_p', 0.0d0)
  prm%body_force = 0.0d0
  CALL get_reals(cfg, 'body_force', dim, prm%body_force(1:dim))
  prm%seed = INT(MAX(get_int(cfg, 'random_seed', 0), 0), c_int64_t)
  IF (prm%dt <= 0.0d0) CALL die('dt must be given (> 0)')

  step_start = MAX(get_int(cfg, 'step_start', 0), 0)
  nsteps = get_int(cfg, 'step_end', 0) - step_start
  IF (nsteps < 0) CALL die('step_end must not be below step_start')
  stat_freq = MAX(get_int(cfg, 'statistic_freq_step', 1), 1)
  out_freq = get_int(cfg, 'output_particles_freq_step', -1)
  stat_file = get_string(cfg, 'statistic_file', 'mcf_statistic.dat')
  out_file = get_string(cfg, 'output_particles_file', 'mcf_particles')
  out_unf = is_unformatted(get_string(cfg, 'output_particles_fmt', 'FORMATTED'))
  restart_file = get_string(cfg, 'restart_particles_file', 'mcf_restart_particles')
  restart_unf = is_unformatted(get_string(cfg, 'restart_particles_fmt', 'FORMATTED'))
  restart_freq = get_int(cfg, 'restart_freq_step', -1)
  conf_out_file = get_string(cfg, 'output_conformation_file', 'mcf_conformation')
  conf_unf = is_unformatted(get_string(cfg, 'output_conformation_fmt', 'FORMATTED'))
  conf_freq = get_int(cfg, 'output_conformation_freq_step', -1)
  restart_conf_file = get_string(cfg, 'restart_conformation_file', 'mcf_restart_conformation')
  restart_phys_file = get_string(cfg, 'restart_physics_file', 'mcf_restart_physics')
  read_conf_file = get_string(cfg, 'read_conformation_file', '')
  read_external = get_logical(cfg, 'read_external', .FALSE.)
  fused = get_logical(cfg, 'fused_step', .FALSE.)

  ! ---- particles ----
  IF (read_external) THEN
     part_file = get_string(cfg, 'read_particles_file', 'mcf_restart_particles.dat')
     CALL read_particles(TRIM(dir)//TRIM(part_file))
  ELSE
     num_part = 1
     CALL get_ints(cfg, 'num_part', dim, num_part(1:dim))
     jitter = get_real(cfg, 'lattice_jitter', 0.0d0)
     n = PRODUCT(num_part(1:dim))
     ALLOCATE(x(dim, n), v(dim, n), m(n), rho(n), id(2, n))
     dx = 1.0d0
     DO d = 1, dim
        dx(d) = (prm%max_phys(d) - prm%min_phys(d)) / num_part(d)
     END DO
     CALL RANDOM_SEED(size=seed_size)
     ALLOCATE(seed_arr(seed_size))
     seed_arr = 12345 + INT(prm%seed)
     CALL RANDOM_SEED(put=seed_arr)
     DO k = 1, n
        idx(1) = MOD(k - 1, num_part(1))
        idx(2) = MOD((k - 1) / num_part(1), num_part(2))
        idx(3) = (k - 1) / (num_part(1) * num_part(2))
        CALL RANDOM_NUMBER(r)
        DO d = 1, dim
           x(d, k) = prm%min_phys(d) + (idx(d) + 0.5d0) * dx(d) + jitter * dx(d) * (2.0d0 * r(d) - 1.0d0)
        END DO
        m(k) = prm%rho * PRODUCT(dx(1:dim))
        id(1, k) = k
        id(2, k) = 0
     END DO
     v = 0.0d0
     rho = prm%rho
  END IF
  WRITE(*, '(A,I0,A,I0,A,I0,A)') 'mcf_sdpd: ', n, ' particles, ', dim, 'D, ', nsteps, ' steps'

  ! Lees-Edwards faces: bcdef = 10 on both faces of y or z
  ! (io_read_physics.F90:1358-1400 reads bcdef and shear_v)
  bcdef = 1
  CALL get_ints(cfg, 'bcdef', 2 * dim, bcdef(1:2 * dim))
  shear_in = 0.0d0
  CALL get_reals(cfg, 'shear_v', 2 * dim * (dim - 1), shear_in(1:2 * dim * (dim - 1)))
  IF (bcdef(1) == 10 .OR. bcdef(2) == 10) CALL die('Lees-Edwards faces of x are not supported')
  le_dim = 0
  DO d = 1, 2 * dim
     IF (bcdef(d) /= 1 .AND. bcdef(d) /= 10) CALL die('only periodic (1) and Lees-Edwards (10) faces')
  END DO
  DO d = 2, dim
     IF (bcdef(2 * d - 1) == 10 .AND. bcdef(2 * d) == 10) THEN
        IF (le_dim /= 0) CALL die('Lees-Edwards faces on one dimension only')
        le_dim = d
     ELSE IF (bcdef(2 * d - 1) == 10 .OR. bcdef(2 * d) == 10) THEN
        CALL die('bcdef = 10 (Lees-Edwards) must be set on both faces of a dimension')
     END IF
  END DO

  ! ---- set-up ----
  rc = sdpd_create(prm, ctx)
  IF (rc /= SDPD_OK) CALL die('sdpd_create failed')
  IF (le_dim > 0) THEN
     ! shear_v lists, per face, the velocity components along the face;
     ! x is the first of them for the y and z faces
     d = le_dim
     v_lo = 0.0d0
     v_hi = 0.0d0
     v_lo(1) = shear_in((2 * d - 2) * (dim - 1) + 1)
     v_hi(1) = shear_in((2 * d - 1) * (dim - 1) + 1)
     IF (dim == 3) THEN
        ! the other component along the face (z for y faces, y for z
        ! faces); the library rejects it unless it is zero
        v_lo(5 - d) = shear_in((2 * d - 2) * 2 + 2)
        v_hi(5 - d) = shear_in((2 * d - 1) * 2 + 2)
     END IF
     CALL check(sdpd_set_lees_edwards(ctx, INT(d, c_int32_t), v_lo, v_hi), 'sdpd_set_lees_edwards')
  END IF
  ! colloids: num_colloid and the coll_* keys, read once per colloid as
  ! io_read_physics.F90:901-1350 does; disks and spheres, explicit
  ! integrators; their boundary particles come with the particle file
  ! (species id = colloid index)
  ncoll = get_int(cfg, 'num_colloid', 0)
  IF (ncoll > 0) THEN
     k = get_int(cfg, 'integrate_colloid_type', 2)
     IF (k /= 1 .AND. k /= 2) CALL die('only the explicit colloid integrators (integrate_colloid_type 1, 2)')
     cprm%noslip = get_int(cfg, 'coll_noslip', 2)
     cprm%rho_type = get_int(cfg, 'coll_rho_type', 0)
     cprm%translate = MERGE(1, 0, get_logical(cfg, 'coll_translate', .FALSE.))
     cprm%rotate = MERGE(1, 0, get_logical(cfg, 'coll_rotate', .FALSE.))
     cprm%integrate_ab = get_int(cfg, 'integrate_colloid_ab', 1)
     cprm%sub_steps = get_int(cfg, 'coll_sub_time_step', 1)
     cprm%body_force_type = get_int(cfg, 'coll_body_force_type', 0)
     cprm%pad = 0
     cprm%body_force = 0.0d0
     CALL get_reals(cfg, 'coll_body_force', dim, cprm%body_force(1:dim))
     ! outer layer 0.05 dx(1) (physics_adjust_parameters.F90:320)
     num_part = 0
     CALL get_ints(cfg, 'num_part', dim, num_part(1:dim))
     IF (num_part(1) > 0) THEN
        dxc = (prm%max_phys(1) - prm%min_phys(1)) / num_part(1)
     ELSE
        dxc = prm%cut_off / 3.0d0
     END IF
     cprm%dout = 0.05d0 * dxc
     ALLOCATE(c_rad(ncoll), c_m(ncoll), c_mmi(3, ncoll), c_x(3, ncoll), c_v(3, ncoll), c_w(3, ncoll))
     c_x = 0.0d0
     c_v = 0.0d0
     c_w = 0.0d0
     c_mmi = 0.0d0
     cpi = ACOS(-1.0d0)
     DO k = 1, ncoll
        rr = 2.0d0
        CALL get_reals_nth(cfg, 'coll_shape', k, 1, rr(1:1))
        IF (NINT(rr(1)) /= 2) CALL die('only disks and spheres (coll_shape = 2) are supported')
        rr = 0.0d0
        CALL get_reals_nth(cfg, 'coll_radius', k, dim, rr(1:dim))
        c_rad(k) = rr(1)
        CALL get_reals_nth(cfg, 'coll_x', k, dim, c_x(1:dim, k))
        CALL get_reals_nth(cfg, 'coll_v', k, dim, c_v(1:dim, k))
        CALL get_reals_nth(cfg, 'coll_omega', k, 3, c_w(1:3, k))
        ! mass and moment of inertia (particles_compute_mass.F90:153-357)
        IF (dim == 2) THEN
           c_m(k) = prm%rho * (cpi * c_rad(k)**2.0d0)
           c_mmi(3, k) = 0.5d0 * c_m(k) * c_rad(k)**2
        ELSE
           c_m(k) = prm%rho * (4.0d0 * cpi * c_rad(k)**3.0d0 / 3.0d0)
           c_mmi(1:3, k) = 0.4d0 * c_m(k) * c_rad(k)**2
        END IF
     END DO
     CALL check(sdpd_set_colloids(ctx, cprm, INT(ncoll, c_int32_t), c_rad, c_m, c_mmi, c_x, c_v, &
          c_w), 'sdpd_set_colloids')
  END IF
  CALL check(sdpd_upload(ctx, x, v, m, id, INT(n, c_int32_t)), 'sdpd_upload')
  CALL check(sdpd_set_step(ctx, INT(step_start, c_int64_t)), 'sdpd_set_step')
  IF (.NOT. newtonian) THEN
     ALLOCATE(ct(dim * dim, n))
     ! initial conformation from a file (io_read_conformation.F90) when the
     ! particles come from one; else the identity of sdpd_upload
     IF (read_external .AND. LEN_TRIM(read_conf_file) > 0) THEN
        CALL read_conformation(TRIM(dir)//TRIM(read_conf_file))
        CALL check(sdpd_upload_ct(ctx, ct), 'sdpd_upload_ct')
     END IF
  END IF
  ! time of step_start: 0, or the time_start of a restart physics file
  ! (marching_marching.F90:325-331)
  time_cur = get_real(cfg, 'time_start', 0.0d0)

  OPEN(newunit=su, file=TRIM(stat_file), status='replace', action='write')

  ! prologue: initial density, pressure and forces (marching_marching.F90:528-1123)
  CALL check(sdpd_build_cells(ctx), 'sdpd_build_cells')
  CALL check(sdpd_density(ctx), 'sdpd_density')
  CALL check(sdpd_forces(ctx), 'sdpd_forces')
  CALL write_statistic(step_start)

  CALL CPU_TIME(t0)
  step = step_start
  DO WHILE (step < step_start + nsteps)
     IF (fused) THEN
        ! whole steps on the device up to the next output of any kind: the
        ! host copies data only when it writes a file
        next_out = step_start + nsteps
        next_out = MIN(next_out, next_multiple(step, stat_freq))
        IF (out_freq > 0) next_out = MIN(next_out, next_multiple(step, out_freq))
        IF (restart_freq > 0) next_out = MIN(next_out, next_multiple(step, restart_freq))
        IF (conf_freq > 0 .AND. .NOT. newtonian) next_out = MIN(next_out, next_multiple(step, conf_freq))
        k = next_out - step
        CALL check(sdpd_step(ctx, INT(k, c_int32_t)), 'sdpd_step')
        DO d = 1, k
           time_cur = time_cur + prm%dt  ! marching_marching.F90:1393
        END DO
        step = step + k
     ELSE
        step = step + 1
        time_cur = time_cur + prm%dt
        CALL check(sdpd_set_step(ctx, INT(step, c_int64_t)), 'sdpd_set_step')
        CALL vv_step()
     END IF
     IF (MOD(step - step_start, stat_freq) == 0) CALL write_statistic(step)
     IF (out_freq > 0) THEN
        IF (MOD(step - step_start, out_freq) == 0) CALL write_particles(step)
     END IF
     IF (conf_freq > 0 .AND. .NOT. newtonian) THEN
        IF (MOD(step - step_start, conf_freq) == 0) CALL write_conformation(step, .FALSE.)
     END IF
     IF (restart_freq > 0) THEN
        IF (MOD(step - step_start, restart_freq) == 0) CALL write_restart(step)
     END IF
  END DO
  CALL check(sdpd_sync(ctx), 'sdpd_sync')
  CALL CPU_TIME(t1)
  CALL check(sdpd_get_step(ctx, cur_step), 'sdpd_get_step')
  CLOSE(su)
  ! the particles of the last step (the reference writes them as a restart)
  CALL write_particles(step)
  CALL write_restart(step)
  WRITE(*, '(A,I0,A,ES12.4,A)') 'mcf_sdpd: reached step ', cur_step, ' in ', t1 - t0, ' s of host time'
  rc32 = sdpd_destroy(ctx)

CONTAINS

  ! One velocity-Verlet step in the order of marching_integrate_VV.F90.
  SUBROUTINE vv_step()
    CALL check(sdpd_kick(ctx, 0.5d0), 'sdpd_kick')                    ! :303
    CALL check(sdpd_drift_wrap(ctx), 'sdpd_drift_wrap')               ! :321, :814
    IF (.NOT. newtonian) CALL check(sdpd_conformation(ctx, 0.5d0), 'sdpd_conformation')  ! :413
    CALL check(sdpd_build_cells(ctx), 'sdpd_build_cells')             ! :968-971
    CALL check(sdpd_density(ctx), 'sdpd_density')                     ! :1002, :1226
    CALL check(sdpd_forces(ctx), 'sdpd_forces')                       ! :1279, :1374, :1496
    IF (.NOT. newtonian) CALL check(sdpd_conformation(ctx, 0.5d0), 'sdpd_conformation')  ! :1511
    CALL check(sdpd_kick(ctx, 0.5d0), 'sdpd_kick')                    ! :1619
  END SUBROUTINE vv_step

  INTEGER FUNCTION next_multiple(st, freq)
    INTEGER, INTENT(IN) :: st, freq
    next_multiple = step_start + ((st - step_start) / freq + 1) * freq
  END FUNCTION next_multiple

  LOGICAL FUNCTION is_unformatted(fmt)
    CHARACTER(len=*), INTENT(IN) :: fmt
    CHARACTER(len=VAL_LEN) :: f
    f = lower(ADJUSTL(fmt))
    is_unformatted = f(1:1) == 'u' .OR. f(1:1) == 'b'
  END FUNCTION is_unformatted

  ! io_write_statistic.F90:224-254: (I9, k E16.8) of step, time, kinetic
  ! energy and momentum
  SUBROUTINE write_statistic(st)
    INTEGER, INTENT(IN) :: st
    CHARACTER(len=32) :: form
    CALL check(sdpd_stats(ctx, mom, ekin), 'sdpd_stats')
    WRITE(form, '(A4,I1,A6)') '(I9,', dim + 2, 'E16.8)'
    WRITE(su, form) st, time_cur, ekin, mom(1:dim)
  END SUBROUTINE write_statistic

  ! Records of a file: FORMATTED, one (k E16.8) record per particle
  ! (ppm_io_write_ascii.F90, list-directed count first when asked);
  ! UNFORMATTED, the value count then all values as two binary records
  ! (ppm_io_write_binary.F90), the count as its own pair first when asked.
  SUBROUTINE write_records(path, dat, count_first, unformatted)
    CHARACTER(len=*), INTENT(IN) :: path
    DOUBLE PRECISION, INTENT(IN) :: dat(:,:)
    LOGICAL, INTENT(IN) :: count_first, unformatted
    CHARACTER(len=32) :: form
    INTEGER :: u, i, ncol
    ncol = SIZE(dat, 1)
    IF (unformatted) THEN
       OPEN(newunit=u, file=path, status='replace', action='write', form='unformatted', &
            access='sequential')
       IF (count_first) THEN
          WRITE(u) 1
          WRITE(u) SIZE(dat, 2)
       END IF
       WRITE(u) SIZE(dat)
       WRITE(u) dat
    ELSE
       OPEN(newunit=u, file=path, status='replace', action='write')
       IF (count_first) WRITE(u, *) SIZE(dat, 2)
       WRITE(form, '(A1,I2,A6)') '(', ncol, 'E16.8)'
       DO i = 1, SIZE(dat, 2)
          WRITE(u, form) dat(:, i)
       END DO
    END IF
    CLOSE(u)
  END SUBROUTINE write_records

  ! x v rho m pid sid per particle (io_write_particles.F90:220-431)
  SUBROUTINE particle_records(dat)
    DOUBLE PRECISION, INTENT(OUT) :: dat(:,:)
    INTEGER :: i
    CALL check(sdpd_download(ctx, c_loc(x), c_loc(v), c_loc(rho), c_null_ptr, c_null_ptr), &
         'sdpd_download')
    DO i = 1, n
       dat(1:dim, i) = x(:, i)
       dat(dim + 1:2 * dim, i) = v(:, i)
       dat(2 * dim + 1, i) = rho(i)
       dat(2 * dim + 2, i) = m(i)
       dat(2 * dim + 3, i) = DBLE(id(1, i))
       dat(2 * dim + 4, i) = DBLE(id(2, i))
    END DO
  END SUBROUTINE particle_records

  ! <output_particles_file>[_r]<step>.out (io_write_particles.F90:224-232)
  SUBROUTINE write_particles(st)
    INTEGER, INTENT(IN) :: st
    CHARACTER(len=VAL_LEN+16) :: name
    DOUBLE PRECISION, ALLOCATABLE :: dat(:,:)
    ALLOCATE(dat(2 * dim + 4, n))
    CALL particle_records(dat)
    IF (read_external) THEN
       WRITE(name, '(2A,I8.8,A)') TRIM(out_file), '_r', st, '.out'
    ELSE
       WRITE(name, '(A,I8.8,A)') TRIM(out_file), st, '.out'
    END IF
    CALL write_records(TRIM(name), dat, .FALSE., out_unf)
  END SUBROUTINE write_particles

  ! <restart_particles_file><step>.dat: the count, then the particle
  ! records (io_write_restart_particles.F90:105-215), and the restart
  ! conformation of Oldroyd-B runs
  SUBROUTINE write_restart(st)
    INTEGER, INTENT(IN) :: st
    CHARACTER(len=VAL_LEN+16) :: name
    DOUBLE PRECISION, ALLOCATABLE :: dat(:,:)
    ALLOCATE(dat(2 * dim + 4, n))
    CALL particle_records(dat)
    WRITE(name, '(A,I8.8,A)') TRIM(restart_file), st, '.dat'
    CALL write_records(TRIM(name), dat, .TRUE., restart_unf)
    IF (.NOT. newtonian) CALL write_conformation(st, .TRUE.)
    CALL write_restart_physics(st)
  END SUBROUTINE write_restart

  ! <restart_physics_file><step>.mcf: the physics keys with step_start and
  ! time_start of this step and the colloids' current state, in the order
  ! and number formats of io_write_restart_physics.F90:236-953 (E16.8
  ! reals, I2..I10 integers, logicals); a resumed run names it as its
  ! mcf_physics_config_file.  Keys this host does not model keep the
  ! values of the run's own files (their defaults when absent).
  SUBROUTINE write_restart_physics(st)
    INTEGER, INTENT(IN) :: st
    CHARACTER(len=VAL_LEN+16) :: name
    INTEGER :: u, kk, nsp, iv(6)
    DOUBLE PRECISION :: rv(12)
    DOUBLE PRECISION, ALLOCATABLE, TARGET :: cx(:,:), cv(:,:), cw(:,:)
    WRITE(name, '(A,I8.8,A)') TRIM(restart_phys_file), st, '.mcf'
    OPEN(newunit=u, file=TRIM(name), status='replace', action='write')
    WRITE(u, '(A)') '#---------------------------------------------'
    WRITE(u, '(A)') '# physics keys of the run at this step, read as'
    WRITE(u, '(A)') '# the keys of physics_config.mcf'
    WRITE(u, '(A)') '#---------------------------------------------'
    WRITE(u, '(A)') ' '
    nsp = 1
    IF (ncoll > 0) nsp = 2
    CALL put_int(u, 'num_species', nsp, 2)
    CALL put_int(u, 'num_dim', dim, 2)
    CALL put_reals(u, 'min_phys', prm%min_phys(1:dim))
    CALL put_reals(u, 'max_phys', prm%max_phys(1:dim))
    iv = 0
    CALL get_ints(cfg, 'num_part', dim, iv(1:dim))
    CALL put_ints(u, 'num_part', iv(1:dim), 4)
    CALL put_real(u, 'cut_off', prm%cut_off)
    CALL put_real(u, 'dt', prm%dt)
    CALL put_int(u, 'step_start', st, 10)
    CALL put_int(u, 'step_end', get_int(cfg, 'step_end', 0), 10)
    CALL put_real(u, 'time_start', time_cur)
    CALL put_real(u, 'time_end', get_real(cfg, 'time_end', 0.0d0))
    CALL put_real(u, 'rho', prm%rho)
    CALL put_real(u, 'eta', prm%eta)
    CALL put_real(u, 'eta_coef', prm%eta_coef)
    CALL put_real(u, 'ksai', get_real(cfg, 'ksai', 0.0d0))
    CALL put_real(u, 'kt', prm%kt)
    CALL put_real(u, 'c', prm%c)
    CALL put_real(u, 'rho_ref', prm%rho_ref)
    CALL put_real(u, 'gamma', prm%gamma)
    CALL put_int(u, 'relax_type', get_int(cfg, 'relax_type', 0), 10)
    CALL put_real(u, 'dt_relax', get_real(cfg, 'dt_relax', 0.0d0))
    CALL put_int(u, 'step_relax', get_int(cfg, 'step_relax', 0), 10)
    CALL put_real(u, 'time_relax', get_real(cfg, 'time_relax', 0.0d0))
    CALL put_real(u, 'disorder_level', get_real(cfg, 'disorder_level', 0.0d0))
    CALL put_real(u, 'kt_relax', get_real(cfg, 'kt_relax', 0.0d0))
    CALL put_real(u, 'c_relax', get_real(cfg, 'c_relax', 0.0d0))
    IF (.NOT. newtonian) THEN
       CALL put_real(u, 'tau', prm%tau)
       CALL put_real(u, 'n_p', prm%n_p)
       CALL put_real(u, 'kt_p', prm%kt_p)
       CALL put_logical(u, 'eigen_dynamics', get_logical(cfg, 'eigen_dynamics', .FALSE.))
       rv = 0.0d0
       CALL get_reals(cfg, 'eval', dim, rv(1:dim))
       CALL put_reals(u, 'eval', rv(1:dim))
       CALL put_real(u, 'eval_tolerance', get_real(cfg, 'eval_tolerance', 0.0d0))
       rv = 0.0d0
       CALL get_reals(cfg, 'evec', dim * dim, rv(1:dim * dim))
       CALL put_reals(u, 'evec', rv(1:dim * dim))
       CALL put_logical(u, 'evec_normalize', get_logical(cfg, 'evec_normalize', .FALSE.))
       CALL put_real(u, 'evec_tolerance', get_real(cfg, 'evec_tolerance', 0.0d0))
    END IF
    CALL put_int(u, 'body_force_type', prm%body_force_type, 3)
    CALL put_reals(u, 'body_force', prm%body_force(1:dim))
    rv = 0.0d0
    CALL get_reals(cfg, 'body_force_d', dim, rv(1:dim))
    CALL put_reals(u, 'body_force_d', rv(1:dim))
    CALL put_int(u, 'flow_direction', get_int(cfg, 'flow_direction', 1), 5)
    CALL put_real(u, 'flow_width', get_real(cfg, 'flow_width', 0.0d0))
    CALL put_real(u, 'flow_v', get_real(cfg, 'flow_v', 0.0d0))
    CALL put_int(u, 'flow_adjust_freq', get_int(cfg, 'flow_adjust_freq', 0), 5)
    CALL put_int(u, 'num_colloid', ncoll, 3)
    IF (ncoll > 0) THEN
       ALLOCATE(cx(3, ncoll), cv(3, ncoll), cw(3, ncoll))
       CALL check(sdpd_get_colloids(ctx, c_loc(cx), c_loc(cv), c_loc(cw), c_null_ptr, c_null_ptr), &
            'sdpd_get_colloids')
       CALL put_int(u, 'coll_rho_type', INT(cprm%rho_type), 3)
       CALL put_logical(u, 'coll_translate', cprm%translate /= 0)
       CALL put_logical(u, 'coll_rotate', cprm%rotate /= 0)
       CALL put_int(u, 'coll_place', get_int(cfg, 'coll_place', 1), 3)
       CALL put_int(u, 'coll_noslip', INT(cprm%noslip), 3)
       CALL put_int(u, 'coll_body_force_type', INT(cprm%body_force_type), 3)
       CALL put_reals(u, 'coll_body_force', cprm%body_force(1:dim))
       CALL put_real(u, 'cc_lub_cut_off', get_real(cfg, 'cc_lub_cut_off', 0.0d0))
       CALL put_real(u, 'cc_lub_cut_on', get_real(cfg, 'cc_lub_cut_on', 0.0d0))
       CALL put_real(u, 'cc_repul_cut_off', get_real(cfg, 'cc_repul_cut_off', 0.0d0))
       CALL put_real(u, 'cc_repul_cut_on', get_real(cfg, 'cc_repul_cut_on', 0.0d0))
       CALL put_real(u, 'cc_repul_F0', get_real(cfg, 'cc_repul_f0', 0.0d0))
       CALL put_real(u, 'cw_lub_cut_off', get_real(cfg, 'cw_lub_cut_off', 0.0d0))
       CALL put_real(u, 'cw_lub_cut_on', get_real(cfg, 'cw_lub_cut_on', 0.0d0))
       CALL put_real(u, 'cw_repul_cut_off', get_real(cfg, 'cw_repul_cut_off', 0.0d0))
       CALL put_real(u, 'cw_repul_cut_on', get_real(cfg, 'cw_repul_cut_on', 0.0d0))
       CALL put_real(u, 'cw_repul_F0', get_real(cfg, 'cw_repul_f0', 0.0d0))
       DO kk = 1, ncoll
          CALL put_int(u, 'coll_shape', 2, 3)
          rv = 0.0d0
          rv(1) = c_rad(kk)
          CALL get_reals_nth(cfg, 'coll_radius', kk, dim, rv(1:dim))
          CALL put_reals(u, 'coll_radius', rv(1:dim))
          iv = 1
          CALL get_ints_nth(cfg, 'coll_freq', kk, 1, iv(1:1))
          CALL put_int(u, 'coll_freq', iv(1), 3)
          CALL put_real(u, 'coll_m', c_m(kk))
          CALL put_reals(u, 'coll_mmi', c_mmi(1:3, kk))
          CALL put_reals(u, 'coll_x', cx(1:dim, kk))
          CALL put_reals(u, 'coll_v', cv(1:dim, kk))
          rv = 0.0d0
          CALL get_reals_nth(cfg, 'coll_rot_vector', kk, 4, rv(1:4))
          CALL put_reals(u, 'coll_rot_vector', rv(1:4))
          rv = 0.0d0
          CALL get_reals_nth(cfg, 'coll_theta', kk, 3, rv(1:3))
          CALL put_reals(u, 'coll_theta', rv(1:3))
          CALL put_reals(u, 'coll_omega', cw(1:3, kk))
       END DO
       ! the keys of the colloid integrator this host reads with them
       CALL put_int(u, 'integrate_colloid_type', get_int(cfg, 'integrate_colloid_type', 2), 3)
       CALL put_int(u, 'integrate_colloid_AB', INT(cprm%integrate_ab), 3)
       CALL put_int(u, 'coll_sub_time_step', INT(cprm%sub_steps), 3)
    END IF
    iv = 1
    CALL get_ints(cfg, 'bcdef', 2 * dim, iv(1:2 * dim))
    CALL put_ints(u, 'bcdef', iv(1:2 * dim), 2)
    iv = 0
    CALL get_ints(cfg, 'shear_type', 2 * dim, iv(1:2 * dim))
    CALL put_ints(u, 'shear_type', iv(1:2 * dim), 2)
    rv = 0.0d0
    CALL get_reals(cfg, 'shear_v', 2 * dim * (dim - 1), rv(1:2 * dim * (dim - 1)))
    CALL put_reals(u, 'shear_v', rv(1:2 * dim * (dim - 1)))
    rv = 0.0d0
    CALL get_reals(cfg, 'shear_freq', 2 * dim, rv(1:2 * dim))
    CALL put_reals(u, 'shear_freq', rv(1:2 * dim))
    CALL put_int(u, 'wall_rho_type', get_int(cfg, 'wall_rho_type', 0), 3)
    CALL put_int(u, 'wall_noslip', get_int(cfg, 'wall_noslip', 0), 2)
    WRITE(u, '(A)') '#---------------------------------------------'
    WRITE(u, '(A)') '# end of the restart physics keys'
    WRITE(u, '(A)') '#---------------------------------------------'
    CLOSE(u)
  END SUBROUTINE write_restart_physics

  ! io_write_conformation.F90:175-245 (records of dim*dim values), or the
  ! restart conformation (count first, io_write_restart_conformation.F90)
  SUBROUTINE write_conformation(st, restart)
    INTEGER, INTENT(IN) :: st
    LOGICAL, INTENT(IN) :: restart
    CHARACTER(len=VAL_LEN+16) :: name
    CALL check(sdpd_download_tensors(ctx, c_loc(ct), c_null_ptr), 'sdpd_download_tensors')
    IF (restart) THEN
       WRITE(name, '(A,I8.8,A)') TRIM(restart_conf_file), st, '.dat'
       CALL write_records(TRIM(name), ct, .TRUE., restart_unf)
    ELSE
       IF (read_external) THEN
          WRITE(name, '(2A,I8.8,A)') TRIM(conf_out_file), '_r', st, '.out'
       ELSE
          WRITE(name, '(A,I8.8,A)') TRIM(conf_out_file), st, '.out'
       END IF
       CALL write_records(TRIM(name), ct, .FALSE., conf_unf)
    END IF
  END SUBROUTINE write_conformation

  ! io_read_conformation.F90: the count, then dim*dim values per particle
  SUBROUTINE read_conformation(path)
    CHARACTER(len=*), INTENT(IN) :: path
    INTEGER :: u, i, ios, nc
    OPEN(newunit=u, file=path, status='old', action='read', iostat=ios)
    IF (ios /= 0) CALL die('cannot open conformation file '//path)
    READ(u, *, iostat=ios) nc
    IF (ios /= 0) CALL die('reading conformation tensor has problem at line 1')
    IF (nc /= n) CALL die('actual number of lines is not equal to number given !')
    DO i = 1, n
       READ(u, *, iostat=ios) ct(:, i)
       IF (ios /= 0) CALL die('actual number of lines is not equal to number given !')
    END DO
    CLOSE(u)
  END SUBROUTINE read_conformation

  SUBROUTINE read_particles(path)
    CHARACTER(len=*), INTENT(IN) :: path
    INTEGER :: u, i, ios
    DOUBLE PRECISION :: pidr, sidr
    OPEN(newunit=u, file=path, status='old', action='read', iostat=ios)
    IF (ios /= 0) CALL die('cannot open particle file '//path)
    READ(u, *, iostat=ios) n
    IF (ios /= 0 .OR. n <= 0) CALL die('reading particles has problem at line 1')
    ALLOCATE(x(dim, n), v(dim, n), m(n), rho(n), id(2, n))
    DO i = 1, n
       READ(u, *, iostat=ios) x(:, i), v(:, i), rho(i), m(i), pidr, sidr
       IF (ios /= 0) CALL die('actual number of lines is not equal to number given !')
       id(1, i) = NINT(pidr)
       id(2, i) = NINT(sidr)
    END DO
    CLOSE(u)
  END SUBROUTINE read_particles

  SUBROUTINE check(code, what)
    INTEGER(c_int), INTENT(IN) :: code
    CHARACTER(len=*), INTENT(IN) :: what
    IF (code /= SDPD_OK) CALL die(what//': '//sdpd_last_error(ctx))
  END SUBROUTINE check

  SUBROUTINE die(msg)
    CHARACTER(len=*), INTENT(IN) :: msg
    WRITE(*, '(A)') 'mcf_sdpd: '//msg
    ERROR STOP 1
  END SUBROUTINE die

END PROGRAM mcf_sdpd
