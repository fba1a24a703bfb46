/*
 * sdpd.h -- C-ABI of the MI355X SDPD step (drop-in for MCF's per-step
 * density / force / integration loops).
 *
 * The reference has no FFI layer: its boundary is the set of Fortran
 * module procedures that marching_integrate_VV calls on TYPE(Particles)
 * and TYPE(Technique).  Each entry point below names the procedure it
 * replaces.  A Fortran host binds them with iso_c_binding
 * (mcf_amd/fortran/sdpd_c.f90).
 *
 * Conventions
 *  - every call returns 0 on success and a negative code on failure,
 *    mirroring the reference's INTEGER stat_info (0 / -1); the message is
 *    available from sdpd_last_error();
 *  - host arrays use the reference's column-major layout: x(num_dim, N),
 *    id(2, N) with pid in row 1 and sid in row 2, ct(num_dim*num_dim, N);
 *  - the context owns every device buffer; host arrays are copied only by
 *    sdpd_upload / sdpd_download / sdpd_stats;
 *  - one host thread and one HIP stream per context, calls are not
 *    reentrant on the same context;
 *  - particles are kept in cell order on the device; downloads return them
 *    in upload order.
 */
#ifndef SDPD_H
#define SDPD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDPD_OK 0
#define SDPD_ERR_ARG -1      /* bad argument or parameter                 */
#define SDPD_ERR_STATE -2    /* call out of order (e.g. no particles yet) */
#define SDPD_ERR_DEVICE -3   /* HIP runtime error                          */
#define SDPD_ERR_PHYSICS -4  /* numerical check failed (e.g. n <= 0)       */

#define SDPD_PRECISION_FP64 0
#define SDPD_PRECISION_MIXED 1

/* Physics and control parameters; names follow the reference's config keys
 * (mcf/mcf_config/ctrl.mcf:44-215, physics_config.mcf). */
typedef struct sdpd_params {
  int32_t num_dim;            /* 2 or 3                                     */
  int32_t symmetry;           /* image convention of the reference's mode   */
  int32_t kernel_type;        /* 1 quintic spline, 2 Lucy                    */
  int32_t rhs_density_type;   /* 1 mass density, 2 number density            */
  int32_t rhs_force_type;     /* 1 Morris (without noise), 2 Espanol, 3
                                 Hu-Adams, 4 Hu-Adams angular; Oldroyd-B:
                                 2, 3 or 4 (4 without noise)                 */
  int32_t state_equation_type;/* 1 Morris, 2 Batchelor (Tait)                */
  int32_t newtonian;          /* 0: Oldroyd-B conformation tensor            */
  int32_t brownian;           /* thermal noise                               */
  int32_t body_force_type;    /* 0 none, 1 constant, 2 +-x on the halves of
                                 the last dimension, 3 sine along y (4 is
                                 refused)                                    */
  int32_t precision;          /* SDPD_PRECISION_FP64 or _MIXED               */
  double min_phys[3];
  double max_phys[3];
  double cut_off;
  double c, rho, rho_ref, gamma;
  double eta, eta_coef, kt, dt;
  double tau, n_p, kt_p;
  double body_force[3];
  uint64_t seed;              /* key of the counter-based noise stream       */
} sdpd_params;

typedef struct sdpd_ctx sdpd_ctx;

/* Replaces mcf.F90 start-up of Particles/Technique for this path
 * (particles_new, technique_init_ppm: mcf.F90:616-649). */
int sdpd_create(const sdpd_params *params, sdpd_ctx **ctx);

/* Replaces particles_init_global_exter / io_read_particles hand-over
 * (io_read_particles.F90:162-240).  x, v: (num_dim, n); m: n;
 * id: (2, n) = (pid, sid): sid 0 for fluid particles, 1..num for the
 * boundary particles of colloid sid when sdpd_set_colloids has been
 * called (mcf_header.F90:118-120). */
int sdpd_upload(sdpd_ctx *ctx, const double *x, const double *v,
                const double *m, const int32_t *id, int32_t n);

/* Oldroyd-B initial conformation ct(num_dim*num_dim, n); the default
 * after sdpd_upload is the identity (particles_init_partial_inter.F90:145-149). */
int sdpd_upload_ct(sdpd_ctx *ctx, const double *ct);

/* Batched ensemble.  The reference runs one box per job; a batched context
 * holds nrep independent periodic boxes of different sizes that share the
 * physics of the params (whose min_phys / max_phys are then unused), so
 * that one launch of each kernel advances all of them.  Each replica is a
 * periodic box of its own: no particle sees another replica's particles.
 * box_lo, box_hi: (3, nrep), 3 columns per replica whatever num_dim.  Any
 * uploaded particles are discarded. */
int sdpd_set_replicas(sdpd_ctx *ctx, int32_t nrep, const double *box_lo,
                      const double *box_hi);

/* sdpd_upload for a batched context: rep(n) is each particle's replica
 * (0 .. nrep-1); particles must lie in their replica's box. */
int sdpd_upload_replicas(sdpd_ctx *ctx, const double *x, const double *v,
                         const double *m, const int32_t *id, const int32_t *rep,
                         int32_t n);

/* Replaces technique_build_list (technique_build_list.F90:1-127 ->
 * ppm_neighlist_clist + ppm_util_rank3d + MkNeighIdx): sorts particles
 * into cells with a stable counting sort. */
int sdpd_build_cells(sdpd_ctx *ctx);

/* Replaces particles_compute_density (particles_compute_density.F90:1-903).
 * Also evaluates the equation of state (particles_compute_pressure,
 * particles_compute_pressure.F90:1-106) in the same pass. */
int sdpd_density(sdpd_ctx *ctx);

/* Replaces particles_compute_pressure when called on its own (the pressure
 * is already current after sdpd_density; this recomputes it from rho). */
int sdpd_eos(sdpd_ctx *ctx);

/* Replaces particles_compute_interaction (particles_compute_interaction.F90)
 * plus particles_apply_body_force (particles_apply_body_force.F90:96-405)
 * and, when non-Newtonian, particles_compute_act
 * (particles_compute_act.F90:112-156).  Noise uses the context step.
 * With colloids, the first call after sdpd_upload also collects their
 * drag and torque and starts their accelerations (marching_marching.F90:
 * 1003-1080). */
int sdpd_forces(sdpd_ctx *ctx);

/* Replaces particles_integrate_velocity (particles_integrate_velocity.F90:84-94):
 * v += lambda * f * dt for fluid particles. */
int sdpd_kick(sdpd_ctx *ctx, double lambda);

/* Replaces particles_integrate_position (order 1) followed by
 * particles_adjust_particles (particles_integrate_position.F90:82-123,
 * particles_adjust_particles.F90:154-161,272-279): x += v*dt for fluid
 * particles, periodic wrap.  With colloids the colloid block of the step
 * runs in between (marching_integrate_VV.F90:455-765): drag and torque
 * from the last forces, rigid-body update, boundary particles carried
 * along. */
int sdpd_drift_wrap(sdpd_ctx *ctx);

/* Replaces particles_integrate_ct (particles_integrate_ct.F90:66-68):
 * ct += lambda * dt * act. */
int sdpd_conformation(sdpd_ctx *ctx, double lambda);

/* Replaces statistic_compute_statistic (statistic_compute_statistic.F90:230-283):
 * kinetic energy and momentum of fluid particles, FP64 compensated sums,
 * plus the colloids' translational share (colloid_compute_statistic.F90). */
int sdpd_stats(sdpd_ctx *ctx, double *mom /*[3]*/, double *ekin);

/* statistic_compute_statistic (statistic_compute_statistic.F90:230-283)
 * for every replica: ekin(nrep), mom(3, nrep); compensated sums in a fixed
 * order. */
int sdpd_stats_replicas(sdpd_ctx *ctx, double *mom, double *ekin);

/* Copies state back in upload order.  Any pointer may be NULL.
 * x, v, f: (num_dim, n); rho, p: n; ct, vgt: (num_dim*num_dim, n). */
int sdpd_download(sdpd_ctx *ctx, double *x, double *v, double *rho,
                  double *p, double *f);
int sdpd_download_tensors(sdpd_ctx *ctx, double *ct, double *vgt);

/* Replaces marching_marching's loop body with integrate_type = 2
 * (marching_integrate.F90:53-61 -> marching_integrate_VV.F90:303-1619):
 * advances nsteps velocity-Verlet steps.  The first call after upload
 * performs the initial evaluation of marching_marching.F90:528-1123
 * unless sdpd_forces has run since. */
int sdpd_step(sdpd_ctx *ctx, int32_t nsteps);

/* Lees-Edwards faces (the reference's bcdef = LE on both faces of one
 * dimension, boundary_update_boundary.F90:102-121): le_dim = 2 (y) or 3
 * (z), 1-based as the reference's face pairs, 0 switches them off;
 * v_lower / v_upper[3] = the face velocities shear_v0(:, 2 le_dim - 1) and
 * shear_v0(:, 2 le_dim).  The faces may only shear along x.  Images across
 * the faces are shifted along x by the shear length (v_upper - v_lower) t
 * and carry the velocity jump (ppm_map_part_ghost_get.F90:777-793,
 * particles_set_boundary_ghost_velocity.F90:175-240); particles crossing
 * them are wrapped the same way (particles_adjust_particles.F90:228-257,
 * 352-381).  t = time_current + dt of the step (the initial evaluation:
 * time_current), time_current accumulating dt from step 0.  One box only,
 * with min_phys = 0. */
int sdpd_set_lees_edwards(sdpd_ctx *ctx, int32_t le_dim, const double *v_lower,
                          const double *v_upper);

/* Rigid colloids made of boundary particles: disks (2D) and spheres (3D)
 * (Class_Colloid.F90; physics_config.mcf coll_* keys).  Fields follow the
 * reference's keys; the per-colloid arrays of sdpd_set_colloids follow
 * Class_Colloid's radius, m, mmi, x, v(:, :, 1), omega(:, :, 1). */
typedef struct sdpd_colloid_params {
  int32_t noslip;          /* coll_noslip: 1 frozen, 2 Morris                */
  int32_t rho_type;        /* coll_rho_type: 0 constant, 1 dynamic           */
  int32_t translate;       /* coll_translate                                 */
  int32_t rotate;          /* coll_rotate                                    */
  int32_t integrate_ab;    /* integrate_colloid_AB: 1..5                     */
  int32_t sub_steps;       /* coll_sub_time_step: 1..8                       */
  int32_t body_force_type; /* coll_body_force_type: 0 none, 1 constant       */
  int32_t pad;
  double body_force[3];    /* coll_body_force (per unit mass)                */
  double dout;             /* outer layer: 0.05 dx(1) in the reference       */
} sdpd_colloid_params;

/* Declares num colloids before sdpd_upload (physics_new colloid set-up,
 * colloid_new.F90).  radius, m: (num); mmi, x, v, omega: (3, num) (the
 * moment of inertia about the principal axes, centre, translational and
 * angular velocity).  One box (no replicas), FP64, Newtonian or Oldroyd-B
 * flow (in the latter both sides of a fluid-boundary pair take the fluid
 * particle's number density and pressure tensor, pp_interaction_fc.inc:
 * 80-113); the boundary particles then come with sdpd_upload (sid = colloid
 * index).
 * Colloid-colloid and colloid-wall forces are not modelled (the
 * reference's pp_interact_cc / cw off).  num = 0 removes the colloids. */
int sdpd_set_colloids(sdpd_ctx *ctx, const sdpd_colloid_params *prm, int32_t num,
                      const double *radius, const double *m, const double *mmi,
                      const double *x, const double *v, const double *omega);

/* Current colloid state and the drag (sum of m f) and torque (sum of
 * r x m f) of the last collection (particles_collect_colloid_interaction
 * .F90:93-136 plus the colloid body force), each (3, num); any pointer
 * may be NULL. */
int sdpd_get_colloids(sdpd_ctx *ctx, double *x, double *v, double *omega, double *drag,
                      double *torque);

/* sdpd_step runs pairs of steps through a captured HIP graph (default on,
 * and off while phase timing is enabled); enable = 0 launches every kernel
 * directly.  Results are identical either way. */
int sdpd_set_graphs(sdpd_ctx *ctx, int32_t enable);

/* Launch shapes of the pair kernels.  No counterpart in the reference
 * (its loops have no tunable shape).  density_shape picks the candidate
 * scan's LDS tile and per-lane list buffer: 0 = 224 records / 16 entries
 * (default), 1 = 128 / 16, 2 = 224 / 12, 3 = 128 / 12; chunk_order 1 lets
 * each XCD work through a contiguous eighth of the chunks (default), 0
 * hands chunk b to workgroup b.  Every choice gives identical results. */
typedef struct sdpd_tuning {
  int32_t density_shape;
  int32_t chunk_order;
} sdpd_tuning;
int sdpd_set_tuning(sdpd_ctx *ctx, const sdpd_tuning *t);
int sdpd_get_tuning(sdpd_ctx *ctx, sdpd_tuning *t);

/* Capacity of the stored neighbour lists, in entries per particle: a
 * multiple of 8 in 8..4096, or 0 for the automatic choice (1.6 times the
 * mean in-cut-off count of the box, plus 32, rounded up to 16).  A chunk
 * with a particle that has more neighbours scans its stencil again in the
 * force pass; every capacity gives identical results.  No counterpart in
 * the reference (its neighbour lists grow as needed).  Set before
 * sdpd_upload; get returns the capacity in use after it. */
int sdpd_set_list_capacity(sdpd_ctx *ctx, int32_t cap);
int sdpd_get_list_capacity(sdpd_ctx *ctx, int32_t *cap);

/* Verlet lists.  No counterpart in the reference (technique_build_list
 * rebuilds PPM's cell list every step, marching_integrate_VV.F90:968-971).
 * The stored lists hold every pair within cut_off + skin of their build;
 * the particles keep their order until the next build, which happens once
 * a particle has moved more than skin / 2 since the last one or, with
 * sheared faces, once twice the largest displacement plus the drift of the
 * sheared copies since the build (|v_upper - v_lower| times the time)
 * exceeds the skin.  Every step still decides in or out of the cut-off on
 * its own positions with the reference's FP64 formula, the image across a
 * sheared face taken from them, so the neighbour sets equal a per-step
 * rebuild's.  Mixed precision with sheared faces rebuilds every step.
 * skin < 0 (the default): automatic, 2 v dt 256 with v the fastest upload
 * speed or five thermal deviations sqrt(kT / m_min), whichever is larger,
 * kept between 1/20 of and the largest skin the sort grid allows (twice its
 * cell edge minus cut_off); 0: lists rebuilt every step; > 0: that skin,
 * refused at sdpd_upload if larger than the grid allows.  Set before
 * sdpd_upload; get returns the skin in use (0: rebuilt every step) after
 * it.
 * sdpd_get_list_builds returns the lists' builds since sdpd_upload (-1 when
 * they are rebuilt every step). */
int sdpd_set_skin(sdpd_ctx *ctx, double skin);
int sdpd_get_skin(sdpd_ctx *ctx, double *skin);
int sdpd_get_list_builds(sdpd_ctx *ctx, int64_t *builds);

/* Autotuner: times reps passes at the current positions for each launch
 * shape (the simulation does not advance), applies the fastest and returns
 * it with its time per density + force pass pair (ms).  Lists rebuilt every
 * step: density + force pass pairs for every shape.  Verlet lists: the
 * chunk order on pass pairs that walk the lists (each step's cost), then
 * the density shape on candidate scans of every chunk (each build's). */
int sdpd_autotune(sdpd_ctx *ctx, int32_t reps, sdpd_tuning *best, double *ms_best);

/* Noise step counter (the reference's marching step). */
int sdpd_set_step(sdpd_ctx *ctx, int64_t step);
int sdpd_get_step(sdpd_ctx *ctx, int64_t *step);

/* PPM's cell list for the current positions (technique_build_list ->
 * ppm_util_rank3d on the particles in upload order, ppm_util_rank3d.F90:
 * 395-531, symmetric layout without the ghost layer): nm[3] cells per
 * dimension (INT(L / (1.0001 cut_off))), lhbx(ncell+1) 0-based offsets,
 * lpdx(n) upload indices, ascending inside a cell.  The step itself sorts
 * on a grid with half this cell edge. */
int sdpd_get_cells(sdpd_ctx *ctx, int32_t *nm, int32_t *lhbx, int32_t *lpdx);

/* Number of in-cutoff ordered pairs found by the last density pass (the
 * force pass of the same step evaluates exactly these) and the number of
 * candidate tests that were re-decided in FP64 inside the cut-off band. */
int sdpd_get_counters(sdpd_ctx *ctx, int64_t *pairs, int64_t *fp64_rechecks);

/* Neighbour-set signature of the current cell list, in upload order: the
 * number of in-cutoff neighbours of each particle and the wrapping sum of
 * a 32-bit hash of (pid_i, pid_j) over them.  Both precision modes share
 * one in/out decision (FP32 test, FP64 inside a band around the cut-off);
 * mixed_decision is kept for callers that name the mode explicitly. */
int sdpd_neighbor_signature(sdpd_ctx *ctx, int32_t mixed_decision,
                            int32_t *count, uint32_t *hash);

/* Counter-based noise primitive, exposed for checking the stream. */
int sdpd_pair_normals(const uint32_t *pid_lo, const uint32_t *pid_hi,
                      int32_t count, uint64_t seed, int64_t step, float *g);

/* Per-phase kernel timing with HIP events on the context's stream (the
 * counterpart of the reference's VV phase timers with debug_flag = 3,
 * marching_integrate_VV.F90:264-277,1679-1709).  Phases:
 * 0 step boundary / kicks, 1 cell sort, 2 density + EOS, 3 pair forces,
 * 4 statistics.  sdpd_get_phase_times waits for the stream, returns the
 * summed milliseconds and launch counts per phase and clears them. */
#define SDPD_NUM_PHASES 5
int sdpd_set_timing(sdpd_ctx *ctx, int32_t enable);
int sdpd_get_phase_times(sdpd_ctx *ctx, double *ms, int64_t *launches);

/* Waits for all work queued on the context's stream. */
int sdpd_sync(sdpd_ctx *ctx);

/* Largest particle count sdpd_evaluate_noalloc accepts. */
#define SDPD_NOALLOC_MAX_PARTICLES 65536

/* Allocation-free evaluation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * computes what sdpd_build_cells, sdpd_density and sdpd_forces compute on
 * an uploaded context (the initial evaluation of marching_marching), for
 * particles already in device memory, with the same kernels and the same
 * results bit for bit.  All arrays are device pointers in the reference's
 * layout: x, v (num_dim, n); m (n); id (2, n) = (pid, sid), sid 0.
 * Outputs in the input order: rho, p (n; either may be NULL) and the
 * accelerations f (num_dim, n).  step keys the thermal noise.  One periodic
 * box of Newtonian fluid (2D or 3D, FP64 or mixed precision, no colloids,
 * no sheared faces); positions must lie in [min_phys, max_phys).  The work
 * is queued on `stream` (a hipStream_t, NULL for the null stream) and the
 * call returns without waiting.  Its working buffers come from one static
 * device array of 32 MiB that every device the library runs on holds until
 * the process ends; each call overwrites it, so calls that overlap on one
 * device race.  More than SDPD_NOALLOC_MAX_PARTICLES particles, or a box
 * whose cell tables and neighbour lists do not fit the array, are refused
 * with SDPD_ERR_ARG and a message on stderr. */
int sdpd_evaluate_noalloc(const sdpd_params *params, int64_t step, const double *x,
                          const double *v, const double *m, const int32_t *id, int32_t n,
                          double *rho, double *p, double *f, void *stream);

const char *sdpd_last_error(const sdpd_ctx *ctx);
int sdpd_destroy(sdpd_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
