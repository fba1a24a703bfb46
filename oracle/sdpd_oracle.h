/*
 * sdpd_oracle.h -- CPU FP64 restatement of MCF's per-step SDPD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product in mcf_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in mcf_amd/ links,
 * imports or calls it.
 *
 * Arrays follow the reference's Fortran layout, 0-based in C:
 *   x(dim, N) column-major  -> x[d + dim*i]
 *   ct/act/vgt(dim*dim, N)  -> t[a + dim*b + dim*dim*i]   (a = row, b = column)
 * Every function returns 0 on success and -1 on failure, like the
 * reference's stat_info.
 */
#ifndef SDPD_ORACLE_H
#define SDPD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct orc_colloids;

typedef struct orc_params {
  int32_t num_dim;          /* 2 or 3                                          */
  int32_t symmetry;         /* 1: half stencil + upper ghost layer (reference default) */
  int32_t kernel_type;      /* 1 quintic spline, 2 Lucy                        */
  int32_t rhs_density_type; /* 1 mass density, 2 number density                */
  int32_t rhs_force_type;   /* 3 Hu-Adams, 4 Hu-Adams angular                  */
  int32_t eos_type;         /* 1 Morris, 2 Batchelor (Tait)                    */
  int32_t newtonian;        /* 0: Oldroyd-B conformation tensor                */
  int32_t brownian;         /* thermal noise on/off                            */
  int32_t body_force_type;  /* 0 none, 1 constant per direction                */
  int32_t pad0;
  double min_phys[3];
  double max_phys[3];
  double cut_off;
  double c, rho, rho_ref, gamma;
  double eta, eta_coef, kt, dt;
  double tau, n_p, kt_p;
  double body_force[3];
  uint64_t seed;
  /* Lees-Edwards faces: le_dim = 0 none, 2 (y) or 3 (z) 1-based as the
   * reference's bcdef index; le_vlo / le_vhi = velocities of the lower and
   * upper face (shear_v0(:, 2i-1), shear_v0(:, 2i)); le_time = the time the
   * shear length is evaluated at: shear_length(:, upper) = (vhi - vlo) *
   * le_time (boundary_update_boundary.F90:102-121).  min_phys must be 0 in
   * every dimension (the reference's MODULO wraps into [0, L)). */
  int32_t le_dim;
  int32_t pad1;
  double le_vlo[3];
  double le_vhi[3];
  double le_time;
  /* colloids (NULL: fluid only); boundary particles carry sid 1..num */
  struct orc_colloids *coll;
} orc_params;

/* Rigid colloids built from boundary particles: disks in 2D, spheres in
 * 3D (Class_Colloid.F90).  Vectors are stored with 3 components per
 * colloid whatever the dimension; histories hold `ab` levels, newest
 * first, as the reference's third index.  The caller owns the arrays;
 * orc_vv_step updates them in place. */
typedef struct orc_colloids {
  int32_t num;             /* colloids                                     */
  int32_t noslip;          /* coll_noslip: 1 frozen, 2 Morris              */
  int32_t rho_type;        /* coll_rho_type: 0 constant, 1 dynamic         */
  int32_t translate;       /* coll_translate                               */
  int32_t rotate;          /* coll_rotate                                  */
  int32_t ab;              /* integrate_colloid_AB, 1..5                   */
  int32_t sub_steps;       /* coll_sub_time_step                           */
  int32_t body_force_type; /* coll_body_force_type: 0 none, 1 constant     */
  double body_force[3];    /* coll_body_force (per unit mass)              */
  double dout;             /* outer layer (0.05 dx(1), physics_adjust_parameters.F90:320) */
  double le_time;          /* shear-length time of the colloid block       */
  int64_t step_start;
  const double *radius;    /* (num)                                        */
  const double *m;         /* (num)                                        */
  const double *mmi;       /* (3, num)                                     */
  double *x;               /* (3, num) centres                             */
  double *v;               /* (3, num, ab)                                 */
  double *omega;           /* (3, num, ab)                                 */
  double *f;               /* (3, num, ab) translational acceleration      */
  double *alpha;           /* (3, num, ab) rotational acceleration         */
  double *rot_matrix;      /* (3, 3, num) rotation of the last sub-step    */
  double *acc_matrix;      /* (3, 3, num) accumulated rotation             */
  double *theta;           /* (3, num)                                     */
  double *drag;            /* (3, num) sum of m f over boundary particles  */
  double *torque;          /* (3, num) sum of r x (m f)                    */
} orc_colloids;

/* smoothing kernel W(r), W'(r) */
int orc_kernel(const orc_params *p, double r, double *w, double *gradw);

/* PPM ppm_util_rank3d / rank2d: stable counting sort of points into boxes.
 * nm[d] are the interior cell counts, ngl[0..dim-1] lower and ngl[dim..2dim-1]
 * upper ghost layers.  lhbx has nbox+1 entries, lpdx receives the point
 * indices (0-based) box by box; both are 0-based.  *count = sorted points.  */
int orc_rank(int32_t dim, const double *x, int32_t n, const double *xmin,
             const double *xmax, const int32_t *nm, const int32_t *ngl,
             int32_t *lhbx, int32_t *lpdx, int32_t *count);

/* interior cell counts Nm = INT(L / (1.0001 cut_off)) */
int orc_cell_counts(const orc_params *p, int32_t *nm);

/* periodic ghost images appended after the real particles (PPM ghost_get,
 * periodic bcdef).  Returns the number of ghosts in *nghost; xg receives
 * (dim, nghost) and owner the real index each ghost copies.  cap bounds
 * both outputs.                                                           */
int orc_ghost_get(const orc_params *p, const double *x, int32_t n,
                  double *xg, int32_t *owner, int32_t cap, int32_t *nghost);

/* In-cutoff pairs visited by the reference's cell-list traversal, mapped
 * back to real owners: pairs[2k], pairs[2k+1].                           */
int orc_neighbor_pairs(const orc_params *p, const double *x, int32_t n,
                       int32_t *pairs, int64_t cap, int64_t *npairs);

/* O(N^2) enumeration of in-cutoff pairs (i < j) with the same periodic
 * image convention.                                                      */
int orc_brute_pairs(const orc_params *p, const double *x, int32_t n,
                    int32_t *pairs, int64_t cap, int64_t *npairs);

/* density pass: rho(N) (number density for type 2, mass density for 1) */
int orc_density(const orc_params *p, const double *x, const double *m,
                const int32_t *sid, int32_t n, double *rho);

/* equation of state p(rho) */
int orc_pressure(const orc_params *p, const double *rho, const double *m,
                 int32_t n, double *press);

/* pair forces (acceleration) f(dim, N) and, when non-Newtonian, the
 * velocity gradient tensor vgt(dim*dim, N).  ct may be NULL when Newtonian. */
int orc_interaction(const orc_params *p, int64_t step, const double *x,
                    const double *v, const double *rho, const double *press,
                    const double *m, const int32_t *pid, const int32_t *sid,
                    const double *ct, int32_t n, double *f, double *vgt);

/* Oldroyd-B: act = c L^T + L c - (c - I)/tau ; ct += lambda*dt*act */
int orc_compute_act(const orc_params *p, const double *ct, const double *vgt,
                    int32_t n, double *act);

/* Counter-based noise stream shared with the device path. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
void orc_pair_normals(uint64_t seed, uint32_t pid_lo, uint32_t pid_hi,
                      int64_t step, float g[4]);
/* The stream's ln u (u in (0, 1]) and cos / sin of 2 pi k / 2^24 over
 * arrays, for checks against libm. */
void orc_noise_log_n(const float *u, int64_t n, float *out);
void orc_pair_normals_n(uint64_t seed, const uint32_t *lo, const uint32_t *hi,
                        int64_t step, int64_t n, float *g);
void orc_noise_sincos_n(const uint32_t *k, int64_t n, float *cs, float *sn);

/* Initial evaluation (marching_marching prologue): density, pressure,
 * forces, Oldroyd-B act.  step = noise step of this evaluation.           */
int orc_prologue(const orc_params *p, int64_t step, int32_t n, double *x,
                 double *v, const double *m, const int32_t *pid,
                 const int32_t *sid, double *rho, double *press, double *f,
                 double *ct, double *act, double *vgt);

/* One marching_integrate_VV step, in place.  step = noise step used by the
 * force pass of this step (the caller increments it first, as
 * marching_marching does).                                                */
int orc_vv_step(const orc_params *p, int64_t step, int32_t n, double *x,
                double *v, const double *m, const int32_t *pid,
                const int32_t *sid, double *rho, double *press, double *f,
                double *ct, double *act, double *vgt);

/* statistic_compute_statistic: kinetic energy and momentum of fluid
 * particles (naive summation, as the reference).                        */
int orc_statistics(const orc_params *p, const double *v, const double *m,
                   const int32_t *sid, int32_t n, double *ekin, double *mom);

/* colloid_in_nearest_image.F90 (disk / sphere): image of colloid sid's
 * centre nearest to xp (d_max = radius + dout), Lees-Edwards images first
 * with the shear length at time t; xim, rx = xp - xim and the image's
 * velocity vim, 3 components each. */
int orc_colloid_nearest(const orc_params *p, double t, const double *xp, int32_t sid,
                        double *xim, double *rx, double *vim);

/* colloid_noslip.F90: velocity given to boundary particle xc (colloid
 * sid) for its pair with fluid particle xf moving at vf; shear length of
 * the Lees-Edwards images at p->le_time. */
int orc_colloid_noslip(const orc_params *p, const double *xf, const double *xc,
                       const double *vf, int32_t sid, double *vc);

/* particles_collect_colloid_interaction.F90:93-136: drag (sum of m f) and
 * torque (sum of r x m f) of every colloid over its boundary particles,
 * into p->coll->drag / torque, with r from the colloid block's shear
 * length time. */
int orc_colloid_collect(const orc_params *p, const double *x, const double *f,
                        const double *m, const int32_t *sid, int32_t n);

/* An injected noise stream for step `step`: count pairs, keys
 * pid_lo << 32 | pid_hi in ascending order, normals(4, count).  The force
 * pass of that step takes its normals from the table (a pair missing from
 * it makes the evaluation fail); count = 0 removes the table.  The arrays
 * are read in place and must outlive the evaluations that use them. */
int orc_set_noise_table(int64_t step, int64_t count, const uint64_t *keys,
                        const float *normals);

/* orc_vv_step in two halves: the opening half kick, drift, ct half update,
 * colloid block and wrap; then the evaluation at the new positions, ct half
 * update and closing half kick. */
int orc_vv_begin(const orc_params *p, int64_t step, int32_t n, double *x,
                 double *v, const double *m, const int32_t *pid,
                 const int32_t *sid, double *rho, double *press, double *f,
                 double *ct, double *act, double *vgt);
int orc_vv_finish(const orc_params *p, int64_t step, int32_t n, double *x,
                  double *v, const double *m, const int32_t *pid,
                  const int32_t *sid, double *rho, double *press, double *f,
                  double *ct, double *act, double *vgt);

/* Threads of the density and force traversals (OpenMP builds): 1, the
 * default, is the reference's serial order; more colour the home cells so
 * that parallel rows never share a particle, which changes the order of
 * the sums.  For timing only (the benchmark's CPU baseline); returns -1
 * when the build has no OpenMP and n > 1. */
int orc_set_threads(int32_t n);

#ifdef __cplusplus
}
#endif
#endif
