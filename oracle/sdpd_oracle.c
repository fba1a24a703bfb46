/*
 * sdpd_oracle.c -- CPU FP64 restatement of MCF's per-step SDPD path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
 * product in mcf_amd/.
 *
 * It restates, in plain C and in the reference's own traversal order:
 *   - the quintic-spline / Lucy kernels
 *       (mcf/src/kernel_new.F90:47-93, kernel_kernel_quintic_spline.F90:1-221,
 *        kernel_kernel_Lucy.F90:1-174);
 *   - PPM's cell list: Nm, ghost layers and the stable counting sort
 *       (ppm.tgz:src/ppm_neighlist_clist.F90:370-393,
 *        ppm_util_rank3d.F90:395-531, ppm_util_rank2d.F90);
 *   - PPM's periodic ghost images and their fold-back
 *       (ppm.tgz:src/ppm_bcdef_periodic_{x,y,z}.inc:37-90,
 *        ppm_map_part_ghost_pop.F90:323-376);
 *   - the half-stencil (symmetric) and full-stencil cell traversal
 *       (ppm.tgz:src/ppm_neighlist_MkNeighIdx.F90:176-270,
 *        mcf/src/particles_compute_density.F90:578-867,
 *        particles_compute_interaction.F90:618-911, pp_interaction.inc:37-49);
 *   - density, equation of state, Morris / Espanol / Hu-Adams / angular /
 *     Oldroyd-B pair forces (rhs_force_ff_Newtonian_Morris.F90:112-136,
 *        rhs_force_ff_Newtonian_Espanol.F90:152-251)
 *       (rhs_density.F90:1-102, stateEquation_compute_pressure.F90:87-180,
 *        stateEquation_new.F90:102, particles_compute_pressure.F90:1-106,
 *        particles_compute_pressure_tensor.F90:94-133,
 *        rhs_force_ff_Newtonian_HuAdams.F90:8-246,
 *        rhs_force_ff_Newtonian_HuAdams_angular.F90:8-247,
 *        rhs_force_ff_non_Newtonian_HuAdams.F90:1-207, pp_vgt_ip.inc,
 *        pp_vgt_jp.inc);
 *   - the Oldroyd-B conformation update (particles_compute_act.F90:112-156,
 *        particles_integrate_ct.F90:66-68);
 *   - the velocity-Verlet step and periodic wrap
 *       (marching_integrate_VV.F90:303-1619,
 *        particles_integrate_velocity.F90:84-94,
 *        particles_integrate_position.F90:82-123,
 *        particles_adjust_particles.F90:154-161,272-279,
 *        particles_apply_body_force.F90:125-353);
 *   - the statistics sums (statistic_compute_statistic.F90:230-283).
 *
 * Thermal noise: the reference draws a sequential RANDOM_NUMBER stream in
 * pair-visit order (random_random.F90:48-73), which no other traversal can
 * reproduce.  Here it is replaced by the counter-based stream that the
 * device path uses too: Philox-4x32-10 (Salmon et al., SC'11) keyed by
 * (seed) with counter (min pid, max pid, step), turned into normals by
 * Box-Muller with log/sin/cos written from their Taylor series in exactly
 * rounded float operations, so both sides produce identical bits.  The
 * reference's symmetric traceless 3x3 matrix times e_ij,
 * (rhs_force_ff_Newtonian_HuAdams.F90:186-219) is a Gaussian vector with
 * covariance I/2 + e e^T/6 (3D) or I/2 (2D); it is drawn here from three
 * (two) normals with that covariance, oriented by the lower pid so that the
 * pair force stays antisymmetric.
 *
 * Parity status of each function (checks in tests/test_oracle_*.py):
 *   kernel           pinned: closed forms, normalisation, finite differences
 *   rank / cells     pinned: std stable sort by key, hand cases at xmax
 *   ghosts + pairs   pinned: O(N^2) brute force (set equality, exactly once)
 *   density          pinned: brute force sum, perfect-lattice closed sum
 *   pressure         pinned: closed forms
 *   interaction      pinned: two-particle closed form, lattice at rest,
 *                    momentum conservation, shear-wave viscous decay rate
 *   noise            pinned: Philox known-answer vectors, normal moments/KS,
 *                    covariance against the reference's 9-normal recipe,
 *                    fluctuation-dissipation temperature
 *   act / VV         pinned: homogeneous-shear steady state, constant-force
 *                    closed form, time reversibility at kT = 0
 *   statistics       pinned: closed sums
 */
#include "sdpd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.141592653589793238
/* mcf_machine_zero as set at start-up: 10^(INT(LOG10(EPSILON(1d0))) + 1)
 * (technique_new.F90:361-363) */
#define ORC_MACHINE_ZERO 1e-14

/* ------------------------------------------------------------------ */
/* kernel                                                              */
/* ------------------------------------------------------------------ */

typedef struct {
  double h, coef, coef_grad;
  int type;
} orc_kern;

static void kern_init(const orc_params *p, orc_kern *k) {
  /* kernel_new.F90:47-93 */
  k->type = p->kernel_type;
  if (p->kernel_type == 1) {
    k->h = p->cut_off / 3.0;
    if (p->num_dim == 2)
      k->coef = 7.0 / (478.0 * ORC_PI * k->h * k->h);
    else
      k->coef = 3.0 / (359.0 * ORC_PI * k->h * k->h * k->h);
  } else {
    k->h = p->cut_off;
    if (p->num_dim == 2)
      k->coef = 5.0 / (ORC_PI * k->h * k->h);
    else
      k->coef = 105.0 / (16.0 * ORC_PI * k->h * k->h * k->h);
  }
  k->coef_grad = -k->coef / k->h;
}

static void kern_eval(const orc_kern *k, double r, double *w, double *gradw) {
  double s = r / k->h;
  if (k->type == 1) {
    /* kernel_kernel_quintic_spline.F90:160-215 */
    double s1 = 1.0 - s, s2 = 2.0 - s, s3 = 3.0 - s;
    double s1_4 = (s1 * s1) * (s1 * s1);
    double s2_4 = (s2 * s2) * (s2 * s2);
    double s3_4 = (s3 * s3) * (s3 * s3);
    double s1_5 = 15.0 * s1_4 * s1;
    double s2_5 = -6.0 * s2_4 * s2;
    double s3_5 = s3_4 * s3;
    double g1 = 75.0 * s1_4, g2 = -30.0 * s2_4, g3 = 5.0 * s3_4;
    if (s < 1.0) {
      *w = k->coef * (s3_5 + s2_5 + s1_5);
      *gradw = k->coef_grad * (g3 + g2 + g1);
    } else if (s < 2.0) {
      *w = k->coef * (s3_5 + s2_5);
      *gradw = k->coef_grad * (g3 + g2);
    } else if (s < 3.0) {
      *w = k->coef * s3_5;
      *gradw = k->coef_grad * g3;
    } else {
      *w = 0.0;
      *gradw = 0.0;
    }
  } else {
    /* kernel_kernel_Lucy.F90 */
    if (s <= 1.0) {
      double s1 = 1.0 - s;
      *w = k->coef * (1.0 + 3.0 * s) * (s1 * s1 * s1);
      *gradw = k->coef_grad * 12.0 * s * (s1 * s1);
    } else {
      *w = 0.0;
      *gradw = 0.0;
    }
  }
}

int orc_kernel(const orc_params *p, double r, double *w, double *gradw) {
  orc_kern k;
  if (r < 0.0) return -1;
  kern_init(p, &k);
  kern_eval(&k, r, w, gradw);
  return 0;
}

/* ------------------------------------------------------------------ */
/* counting sort of points into boxes (ppm_util_rank3d.F90:395-531)     */
/* ------------------------------------------------------------------ */

int orc_rank(int32_t dim, const double *x, int32_t n, const double *xmin,
             const double *xmax, const int32_t *nm, const int32_t *ngl,
             int32_t *lhbx, int32_t *lpdx, int32_t *count) {
  int32_t nmt[3] = {1, 1, 1};
  double rd[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    if (nm[d] <= 0 || xmax[d] <= xmin[d]) return -1;
    nmt[d] = nm[d] + ngl[d] + ngl[dim + d];
    rd[d] = (double)nm[d] / (xmax[d] - xmin[d]);
  }
  int64_t nbox = (int64_t)nmt[0] * nmt[1] * nmt[2];
  int32_t *pbox = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t *npbx = (int32_t *)calloc((size_t)nbox, sizeof(int32_t));
  int32_t *cbox = (int32_t *)malloc(sizeof(int32_t) * (size_t)nbox);
  if (!pbox || !npbx || !cbox) {
    free(pbox); free(npbx); free(cbox);
    return -1;
  }
  int32_t icount = 0;
  for (int32_t ip = 0; ip < n; ++ip) {
    int32_t c[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
      double xv = x[d + (size_t)dim * ip];
      int32_t k = (int32_t)floor((xv - xmin[d]) * rd[d]) + ngl[d];
      /* index corrections at the upper face (rank3d:433-459) */
      if (xv >= xmax[d] && k < nm[d] + ngl[d]) k = nm[d] + ngl[d];
      if (xv < xmax[d] && k >= nm[d] + ngl[d]) k = nm[d] + ngl[d] - 1;
      c[d] = k;
    }
    int inside = 1;
    for (int d = 0; d < dim; ++d)
      if (c[d] < 0 || c[d] >= nmt[d]) inside = 0;
    if (inside) {
      ++icount;
      pbox[ip] = c[0] + nmt[0] * (c[1] + nmt[1] * c[2]);
    } else {
      pbox[ip] = -1;
    }
  }
  for (int32_t ip = 0; ip < n; ++ip)
    if (pbox[ip] >= 0) npbx[pbox[ip]]++;
  cbox[0] = 0;
  lhbx[0] = 0;
  for (int64_t b = 1; b < nbox; ++b) {
    cbox[b] = cbox[b - 1] + npbx[b - 1];
    lhbx[b] = cbox[b];
  }
  lhbx[nbox] = lhbx[nbox - 1] + npbx[nbox - 1];
  for (int32_t ip = 0; ip < n; ++ip) {
    int32_t b = pbox[ip];
    if (b >= 0) lpdx[cbox[b]++] = ip;
  }
  *count = icount;
  free(pbox); free(npbx); free(cbox);
  return 0;
}

int orc_cell_counts(const orc_params *p, int32_t *nm) {
  /* ppm_neighlist_clist.F90:370-376 with cutoff = 1.0001 cut_off
   * (mcf.F90:617, technique_build_list.F90:94) */
  double cell = p->cut_off * 1.0001;
  for (int d = 0; d < p->num_dim; ++d) {
    int32_t v = (int32_t)((p->max_phys[d] - p->min_phys[d]) / cell);
    nm[d] = v < 1 ? 1 : v;
  }
  for (int d = p->num_dim; d < 3; ++d) nm[d] = 1;
  return 0;
}

/* ------------------------------------------------------------------ */
/* periodic ghosts (ppm_bcdef_periodic_{x,y,z}.inc:37-90)              */
/* ------------------------------------------------------------------ */

static int ghost_count_bound(const orc_params *p, int32_t n) {
  (void)p;
  return 8 * n + 16; /* at most 7 images per particle in 3D (3 in 2D) */
}

/* Fortran MODULO(a, p) for reals: the exact remainder with the sign of p
 * (computed through fmod, which is exact, as gfortran does). */
static double f_modulo(double a, double p) {
  double r = fmod(a, p);
  if (r != 0.0 && ((r < 0.0) != (p < 0.0))) r = r + p;
  return r;
}

/* index of the Lees-Edwards dimension (0-based), -1 when none */
static int le_index(const orc_params *p) {
  return p->le_dim >= 1 && p->le_dim <= p->num_dim ? p->le_dim - 1 : -1;
}

int orc_ghost_get(const orc_params *p, const double *x, int32_t n, double *xg,
                  int32_t *owner, int32_t cap, int32_t *nghost) {
  const int dim = p->num_dim;
  const double g = p->cut_off * 1.0001;
  /* work list = real particles followed by ghosts created so far */
  int32_t total = n;
  int32_t bound = ghost_count_bound(p, n);
  double *xt = (double *)malloc(sizeof(double) * (size_t)dim * bound);
  int32_t *ig = (int32_t *)malloc(sizeof(int32_t) * (size_t)bound);
  if (!xt || !ig) { free(xt); free(ig); return -1; }
  memcpy(xt, x, sizeof(double) * (size_t)dim * n);
  for (int32_t i = 0; i < n; ++i) ig[i] = i;
  const int gle = le_index(p);
  if (gle >= 0) {
    /* Lees-Edwards images first (ppm_map_part_ghost_get.F90:777-793,
     * ppm_bcdef_le_y.inc): across the lower face the copy moves up by L
     * and the other coordinates by MODULO(x + shear_length(:,upper), L);
     * without symmetry also across the upper face, moved down by L and by
     * MODULO(x + shear_length(:,lower), L), shear_length(:,lower) = -S. */
    const int d = gle;
    const double L = p->max_phys[d] - p->min_phys[d];
    int32_t nplus = total;
    double lo_f = p->min_phys[d], lo_i = p->min_phys[d] + g;
    for (int32_t i = 0; i < nplus; ++i) {
      double xv = xt[d + (size_t)dim * i];
      if (xv >= lo_f && xv < lo_i) {
        for (int e = 0; e < dim; ++e) {
          double Le = p->max_phys[e] - p->min_phys[e];
          double sl = (p->le_vhi[e] - p->le_vlo[e]) * p->le_time;
          xt[e + (size_t)dim * total] =
              e == d ? xv + L : f_modulo(xt[e + (size_t)dim * i] + sl, Le);
        }
        ig[total] = ig[i];
        ++total;
      }
    }
    if (!p->symmetry) {
      double hi_f = p->max_phys[d], hi_i = p->max_phys[d] - g;
      for (int32_t i = 0; i < nplus; ++i) {
        double xv = xt[d + (size_t)dim * i];
        if (xv > hi_i && xv <= hi_f) {
          for (int e = 0; e < dim; ++e) {
            double Le = p->max_phys[e] - p->min_phys[e];
            double sl = (p->le_vlo[e] - p->le_vhi[e]) * p->le_time;
            xt[e + (size_t)dim * total] =
                e == d ? xv - L : f_modulo(xt[e + (size_t)dim * i] + sl, Le);
          }
          ig[total] = ig[i];
          ++total;
        }
      }
    }
  }
  for (int d = 0; d < dim; ++d) {
    if (d == gle) continue;
    double L = p->max_phys[d] - p->min_phys[d];
    int32_t nplus = total;
    double lo_f = p->min_phys[d], lo_i = p->min_phys[d] + g;
    for (int32_t i = 0; i < nplus; ++i) {
      double xv = xt[d + (size_t)dim * i];
      if (xv >= lo_f && xv < lo_i) {
        for (int e = 0; e < dim; ++e)
          xt[e + (size_t)dim * total] = xt[e + (size_t)dim * i];
        xt[d + (size_t)dim * total] = xv + L;
        ig[total] = ig[i];
        ++total;
      }
    }
    if (!p->symmetry) {
      double hi_f = p->max_phys[d], hi_i = p->max_phys[d] - g;
      for (int32_t i = 0; i < nplus; ++i) {
        double xv = xt[d + (size_t)dim * i];
        if (xv > hi_i && xv <= hi_f) {
          for (int e = 0; e < dim; ++e)
            xt[e + (size_t)dim * total] = xt[e + (size_t)dim * i];
          xt[d + (size_t)dim * total] = xv - L;
          ig[total] = ig[i];
          ++total;
        }
      }
    }
  }
  int32_t ng = total - n;
  if (ng > cap) { free(xt); free(ig); return -1; }
  memcpy(xg, xt + (size_t)dim * n, sizeof(double) * (size_t)dim * ng);
  memcpy(owner, ig + n, sizeof(int32_t) * (size_t)ng);
  *nghost = ng;
  free(xt);
  free(ig);
  return 0;
}

/* ------------------------------------------------------------------ */
/* cell structure over real + ghost particles                           */
/* ------------------------------------------------------------------ */

typedef struct {
  int dim;
  int32_t nreal, nall;
  double *xall;     /* (dim, nall) */
  int32_t *owner;   /* nall */
  int32_t nmt[3];
  int64_t nbox;
  int32_t *lhbx, *lpdx;
  int nnp;
  int inp[27][3], jnp[27][3];
  int cstart;       /* first home cell index in every dimension */
  int cend[3];      /* last home cell index (inclusive)           */
} orc_cells;

static void cells_free(orc_cells *c) {
  free(c->xall); free(c->owner); free(c->lhbx); free(c->lpdx);
  memset(c, 0, sizeof(*c));
}

static void make_stencil(int dim, int symmetry, orc_cells *c) {
  /* ppm_neighlist_MkNeighIdx.F90:176-270 */
  memset(c->inp, 0, sizeof(c->inp));
  memset(c->jnp, 0, sizeof(c->jnp));
  if (symmetry) {
    if (dim == 2) {
      c->nnp = 5;
      c->jnp[1][0] = 1;                     /* 0 -- 1 */
      c->inp[2][1] = 1;                     /* 0 -- 3 */
      c->jnp[3][0] = 1; c->jnp[3][1] = 1;   /* 0 -- 4 */
      c->inp[4][0] = 1; c->jnp[4][1] = 1;   /* 1 -- 3 */
    } else {
      c->nnp = 14;
      c->jnp[1][0] = 1;
      c->inp[2][1] = 1;
      c->jnp[3][0] = 1; c->jnp[3][1] = 1;
      c->jnp[4][2] = 1;
      c->jnp[5][0] = 1; c->jnp[5][2] = 1;
      c->jnp[6][1] = 1; c->jnp[6][2] = 1;
      c->jnp[7][0] = 1; c->jnp[7][1] = 1; c->jnp[7][2] = 1;
      c->inp[8][0] = 1; c->jnp[8][1] = 1;
      c->inp[9][0] = 1; c->jnp[9][2] = 1;
      c->inp[10][0] = 1; c->jnp[10][1] = 1; c->jnp[10][2] = 1;
      c->inp[11][1] = 1; c->jnp[11][2] = 1;
      c->inp[12][1] = 1; c->jnp[12][0] = 1; c->jnp[12][2] = 1;
      c->inp[13][0] = 1; c->inp[13][1] = 1; c->jnp[13][2] = 1;
    }
  } else {
    int nz = dim == 3 ? 1 : 0, l = 0;
    for (int k = -nz; k <= nz; ++k)
      for (int j = -1; j <= 1; ++j)
        for (int i = -1; i <= 1; ++i) {
          c->jnp[l][0] = i; c->jnp[l][1] = j; c->jnp[l][2] = k;
          ++l;
        }
    c->nnp = l;
  }
}

static int cells_build(const orc_params *p, const double *x, int32_t n,
                       orc_cells *c) {
  memset(c, 0, sizeof(*c));
  const int dim = p->num_dim;
  c->dim = dim;
  c->nreal = n;
  int32_t cap = ghost_count_bound(p, n);
  c->xall = (double *)malloc(sizeof(double) * (size_t)dim * (n + cap));
  c->owner = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n + cap));
  if (!c->xall || !c->owner) { cells_free(c); return -1; }
  memcpy(c->xall, x, sizeof(double) * (size_t)dim * n);
  for (int32_t i = 0; i < n; ++i) c->owner[i] = i;
  int32_t ng = 0;
  if (orc_ghost_get(p, x, n, c->xall + (size_t)dim * n, c->owner + n, cap, &ng)) {
    cells_free(c);
    return -1;
  }
  c->nall = n + ng;
  int32_t nm[3], ngl[6] = {0, 0, 0, 0, 0, 0};
  orc_cell_counts(p, nm);
  for (int d = 0; d < dim; ++d) {
    if (p->symmetry) {
      ngl[dim + d] = 1;
    } else {
      ngl[d] = 1;
      ngl[dim + d] = 1;
    }
  }
  c->nmt[0] = c->nmt[1] = c->nmt[2] = 1;
  for (int d = 0; d < dim; ++d) c->nmt[d] = nm[d] + ngl[d] + ngl[dim + d];
  c->nbox = (int64_t)c->nmt[0] * c->nmt[1] * c->nmt[2];
  c->lhbx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(c->nbox + 1));
  c->lpdx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(c->nall > 0 ? c->nall : 1));
  if (!c->lhbx || !c->lpdx) { cells_free(c); return -1; }
  int32_t count = 0;
  if (orc_rank(dim, c->xall, c->nall, p->min_phys, p->max_phys, nm, ngl,
               c->lhbx, c->lpdx, &count)) {
    cells_free(c);
    return -1;
  }
  make_stencil(dim, p->symmetry, c);
  /* home cells: 0..Nmt-2 (symmetric) or 1..Nmt-2 (full)
   * (particles_compute_density.F90:583-605) */
  c->cstart = p->symmetry ? 0 : 1;
  c->cend[0] = c->nmt[0] - 2;
  c->cend[1] = c->nmt[1] - 2;
  c->cend[2] = dim == 3 ? c->nmt[2] - 2 : c->cstart;
  return 0;
}

/* squared distance with the reference's operation order (no contraction:
 * this file is compiled with -ffp-contract=off) */
static inline double dist2(int dim, const double *xi, const double *xj,
                           double *rij) {
  double d2 = 0.0;
  for (int d = 0; d < dim; ++d) {
    rij[d] = xi[d] - xj[d];
    d2 = d2 + rij[d] * rij[d];
  }
  return d2;
}

/* Visit every candidate pair in the reference order; call visit() on the
 * in-cutoff ones.  visit gets (ip, jp) as indices into xall, r, rij. */
typedef void (*visit_fn)(void *ctx, int32_t ip, int32_t jp, double r,
                         const double *rij);

/* the candidate pairs of home cell (i, j, k) in the reference order */
static void traverse_cell(const orc_params *p, const orc_cells *c, visit_fn visit,
                          void *ctx, int i, int j, int k) {
  const int dim = c->dim;
  const double cut2 = p->cut_off * p->cut_off;
  const int64_t n1 = c->nmt[0];
  const int64_t n2 = dim == 3 ? (int64_t)c->nmt[0] * c->nmt[1] : 0;
  double rij[3];
  int64_t ccell = i + n1 * j + n2 * k;
  for (int it = 0; it < c->nnp; ++it) {
    int64_t icell = ccell + c->inp[it][0] + n1 * c->inp[it][1] +
                    n2 * c->inp[it][2];
    int64_t jcell = ccell + c->jnp[it][0] + n1 * c->jnp[it][1] +
                    n2 * c->jnp[it][2];
    int32_t istart = c->lhbx[icell], iend = c->lhbx[icell + 1];
    int32_t jstart = c->lhbx[jcell], jend = c->lhbx[jcell + 1];
    for (int32_t ipart = istart; ipart < iend; ++ipart) {
      int32_t ip = c->lpdx[ipart];
      if (jcell == icell && p->symmetry) jstart = ipart + 1;
      for (int32_t jpart = jstart; jpart < jend; ++jpart) {
        if (jcell == icell && jpart == ipart) continue;
        int32_t jp = c->lpdx[jpart];
        double d2 = dist2(dim, c->xall + (size_t)dim * ip,
                          c->xall + (size_t)dim * jp, rij);
        if (d2 >= cut2) continue;
        visit(ctx, ip, jp, sqrt(d2), rij);
      }
    }
  }
}

/* Threads of the density and force traversals.  1 (the default) keeps the
 * reference's serial order; only the CPU baseline of the benchmark asks
 * for more. */
static int g_threads = 1;

int orc_set_threads(int32_t n) {
#ifdef _OPENMP
  g_threads = n < 1 ? 1 : n;
  return 0;
#else
  g_threads = 1;
  return n <= 1 ? 0 : -1;
#endif
}

#ifdef _OPENMP
/* Home cells split by the parity of their two outer indices (k, j in 3D;
 * j, i in 2D) into four colours.  The half stencil reaches offsets 0 and
 * +1 only (the full stencil writes the home particle only), so two home
 * rows of one colour never touch the same particle, and a colour's rows
 * run in parallel.  Sums then accumulate in another order than the serial
 * traversal's: a timing path, not the checker's. */
static void traverse_colored(const orc_params *p, const orc_cells *c, visit_fn visit,
                             void *ctx) {
  const int dim = c->dim;
  const int a0 = c->cstart, a1 = dim == 3 ? c->cend[2] : c->cend[1];
  const int b0 = c->cstart, b1 = dim == 3 ? c->cend[1] : c->cend[0];
  for (int color = 0; color < 4; ++color) {
    const int pa = color & 1, pb = color >> 1;
    const int na = a1 - a0 + 1 - pa > 0 ? (a1 - a0 + 1 - pa + 1) / 2 : 0;
    const int nb = b1 - b0 + 1 - pb > 0 ? (b1 - b0 + 1 - pb + 1) / 2 : 0;
#pragma omp parallel for collapse(2) schedule(dynamic) num_threads(g_threads)
    for (int ia = 0; ia < na; ++ia)
      for (int ib = 0; ib < nb; ++ib) {
        const int a = a0 + pa + 2 * ia, b = b0 + pb + 2 * ib;
        if (dim == 3) {
          for (int i = c->cstart; i <= c->cend[0]; ++i) traverse_cell(p, c, visit, ctx, i, b, a);
        } else {
          traverse_cell(p, c, visit, ctx, b, a, 0);
        }
      }
  }
}
#endif

/* shared: visit writes only the sums of ip and jp (the pair-list visitor
 * appends to one list and runs serially) */
static void traverse(const orc_params *p, const orc_cells *c, visit_fn visit,
                     void *ctx, int shared) {
  (void)shared;
#ifdef _OPENMP
  if (g_threads > 1 && shared) {
    traverse_colored(p, c, visit, ctx);
    return;
  }
#endif
  for (int k = c->cstart; k <= c->cend[2]; ++k)
    for (int j = c->cstart; j <= c->cend[1]; ++j)
      for (int i = c->cstart; i <= c->cend[0]; ++i)
        traverse_cell(p, c, visit, ctx, i, j, k);
}

/* ------------------------------------------------------------------ */
/* neighbour pair sets                                                  */
/* ------------------------------------------------------------------ */

typedef struct {
  const orc_cells *c;
  int32_t *pairs;
  int64_t cap, n;
  int overflow;
} pairs_ctx;

static void visit_pairs(void *vctx, int32_t ip, int32_t jp, double r,
                        const double *rij) {
  (void)r; (void)rij;
  pairs_ctx *pc = (pairs_ctx *)vctx;
  if (pc->n >= pc->cap) { pc->overflow = 1; return; }
  pc->pairs[2 * pc->n] = pc->c->owner[ip];
  pc->pairs[2 * pc->n + 1] = pc->c->owner[jp];
  pc->n++;
}

int orc_neighbor_pairs(const orc_params *p, const double *x, int32_t n,
                       int32_t *pairs, int64_t cap, int64_t *npairs) {
  orc_cells c;
  if (cells_build(p, x, n, &c)) return -1;
  pairs_ctx pc = {&c, pairs, cap, 0, 0};
  traverse(p, &c, visit_pairs, &pc, 0);
  cells_free(&c);
  *npairs = pc.n;
  return pc.overflow ? -1 : 0;
}

int orc_brute_pairs(const orc_params *p, const double *x, int32_t n,
                    int32_t *pairs, int64_t cap, int64_t *npairs) {
  /* Minimum image with the reference's rounding: the particle on the
   * lower side of a periodic face is the one shifted by +L (symmetric
   * mode), or the partner is shifted (full mode). */
  const int dim = p->num_dim;
  const double cut2 = p->cut_off * p->cut_off;
  int64_t m = 0;
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = i + 1; j < n; ++j) {
      double best = 1e300;
      /* try every image combination, keep the in-cutoff one if unique */
      int nimg = dim == 3 ? 27 : 9, found = 0;
      for (int img = 0; img < nimg; ++img) {
        int s[3] = {img % 3 - 1, (img / 3) % 3 - 1, dim == 3 ? img / 9 - 1 : 0};
        double d2 = 0.0;
        const int gle = le_index(p);
        for (int d = 0; d < dim; ++d) {
          double L = p->max_phys[d] - p->min_phys[d];
          double xi = x[d + (size_t)dim * i], xj = x[d + (size_t)dim * j];
          if (gle >= 0 && d != gle && s[gle] != 0) {
            /* across a Lees-Edwards face the image of the shifted particle
             * (j above: j's upper copy; below: i's upper copy with
             * symmetry, j's lower copy without) is moved in the other
             * coordinates by MODULO(x + shear_length, L) first */
            double up = (p->le_vhi[d] - p->le_vlo[d]) * p->le_time;
            double dn = (p->le_vlo[d] - p->le_vhi[d]) * p->le_time;
            if (s[gle] > 0)
              xj = f_modulo(xj + up, L);
            else if (p->symmetry)
              xi = f_modulo(xi + up, L);
            else
              xj = f_modulo(xj + dn, L);
          }
          double r;
          if (s[d] == 0) {
            r = xi - xj;
          } else if (p->symmetry) {
            /* s = +1: j seen shifted by +L, i.e. i lies above j */
            r = s[d] > 0 ? xi - (xj + L) : (xi + L) - xj;
          } else {
            r = s[d] > 0 ? xi - (xj + L) : xi - (xj - L);
          }
          d2 = d2 + r * r;
        }
        if (d2 < cut2) {
          ++found;
          if (d2 < best) best = d2;
        }
      }
      if (found > 1) return -1; /* box too small for a unique image */
      if (found) {
        if (m >= cap) return -1;
        pairs[2 * m] = i;
        pairs[2 * m + 1] = j;
        ++m;
      }
    }
  *npairs = m;
  return 0;
}

/* ------------------------------------------------------------------ */
/* density and equation of state                                        */
/* ------------------------------------------------------------------ */

typedef struct {
  const orc_params *p;
  const orc_cells *c;
  orc_kern k;
  const double *mall;
  const int32_t *sall;   /* species ids over real particles and ghosts */
  int coll_const;        /* colloid boundary particles keep a constant density */
  double *rho;
} dens_ctx;

static void visit_density(void *vctx, int32_t ip, int32_t jp, double r,
                          const double *rij) {
  (void)rij;
  dens_ctx *dc = (dens_ctx *)vctx;
  const int32_t si = dc->sall[ip], sj = dc->sall[jp];
  /* boundary particles with constant density do not contribute to each
   * other (particles_compute_density.F90:766-771) */
  if (si != 0 && sj != 0 && dc->coll_const) return;
  double w, gw;
  kern_eval(&dc->k, r, &w, &gw);
  double rhoi, rhoj;
  if (dc->p->rhs_density_type == 1) {
    rhoi = dc->mall[jp] * w;  /* rhs_density.F90:69-70 */
    rhoj = dc->mall[ip] * w;
  } else {
    rhoi = w;                 /* rhs_density.F90:95-96 */
    rhoj = w;
  }
  /* densities are summed for fluid particles and for boundary particles
   * of dynamic density (:812-848) */
  if (si == 0 || !dc->coll_const) dc->rho[ip] = dc->rho[ip] + rhoi;
  if (dc->p->symmetry && (sj == 0 || !dc->coll_const)) dc->rho[jp] = dc->rho[jp] + rhoj;
}

static int density_on_cells(const orc_params *p, const orc_cells *c,
                            const double *m, const int32_t *sid, double *rho_real) {
  dens_ctx dc;
  dc.p = p;
  dc.c = c;
  kern_init(p, &dc.k);
  double *mall = (double *)malloc(sizeof(double) * (size_t)c->nall);
  int32_t *sall = (int32_t *)malloc(sizeof(int32_t) * (size_t)c->nall);
  double *rho = (double *)calloc((size_t)c->nall, sizeof(double));
  if (!mall || !sall || !rho) { free(mall); free(sall); free(rho); return -1; }
  for (int32_t i = 0; i < c->nall; ++i) {
    mall[i] = m[c->owner[i]];
    sall[i] = sid[c->owner[i]];
  }
  double w0, g0;
  kern_eval(&dc.k, 0.0, &w0, &g0);
  dc.coll_const = p->coll && p->coll->rho_type == 0;
  /* self contribution (particles_compute_density.F90:369-558): fluid and
   * dynamic boundary particles start from W(0); constant boundary
   * particles take the initial density rho (number density rho / m) */
  for (int32_t i = 0; i < c->nreal; ++i) {
    if (sid[i] > 0 && dc.coll_const)
      rho[i] = p->rhs_density_type == 1 ? p->rho : p->rho / m[i];
    else
      rho[i] = p->rhs_density_type == 1 ? w0 * m[i] : w0;
  }
  dc.mall = mall;
  dc.sall = sall;
  dc.rho = rho;
  traverse(p, c, visit_density, &dc, 1);
  /* ghost contributions back to their owners (map_ghost_put) */
  if (p->symmetry)
    for (int32_t g = c->nreal; g < c->nall; ++g)
      rho[c->owner[g]] = rho[c->owner[g]] + rho[g];
  memcpy(rho_real, rho, sizeof(double) * (size_t)c->nreal);
  free(mall);
  free(sall);
  free(rho);
  return 0;
}

/* species ids: fluid (0) or a colloid of p->coll (1..num) */
static int check_sid(const orc_params *p, const int32_t *sid, int32_t n) {
  const int32_t nc = p->coll ? p->coll->num : 0;
  for (int32_t i = 0; i < n; ++i)
    if (sid[i] < 0 || sid[i] > nc) return -1;
  return 0;
}

int orc_density(const orc_params *p, const double *x, const double *m,
                const int32_t *sid, int32_t n, double *rho) {
  if (check_sid(p, sid, n)) return -1;
  orc_cells c;
  if (cells_build(p, x, n, &c)) return -1;
  int rc = density_on_cells(p, &c, m, sid, rho);
  cells_free(&c);
  return rc;
}

static double eos_p0(const orc_params *p) {
  /* stateEquation_new.F90:102 */
  return p->c * p->c * p->rho * pow(p->rho_ref / p->rho, p->gamma) / p->gamma;
}

int orc_pressure(const orc_params *p, const double *rho, const double *m,
                 int32_t n, double *press) {
  /* particles_compute_pressure.F90: number density is turned into mass
   * density first; stateEquation_compute_pressure.F90:158,176 */
  if (p->eos_type == 2 && (p->rho_ref == 0.0 || p->gamma == 0.0)) return -1;
  double p0 = p->eos_type == 2 ? eos_p0(p) : 0.0;
  for (int32_t i = 0; i < n; ++i) {
    double r = p->rhs_density_type == 2 ? rho[i] * m[i] : rho[i];
    if (p->eos_type == 1)
      press[i] = p->c * p->c * (pow(r, p->gamma) - p->rho_ref);
    else
      press[i] = p0 * (pow(r / p->rho_ref, p->gamma) - 1.0);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* counter-based noise                                                  */
/* ------------------------------------------------------------------ */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* natural log of u in (0, 1]: exponent split, m in [1/sqrt2, sqrt2], then
 * ln m = x P(x), x = m - 1 (exact), P of degree 9 fitted to ln(1+x)/x on
 * [1/sqrt2 - 1, sqrt2 - 1] for relative error, coefficients rounded to
 * float (Horner with fmaf, no division) */
static float noise_log(float u) {
  uint32_t b = f2u(u);
  int e = (int)(b >> 23) - 127;
  float m = u2f((b & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e = e + 1;
  }
  const float x = m - 1.0f;
  float p = fmaf(x, -0x1.38158ep-4f, 0x1.08442ap-3f);
  p = fmaf(x, p, -0x1.0f3b4ep-3f);
  p = fmaf(x, p, 0x1.2272a8p-3f);
  p = fmaf(x, p, -0x1.542702p-3f);
  p = fmaf(x, p, 0x1.99a3dcp-3f);
  p = fmaf(x, p, -0x1.000428p-2f);
  p = fmaf(x, p, 0x1.55555p-2f);
  p = fmaf(x, p, -0x1.fffff8p-2f);
  p = fmaf(x, p, 1.0f);
  return fmaf((float)e, 0.693147182f, x * p);
}

/* cos and sin of 2 pi k / 2^24, k in [0, 2^24): quadrant reduction is
 * exact in integers, the remainder |theta| <= pi/4 uses Taylor series */
static void noise_sincos(uint32_t k, float *cs, float *sn) {
  uint32_t qn = (k + (1u << 21)) >> 22;
  int32_t rem = (int32_t)k - (int32_t)(qn << 22);
  float th = (float)rem * 0x1.921fb6p-22f;
  float s = th * th;
  float ps = fmaf(s, 2.75573192e-6f, -1.98412698e-4f);
  ps = fmaf(s, ps, 8.33333333e-3f);
  ps = fmaf(s, ps, -1.66666667e-1f);
  float sv = fmaf(th * s, ps, th);
  float pc = fmaf(s, -2.75573192e-7f, 2.48015873e-5f);
  pc = fmaf(s, pc, -1.38888889e-3f);
  pc = fmaf(s, pc, 4.16666667e-2f);
  pc = fmaf(s, pc, -0.5f);
  float cv = fmaf(s, pc, 1.0f);
  switch (qn & 3u) {
    case 0: *cs = cv; *sn = sv; break;
    case 1: *cs = -sv; *sn = cv; break;
    case 2: *cs = -cv; *sn = -sv; break;
    default: *cs = sv; *sn = -cv; break;
  }
}

void orc_noise_log_n(const float *u, int64_t n, float *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = noise_log(u[i]);
}

void orc_pair_normals_n(uint64_t seed, const uint32_t *lo, const uint32_t *hi,
                        int64_t step, int64_t n, float *g) {
  for (int64_t i = 0; i < n; ++i) orc_pair_normals(seed, lo[i], hi[i], step, g + 4 * i);
}

void orc_noise_sincos_n(const uint32_t *k, int64_t n, float *cs, float *sn) {
  for (int64_t i = 0; i < n; ++i) noise_sincos(k[i], &cs[i], &sn[i]);
}

void orc_pair_normals(uint64_t seed, uint32_t pid_lo, uint32_t pid_hi,
                      int64_t step, float g[4]) {
  uint32_t ctr[4] = {pid_lo, pid_hi, (uint32_t)(uint64_t)step,
                     (uint32_t)((uint64_t)step >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  orc_philox4x32_10(ctr, key, w);
  for (int h = 0; h < 2; ++h) {
    float u = (float)((w[2 * h] >> 8) + 1u) * 0x1p-24f;
    float r = sqrtf(-2.0f * noise_log(u));
    float cs, sn;
    noise_sincos(w[2 * h + 1] >> 8, &cs, &sn);
    g[2 * h] = r * cs;
    g[2 * h + 1] = r * sn;
  }
}

/* ------------------------------------------------------------------ */
/* colloids                                                             */
/* ------------------------------------------------------------------ */

/* mcf_colloid_dist_ratio_max (mcf_header.F90:152) */
#define ORC_COLL_RATIO_MAX 0.5

static void cross3(const double *a, const double *b, double *c) {
  /* tool_cross_product.F90 */
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

int orc_colloid_nearest(const orc_params *p, double t, const double *xp, int32_t sid,
                        double *xim, double *rx, double *vim) {
  const orc_colloids *cp = p->coll;
  if (!cp || sid < 1 || sid > cp->num) return -1;
  const int dim = p->num_dim;
  const int32_t k = sid - 1;
  const double dmax = cp->radius[k] + cp->dout;
  double L[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) L[d] = p->max_phys[d] - p->min_phys[d];
  for (int d = 0; d < 3; ++d) {
    xim[d] = cp->x[d + 3 * (size_t)k];
    vim[d] = cp->v[d + 3 * (size_t)k]; /* newest level */
    rx[d] = 0.0;
  }
  for (int d = 0; d < dim; ++d) rx[d] = xp[d] - xim[d];
  const int g = le_index(p);
  if (g >= 0) {
    /* colloid_in_nearest_image.F90: the Lees-Edwards face pair first,
     * shear_length(:, upper) = (vhi - vlo) t, (:, lower) = (vlo - vhi) t */
    if (rx[g] > dmax) {
      for (int q = 0; q < dim; ++q)
        if (q == g) {
          xim[q] = xim[q] + L[q];
        } else {
          xim[q] = f_modulo(xim[q] + (p->le_vhi[q] - p->le_vlo[q]) * t, L[q]);
          vim[q] = vim[q] + (p->le_vhi[q] - p->le_vlo[q]);
        }
    } else if (-rx[g] > dmax) {
      for (int q = 0; q < dim; ++q)
        if (q == g) {
          xim[q] = xim[q] - L[q];
        } else {
          xim[q] = f_modulo(xim[q] + (p->le_vlo[q] - p->le_vhi[q]) * t, L[q]);
          vim[q] = vim[q] + (p->le_vlo[q] - p->le_vhi[q]);
        }
    }
    for (int d = 0; d < dim; ++d) rx[d] = xp[d] - xim[d];
  }
  for (int d = 0; d < dim; ++d) {
    if (rx[d] > dmax)
      xim[d] = xim[d] + L[d];
    else if (-rx[d] > dmax)
      xim[d] = xim[d] - L[d];
  }
  for (int d = 0; d < dim; ++d) rx[d] = xp[d] - xim[d];
  return 0;
}

int orc_colloid_noslip(const orc_params *p, const double *xf, const double *xc,
                       const double *vf, int32_t sid, double *vc) {
  const orc_colloids *cp = p->coll;
  if (!cp) return -1;
  const int dim = p->num_dim;
  double xcoll[3], r_xc[3], vcoll[3];
  if (orc_colloid_nearest(p, p->le_time, xc, sid, xcoll, r_xc, vcoll)) return -1;
  const int32_t k = sid - 1;
  const double *om = cp->omega + 3 * (size_t)k; /* newest level */
  for (int d = 0; d < 3; ++d) vc[d] = 0.0;
  if (cp->noslip == 1) {
    /* colloid_noslip_frozen.F90 */
    for (int d = 0; d < dim; ++d) vc[d] = vcoll[d];
    if (cp->rotate) {
      double r_vc[3];
      cross3(om, r_xc, r_vc);
      for (int d = 0; d < dim; ++d) vc[d] = vc[d] + r_vc[d];
    }
    return 0;
  }
  if (cp->noslip != 2) return -1;
  /* colloid_noslip_Morris_sphere.F90:129-300 (the disk routine has the
   * same steps) */
  const double R = cp->radius[k];
  double r_xf[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) r_xf[d] = xf[d] - xcoll[d];
  double s2 = 0.0;
  for (int d = 0; d < dim; ++d) s2 = s2 + r_xf[d] * r_xf[d];
  const double d_fcoll = sqrt(s2);
  if (d_fcoll <= R) {
    /* the fluid particle is inside: rigid velocity at the boundary particle */
    for (int d = 0; d < dim; ++d) vc[d] = vcoll[d];
    if (cp->rotate) {
      double r_vc[3];
      cross3(om, r_xc, r_vc);
      for (int d = 0; d < dim; ++d) vc[d] = vc[d] + r_vc[d];
    }
    return 0;
  }
  double c2 = 0.0;
  for (int d = 0; d < dim; ++d) c2 = c2 + r_xc[d] * r_xc[d];
  if (sqrt(c2) > R + cp->dout) return -1; /* boundary particle left its colloid */
  double nv[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) nv[d] = r_xf[d] / d_fcoll;
  double d_cn = 0.0;
  for (int d = 0; d < dim; ++d) d_cn = d_cn + r_xc[d] * nv[d];
  const double d_cs = R - d_cn;
  double d_fs = d_fcoll - R;
  double r_xs[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) r_xs[d] = R / d_fcoll * r_xf[d];
  double corr;
  if (fabs(d_fs) < ORC_MACHINE_ZERO) {
    d_fs = ORC_MACHINE_ZERO;
    corr = ORC_COLL_RATIO_MAX;
  } else {
    corr = d_cs / d_fs;
    if (corr > ORC_COLL_RATIO_MAX)
      corr = ORC_COLL_RATIO_MAX;
    else if (corr < -ORC_COLL_RATIO_MAX)
      corr = -ORC_COLL_RATIO_MAX;
  }
  double r_vs[3] = {0, 0, 0};
  if (cp->rotate) cross3(om, r_xs, r_vs);
  for (int d = 0; d < dim; ++d)
    vc[d] = -corr * (vf[d] - vcoll[d] - r_vs[d]) + vcoll[d] + r_vs[d];
  return 0;
}

int orc_colloid_collect(const orc_params *p, const double *x, const double *f,
                        const double *m, const int32_t *sid, int32_t n) {
  orc_colloids *cp = p->coll;
  if (!cp) return -1;
  const int dim = p->num_dim;
  for (int32_t q = 0; q < 3 * cp->num; ++q) {
    cp->drag[q] = 0.0;
    cp->torque[q] = 0.0;
  }
  for (int32_t j = 0; j < n; ++j) {
    if (sid[j] <= 0) continue;
    if (sid[j] > cp->num) return -1;
    const int32_t k = sid[j] - 1;
    double F[3] = {0, 0, 0}, xim[3], rx[3], vim[3], T[3];
    for (int d = 0; d < dim; ++d) F[d] = f[d + (size_t)dim * j] * m[j];
    for (int d = 0; d < dim; ++d) cp->drag[d + 3 * (size_t)k] += F[d];
    if (orc_colloid_nearest(p, cp->le_time, x + (size_t)dim * j, sid[j], xim, rx, vim)) return -1;
    cross3(rx, F, T);
    for (int d = 0; d < 3; ++d) cp->torque[d + 3 * (size_t)k] += T[d];
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* pair forces                                                          */
/* ------------------------------------------------------------------ */

typedef struct {
  const orc_params *p;
  const orc_cells *c;
  orc_kern k;
  int64_t step;
  const double *v, *rho, *press, *m, *ct; /* over xall */
  const int32_t *pid, *sid;
  double *f, *vgt;                         /* over xall */
  double *pt;                              /* (dim,dim,nall) */
  int fail;
} force_ctx;

/* An injected noise stream: the normals of one step's pairs, sorted by
 * key = pid_lo << 32 | pid_hi.  While a table is installed for step s the
 * force pass of step s takes its normals from it (a pair missing from the
 * table fails the evaluation), so that a check can run on the normals
 * another generator drew; otherwise the counter-based stream above. */
static int64_t g_noise_step = 0;
static int64_t g_noise_count = 0;
static const uint64_t *g_noise_keys = NULL;
static const float *g_noise_normals = NULL;
static int g_noise_missing = 0;

int orc_set_noise_table(int64_t step, int64_t count, const uint64_t *keys,
                        const float *normals) {
  if (count < 0 || (count > 0 && (!keys || !normals))) return -1;
  g_noise_step = step;
  g_noise_count = count;
  g_noise_keys = count > 0 ? keys : NULL;
  g_noise_normals = count > 0 ? normals : NULL;
  g_noise_missing = 0;
  return 0;
}

static void force_normals(uint64_t seed, uint32_t lo, uint32_t hi, int64_t step, float g[4]) {
  if (g_noise_keys && step == g_noise_step) {
    const uint64_t key = ((uint64_t)lo << 32) | hi;
    int64_t a = 0, b = g_noise_count;
    while (a < b) {
      const int64_t mid = a + (b - a) / 2;
      if (g_noise_keys[mid] < key) a = mid + 1; else b = mid;
    }
    if (a < g_noise_count && g_noise_keys[a] == key) {
      for (int q = 0; q < 4; ++q) g[q] = g_noise_normals[4 * a + q];
      return;
    }
    g_noise_missing = 1;
    for (int q = 0; q < 4; ++q) g[q] = 0.0f;
    return;
  }
  orc_pair_normals(seed, lo, hi, step, g);
}

/* Random force direction for the ordered pair (ip -> jp): a Gaussian
 * vector with the covariance of the reference's symmetric traceless
 * dW . e_ij, oriented by the lower pid (antisymmetric in ip <-> jp). */
static void noise_vector(const force_ctx *fc, int32_t ip, int32_t jp,
                         const double *eij, double *we) {
  const int dim = fc->p->num_dim;
  uint32_t a = (uint32_t)fc->pid[ip], b = (uint32_t)fc->pid[jp];
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  float g[4];
  force_normals(fc->p->seed, lo, hi, fc->step, g);
  double sgn = a < b ? 1.0 : -1.0;
  double ec[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) ec[d] = sgn * eij[d];
  const double ca = 0.70710678118654752440; /* sqrt(1/2) */
  if (dim == 3) {
    const double cb = 0.81649658092772603273 - 0.70710678118654752440;
    double ge = (double)g[0] * ec[0] + (double)g[1] * ec[1] + (double)g[2] * ec[2];
    for (int d = 0; d < 3; ++d)
      we[d] = sgn * (ca * (double)g[d] + cb * ge * ec[d]);
  } else {
    for (int d = 0; d < 2; ++d) we[d] = sgn * (ca * (double)g[d]);
  }
}

/* rhs_force_ff_Newtonian_Espanol.F90:198-249: the reference draws a full
 * dW, takes its trace / num_dim and its symmetric traceless part S.  The
 * two are independent Gaussians: S e has the covariance noise_vector
 * draws, and the trace is N(0, 1 / num_dim), here the fourth normal of the
 * same pair block over sqrt(num_dim). */
static double noise_trace(const force_ctx *fc, int32_t ip, int32_t jp) {
  uint32_t a = (uint32_t)fc->pid[ip], b = (uint32_t)fc->pid[jp];
  float g[4];
  force_normals(fc->p->seed, a < b ? a : b, a < b ? b : a, fc->step, g);
  return (double)g[3] / sqrt((double)fc->p->num_dim);
}

static void visit_force(void *vctx, int32_t ip, int32_t jp, double dij,
                        const double *rij) {
  force_ctx *fc = (force_ctx *)vctx;
  const orc_params *p = fc->p;
  const int dim = p->num_dim;
  const double *vi = fc->v + (size_t)dim * ip, *vj = fc->v + (size_t)dim * jp;
  double pi = fc->press[ip], pj = fc->press[jp];
  double vbuf[3];
  const int32_t si = fc->sid[ip], sj = fc->sid[jp];
  if (si != 0 || sj != 0) {
    /* pp_interaction.inc: two boundary particles do not interact (same
     * colloid: CYCLE; different colloids: pp_interact_cc off) */
    if (si != 0 && sj != 0) return;
    const double *xi = fc->c->xall + (size_t)dim * ip;
    const double *xj = fc->c->xall + (size_t)dim * jp;
    if (sj != 0) {
      /* pp_interaction_fc.inc: jp takes the no-slip velocity and ip's
       * pressure */
      if (orc_colloid_noslip(p, xi, xj, vi, sj, vbuf)) { fc->fail = 1; return; }
      vj = vbuf;
      pj = pi;
    } else {
      /* pp_interaction_cf.inc: the same with the roles swapped */
      if (orc_colloid_noslip(p, xj, xi, vj, si, vbuf)) { fc->fail = 1; return; }
      vi = vbuf;
      pi = pj;
    }
  }
  double w, gradw;
  kern_eval(&fc->k, dij, &w, &gradw);
  double numi = fc->rho[ip], numj = fc->rho[jp];
  double mi = fc->m[ip], mj = fc->m[jp];
  if (mi < ORC_MACHINE_ZERO || mj < ORC_MACHINE_ZERO ||
      numi < ORC_MACHINE_ZERO || numj < ORC_MACHINE_ZERO) {
    fc->fail = 1;
    return;
  }
  double eij[3], vij[3], fi[3] = {0, 0, 0}, fj[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    eij[d] = rij[d] / dij;
    vij[d] = vi[d] - vj[d];
  }
  const int brown = p->brownian && p->kt > ORC_MACHINE_ZERO;
  if (!p->newtonian) {
    const int d2n = dim * dim;
    const double *pti = fc->pt + (size_t)d2n * ip, *ptj = fc->pt + (size_t)d2n * jp;
    if (si != 0 || sj != 0) {
      /* pp_interaction_fc.inc:80-113, pp_interaction_cf.inc:80-110: in a
       * non-Newtonian flow both sides of a fluid-boundary pair take the
       * fluid particle's number density (rho_ip = rho_jp) and pressure
       * tensor (pt of the boundary particle := pt of the fluid one), for
       * the velocity gradients and the force alike */
      if (sj != 0) {
        numj = numi;
        ptj = pti;
      } else {
        numi = numj;
        pti = ptj;
      }
    }
    /* pp_vgt_ip.inc / pp_vgt_jp.inc */
    const double *xi = fc->c->xall + (size_t)dim * ip;
    const double *xj = fc->c->xall + (size_t)dim * jp;
    for (int bb = 0; bb < dim; ++bb)
      for (int aa = 0; aa < dim; ++aa)
        fc->vgt[aa + dim * bb + (size_t)d2n * ip] +=
            (vj[aa] - vi[aa]) * (xi[bb] - xj[bb]) * gradw / numi / dij;
    if (p->symmetry)
      for (int bb = 0; bb < dim; ++bb)
        for (int aa = 0; aa < dim; ++aa)
          fc->vgt[aa + dim * bb + (size_t)d2n * jp] +=
              (vi[aa] - vj[aa]) * (xj[bb] - xi[bb]) * gradw / numj / dij;
    /* rhs_force_ff_non_Newtonian_HuAdams.F90:105-131 (type 3) and
     * rhs_force_ff_non_Newtonian_Espanol.F90 (type 2, the same but for its
     * thermal force); rhs_force_ff_non_Newtonian_HuAdams_angular.F90 (type
     * 4) averages the diagonal, (Pi_i + Pi_j) / 2 (1/n_i^2 + 1/n_j^2), and
     * dissipates along e */
    const int ang = p->rhs_force_type == 4;
    double fcv[3] = {0, 0, 0};
    for (int jj = 0; jj < dim; ++jj)
      for (int ii = 0; ii < dim; ++ii) {
        if (ang && ii == jj) {
          double pij = (pti[ii + dim * jj] + ptj[ii + dim * jj]) / 2.0;
          fcv[jj] = fcv[jj] + pij * eij[ii] / (numi * numi) + pij * eij[ii] / (numj * numj);
        } else {
          fcv[jj] = fcv[jj] + pti[ii + dim * jj] * eij[ii] / (numi * numi) +
                    ptj[ii + dim * jj] * eij[ii] / (numj * numj);
        }
      }
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] - fcv[d] * gradw / mi;
      fj[d] = fj[d] + fcv[d] * gradw / mj;
    }
    if (ang) {
      double ev = 0.0;
      for (int d = 0; d < dim; ++d) ev = ev + eij[d] * vij[d];
      double f_da = p->eta_coef * p->eta * (1.0 / (numi * numi) + 1.0 / (numj * numj)) *
                    gradw * ev / dij;
      for (int d = 0; d < dim; ++d) {
        fi[d] = fi[d] + f_da * eij[d] / mi;
        fj[d] = fj[d] - f_da * eij[d] / mj;
      }
      if (brown) { fc->fail = 1; return; } /* "Brownian force is not available" */
      goto accumulate;
    }
    double f_d = p->eta * (1.0 / (numi * numi) + 1.0 / (numj * numj)) * gradw / dij;
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] + f_d * vij[d] / mi;
      fj[d] = fj[d] - f_d * vij[d] / mj;
    }
    if (brown && p->rhs_force_type == 2) {
      double we[3];
      noise_vector(fc, ip, jp, eij, we);
      const double tr = noise_trace(fc, ip, jp);
      double zij = 4.0 * p->kt * gradw / dij / numi / numj;
      double a = 5.0 * p->eta / 3.0;
      double b = ((double)dim + 2.0) * p->eta / 3.0;
      double aij = sqrt(-zij * a);
      double bij = sqrt(-zij * (double)dim / 2.0 * (b + a * (2.0 / (double)dim - 1.0)));
      for (int d = 0; d < dim; ++d) {
        double f_r = (aij * we[d] + bij * tr * eij[d]) / sqrt(p->dt);
        fi[d] = fi[d] + f_r / mi;
        fj[d] = fj[d] - f_r / mj;
      }
    } else if (brown) {
      if (f_d > 0.0) { fc->fail = 1; return; }
      double we[3];
      noise_vector(fc, ip, jp, eij, we);
      double f_r = sqrt(-4.0 * p->kt * f_d) / sqrt(p->dt);
      for (int d = 0; d < dim; ++d) {
        fi[d] = fi[d] + f_r * we[d] / mi;
        fj[d] = fj[d] - f_r * we[d] / mj;
      }
    }
  } else if (p->rhs_force_type == 3) {
    /* rhs_force_ff_Newtonian_HuAdams.F90:140-237 */
    double f_c = (pi / (numi * numi) + pj / (numj * numj)) * gradw;
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] - f_c * eij[d] / mi;
      fj[d] = fj[d] + f_c * eij[d] / mj;
    }
    double f_d = p->eta * (1.0 / (numi * numi) + 1.0 / (numj * numj)) * gradw / dij;
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] + f_d * vij[d] / mi;
      fj[d] = fj[d] - f_d * vij[d] / mj;
    }
    if (brown) {
      if (f_d > 0.0) { fc->fail = 1; return; }
      double we[3];
      noise_vector(fc, ip, jp, eij, we);
      double f_r = sqrt(-4.0 * p->kt * f_d) / sqrt(p->dt);
      for (int d = 0; d < dim; ++d) {
        fi[d] = fi[d] + f_r * we[d] / mi;
        fj[d] = fj[d] - f_r * we[d] / mj;
      }
    }
  } else if (p->rhs_force_type == 1) {
    /* rhs_force_ff_Newtonian_Morris.F90:112-136: accelerations carry the
     * partner's mass; rho is the density pass's density */
    double f_c = (pi / (numi * numi) + pj / (numj * numj)) * gradw;
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] - mj * f_c * eij[d];
      fj[d] = fj[d] + mi * f_c * eij[d];
    }
    double f_d = 2.0 * p->eta * gradw / numi / numj / dij;
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] + mj * f_d * vij[d];
      fj[d] = fj[d] - mi * f_d * vij[d];
    }
  } else if (p->rhs_force_type == 2) {
    /* rhs_force_ff_Newtonian_Espanol.F90:152-251 */
    double ev = 0.0;
    for (int d = 0; d < dim; ++d) ev = ev + eij[d] * vij[d];
    double f_c = (pi / (numi * numi) + pj / (numj * numj)) * gradw;
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] - f_c * eij[d] / mi;
      fj[d] = fj[d] + f_c * eij[d] / mj;
    }
    double f_d1 = (2.0 - 1.0 / dim) * p->eta * gradw / numi / numj / dij;
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] + f_d1 * vij[d] / mi;
      fj[d] = fj[d] - f_d1 * vij[d] / mj;
    }
    double f_d2 = p->eta * (dim + 2) / dim * gradw * ev / numi / numj / dij;
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] + f_d2 * eij[d] / mi;
      fj[d] = fj[d] - f_d2 * eij[d] / mj;
    }
    if (brown) {
      double we[3];
      noise_vector(fc, ip, jp, eij, we);
      const double tr = noise_trace(fc, ip, jp);
      double zij = 4.0 * p->kt * gradw / dij / numi / numj;
      double a = 5.0 * p->eta / 3.0;
      double b = ((double)dim + 2.0) * p->eta / 3.0;
      double aij = sqrt(-zij * a);
      double bij = sqrt(-zij * (double)dim / 2.0 * (b + a * (2.0 / (double)dim - 1.0)));
      for (int d = 0; d < dim; ++d) {
        double f_r = (aij * we[d] + bij * tr * eij[d]) / sqrt(p->dt);
        fi[d] = fi[d] + f_r / mi;
        fj[d] = fj[d] - f_r / mj;
      }
    }
  } else {
    /* rhs_force_ff_Newtonian_HuAdams_angular.F90:166-229 */
    double ev = 0.0;
    for (int d = 0; d < dim; ++d) ev = ev + eij[d] * vij[d];
    double pij = (pi + pj) / 2.0;
    double aij = (1.0 / (numi * numi) + 1.0 / (numj * numj)) * gradw;
    double f_c = aij * pij;
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] - f_c * eij[d] / mi;
      fj[d] = fj[d] + f_c * eij[d] / mj;
    }
    double f_d = p->eta_coef * p->eta * aij * ev / dij;
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] + f_d * eij[d] / mi;
      fj[d] = fj[d] - f_d * eij[d] / mj;
    }
    if (brown) {
      uint32_t a = (uint32_t)fc->pid[ip], b = (uint32_t)fc->pid[jp];
      float g[4];
      force_normals(p->seed, a < b ? a : b, a < b ? b : a, fc->step, g);
      double eps = (double)g[0];
      double f_r = eps * sqrt(2.0 * p->kt * p->eta) *
                   sqrt(-p->eta_coef * aij / dij) / sqrt(p->dt);
      for (int d = 0; d < dim; ++d) {
        fi[d] = fi[d] + f_r * eij[d] / mi;
        fj[d] = fj[d] - f_r * eij[d] / mj;
      }
    }
  }
accumulate:;
  double *f = fc->f;
  for (int d = 0; d < dim; ++d) f[d + (size_t)dim * ip] += fi[d];
  if (p->symmetry)
    for (int d = 0; d < dim; ++d) f[d + (size_t)dim * jp] += fj[d];
}

static int interaction_on_cells(const orc_params *p, const orc_cells *c,
                                int64_t step, const double *v,
                                const double *rho, const double *press,
                                const double *m, const int32_t *pid,
                                const int32_t *sid, const double *ct,
                                double *f_real, double *vgt_real) {
  const int dim = p->num_dim, d2n = dim * dim;
  const int32_t na = c->nall;
  force_ctx fc;
  memset(&fc, 0, sizeof(fc));
  fc.p = p;
  fc.c = c;
  fc.step = step;
  kern_init(p, &fc.k);
  double *va = (double *)malloc(sizeof(double) * (size_t)dim * na);
  double *ra = (double *)malloc(sizeof(double) * (size_t)na);
  double *pa = (double *)malloc(sizeof(double) * (size_t)na);
  double *ma = (double *)malloc(sizeof(double) * (size_t)na);
  int32_t *ia = (int32_t *)malloc(sizeof(int32_t) * (size_t)na);
  int32_t *sa = (int32_t *)malloc(sizeof(int32_t) * (size_t)na);
  double *fa = (double *)calloc((size_t)dim * na, sizeof(double));
  double *vg = (double *)calloc((size_t)d2n * na, sizeof(double));
  double *pt = (double *)calloc((size_t)d2n * na, sizeof(double));
  if (!va || !ra || !pa || !ma || !ia || !sa || !fa || !vg || !pt) {
    free(va); free(ra); free(pa); free(ma); free(ia); free(sa); free(fa); free(vg); free(pt);
    return -1;
  }
  /* map_ghost_get(x, rho, v, ct): ghosts carry their owner's values */
  for (int32_t i = 0; i < na; ++i) {
    int32_t o = c->owner[i];
    for (int d = 0; d < dim; ++d) va[d + (size_t)dim * i] = v[d + (size_t)dim * o];
    ra[i] = rho[o];
    pa[i] = press[o];
    ma[i] = m[o];
    ia[i] = pid[o];
    sa[i] = sid[o];
  }
  {
    /* Lees-Edwards images move with the face they crossed
     * (particles_set_boundary_ghost_velocity.F90:175-240) */
    const int gle = le_index(p);
    if (gle >= 0)
      for (int32_t i = c->nreal; i < na; ++i) {
        double xg = c->xall[gle + (size_t)dim * i];
        if (xg < p->min_phys[gle]) {
          for (int k = 0; k < dim; ++k)
            if (k != gle)
              va[k + (size_t)dim * i] = va[k + (size_t)dim * i] + p->le_vlo[k] - p->le_vhi[k];
        } else if (xg >= p->max_phys[gle]) {
          for (int k = 0; k < dim; ++k)
            if (k != gle)
              va[k + (size_t)dim * i] = va[k + (size_t)dim * i] + p->le_vhi[k] - p->le_vlo[k];
        }
      }
  }
  if (!p->newtonian) {
    /* particles_compute_pressure_tensor.F90:114-133, G = n_p kt_p */
    double G = p->n_p * p->kt_p;
    for (int32_t i = 0; i < na; ++i) {
      int32_t o = c->owner[i];
      for (int jj = 0; jj < dim; ++jj)
        for (int ii = 0; ii < dim; ++ii) {
          double val = -G * ct[ii + dim * jj + (size_t)d2n * o];
          if (ii == jj) val = val + pa[i] + G;
          pt[ii + dim * jj + (size_t)d2n * i] = val;
        }
    }
  }
  fc.v = va; fc.rho = ra; fc.press = pa; fc.m = ma; fc.pid = ia; fc.sid = sa;
  fc.ct = ct; fc.f = fa; fc.vgt = vg; fc.pt = pt;
  g_noise_missing = 0;
  traverse(p, c, visit_force, &fc, 1);
  if (g_noise_missing) fc.fail = 1;
  if (p->symmetry)
    for (int32_t g = c->nreal; g < na; ++g) {
      int32_t o = c->owner[g];
      for (int d = 0; d < dim; ++d)
        fa[d + (size_t)dim * o] += fa[d + (size_t)dim * g];
      if (!p->newtonian)
        for (int q = 0; q < d2n; ++q)
          vg[q + (size_t)d2n * o] += vg[q + (size_t)d2n * g];
    }
  /* particles_apply_body_force.F90:125-193 (type 1), fluid particles only */
  if (p->body_force_type == 1)
    for (int d = 0; d < dim; ++d)
      if (fabs(p->body_force[d]) > ORC_MACHINE_ZERO)
        for (int32_t i = 0; i < c->nreal; ++i)
          if (sid[i] == 0) fa[d + (size_t)dim * i] += p->body_force[d];
  /* :196-273 (type 2): +g_x where the last coordinate lies below
   * max_phys(num_dim) / 2, -g_x above it; :276-353 (type 3): g_x sin(k y)
   * with k = 2 pi / (max_phys(2) - min_phys(2)); fluid particles only */
  if ((p->body_force_type == 2 || p->body_force_type == 3) &&
      fabs(p->body_force[0]) > ORC_MACHINE_ZERO) {
    const double half = p->max_phys[dim - 1] * 0.5;
    const double k = 2.0 * ORC_PI / (p->max_phys[1] - p->min_phys[1]);
    for (int32_t i = 0; i < c->nreal; ++i) {
      if (sid[i] != 0) continue;
      const double *xi = c->xall + (size_t)dim * i;
      double *fi = fa + (size_t)dim * i;
      if (p->body_force_type == 2) {
        if (xi[dim - 1] < half)
          fi[0] += p->body_force[0];
        else if (xi[dim - 1] > half)
          fi[0] -= p->body_force[0];
      } else {
        fi[0] += p->body_force[0] * sin(k * xi[1]);
      }
    }
  }
  memcpy(f_real, fa, sizeof(double) * (size_t)dim * c->nreal);
  if (!p->newtonian && vgt_real)
    memcpy(vgt_real, vg, sizeof(double) * (size_t)d2n * c->nreal);
  int fail = fc.fail;
  free(va); free(ra); free(pa); free(ma); free(ia); free(sa); free(fa); free(vg); free(pt);
  return fail ? -1 : 0;
}

int orc_interaction(const orc_params *p, int64_t step, const double *x,
                    const double *v, const double *rho, const double *press,
                    const double *m, const int32_t *pid, const int32_t *sid,
                    const double *ct, int32_t n, double *f, double *vgt) {
  if (check_sid(p, sid, n)) return -1;
  if (p->rhs_force_type < 1 || p->rhs_force_type > 4) return -1;
  if (!p->newtonian && p->rhs_force_type == 1) return -1;
  orc_cells c;
  if (cells_build(p, x, n, &c)) return -1;
  int rc = interaction_on_cells(p, &c, step, v, rho, press, m, pid, sid, ct, f, vgt);
  cells_free(&c);
  return rc;
}

/* ------------------------------------------------------------------ */
/* Oldroyd-B, time integration, statistics                             */
/* ------------------------------------------------------------------ */

int orc_compute_act(const orc_params *p, const double *ct, const double *vgt,
                    int32_t n, double *act) {
  const int dim = p->num_dim, d2n = dim * dim;
  if (fabs(p->tau) < ORC_MACHINE_ZERO) return -1;
  for (int32_t q = 0; q < n; ++q) {
    const double *c = ct + (size_t)d2n * q, *L = vgt + (size_t)d2n * q;
    for (int j = 0; j < dim; ++j)
      for (int i = 0; i < dim; ++i) {
        double t = 0.0;
        for (int k = 0; k < dim; ++k)
          t = t + c[i + dim * k] * L[j + dim * k] + L[i + dim * k] * c[k + dim * j];
        double id = i == j ? 1.0 : 0.0;
        act[i + dim * j + (size_t)d2n * q] = t - (c[i + dim * j] - id) / p->tau;
      }
  }
  return 0;
}

static void integrate_velocity(const orc_params *p, int32_t n, double *v,
                               const double *f, const int32_t *sid,
                               double lambda) {
  const int dim = p->num_dim;
  for (int32_t i = 0; i < n; ++i)
    if (sid[i] == 0)
      for (int d = 0; d < dim; ++d)
        v[d + (size_t)dim * i] = v[d + (size_t)dim * i] +
                                 lambda * f[d + (size_t)dim * i] * p->dt;
}

static void integrate_position(const orc_params *p, int32_t n, double *x,
                               const double *v, const int32_t *sid) {
  const int dim = p->num_dim;
  for (int32_t i = 0; i < n; ++i)
    if (sid[i] == 0)
      for (int d = 0; d < dim; ++d)
        x[d + (size_t)dim * i] = x[d + (size_t)dim * i] + v[d + (size_t)dim * i] * p->dt;
}

/* particles_adjust_particles.F90:137-389, periodic and Lees-Edwards
 * faces: crossing a Lees-Edwards face moves the other coordinates by
 * MODULO(x - shear_length, L) and adds the face velocity jump, and ends
 * the particle's check (EXIT). */
static void adjust_particles(const orc_params *p, int32_t n, double *x, double *v) {
  const int dim = p->num_dim;
  const int gle = le_index(p);
  for (int32_t i = 0; i < n; ++i)
    for (int d = 0; d < dim; ++d) {
      double L = p->max_phys[d] - p->min_phys[d];
      double *xv = x + d + (size_t)dim * i;
      if (*xv < p->min_phys[d]) {
        if (d == gle) {
          *xv = *xv + L;
          for (int k = 0; k < dim; ++k)
            if (k != d) {
              double Lk = p->max_phys[k] - p->min_phys[k];
              double sl = (p->le_vlo[k] - p->le_vhi[k]) * p->le_time; /* shear_length(k, lower) */
              x[k + (size_t)dim * i] = f_modulo(x[k + (size_t)dim * i] - sl, Lk);
              v[k + (size_t)dim * i] = v[k + (size_t)dim * i] + (p->le_vhi[k] - p->le_vlo[k]);
            }
          break;
        }
        *xv = *xv + L;
      }
      if (*xv >= p->max_phys[d]) {
        if (d == gle) {
          *xv = *xv - L;
          for (int k = 0; k < dim; ++k)
            if (k != d) {
              double Lk = p->max_phys[k] - p->min_phys[k];
              double sl = (p->le_vhi[k] - p->le_vlo[k]) * p->le_time; /* shear_length(k, upper) */
              x[k + (size_t)dim * i] = f_modulo(x[k + (size_t)dim * i] - sl, Lk);
              v[k + (size_t)dim * i] = v[k + (size_t)dim * i] + (p->le_vlo[k] - p->le_vhi[k]);
            }
          break;
        }
        *xv = *xv - L;
      }
    }
}

static void integrate_ct(const orc_params *p, int32_t n, double *ct,
                         const double *act, double lambda) {
  const size_t tot = (size_t)p->num_dim * p->num_dim * n;
  for (size_t q = 0; q < tot; ++q) ct[q] = ct[q] + act[q] * p->dt * lambda;
}

/* Adams-Bashforth increment of order `order` from a history (3, num, ab),
 * newest first, for colloid k (colloid_integrate_translate_position.F90
 * and its siblings use these coefficients). */
static void ab_increment(const double *h, int32_t num, int32_t k, int order, double dt,
                         double *out) {
  const size_t lv = 3 * (size_t)num;
  for (int d = 0; d < 3; ++d) {
    const double *a = h + d + 3 * (size_t)k;
    switch (order) {
      case 1: out[d] = a[0] * dt; break;
      case 2: out[d] = (3.0 * a[0] - a[lv]) * dt / 2.0; break;
      case 3: out[d] = (23.0 * a[0] - 16.0 * a[lv] + 5.0 * a[2 * lv]) * dt / 12.0; break;
      case 4:
        out[d] = (55.0 * a[0] - 59.0 * a[lv] + 37.0 * a[2 * lv] - 9.0 * a[3 * lv]) * dt / 24.0;
        break;
      default:
        out[d] = (1901.0 * a[0] - 2774.0 * a[lv] + 2616.0 * a[2 * lv] - 1274.0 * a[3 * lv] +
                  251.0 * a[4 * lv]) * dt / 720.0;
        break;
    }
  }
}

/* history levels move one back: h(:,:,i) = h(:,:,i-1), i = ab..2 */
static void history_shift(double *h, int32_t num, int32_t ab) {
  const size_t lv = 3 * (size_t)num;
  for (int32_t i = ab - 1; i >= 1; --i) memcpy(h + i * lv, h + (i - 1) * lv, lv * sizeof(double));
}

/* tool_rotation_matrix.F90: rotation by phi about the unit axis ax,
 * column-major rm(3,3) */
static void rotation_matrix(const double *ax, double phi, double *rm) {
  const double cp = cos(phi), cp1 = 1.0 - cp, sp = sin(phi);
  rm[0 + 3 * 0] = ax[0] * ax[0] * cp1 + cp;
  rm[0 + 3 * 1] = ax[0] * ax[1] * cp1 - ax[2] * sp;
  rm[0 + 3 * 2] = ax[0] * ax[2] * cp1 + ax[1] * sp;
  rm[1 + 3 * 0] = ax[1] * ax[0] * cp1 + ax[2] * sp;
  rm[1 + 3 * 1] = ax[1] * ax[1] * cp1 + cp;
  rm[1 + 3 * 2] = ax[1] * ax[2] * cp1 - ax[0] * sp;
  rm[2 + 3 * 0] = ax[2] * ax[0] * cp1 - ax[1] * sp;
  rm[2 + 3 * 1] = ax[2] * ax[1] * cp1 + ax[0] * sp;
  rm[2 + 3 * 2] = ax[2] * ax[2] * cp1 + cp;
}

/* colloid_adjust_colloid.F90: centres leaving the box (periodic and
 * Lees-Edwards faces, shear length at the colloid block's time) */
static void adjust_colloids(const orc_params *p) {
  orc_colloids *cp = p->coll;
  const int dim = p->num_dim, g = le_index(p);
  const double t = cp->le_time;
  for (int32_t j = 0; j < cp->num; ++j) {
    double *X = cp->x + 3 * (size_t)j, *V = cp->v + 3 * (size_t)j;
    for (int d = 0; d < dim; ++d) {
      const double L = p->max_phys[d] - p->min_phys[d];
      if (X[d] < p->min_phys[d]) {
        if (d == g) {
          X[d] = X[d] + L;
          for (int q = 0; q < dim; ++q)
            if (q != d) {
              double Lq = p->max_phys[q] - p->min_phys[q];
              X[q] = f_modulo(X[q] - (p->le_vlo[q] - p->le_vhi[q]) * t, Lq);
              V[q] = V[q] + (p->le_vhi[q] - p->le_vlo[q]);
            }
          break;
        }
        const double cross = (p->max_phys[d] - X[d]) / L;
        X[d] = X[d] + L * (double)(int)cross;
      }
      if (X[d] >= p->max_phys[d]) {
        if (d == g) {
          X[d] = X[d] - L;
          for (int q = 0; q < dim; ++q)
            if (q != d) {
              double Lq = p->max_phys[q] - p->min_phys[q];
              X[q] = f_modulo(X[q] - (p->le_vhi[q] - p->le_vlo[q]) * t, Lq);
              V[q] = V[q] + (p->le_vlo[q] - p->le_vhi[q]);
            }
          break;
        }
        const double cross = (X[d] - p->min_phys[d]) / L;
        X[d] = X[d] - L * (double)(int)cross;
      }
    }
  }
}

/* The colloid block of marching_integrate_VV.F90:455-765 with the
 * explicit integrators (integrate_colloid_type 1, 2) and no colloid-colloid
 * forces: drag and torque from the boundary particles' last forces, then
 * per sub-step the rigid-body update of the centres, velocities and
 * rotations, carrying the boundary particles along. */
static int colloid_block(const orc_params *p, int64_t step, int32_t n, double *x,
                         const double *f, const double *m, const int32_t *sid) {
  orc_colloids *cp = p->coll;
  if (orc_colloid_collect(p, x, f, m, sid, n)) return -1;
  if (!cp->translate && !cp->rotate) return 0;
  const int dim = p->num_dim;
  const int32_t nc = cp->num, ab = cp->ab;
  if (ab < 1 || ab > 5 || cp->sub_steps < 1) return -1;
  const size_t lv = 3 * (size_t)nc;
  double *drag0 = (double *)malloc(lv * sizeof(double));
  double *torque0 = (double *)malloc(lv * sizeof(double));
  if (!drag0 || !torque0) { free(drag0); free(torque0); return -1; }
  memcpy(drag0, cp->drag, lv * sizeof(double));
  memcpy(torque0, cp->torque, lv * sizeof(double));
  const double dts = p->dt / cp->sub_steps;
  int rc = 0;
  for (int i = 1; i <= cp->sub_steps && !rc; ++i) {
    /* the AB order ramps up with the step index passed by the caller;
     * the rotating velocity is passed step - 1 + i + step_start */
    const int64_t st = step - 1 + i - cp->step_start;
    const int64_t sw = step - 1 + i + cp->step_start;
    const int ord = (int)(st < ab ? st : ab), ordw = (int)(sw < ab ? sw : ab);
    if (ord < 1 || ordw < 1) { rc = -1; break; }
    if (cp->rotate)
      for (int32_t k = 0; k < nc; ++k) {
        /* colloid_compute_rotation_vector.F90 */
        double axis[3];
        ab_increment(cp->omega, nc, k, ord, dts, axis);
        double phi = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
        if (phi < ORC_MACHINE_ZERO) {
          phi = 0.0;
          axis[0] = 1.0;
          axis[1] = axis[2] = 0.0;
        } else {
          for (int d = 0; d < 3; ++d) axis[d] = axis[d] / phi;
        }
        cp->theta[2 + 3 * (size_t)k] += phi;
        /* colloid_compute_rotation_matrix.F90, _accumulation_matrix.F90 */
        double *R = cp->rot_matrix + 9 * (size_t)k, *A = cp->acc_matrix + 9 * (size_t)k;
        rotation_matrix(axis, phi, R);
        double RA[9];
        for (int c = 0; c < 3; ++c)
          for (int r = 0; r < 3; ++r)
            RA[r + 3 * c] = R[r + 0] * A[0 + 3 * c] + R[r + 3] * A[1 + 3 * c] + R[r + 6] * A[2 + 3 * c];
        memcpy(A, RA, sizeof(RA));
      }
    /* particles_compute_colloid_relative_position.F90 */
    for (int32_t j = 0; j < n && !rc; ++j) {
      if (sid[j] <= 0) continue;
      double xim[3], rx[3], vim[3];
      double *xj = x + (size_t)dim * j;
      if (orc_colloid_nearest(p, cp->le_time, xj, sid[j], xim, rx, vim)) { rc = -1; break; }
      if (cp->rotate) {
        const double *R = cp->rot_matrix + 9 * (size_t)(sid[j] - 1);
        for (int r = 0; r < dim; ++r) {
          double t = 0.0;
          for (int c = 0; c < dim; ++c) t = t + R[r + 3 * c] * rx[c];
          xj[r] = t;
        }
      } else {
        for (int d = 0; d < dim; ++d) xj[d] = rx[d];
      }
    }
    if (rc) break;
    if (cp->rotate) {
      /* colloid_integrate_rotate_velocity.F90 */
      history_shift(cp->omega, nc, ab);
      for (int32_t k = 0; k < nc; ++k) {
        double inc[3];
        ab_increment(cp->alpha, nc, k, ordw, dts, inc);
        for (int d = 0; d < 3; ++d) cp->omega[d + 3 * (size_t)k] += inc[d];
      }
    }
    if (cp->translate) {
      /* colloid_integrate_translate_position.F90, _velocity.F90 */
      for (int32_t k = 0; k < nc; ++k) {
        double inc[3];
        ab_increment(cp->v, nc, k, ord, dts, inc);
        for (int d = 0; d < dim; ++d) cp->x[d + 3 * (size_t)k] += inc[d];
      }
      history_shift(cp->v, nc, ab);
      for (int32_t k = 0; k < nc; ++k) {
        double inc[3];
        ab_increment(cp->f, nc, k, ord, dts, inc);
        for (int d = 0; d < dim; ++d) cp->v[d + 3 * (size_t)k] += inc[d];
      }
    }
    /* colloid_compute_interaction.F90 (no colloid-colloid or wall forces)
     * and colloid_apply_body_force.F90 */
    memcpy(cp->drag, drag0, lv * sizeof(double));
    memcpy(cp->torque, torque0, lv * sizeof(double));
    if (cp->body_force_type == 1)
      for (int32_t k = 0; k < nc; ++k)
        for (int d = 0; d < dim; ++d) cp->drag[d + 3 * (size_t)k] += cp->body_force[d] * cp->m[k];
    /* colloid_compute_translate_acceleration.F90 */
    history_shift(cp->f, nc, ab);
    for (int32_t k = 0; k < nc; ++k)
      for (int d = 0; d < dim; ++d)
        cp->f[d + 3 * (size_t)k] = cp->translate ? cp->drag[d + 3 * (size_t)k] / cp->m[k] : 0.0;
    adjust_colloids(p);
    /* colloid_compute_rotate_acceleration.F90 (disks and spheres) */
    history_shift(cp->alpha, nc, ab);
    for (int32_t k = 0; k < nc; ++k) {
      double *al = cp->alpha + 3 * (size_t)k;
      const double *tq = cp->torque + 3 * (size_t)k, *mi = cp->mmi + 3 * (size_t)k;
      if (!cp->rotate) {
        al[0] = al[1] = al[2] = 0.0;
      } else if (dim == 2) {
        al[2] = tq[2] / mi[2];
      } else {
        for (int d = 0; d < 3; ++d) al[d] = tq[d] / mi[d];
      }
    }
    /* particles_compute_colloid_absolute_position.F90 */
    for (int32_t j = 0; j < n; ++j) {
      if (sid[j] <= 0) continue;
      const double *X = cp->x + 3 * (size_t)(sid[j] - 1);
      for (int d = 0; d < dim; ++d) x[d + (size_t)dim * j] += X[d];
    }
  }
  free(drag0);
  free(torque0);
  return rc;
}

static int evaluate(const orc_params *p, int64_t step, int32_t n,
                    const double *x, const double *v, const double *m,
                    const int32_t *pid, double *rho, double *press, double *f,
                    const int32_t *sid, double *ct, double *act, double *vgt) {
  orc_cells c;
  if (cells_build(p, x, n, &c)) return -1;
  int rc = density_on_cells(p, &c, m, sid, rho);
  if (!rc) rc = orc_pressure(p, rho, m, n, press);
  if (!rc)
    rc = interaction_on_cells(p, &c, step, v, rho, press, m, pid, sid, ct, f, vgt);
  cells_free(&c);
  if (!rc && !p->newtonian) rc = orc_compute_act(p, ct, vgt, n, act);
  return rc;
}

int orc_prologue(const orc_params *p, int64_t step, int32_t n, double *x,
                 double *v, const double *m, const int32_t *pid,
                 const int32_t *sid, double *rho, double *press, double *f,
                 double *ct, double *act, double *vgt) {
  if (check_sid(p, sid, n)) return -1;
  int rc = evaluate(p, step, n, x, v, m, pid, rho, press, f, sid, ct, act, vgt);
  if (rc || !p->coll) return rc;
  /* marching_marching.F90:1003-1116: drag and torque of the initial
   * forces start the colloids' acceleration histories */
  orc_colloids *cp = p->coll;
  const int dim = p->num_dim;
  const double t_block = cp->le_time;
  cp->le_time = p->le_time;
  rc = orc_colloid_collect(p, x, f, m, sid, n);
  cp->le_time = t_block;
  if (rc) return rc;
  if (cp->body_force_type == 1)
    for (int32_t k = 0; k < cp->num; ++k)
      for (int d = 0; d < dim; ++d) cp->drag[d + 3 * (size_t)k] += cp->body_force[d] * cp->m[k];
  history_shift(cp->f, cp->num, cp->ab);
  history_shift(cp->alpha, cp->num, cp->ab);
  for (int32_t k = 0; k < cp->num; ++k) {
    double *fk = cp->f + 3 * (size_t)k, *al = cp->alpha + 3 * (size_t)k;
    const double *tq = cp->torque + 3 * (size_t)k, *mi = cp->mmi + 3 * (size_t)k;
    for (int d = 0; d < dim; ++d) fk[d] = cp->translate ? cp->drag[d + 3 * (size_t)k] / cp->m[k] : 0.0;
    if (!cp->rotate) {
      al[0] = al[1] = al[2] = 0.0;
    } else if (dim == 2) {
      al[2] = tq[2] / mi[2];
    } else {
      for (int d = 0; d < 3; ++d) al[d] = tq[d] / mi[d];
    }
  }
  return 0;
}

/* marching_integrate_VV.F90:303-1619, in two halves around the point where
 * the positions of the step's evaluation are known (a check may install
 * the step's injected noise there): the opening half kick, drift, ct half
 * update, colloid block and wrap; then the evaluation, ct half update and
 * closing half kick. */
int orc_vv_begin(const orc_params *p, int64_t step, int32_t n, double *x,
                 double *v, const double *m, const int32_t *pid,
                 const int32_t *sid, double *rho, double *press, double *f,
                 double *ct, double *act, double *vgt) {
  (void)pid; (void)rho; (void)press; (void)vgt;
  if (check_sid(p, sid, n)) return -1;
  integrate_velocity(p, n, v, f, sid, 0.5);
  integrate_position(p, n, x, v, sid);
  if (!p->newtonian) integrate_ct(p, n, ct, act, 0.5);
  if (p->coll && colloid_block(p, step, n, x, f, m, sid)) return -1;
  adjust_particles(p, n, x, v);
  return 0;
}

int orc_vv_finish(const orc_params *p, int64_t step, int32_t n, double *x,
                  double *v, const double *m, const int32_t *pid,
                  const int32_t *sid, double *rho, double *press, double *f,
                  double *ct, double *act, double *vgt) {
  int rc = evaluate(p, step, n, x, v, m, pid, rho, press, f, sid, ct, act, vgt);
  if (rc) return rc;
  if (!p->newtonian) integrate_ct(p, n, ct, act, 0.5);
  integrate_velocity(p, n, v, f, sid, 0.5);
  return 0;
}

int orc_vv_step(const orc_params *p, int64_t step, int32_t n, double *x,
                double *v, const double *m, const int32_t *pid,
                const int32_t *sid, double *rho, double *press, double *f,
                double *ct, double *act, double *vgt) {
  int rc = orc_vv_begin(p, step, n, x, v, m, pid, sid, rho, press, f, ct, act, vgt);
  if (rc) return rc;
  return orc_vv_finish(p, step, n, x, v, m, pid, sid, rho, press, f, ct, act, vgt);
}

int orc_statistics(const orc_params *p, const double *v, const double *m,
                   const int32_t *sid, int32_t n, double *ekin, double *mom) {
  const int dim = p->num_dim;
  double e = 0.0, mm[3] = {0, 0, 0};
  for (int32_t j = 0; j < n; ++j)
    if (sid[j] == 0) {
      double v2 = 0.0;
      for (int d = 0; d < dim; ++d) v2 = v2 + v[d + (size_t)dim * j] * v[d + (size_t)dim * j];
      e = e + 0.5 * m[j] * v2;
      for (int d = 0; d < dim; ++d) mm[d] = mm[d] + m[j] * v[d + (size_t)dim * j];
    }
  if (p->coll) {
    /* colloid_compute_statistic.F90: translational kinetic energy and
     * momentum of the colloids, added by statistic_compute_statistic.F90:465-475 */
    const orc_colloids *cp = p->coll;
    double ek = 0.0, cm[3] = {0, 0, 0};
    for (int32_t k = 0; k < cp->num; ++k) {
      double v2 = 0.0, mk[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) {
        const double vk = cp->v[d + 3 * (size_t)k];
        v2 = v2 + vk * vk;
        mk[d] = mk[d] + cp->m[k] * vk;
      }
      ek = ek + (0.0 + 0.5 * cp->m[k] * v2);
      for (int d = 0; d < dim; ++d) cm[d] = cm[d] + mk[d];
    }
    e = e + ek;
    for (int d = 0; d < dim; ++d) mm[d] = mm[d] + cm[d];
  }
  *ekin = e;
  for (int d = 0; d < dim; ++d) mom[d] = mm[d];
  return 0;
}
