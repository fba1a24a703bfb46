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
 *   - density, equation of state, Hu-Adams / angular / Oldroyd-B pair forces
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
 *        particles_apply_body_force.F90:125-193);
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
#define ORC_MACHINE_ZERO 4.440892098500626e-16 /* 2*EPSILON(1d0) */

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

static void traverse(const orc_params *p, const orc_cells *c, visit_fn visit,
                     void *ctx) {
  const int dim = c->dim;
  const double cut2 = p->cut_off * p->cut_off;
  const int64_t n1 = c->nmt[0];
  const int64_t n2 = dim == 3 ? (int64_t)c->nmt[0] * c->nmt[1] : 0;
  double rij[3];
  for (int k = c->cstart; k <= c->cend[2]; ++k)
    for (int j = c->cstart; j <= c->cend[1]; ++j)
      for (int i = c->cstart; i <= c->cend[0]; ++i) {
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
  traverse(p, &c, visit_pairs, &pc);
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
  double *rho;
} dens_ctx;

static void visit_density(void *vctx, int32_t ip, int32_t jp, double r,
                          const double *rij) {
  (void)rij;
  dens_ctx *dc = (dens_ctx *)vctx;
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
  dc->rho[ip] = dc->rho[ip] + rhoi;
  if (dc->p->symmetry) dc->rho[jp] = dc->rho[jp] + rhoj;
}

static int density_on_cells(const orc_params *p, const orc_cells *c,
                            const double *m, double *rho_real) {
  dens_ctx dc;
  dc.p = p;
  dc.c = c;
  kern_init(p, &dc.k);
  double *mall = (double *)malloc(sizeof(double) * (size_t)c->nall);
  double *rho = (double *)calloc((size_t)c->nall, sizeof(double));
  if (!mall || !rho) { free(mall); free(rho); return -1; }
  for (int32_t i = 0; i < c->nall; ++i) mall[i] = m[c->owner[i]];
  double w0, g0;
  kern_eval(&dc.k, 0.0, &w0, &g0);
  /* self contribution (particles_compute_density.F90:369-558, fluid) */
  for (int32_t i = 0; i < c->nreal; ++i)
    rho[i] = p->rhs_density_type == 1 ? w0 * m[i] : w0;
  dc.mall = mall;
  dc.rho = rho;
  traverse(p, c, visit_density, &dc);
  /* ghost contributions back to their owners (map_ghost_put) */
  if (p->symmetry)
    for (int32_t g = c->nreal; g < c->nall; ++g)
      rho[c->owner[g]] = rho[c->owner[g]] + rho[g];
  memcpy(rho_real, rho, sizeof(double) * (size_t)c->nreal);
  free(mall);
  free(rho);
  return 0;
}

int orc_density(const orc_params *p, const double *x, const double *m,
                const int32_t *sid, int32_t n, double *rho) {
  for (int32_t i = 0; i < n; ++i)
    if (sid[i] != 0) return -1; /* fluid particles only */
  orc_cells c;
  if (cells_build(p, x, n, &c)) return -1;
  int rc = density_on_cells(p, &c, m, rho);
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

/* natural log of u in (0, 1]: exponent split plus
 * ln m = 2 atanh(t), t = (m-1)/(m+1), m in [1/sqrt2, sqrt2] (Taylor) */
static float noise_log(float u) {
  uint32_t b = f2u(u);
  int e = (int)(b >> 23) - 127;
  float m = u2f((b & 0x7fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e = e + 1;
  }
  float t = (m - 1.0f) / (m + 1.0f);
  float s = t * t;
  float q = fmaf(s, 0.111111111f, 0.142857143f);
  q = fmaf(s, q, 0.2f);
  q = fmaf(s, q, 0.333333333f);
  q = fmaf(s, q, 1.0f);
  float lnm = (2.0f * t) * q;
  return fmaf((float)e, 0.693147182f, lnm);
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
/* pair forces                                                          */
/* ------------------------------------------------------------------ */

typedef struct {
  const orc_params *p;
  const orc_cells *c;
  orc_kern k;
  int64_t step;
  const double *v, *rho, *press, *m, *ct; /* over xall */
  const int32_t *pid;
  double *f, *vgt;                         /* over xall */
  double *pt;                              /* (dim,dim,nall) */
  int fail;
} force_ctx;

/* Random force direction for the ordered pair (ip -> jp): a Gaussian
 * vector with the covariance of the reference's symmetric traceless
 * dW . e_ij, oriented by the lower pid (antisymmetric in ip <-> jp). */
static void noise_vector(const force_ctx *fc, int32_t ip, int32_t jp,
                         const double *eij, double *we) {
  const int dim = fc->p->num_dim;
  uint32_t a = (uint32_t)fc->pid[ip], b = (uint32_t)fc->pid[jp];
  uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
  float g[4];
  orc_pair_normals(fc->p->seed, lo, hi, fc->step, g);
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

static void visit_force(void *vctx, int32_t ip, int32_t jp, double dij,
                        const double *rij) {
  force_ctx *fc = (force_ctx *)vctx;
  const orc_params *p = fc->p;
  const int dim = p->num_dim;
  double w, gradw;
  kern_eval(&fc->k, dij, &w, &gradw);
  const double *vi = fc->v + (size_t)dim * ip, *vj = fc->v + (size_t)dim * jp;
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
    /* pp_vgt_ip.inc / pp_vgt_jp.inc */
    const double *xi = fc->c->xall + (size_t)dim * ip;
    const double *xj = fc->c->xall + (size_t)dim * jp;
    const int d2n = dim * dim;
    for (int bb = 0; bb < dim; ++bb)
      for (int aa = 0; aa < dim; ++aa)
        fc->vgt[aa + dim * bb + (size_t)d2n * ip] +=
            (vj[aa] - vi[aa]) * (xi[bb] - xj[bb]) * gradw / numi / dij;
    if (p->symmetry)
      for (int bb = 0; bb < dim; ++bb)
        for (int aa = 0; aa < dim; ++aa)
          fc->vgt[aa + dim * bb + (size_t)d2n * jp] +=
              (vi[aa] - vj[aa]) * (xj[bb] - xi[bb]) * gradw / numj / dij;
    /* rhs_force_ff_non_Newtonian_HuAdams.F90:105-131 */
    const double *pti = fc->pt + (size_t)d2n * ip, *ptj = fc->pt + (size_t)d2n * jp;
    double fcv[3] = {0, 0, 0};
    for (int jj = 0; jj < dim; ++jj)
      for (int ii = 0; ii < dim; ++ii)
        fcv[jj] = fcv[jj] + pti[ii + dim * jj] * eij[ii] / (numi * numi) +
                  ptj[ii + dim * jj] * eij[ii] / (numj * numj);
    for (int d = 0; d < dim; ++d) {
      fi[d] = fi[d] - fcv[d] * gradw / mi;
      fj[d] = fj[d] + fcv[d] * gradw / mj;
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
  } else if (p->rhs_force_type == 3) {
    /* rhs_force_ff_Newtonian_HuAdams.F90:140-237 */
    double pi = fc->press[ip], pj = fc->press[jp];
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
  } else {
    /* rhs_force_ff_Newtonian_HuAdams_angular.F90:166-229 */
    double pi = fc->press[ip], pj = fc->press[jp];
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
      orc_pair_normals(p->seed, a < b ? a : b, a < b ? b : a, fc->step, g);
      double eps = (double)g[0];
      double f_r = eps * sqrt(2.0 * p->kt * p->eta) *
                   sqrt(-p->eta_coef * aij / dij) / sqrt(p->dt);
      for (int d = 0; d < dim; ++d) {
        fi[d] = fi[d] + f_r * eij[d] / mi;
        fj[d] = fj[d] - f_r * eij[d] / mj;
      }
    }
  }
  double *f = fc->f;
  for (int d = 0; d < dim; ++d) f[d + (size_t)dim * ip] += fi[d];
  if (p->symmetry)
    for (int d = 0; d < dim; ++d) f[d + (size_t)dim * jp] += fj[d];
}

static int interaction_on_cells(const orc_params *p, const orc_cells *c,
                                int64_t step, const double *v,
                                const double *rho, const double *press,
                                const double *m, const int32_t *pid,
                                const double *ct, double *f_real,
                                double *vgt_real) {
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
  double *fa = (double *)calloc((size_t)dim * na, sizeof(double));
  double *vg = (double *)calloc((size_t)d2n * na, sizeof(double));
  double *pt = (double *)calloc((size_t)d2n * na, sizeof(double));
  if (!va || !ra || !pa || !ma || !ia || !fa || !vg || !pt) {
    free(va); free(ra); free(pa); free(ma); free(ia); free(fa); free(vg); free(pt);
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
  fc.v = va; fc.rho = ra; fc.press = pa; fc.m = ma; fc.pid = ia;
  fc.ct = ct; fc.f = fa; fc.vgt = vg; fc.pt = pt;
  traverse(p, c, visit_force, &fc);
  if (p->symmetry)
    for (int32_t g = c->nreal; g < na; ++g) {
      int32_t o = c->owner[g];
      for (int d = 0; d < dim; ++d)
        fa[d + (size_t)dim * o] += fa[d + (size_t)dim * g];
      if (!p->newtonian)
        for (int q = 0; q < d2n; ++q)
          vg[q + (size_t)d2n * o] += vg[q + (size_t)d2n * g];
    }
  /* particles_apply_body_force.F90:125-193 (type 1) */
  if (p->body_force_type == 1)
    for (int d = 0; d < dim; ++d)
      if (fabs(p->body_force[d]) > ORC_MACHINE_ZERO)
        for (int32_t i = 0; i < c->nreal; ++i)
          fa[d + (size_t)dim * i] += p->body_force[d];
  memcpy(f_real, fa, sizeof(double) * (size_t)dim * c->nreal);
  if (!p->newtonian && vgt_real)
    memcpy(vgt_real, vg, sizeof(double) * (size_t)d2n * c->nreal);
  int fail = fc.fail;
  free(va); free(ra); free(pa); free(ma); free(ia); free(fa); free(vg); free(pt);
  return fail ? -1 : 0;
}

int orc_interaction(const orc_params *p, int64_t step, const double *x,
                    const double *v, const double *rho, const double *press,
                    const double *m, const int32_t *pid, const int32_t *sid,
                    const double *ct, int32_t n, double *f, double *vgt) {
  for (int32_t i = 0; i < n; ++i)
    if (sid[i] != 0) return -1;
  if (p->rhs_force_type != 3 && p->rhs_force_type != 4) return -1;
  if (!p->newtonian && p->rhs_force_type != 3) return -1;
  orc_cells c;
  if (cells_build(p, x, n, &c)) return -1;
  int rc = interaction_on_cells(p, &c, step, v, rho, press, m, pid, ct, f, vgt);
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

static int evaluate(const orc_params *p, int64_t step, int32_t n,
                    const double *x, const double *v, const double *m,
                    const int32_t *pid, double *rho, double *press, double *f,
                    double *ct, double *act, double *vgt) {
  orc_cells c;
  if (cells_build(p, x, n, &c)) return -1;
  int rc = density_on_cells(p, &c, m, rho);
  if (!rc) rc = orc_pressure(p, rho, m, n, press);
  if (!rc)
    rc = interaction_on_cells(p, &c, step, v, rho, press, m, pid, ct, f, vgt);
  cells_free(&c);
  if (!rc && !p->newtonian) rc = orc_compute_act(p, ct, vgt, n, act);
  return rc;
}

int orc_prologue(const orc_params *p, int64_t step, int32_t n, double *x,
                 double *v, const double *m, const int32_t *pid,
                 const int32_t *sid, double *rho, double *press, double *f,
                 double *ct, double *act, double *vgt) {
  for (int32_t i = 0; i < n; ++i)
    if (sid[i] != 0) return -1;
  (void)sid;
  return evaluate(p, step, n, x, v, m, pid, rho, press, f, ct, act, vgt);
}

int orc_vv_step(const orc_params *p, int64_t step, int32_t n, double *x,
                double *v, const double *m, const int32_t *pid,
                const int32_t *sid, double *rho, double *press, double *f,
                double *ct, double *act, double *vgt) {
  /* marching_integrate_VV.F90:303-1619 */
  integrate_velocity(p, n, v, f, sid, 0.5);
  integrate_position(p, n, x, v, sid);
  if (!p->newtonian) integrate_ct(p, n, ct, act, 0.5);
  adjust_particles(p, n, x, v);
  int rc = evaluate(p, step, n, x, v, m, pid, rho, press, f, ct, act, vgt);
  if (rc) return rc;
  if (!p->newtonian) integrate_ct(p, n, ct, act, 0.5);
  integrate_velocity(p, n, v, f, sid, 0.5);
  return 0;
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
  *ekin = e;
  for (int d = 0; d < dim; ++d) mom[d] = mm[d];
  return 0;
}
