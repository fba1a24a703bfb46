// sdpd_device.h -- device-side parameters and pair arithmetic shared by the
// SDPD kernels (gfx950).  Everything here is FP64 unless the name says f32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sdpd {

// Launch-time constants of one box.  Passed by value to every kernel.
struct DevParams {
  int32_t dim;
  int32_t symmetry;
  int32_t kernel_type;
  int32_t density_type;
  int32_t force_type;
  int32_t eos_type;
  int32_t newtonian;
  int32_t noise;          // Brownian and kt > machine zero
  int32_t body_force;     // body force type present: 1 constant, 2 two-way, 3 sine
  int32_t n;              // particles
  // sort / gather grid: half of PPM's cell edge, so the stencil is 5 cells
  // wide per dimension (at least 6 cells per dimension)
  // (single box: nm, length, ... below; batches: RepGeom per replica)
  int32_t nm[3];
  int32_t ncell;          // cells of all replicas
  // PPM's own grid, cell edge >= 1.0001 cut_off (at least 3 per dimension)
  int32_t nm_ref[3];
  int32_t ncell_ref;
  double min_phys[3];
  double length[3];
  double rd[3];           // nm / length
  double rd_ref[3];       // nm_ref / length (PPM's rdx)
  double cs[3];           // cell edge length / nm
  double cut_off, cut2;
  double h, inv_h, coef, coef_grad, w0;
  double c2, rho_ref, gamma, p0;
  double eta, eta_coef, kt, dt, sqrt_dt, inv_sqrt_dt, sqrt_2kt_eta;
  double tau, G;
  double bf[3];
  double bf_split;        // type 2: half of max_phys along the last dimension
  double bf_k;            // type 3: 2 pi / (max_phys(2) - min_phys(2))
  uint64_t seed;
  int64_t step;
  int32_t mixed;          // FP32 pair arithmetic
  int32_t list_cap;       // stored neighbour entries per particle
  int32_t nrep;           // boxes (replicas) in the context
  int32_t nrows;          // rows of the sort grids of all replicas
  // tuned shapes: density scan (0: tile 224, buffer 16; 1: 128, 16;
  // 2: 224, 12; 3: 128, 12) and chunk order (1: XCD-contiguous)
  int32_t density_shape;
  int32_t xcd_order;
  // Lees-Edwards faces: le_g = 0-based dimension (1 or 2) whose faces shear
  // along x, -1 when none; face velocities (x components) and their jump
  int32_t le_g;
  double le_vlo, le_vhi, le_u;  // le_u = le_vhi - le_vlo
  float le_uf;
  // band around cut2 of the FP32 candidate test; inside it the decision is
  // taken in FP64.  FP32 positions are relative to the own cell corner.
  float cut2_lo_f32, cut2_hi_f32;
  float lengthf[3];
  float csf[3];           // FP32 cell edges of the sort grid
  // FP32 copies of the constants (mixed precision)
  float inv_hf, coeff, coef_gradf;
  float etaf, eta_coeff, ktf, inv_sqrt_dtf, sqrt_2kt_etaf;
  // colloids: boundary particles carry sid 1..ncoll (0: fluid only); the
  // reference's coll_noslip, coll_rho_type, coll_translate, coll_rotate,
  // integrate_colloid_AB, coll_sub_time_step, coll_body_force(_type)
  int32_t ncoll;
  int32_t coll_noslip;    // 1 frozen, 2 Morris
  int32_t coll_rho_type;  // 0 constant, 1 dynamic
  int32_t coll_translate, coll_rotate, coll_ab, coll_sub, coll_bf_type;
  double coll_dout;       // outer layer of a colloid (0.05 dx)
  double coll_bf[3];
  double rho0;            // initial density (physics key rho)
  int64_t coll_step0;     // step of the initial evaluation (the run's step_start)
  // boxes with sheared faces and colloids: rows of cells within this many
  // cells of a face take the Lees-Edwards force kernel (the stencil's reach
  // plus a colloid's radius, outer layer and the cut-off: the nearest image
  // of a colloid seen from there may lie across the face)
  int32_t coll_le_cells;
  // Verlet lists: the stored lists hold every candidate within rc + skin
  // at their build, the state keeps its order until the next build, and a
  // build happens once a particle has moved more than skin / 2 since the
  // last one (with sheared faces: once twice that displacement plus the
  // sheared copies' drift exceeds the skin).  Every step decides in / out
  // of the cut-off afresh on the current positions.
  // verlet = 0: lists of in-cut-off pairs, rebuilt every step.
  int32_t verlet;
  double skin;
  double half_skin2;      // (skin / 2)^2, FP64 displacement test
  float skin2_hi_f32;     // FP32 acceptance of the build scan, (rc + skin)^2 and the band
};

// Colloid state read by the pair kernels (disks in 2D, spheres in 3D):
// centre, translational and angular velocity (newest), radius.
struct CollHot {
  double x[3];
  double v[3];
  double w[3];
  double radius;
  double pad[2];
};

// Full rigid-body state of a colloid.  Histories hold the Adams-Bashforth
// levels, newest first (the reference's third index).  xs / rot describe
// the last colloid block: the centre before each sub-step and after the
// last, and each sub-step's rotation (column-major 3 x 3).
constexpr int kCollHist = 5;
constexpr int kCollSub = 8;
struct CollState {
  double x[3];
  double v[kCollHist][3];
  double w[kCollHist][3];
  double f[kCollHist][3];
  double al[kCollHist][3];
  double acc[9];
  double theta[3];
  double drag[3];
  double torque[3];
  double radius, m;
  double mmi[3];
  double xs[kCollSub + 1][3];
  double rot[kCollSub][9];
};

// Geometry of one periodic box of a context (a context holds one box or a
// batch of independent replicas of different sizes sharing the physics).
// Cells of all replicas are numbered globally: replica q owns cells
// [cell_base, cell_base + nm0 nm1 nm2) and rows (lines of cells along x)
// [row_base, row_base + nm1 nm2).
struct RepGeom {
  int32_t nm[3];          // sort-grid cells per dimension
  int32_t cell_base;
  int32_t row_base;
  int32_t pad;
  double lo[3];           // lower corner
  double L[3];            // edge lengths
  double rd[3];           // nm / L
  double cs[3];           // cell edges
  float csf[3];
  float pad2;
};

// cell index along one dimension: PPM's FLOOR((x - x0) * rdx) with its
// upper-face corrections (ppm_util_rank3d.F90:433-459), clamped
__device__ __forceinline__ int cell_coord(double xv, double x0, double rd, double xmax, int nm) {
  int k = (int)floor((xv - x0) * rd);
  if (xv >= xmax && k < nm) k = nm;
  if (xv < xmax && k >= nm) k = nm - 1;
  // particles are wrapped into [lo, lo + L) before every sort; clamp so a
  // stray value can never address outside the cell arrays
  if (k < 0) k = 0;
  if (k >= nm) k = nm - 1;
  return k;
}

// Fortran MODULO(a, p) for reals: the exact remainder (fmod) carrying the
// sign of p.
__device__ __forceinline__ double f_modulo(double a, double p) {
  double r = fmod(a, p);
  if (r != 0.0 && ((r < 0.0) != (p < 0.0))) r = r + p;
  return r;
}

// Histogram slot of key k, aggregated per wave (every lane of the wave
// calls it; lanes with act false count nothing and get 0).  Equal keys in
// runs of neighbouring lanes -- particles stored in cell order -- cost one
// atomic per run: the run's first lane adds the run length and the run
// shares the old count.  The order of slots inside a cell is not
// deterministic; consumers rank by index.
__device__ __forceinline__ int32_t histogram_slot(int32_t* __restrict__ count, int32_t k, bool act) {
  const int lane = threadIdx.x & 63;
  const int32_t key = act ? k : -1 - lane;  // inactive lanes: runs of their own
  const int32_t kprev = __shfl_up(key, 1);
  const bool head = lane == 0 || kprev != key;
  const uint64_t heads = __ballot(head);
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int first = 63 - __clzll(heads & upto);
  const uint64_t after = heads & ~upto;
  const int end = after ? __ffsll((long long)after) - 1 : 64;
  int32_t base = 0;
  if (head && act) base = atomicAdd(&count[key], end - first);
  base = __shfl(base, first);
  return act ? base + (lane - first) : 0;
}

// global cell key of a position in replica geometry g
__device__ __forceinline__ int32_t cell_key_geo(const RepGeom& g, int dim, double x0, double x1,
                                                double x2) {
  const int cx = cell_coord(x0, g.lo[0], g.rd[0], g.lo[0] + g.L[0], g.nm[0]);
  const int cy = cell_coord(x1, g.lo[1], g.rd[1], g.lo[1] + g.L[1], g.nm[1]);
  const int cz = dim == 3 ? cell_coord(x2, g.lo[2], g.rd[2], g.lo[2] + g.L[2], g.nm[2]) : 0;
  return g.cell_base + cx + g.nm[0] * (cy + g.nm[1] * cz);
}

// ---------------------------------------------------------------------
// smoothing kernel (kernel_kernel_quintic_spline.F90:160-215,
// kernel_kernel_Lucy.F90), same operation order as the reference
// ---------------------------------------------------------------------
// W and dW/dr at s = r/h without branches: with t_k = max(k - s, 0) the
// piecewise quintic spline is coef (t3^5 - 6 t2^5 + 15 t1^5) and its
// derivative coef_grad (5 t3^4 - 30 t2^4 + 75 t1^4), coef_grad = -coef/h
// (kernel_kernel_quintic_spline.F90:1-221); Lucy: coef (1 + 3s) t1^3 and
// coef_grad 12 s t1^2 (kernel_kernel_Lucy.F90:1-174).  Zero from s = 3
// (quintic) or s = 1 (Lucy) on.
__device__ __forceinline__ void kernel_eval(const DevParams& P, double s, double& w,
                                            double& gw) {
  if (P.kernel_type == 1) {
    const double t1 = fmax(1.0 - s, 0.0), t2 = fmax(2.0 - s, 0.0), t3 = fmax(3.0 - s, 0.0);
    const double q1 = (t1 * t1) * (t1 * t1);
    const double q2 = (t2 * t2) * (t2 * t2);
    const double q3 = (t3 * t3) * (t3 * t3);
    w = P.coef * (q3 * t3 - 6.0 * (q2 * t2) + 15.0 * (q1 * t1));
    gw = P.coef_grad * (5.0 * q3 - 30.0 * q2 + 75.0 * q1);
  } else {
    const double t = fmax(1.0 - s, 0.0);
    w = P.coef * (1.0 + 3.0 * s) * (t * t * t);
    gw = P.coef_grad * 12.0 * s * (t * t);
  }
}

__device__ __forceinline__ void kernel_w_gradw_f32(int kernel_type, float s, float coef,
                                                   float coef_grad, float& w, float& gw) {
  if (kernel_type == 1) {
    const float t1 = fmaxf(1.0f - s, 0.0f), t2 = fmaxf(2.0f - s, 0.0f), t3 = fmaxf(3.0f - s, 0.0f);
    const float q1 = (t1 * t1) * (t1 * t1);
    const float q2 = (t2 * t2) * (t2 * t2);
    const float q3 = (t3 * t3) * (t3 * t3);
    w = coef * (q3 * t3 - 6.0f * (q2 * t2) + 15.0f * (q1 * t1));
    gw = coef_grad * (5.0f * q3 - 30.0f * q2 + 75.0f * q1);
  } else {
    const float t = fmaxf(1.0f - s, 0.0f);
    w = coef * (1.0f + 3.0f * s) * (t * t * t);
    gw = coef_grad * 12.0f * s * (t * t);
  }
}

// 1/sqrt(x) for normal x > 0: the hardware estimate (relative error below
// 2^-24 on gfx950) refined by one Newton step y <- y + y (1/2 - x y^2 / 2),
// which squares the error: relative error below 2^-46.
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-(0.5 * x) * y, y, 0.5);
  return fma(y, e, y);
}

// sqrt(x) for x >= 0 through rsqrt_nr (0 for x = 0).
__device__ __forceinline__ double sqrt_nr(double x) { return x > 0.0 ? x * rsqrt_nr(x) : 0.0; }

// ---------------------------------------------------------------------
// equation of state (stateEquation_compute_pressure.F90:158,176); rho is
// the mass density (number density times mass for density type 2)
// ---------------------------------------------------------------------
__device__ __forceinline__ double eos_pressure(const DevParams& P, double rho) {
  if (P.eos_type == 1) return P.c2 * (pow(rho, P.gamma) - P.rho_ref);
  return P.p0 * (pow(rho / P.rho_ref, P.gamma) - 1.0);
}

// ---------------------------------------------------------------------
// counter-based noise: Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// keyed by the seed, counter (min pid, max pid, step); normals by
// Box-Muller with the hardware's log2, square root and sine / cosine
// (within about an ulp each).  The stream is the one sdpd_pair_normals
// returns; checks against a CPU path take their normals from it.
// ---------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                              uint32_t c3, uint32_t k0, uint32_t k1,
                                              uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    // three-input xor in one instruction (truth table 0x96)
    uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
    uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
    c0 = n0;
    c1 = (uint32_t)p1;
    c2 = n2;
    c3 = (uint32_t)p0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Four N(0,1) deviates for the unordered pair (lo, hi) at a step: two
// Box-Muller transforms of 24-bit uniforms, u in (0, 1] (radius) and the
// angle 2 pi k / 2^24 (the sine and cosine instructions take revolutions).
// -2 ln u = -2 ln 2 log2 u, clamped at 0 against a rounding above zero for
// u next to 1.
__device__ __forceinline__ void pair_normals(uint64_t seed, uint32_t lo, uint32_t hi,
                                             int64_t step, float g[4]) {
  uint32_t w[4];
  philox4x32_10(lo, hi, (uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32),
                (uint32_t)seed, (uint32_t)(seed >> 32), w);
#pragma unroll
  for (int hh = 0; hh < 2; ++hh) {
    const float u = (float)((w[2 * hh] >> 8) + 1u) * 0x1p-24f;
    const float r =
        __builtin_amdgcn_sqrtf(fmaxf(-1.38629436112f * __builtin_amdgcn_logf(u), 0.0f));
    const float rev = (float)(w[2 * hh + 1] >> 8) * 0x1p-24f;
    g[2 * hh] = r * __builtin_amdgcn_cosf(rev);
    g[2 * hh + 1] = r * __builtin_amdgcn_sinf(rev);
  }
}

// Squared distance with two rounded operations per term (no FMA), so the
// in/out decision equals the reference's DOT_PRODUCT on the same values
// (pp_interaction.inc:37-46).
__device__ __forceinline__ double dist2_exact(double a, double b, double c) {
#pragma clang fp contract(off)
  double s = a * a;
  s = s + b * b;
  s = s + c * c;
  return s;
}

__device__ __forceinline__ double dist2_exact2(double a, double b) {
#pragma clang fp contract(off)
  double s = a * a;
  s = s + b * b;
  return s;
}

// Per-dimension image of a neighbour cell: returns the wrapped cell index
// and how the coordinate difference is formed.
//   mode 0: xi - xj
//   mode 1: xi - (xj + L)   (neighbour across the upper face)
//   mode 2: (xi + L) - xj   (symmetric mode, home particle on the lower side)
//   mode 3: xi - (xj - L)   (full mode, neighbour across the lower face)
// The symmetric convention reproduces PPM's upper-only ghost layer, where
// the particle near the lower face is the one shifted by +L
// (ppm_bcdef_periodic_x.inc:37-60); the full convention reproduces ghosts
// on both faces (ppm_bcdef_periodic_x.inc:61-90).
__device__ __forceinline__ int wrap_cell(int c, int nm, int symmetry, int& mode) {
  if (c < 0) {
    mode = symmetry ? 2 : 3;
    return c + nm;
  }
  if (c >= nm) {
    mode = 1;
    return c - nm;
  }
  mode = 0;
  return c;
}

// colloid_in_nearest_image.F90 (disks, spheres): the image of the centre of
// colloid C nearest to xp, within d_max = radius + dout per dimension,
// Lees-Edwards faces first (shear length at time t, the centre's image
// velocity takes the face velocity jump), then periodic faces.  rx = xp -
// image, vim = the image's velocity; 3 components each.
template <int DIM>
__device__ __forceinline__ void coll_nearest(const DevParams& P, const CollHot& C,
                                             const double* xp, double t, double* rx,
                                             double* vim) {
  // written without run-time indices into arrays, which would put the
  // kernel's parameters on the stack
  double c0 = C.x[0], c1 = C.x[1], c2 = C.x[2];
  const double dmax = C.radius + P.coll_dout;
  vim[0] = C.v[0];
  vim[1] = C.v[1];
  vim[2] = C.v[2];
  if (P.le_g >= 0) {
    const bool gy = P.le_g == 1;
    const double rg = gy ? xp[1] - c1 : (DIM == 3 ? xp[2] - c2 : 0.0);
    const double Lg = gy ? P.length[1] : P.length[2];
    // the other coordinates are taken MODULO the box as well (a plain wrap
    // for the one the faces do not shear)
    if (rg > dmax) {
      if (gy) c1 = c1 + Lg; else c2 = c2 + Lg;
      c0 = f_modulo(c0 + P.le_u * t, P.length[0]);
      if (DIM == 3) {
        if (gy) c2 = f_modulo(c2, P.length[2]); else c1 = f_modulo(c1, P.length[1]);
      }
      vim[0] = vim[0] + P.le_u;
    } else if (-rg > dmax) {
      if (gy) c1 = c1 - Lg; else c2 = c2 - Lg;
      c0 = f_modulo(c0 + (P.le_vlo - P.le_vhi) * t, P.length[0]);
      if (DIM == 3) {
        if (gy) c2 = f_modulo(c2, P.length[2]); else c1 = f_modulo(c1, P.length[1]);
      }
      vim[0] = vim[0] + (P.le_vlo - P.le_vhi);
    }
  }
  double r0 = xp[0] - c0, r1 = xp[1] - c1, r2 = DIM == 3 ? xp[2] - c2 : 0.0;
  if (r0 > dmax) c0 = c0 + P.length[0]; else if (-r0 > dmax) c0 = c0 - P.length[0];
  if (r1 > dmax) c1 = c1 + P.length[1]; else if (-r1 > dmax) c1 = c1 - P.length[1];
  if (DIM == 3) {
    if (r2 > dmax) c2 = c2 + P.length[2]; else if (-r2 > dmax) c2 = c2 - P.length[2];
  }
  rx[0] = xp[0] - c0;
  rx[1] = xp[1] - c1;
  rx[2] = DIM == 3 ? xp[2] - c2 : 0.0;
}

__device__ __forceinline__ void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// colloid_noslip.F90: velocity a boundary particle of colloid C shows to a
// fluid particle moving at vf.  r_xc = boundary particle - nearest centre
// image (moving at vcoll), r_xf = fluid particle - that image.  Morris
// (colloid_noslip_Morris_sphere.F90:129-300): the fluid velocity relative
// to the surface point r_s = radius r_xf / |r_xf| is mirrored with the
// ratio of the distances to the tangent plane, clamped to +-0.5; a fluid
// particle inside the colloid and the frozen rule take the rigid velocity
// at the boundary particle.
template <int DIM>
__device__ __forceinline__ void coll_noslip(const DevParams& P, const CollHot& C,
                                            const double* r_xc, const double* vcoll,
                                            const double* r_xf, const double* vf,
                                            double* vc) {
  const double R = C.radius;
  double s2 = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) s2 = s2 + r_xf[d] * r_xf[d];
  const double d_fcoll = sqrt(s2);
  if (P.coll_noslip == 1 || d_fcoll <= R) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) vc[d] = vcoll[d];
    if (P.coll_rotate) {
      double r_vc[3];
      cross3(C.w, r_xc, r_vc);
#pragma unroll
      for (int d = 0; d < DIM; ++d) vc[d] = vc[d] + r_vc[d];
    }
    return;
  }
  double d_cn = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) d_cn = d_cn + r_xc[d] * (r_xf[d] / d_fcoll);
  const double d_cs = R - d_cn;
  const double d_fs = d_fcoll - R;
  double corr;
  if (fabs(d_fs) < 1e-14) {
    corr = 0.5;
  } else {
    corr = d_cs / d_fs;
    corr = corr > 0.5 ? 0.5 : (corr < -0.5 ? -0.5 : corr);
  }
  double r_vs[3] = {0.0, 0.0, 0.0};
  if (P.coll_rotate) {
    double r_xs[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) r_xs[d] = R / d_fcoll * r_xf[d];
    cross3(C.w, r_xs, r_vs);
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) vc[d] = -corr * (vf[d] - vcoll[d] - r_vs[d]) + vcoll[d] + r_vs[d];
}

// x_i - x_j with the reference's periodic image chosen from the positions:
// the neighbour within half a box.  Across a face the particle on the low
// side is the one shifted by +L in symmetric mode ((x_i + L) - x_j, PPM's
// upper ghost layer), the neighbour by -L otherwise (x_i - (x_j - L)); a
// neighbour on the low side is seen at x_j + L (ppm_bcdef_periodic_x.inc:
// 37-90).  Two rounded operations, as the reference forms the ghost
// coordinate first.  Valid for pairs closer than L / 2.
__device__ __forceinline__ double image_diff_pos(double xi, double xj, double L, int symmetry) {
#pragma clang fp contract(off)
  const double raw = xi - xj;
  if (fabs(raw) <= 0.5 * L) return raw;
  if (raw > 0.0) return xi - (xj + L);
  return symmetry ? (xi + L) - xj : xi - (xj - L);
}

__device__ __forceinline__ double image_diff(double xi, double xj, double L, int mode) {
#pragma clang fp contract(off)
  switch (mode) {
    case 1: return xi - (xj + L);
    case 2: return (xi + L) - xj;
    case 3: return xi - (xj - L);
    default: return xi - xj;
  }
}

}  // namespace sdpd
