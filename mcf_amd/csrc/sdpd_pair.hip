// sdpd_pair.hip -- K2 density + equation of state and K4 pair forces, FP64
// gather kernels (gfx950).
//
// Replaces the two hot loops of the reference: particles_compute_density
// (particles_compute_density.F90:578-867) and particles_compute_interaction
// (particles_compute_interaction.F90:618-911 with pp_interaction.inc,
// pp_interaction_ff.inc and rhs_force_ff_*).  The reference visits each
// unordered pair once in a half stencil and adds the result to both
// particles (and to periodic ghost copies that are folded back later).
// Here each particle gathers over the 27 cells around its own cell, with
// the periodic image applied per stencil cell, and writes only its own
// sums: no atomics, no ghosts, identical results run to run.  Every pair
// term is formed with the same operands and operation order as the
// reference's term for that particle, so only the summation order differs.
#include "sdpd_kernels.h"

namespace sdpd {

struct CellCoord {
  int cx, cy, cz;
};

__device__ __forceinline__ CellCoord decode_key(const DevParams& P, int32_t k) {
  CellCoord c;
  c.cx = k % P.nm[0];
  int r = k / P.nm[0];
  c.cy = r % P.nm[1];
  c.cz = r / P.nm[1];
  return c;
}

// Visit the in-cutoff neighbours of particle i (FP64 decision), calling
// body(j, r, d0, d1, d2) for each.  Cells are visited z, y, x with x
// fastest; particles within a cell in cell order.
template <int DIM, typename Body>
__device__ __forceinline__ void for_each_neighbor_f64(const DevParams& P,
                                                      const double* __restrict__ x,
                                                      const int32_t* __restrict__ start, int i,
                                                      double xi0, double xi1, double xi2,
                                                      CellCoord c, Body&& body) {
  const int n = P.n;
  const double* __restrict__ x0 = x;
  const double* __restrict__ x1 = x + n;
  const double* __restrict__ x2 = x + 2 * (int64_t)n;
  const int zlo = DIM == 3 ? -1 : 0, zhi = DIM == 3 ? 1 : 0;
  for (int oz = zlo; oz <= zhi; ++oz) {
    int mz = 0;
    int nz = DIM == 3 ? wrap_cell(c.cz + oz, P.nm[2], P.symmetry, mz) : 0;
    for (int oy = -1; oy <= 1; ++oy) {
      int my;
      int ny = wrap_cell(c.cy + oy, P.nm[1], P.symmetry, my);
      for (int ox = -1; ox <= 1; ++ox) {
        int mx;
        int nx = wrap_cell(c.cx + ox, P.nm[0], P.symmetry, mx);
        int nc = nx + P.nm[0] * (ny + P.nm[1] * nz);
        int s = start[nc], e = start[nc + 1];
        for (int j = s; j < e; ++j) {
          if (j == i) continue;
          double d0 = image_diff(xi0, x0[j], P.length[0], mx);
          double d1 = image_diff(xi1, x1[j], P.length[1], my);
          double d2 = DIM == 3 ? image_diff(xi2, x2[j], P.length[2], mz) : 0.0;
          double r2 = DIM == 3 ? dist2_exact(d0, d1, d2) : dist2_exact2(d0, d1);
          if (r2 >= P.cut2) continue;
          body(j, sqrt(r2), d0, d1, d2);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------
// K2: density + EOS
// ---------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_density_f64(DevParams P, DeviceState S) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int n = P.n;
  double xi0 = S.x[i], xi1 = S.x[n + i], xi2 = DIM == 3 ? S.x[2 * (int64_t)n + i] : 0.0;
  double mi = S.m[i];
  CellCoord c = decode_key(P, S.key_sorted[i]);
  // self term first (particles_compute_density.F90:369-558)
  double acc = P.density_type == 1 ? P.w0 * mi : P.w0;
  const double* __restrict__ m = S.m;
  const int type1 = P.density_type == 1;
  for_each_neighbor_f64<DIM>(P, S.x, S.start, i, xi0, xi1, xi2, c,
                             [&](int j, double r, double, double, double) {
                               double w, gw;
                               kernel_w_gradw(P, r, w, gw);
                               acc = acc + (type1 ? m[j] * w : w);
                             });
  S.rho[i] = acc;
  S.p[i] = eos_pressure(P, P.density_type == 2 ? acc * mi : acc);
}

__global__ void k_eos(DevParams P, DeviceState S) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  double r = S.rho[i];
  S.p[i] = eos_pressure(P, P.density_type == 2 ? r * S.m[i] : r);
}

// ---------------------------------------------------------------------
// K4: pair forces
// ---------------------------------------------------------------------

// Random force direction seen by particle i for neighbour j: a Gaussian
// vector with the covariance of the reference's symmetric traceless
// dW . e_ij (I/2 + e e^T/6 in 3D, I/2 in 2D), oriented by the lower pid.
template <int DIM>
__device__ __forceinline__ void noise_vector_f64(const DevParams& P, uint32_t pidi,
                                                 uint32_t pidj, double e0, double e1, double e2,
                                                 double& w0, double& w1, double& w2) {
  uint32_t lo = pidi < pidj ? pidi : pidj, hi = pidi < pidj ? pidj : pidi;
  float g[4];
  pair_normals(P.seed, lo, hi, P.step, g);
  double sgn = pidi < pidj ? 1.0 : -1.0;
  const double ca = 0.70710678118654752440;
  if (DIM == 3) {
    const double cb = 0.81649658092772603273 - 0.70710678118654752440;
    double c0 = sgn * e0, c1 = sgn * e1, c2 = sgn * e2;
    double ge = (double)g[0] * c0 + (double)g[1] * c1 + (double)g[2] * c2;
    w0 = sgn * (ca * (double)g[0] + cb * ge * c0);
    w1 = sgn * (ca * (double)g[1] + cb * ge * c1);
    w2 = sgn * (ca * (double)g[2] + cb * ge * c2);
  } else {
    w0 = sgn * (ca * (double)g[0]);
    w1 = sgn * (ca * (double)g[1]);
    w2 = 0.0;
  }
}

template <int DIM>
__global__ __launch_bounds__(256) void k_forces_f64(DevParams P, DeviceState S) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int n = P.n;
  const int64_t n2 = 2 * (int64_t)n;
  double xi0 = S.x[i], xi1 = S.x[n + i], xi2 = DIM == 3 ? S.x[n2 + i] : 0.0;
  double vi0 = S.v[i], vi1 = S.v[n + i], vi2 = DIM == 3 ? S.v[n2 + i] : 0.0;
  double ni = S.rho[i], pi = S.p[i], mi = S.m[i];
  uint32_t pidi = (uint32_t)S.pid[i];
  CellCoord c = decode_key(P, S.key_sorted[i]);
  const double* __restrict__ v = S.v;
  const double* __restrict__ rho = S.rho;
  const double* __restrict__ pr = S.p;
  const int32_t* __restrict__ pid = S.pid;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  double L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // vgt, column-major a + DIM*b
  double pti[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int newtonian = P.newtonian;
  if (!newtonian) {
    // particles_compute_pressure_tensor.F90:114-133
    for (int b = 0; b < DIM; ++b)
      for (int a = 0; a < DIM; ++a) {
        double val = -P.G * S.ct[(int64_t)(a + DIM * b) * n + i];
        if (a == b) val = val + pi + P.G;
        pti[a + DIM * b] = val;
      }
  }
  unsigned long long cnt = 0;
  for_each_neighbor_f64<DIM>(
      P, S.x, S.start, i, xi0, xi1, xi2, c,
      [&](int j, double r, double d0, double d1, double d2) {
        ++cnt;
        double w, gw;
        kernel_w_gradw(P, r, w, gw);
        double e0 = d0 / r, e1 = d1 / r, e2 = DIM == 3 ? d2 / r : 0.0;
        double vj0 = v[j], vj1 = v[n + j], vj2 = DIM == 3 ? v[n2 + j] : 0.0;
        double u0 = vi0 - vj0, u1 = vi1 - vj1, u2 = vi2 - vj2;
        double nj = rho[j];
        double fi0, fi1, fi2;
        if (!newtonian) {
          // pp_vgt_ip.inc:16-20
          double du[3] = {vj0 - vi0, vj1 - vi1, vj2 - vi2};
          double dx[3] = {d0, d1, d2};
          for (int b = 0; b < DIM; ++b)
            for (int a = 0; a < DIM; ++a)
              L[a + DIM * b] = L[a + DIM * b] + du[a] * dx[b] * gw / ni / r;
          // rhs_force_ff_non_Newtonian_HuAdams.F90:105-131
          double ptj[9];
          double pj = pr[j];
          for (int b = 0; b < DIM; ++b)
            for (int a = 0; a < DIM; ++a) {
              double val = -P.G * S.ct[(int64_t)(a + DIM * b) * n + j];
              if (a == b) val = val + pj + P.G;
              ptj[a + DIM * b] = val;
            }
          double ev[3] = {e0, e1, e2};
          double fc[3] = {0.0, 0.0, 0.0};
          for (int jj = 0; jj < DIM; ++jj)
            for (int ii = 0; ii < DIM; ++ii)
              fc[jj] = fc[jj] + pti[ii + DIM * jj] * ev[ii] / (ni * ni) +
                       ptj[ii + DIM * jj] * ev[ii] / (nj * nj);
          fi0 = 0.0 - fc[0] * gw / mi;
          fi1 = 0.0 - fc[1] * gw / mi;
          fi2 = DIM == 3 ? 0.0 - fc[2] * gw / mi : 0.0;
          double f_d = P.eta * (1.0 / (ni * ni) + 1.0 / (nj * nj)) * gw / r;
          fi0 = fi0 + f_d * u0 / mi;
          fi1 = fi1 + f_d * u1 / mi;
          if (DIM == 3) fi2 = fi2 + f_d * u2 / mi;
          if (P.noise) {
            double w0, w1, w2;
            noise_vector_f64<DIM>(P, pidi, (uint32_t)pid[j], e0, e1, e2, w0, w1, w2);
            double f_r = sqrt(-4.0 * P.kt * f_d) / P.sqrt_dt;
            fi0 = fi0 + f_r * w0 / mi;
            fi1 = fi1 + f_r * w1 / mi;
            if (DIM == 3) fi2 = fi2 + f_r * w2 / mi;
          }
        } else if (P.force_type == 3) {
          // rhs_force_ff_Newtonian_HuAdams.F90:140-237
          double pj = pr[j];
          double f_c = (pi / (ni * ni) + pj / (nj * nj)) * gw;
          fi0 = 0.0 - f_c * e0 / mi;
          fi1 = 0.0 - f_c * e1 / mi;
          fi2 = DIM == 3 ? 0.0 - f_c * e2 / mi : 0.0;
          double f_d = P.eta * (1.0 / (ni * ni) + 1.0 / (nj * nj)) * gw / r;
          fi0 = fi0 + f_d * u0 / mi;
          fi1 = fi1 + f_d * u1 / mi;
          if (DIM == 3) fi2 = fi2 + f_d * u2 / mi;
          if (P.noise) {
            double w0, w1, w2;
            noise_vector_f64<DIM>(P, pidi, (uint32_t)pid[j], e0, e1, e2, w0, w1, w2);
            double f_r = sqrt(-4.0 * P.kt * f_d) / P.sqrt_dt;
            fi0 = fi0 + f_r * w0 / mi;
            fi1 = fi1 + f_r * w1 / mi;
            if (DIM == 3) fi2 = fi2 + f_r * w2 / mi;
          }
        } else {
          // rhs_force_ff_Newtonian_HuAdams_angular.F90:166-229
          double pj = pr[j];
          double ev = e0 * u0 + e1 * u1;
          if (DIM == 3) ev = ev + e2 * u2;
          double pij = (pi + pj) / 2.0;
          double aij = (1.0 / (ni * ni) + 1.0 / (nj * nj)) * gw;
          double f_c = aij * pij;
          fi0 = 0.0 - f_c * e0 / mi;
          fi1 = 0.0 - f_c * e1 / mi;
          fi2 = DIM == 3 ? 0.0 - f_c * e2 / mi : 0.0;
          double f_d = P.eta_coef * P.eta * aij * ev / r;
          fi0 = fi0 + f_d * e0 / mi;
          fi1 = fi1 + f_d * e1 / mi;
          if (DIM == 3) fi2 = fi2 + f_d * e2 / mi;
          if (P.noise) {
            uint32_t pj_ = (uint32_t)pid[j];
            float g[4];
            pair_normals(P.seed, pidi < pj_ ? pidi : pj_, pidi < pj_ ? pj_ : pidi, P.step, g);
            double f_r = (double)g[0] * sqrt(2.0 * P.kt * P.eta) *
                         sqrt(-P.eta_coef * aij / r) / P.sqrt_dt;
            fi0 = fi0 + f_r * e0 / mi;
            fi1 = fi1 + f_r * e1 / mi;
            if (DIM == 3) fi2 = fi2 + f_r * e2 / mi;
          }
        }
        f0 = f0 + fi0;
        f1 = f1 + fi1;
        if (DIM == 3) f2 = f2 + fi2;
      });
  if (P.body_force) {
    if (fabs(P.bf[0]) > 4.440892098500626e-16) f0 = f0 + P.bf[0];
    if (fabs(P.bf[1]) > 4.440892098500626e-16) f1 = f1 + P.bf[1];
    if (DIM == 3 && fabs(P.bf[2]) > 4.440892098500626e-16) f2 = f2 + P.bf[2];
  }
  S.f[i] = f0;
  S.f[n + i] = f1;
  if (DIM == 3) S.f[n2 + i] = f2;
  if (!newtonian) {
    // particles_compute_act.F90:140-155 on this particle's own ct and vgt
    const int d2n = DIM * DIM;
    double cc[9];
    for (int q = 0; q < d2n; ++q) {
      cc[q] = S.ct[(int64_t)q * n + i];
      S.vgt[(int64_t)q * n + i] = L[q];
    }
    for (int b = 0; b < DIM; ++b)
      for (int a = 0; a < DIM; ++a) {
        double t = 0.0;
        for (int k = 0; k < DIM; ++k)
          t = t + cc[a + DIM * k] * L[b + DIM * k] + L[a + DIM * k] * cc[k + DIM * b];
        double id = a == b ? 1.0 : 0.0;
        S.act[(int64_t)(a + DIM * b) * n + i] = t - (cc[a + DIM * b] - id) / P.tau;
      }
  }
  S.nbr_count[i] = (int32_t)cnt;
}

// Neighbour-set signature: per particle, the number of in-cutoff
// neighbours and the wrapping sum of a hash of (pid_i, pid_j).
__device__ __forceinline__ uint32_t pair_hash(uint32_t a, uint32_t b) {
  uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h;
}

template <int DIM>
__global__ __launch_bounds__(256) void k_signature_f64(DevParams P, DeviceState S) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int n = P.n;
  double xi0 = S.x[i], xi1 = S.x[n + i], xi2 = DIM == 3 ? S.x[2 * (int64_t)n + i] : 0.0;
  CellCoord c = decode_key(P, S.key_sorted[i]);
  uint32_t pidi = (uint32_t)S.pid[i];
  int32_t cnt = 0;
  uint32_t hs = 0;
  const int32_t* __restrict__ pid = S.pid;
  for_each_neighbor_f64<DIM>(P, S.x, S.start, i, xi0, xi1, xi2, c,
                             [&](int j, double, double, double, double) {
                               ++cnt;
                               hs += pair_hash(pidi, (uint32_t)pid[j]);
                             });
  S.nbr_count[i] = cnt;
  S.nbr_hash[i] = hs;
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
static inline int nblk(int n) { return (n + 255) / 256; }

void launch_density_f64(hipStream_t s, const DevParams& P, const DeviceState& S) {
  if (P.dim == 3)
    hipLaunchKernelGGL(k_density_f64<3>, dim3(nblk(P.n)), dim3(256), 0, s, P, S);
  else
    hipLaunchKernelGGL(k_density_f64<2>, dim3(nblk(P.n)), dim3(256), 0, s, P, S);
}

void launch_eos(hipStream_t s, const DevParams& P, const DeviceState& S) {
  hipLaunchKernelGGL(k_eos, dim3(nblk(P.n)), dim3(256), 0, s, P, S);
}

void launch_forces_f64(hipStream_t s, const DevParams& P, const DeviceState& S) {
  if (P.dim == 3)
    hipLaunchKernelGGL(k_forces_f64<3>, dim3(nblk(P.n)), dim3(256), 0, s, P, S);
  else
    hipLaunchKernelGGL(k_forces_f64<2>, dim3(nblk(P.n)), dim3(256), 0, s, P, S);
}

void launch_signature_f64(hipStream_t s, const DevParams& P, const DeviceState& S) {
  if (P.dim == 3)
    hipLaunchKernelGGL(k_signature_f64<3>, dim3(nblk(P.n)), dim3(256), 0, s, P, S);
  else
    hipLaunchKernelGGL(k_signature_f64<2>, dim3(nblk(P.n)), dim3(256), 0, s, P, S);
}

}  // namespace sdpd
