// sdpd_pair.hip -- K2 density + equation of state and K4 pair forces
// (gfx950), FP64 and mixed precision.
//
// Replaces the two hot loops of the reference: particles_compute_density
// (particles_compute_density.F90:578-867) and particles_compute_interaction
// (particles_compute_interaction.F90:618-911 with pp_interaction.inc,
// pp_interaction_ff.inc, pp_vgt_ip.inc and rhs_force_ff_*).  The reference
// visits each unordered pair once in a half stencil and adds the result to
// both particles and to periodic ghost copies that are folded back later.
// Here every particle gathers over the 27 (2D: 9) cells around its own
// cell with the periodic image applied per stencil cell and writes only
// its own sums: no atomics, no ghosts, identical results run to run.
//
// Two phases per wavefront, so that the expensive pair body runs on full
// lanes: (A) each lane scans its candidates (cells in fixed order, cell
// members in cell order) and appends the in-cutoff ones to a private LDS
// list; (B) once any lane's list is full, or all candidates are scanned,
// every lane evaluates its list.  The order in which a particle adds its
// pair terms is therefore the fixed scan order.
//
// Cut-off decision: FP64 mode forms r^2 exactly as the reference does
// (pp_interaction.inc:37-46).  Mixed mode tests r^2 in FP32 on positions
// relative to the cell corner and re-decides in FP64, with the reference's
// image convention, every candidate within a relative band of 2e-5 around
// rc^2 -- wider than the FP32 error bound -- so neighbour sets stay bit-exact.
#include "sdpd_kernels.h"

namespace sdpd {

constexpr int kWave = 64;
constexpr int kBuf = 32;          // list entries per lane between flushes
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
constexpr int kCodeShift = 26;    // list entry = j | (stencil code << 26)
constexpr uint32_t kIndexMask = (1u << kCodeShift) - 1u;

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

// One stencil cell seen from a home cell: wrapped index, per-dimension
// image mode (see wrap_cell) and the FP32 offset of its corner.
struct CellInfo {
  int nc;
  int m0, m1, m2;
  float o0, o1, o2;
};

template <int DIM>
__device__ __forceinline__ CellInfo cell_info(const DevParams& P, const CellCoord& c, int code) {
  CellInfo ci;
  int ox = code % 3 - 1;
  int oy = (code / 3) % 3 - 1;
  int oz = DIM == 3 ? code / 9 - 1 : 0;
  int nx = wrap_cell(c.cx + ox, P.nm[0], P.symmetry, ci.m0);
  int ny = wrap_cell(c.cy + oy, P.nm[1], P.symmetry, ci.m1);
  int nz = 0;
  ci.m2 = 0;
  if (DIM == 3) nz = wrap_cell(c.cz + oz, P.nm[2], P.symmetry, ci.m2);
  ci.nc = nx + P.nm[0] * (ny + P.nm[1] * nz);
  ci.o0 = (float)ox * P.csf[0];
  ci.o1 = (float)oy * P.csf[1];
  ci.o2 = (float)oz * P.csf[2];
  return ci;
}

// Exact FP64 difference and squared distance for a candidate.
template <int DIM>
__device__ __forceinline__ double exact_d2(const DevParams& P, const double* __restrict__ x, int j,
                                           double xi0, double xi1, double xi2, const CellInfo& ci,
                                           double& d0, double& d1, double& d2) {
  const int64_t n = P.n;
  d0 = image_diff(xi0, x[j], P.length[0], ci.m0);
  d1 = image_diff(xi1, x[n + j], P.length[1], ci.m1);
  d2 = DIM == 3 ? image_diff(xi2, x[2 * n + j], P.length[2], ci.m2) : 0.0;
  return DIM == 3 ? dist2_exact(d0, d1, d2) : dist2_exact2(d0, d1);
}

// ---------------------------------------------------------------------
// the two-phase gather driver
// ---------------------------------------------------------------------
// Pol provides:
//   bool test(int j, const CellInfo&)   in-cutoff decision for candidate j
//   void body(int j, const CellInfo&)   accumulate the pair term
template <int DIM, class Pol>
__device__ __forceinline__ void gather(const DevParams& P, const int32_t* __restrict__ start, int i,
                                       bool active, const CellCoord& c, Pol& pol,
                                       uint32_t* __restrict__ buf) {
  const int lane = threadIdx.x & (kWave - 1);
  const int ncode = DIM == 3 ? 27 : 9;
  int code = -1, j = 0, e = 0;
  bool done = !active;
  CellInfo ci;
  auto advance = [&]() {
    while (j >= e) {
      ++code;
      if (code >= ncode) {
        done = true;
        return;
      }
      ci = cell_info<DIM>(P, c, code);
      j = start[ci.nc];
      e = start[ci.nc + 1];
    }
  };
  if (!done) advance();
  int cnt = 0;
  while (true) {
    // phase A: scan until a lane's list is full or every lane is done
    while (true) {
      if (__ballot(cnt >= kBuf) != 0ull) break;
      if (__ballot(!done) == 0ull) break;
      if (!done) {
        if (j != i && pol.test(j, ci)) {
          buf[cnt * kWave + lane] = (uint32_t)j | ((uint32_t)code << kCodeShift);
          ++cnt;
        }
        ++j;
        if (j >= e) advance();
      }
    }
    // phase B: every lane evaluates its list
    for (int q = 0; __ballot(q < cnt) != 0ull; ++q) {
      if (q < cnt) {
        uint32_t ent = buf[q * kWave + lane];
        int jj = (int)(ent & kIndexMask);
        CellInfo cj = cell_info<DIM>(P, c, (int)(ent >> kCodeShift));
        pol.body(jj, cj);
      }
    }
    cnt = 0;
    if (__ballot(!done) == 0ull) break;
  }
}

// ---------------------------------------------------------------------
// cut-off decisions
// ---------------------------------------------------------------------
template <int DIM>
struct DecideF64 {
  const DevParams& P;
  const double* __restrict__ x;
  double xi0, xi1, xi2;
  __device__ __forceinline__ bool operator()(int j, const CellInfo& ci) const {
    double d0, d1, d2;
    return exact_d2<DIM>(P, x, j, xi0, xi1, xi2, ci, d0, d1, d2) < P.cut2;
  }
};

template <int DIM>
struct DecideMixed {
  const DevParams& P;
  const double* __restrict__ x;
  const float4* __restrict__ xr;
  double xi0, xi1, xi2;
  float a0, a1, a2;  // own position relative to the own cell corner
  int rechecks;
  __device__ __forceinline__ float d2f(int j, const CellInfo& ci) const {
    float4 q = xr[j];
    float d0 = a0 - (q.x + ci.o0);
    float d1 = a1 - (q.y + ci.o1);
    float s = d0 * d0 + d1 * d1;
    if (DIM == 3) {
      float d2 = a2 - (q.z + ci.o2);
      s = s + d2 * d2;
    }
    return s;
  }
  __device__ __forceinline__ bool operator()(int j, const CellInfo& ci) {
    float s = d2f(j, ci);
    if (s < P.cut2_lo_f32) return true;
    if (s >= P.cut2_hi_f32) return false;
    ++rechecks;
    double d0, d1, d2;
    return exact_d2<DIM>(P, x, j, xi0, xi1, xi2, ci, d0, d1, d2) < P.cut2;
  }
};

// ---------------------------------------------------------------------
// K2: density + EOS
// ---------------------------------------------------------------------
template <int DIM>
struct DensityF64 {
  const DevParams& P;
  DecideF64<DIM> dec;
  const double* __restrict__ m;
  double acc;
  int cnt;
  __device__ __forceinline__ bool test(int j, const CellInfo& ci) { return dec(j, ci); }
  __device__ __forceinline__ void body(int j, const CellInfo& ci) {
    double d0, d1, d2;
    double r2 = exact_d2<DIM>(P, dec.x, j, dec.xi0, dec.xi1, dec.xi2, ci, d0, d1, d2);
    double w, gw;
    kernel_w_gradw(P, sqrt(r2), w, gw);
    acc = acc + (P.density_type == 1 ? m[j] * w : w);
    ++cnt;
  }
};

// compensated FP32 accumulator (Kahan)
struct KahanF {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

template <int DIM>
struct DensityMixed {
  const DevParams& P;
  DecideMixed<DIM> dec;
  const double* __restrict__ m;
  KahanF acc;
  int cnt;
  __device__ __forceinline__ bool test(int j, const CellInfo& ci) { return dec(j, ci); }
  __device__ __forceinline__ void body(int j, const CellInfo& ci) {
    float r = __builtin_sqrtf(dec.d2f(j, ci));
    float w, gw;
    kernel_w_gradw_f32(P.kernel_type, r * P.inv_hf, P.coeff, P.coef_gradf, w, gw);
    acc.add(P.density_type == 1 ? (float)m[j] * w : w);
    ++cnt;
  }
};

template <int DIM, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_density(DevParams P, DeviceState S) {
  __shared__ uint32_t sbuf[kWaves][kBuf * kWave];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool active = i < P.n;
  const int ii = active ? i : 0;
  const int64_t n = P.n;
  double xi0 = S.x[ii], xi1 = S.x[n + ii], xi2 = DIM == 3 ? S.x[2 * n + ii] : 0.0;
  const double mi = S.m[ii];
  CellCoord c = decode_key(P, S.key_sorted[ii]);
  uint32_t* buf = sbuf[threadIdx.x / kWave];
  double rho;
  int cnt, rechecks = 0;
  if (MIXED) {
    float4 a = S.xr[ii];
    DensityMixed<DIM> pol{P, DecideMixed<DIM>{P, S.x, S.xr, xi0, xi1, xi2, a.x, a.y, a.z, 0}, S.m,
                          KahanF{}, 0};
    pol.acc.add(P.density_type == 1 ? (float)(P.w0 * mi) : (float)P.w0);
    gather<DIM>(P, S.start, i, active, c, pol, buf);
    rho = pol.acc.value();
    cnt = pol.cnt;
    rechecks = pol.dec.rechecks;
  } else {
    // self term first (particles_compute_density.F90:369-558)
    DensityF64<DIM> pol{P, DecideF64<DIM>{P, S.x, xi0, xi1, xi2}, S.m,
                        P.density_type == 1 ? P.w0 * mi : P.w0, 0};
    gather<DIM>(P, S.start, i, active, c, pol, buf);
    rho = pol.acc;
    cnt = pol.cnt;
  }
  if (!active) return;
  const double press = eos_pressure(P, P.density_type == 2 ? rho * mi : rho);
  S.rho[i] = rho;
  S.p[i] = press;
  // per-particle factors of the pair forces: p/n^2 and 1/n^2
  const double inv_n2 = 1.0 / (rho * rho);
  S.an[i] = press / (rho * rho);
  S.bn[i] = inv_n2;
  S.nbr_count[i] = cnt;
  S.recheck[i] = rechecks;
  if (MIXED) {
    S.vr[i] = make_float4((float)S.v[i], (float)S.v[n + i], DIM == 3 ? (float)S.v[2 * n + i] : 0.f,
                          (float)(press / (rho * rho)));
    S.pb[i] = make_float2((float)inv_n2, (float)press);
  }
}

__global__ void k_eos(DevParams P, DeviceState S) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  double r = S.rho[i];
  double press = eos_pressure(P, P.density_type == 2 ? r * S.m[i] : r);
  S.p[i] = press;
  S.an[i] = press / (r * r);
  S.bn[i] = 1.0 / (r * r);
}

// ---------------------------------------------------------------------
// K4: pair forces
// ---------------------------------------------------------------------

// Random force direction seen by particle i for neighbour j: a Gaussian
// vector with the covariance of the reference's symmetric traceless
// dW . e_ij (I/2 + e e^T/6 in 3D, I/2 in 2D), oriented by the lower pid.
template <int DIM, typename T>
__device__ __forceinline__ void noise_vector(const DevParams& P, uint32_t pidi, uint32_t pidj,
                                             T e0, T e1, T e2, T& w0, T& w1, T& w2) {
  uint32_t lo = pidi < pidj ? pidi : pidj, hi = pidi < pidj ? pidj : pidi;
  float g[4];
  pair_normals(P.seed, lo, hi, P.step, g);
  const T sgn = pidi < pidj ? T(1) : T(-1);
  const T ca = T(0.70710678118654752440);
  if (DIM == 3) {
    const T cb = T(0.81649658092772603273) - T(0.70710678118654752440);
    T c0 = sgn * e0, c1 = sgn * e1, c2 = sgn * e2;
    T ge = (T)g[0] * c0 + (T)g[1] * c1 + (T)g[2] * c2;
    w0 = sgn * (ca * (T)g[0] + cb * ge * c0);
    w1 = sgn * (ca * (T)g[1] + cb * ge * c1);
    w2 = sgn * (ca * (T)g[2] + cb * ge * c2);
  } else {
    w0 = sgn * (ca * (T)g[0]);
    w1 = sgn * (ca * (T)g[1]);
    w2 = T(0);
  }
}

// force types: 3 Hu-Adams, 4 Hu-Adams angular, 5 Oldroyd-B (non-Newtonian Hu-Adams)
template <int DIM, int FT>
struct ForceF64 {
  const DevParams& P;
  DecideF64<DIM> dec;
  const DeviceState& S;
  double vi0, vi1, vi2, ai, bi, pi, inv_mi, ni;
  uint32_t pidi;
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  double L[9];
  double pti[9];
  int cnt = 0;
  __device__ __forceinline__ bool test(int j, const CellInfo& ci) { return dec(j, ci); }
  __device__ __forceinline__ void body(int j, const CellInfo& ci) {
    const int64_t n = P.n;
    double d0, d1, d2;
    double r2 = exact_d2<DIM>(P, dec.x, j, dec.xi0, dec.xi1, dec.xi2, ci, d0, d1, d2);
    double r = sqrt(r2);
    double w, gw;
    kernel_w_gradw(P, r, w, gw);
    const double ir = 1.0 / r;
    double e0 = d0 * ir, e1 = d1 * ir, e2 = DIM == 3 ? d2 * ir : 0.0;
    double vj0 = S.v[j], vj1 = S.v[n + j], vj2 = DIM == 3 ? S.v[2 * n + j] : 0.0;
    double u0 = vi0 - vj0, u1 = vi1 - vj1, u2 = vi2 - vj2;
    double bj = S.bn[j];
    double fi0, fi1, fi2;
    ++cnt;
    if (FT == 5) {
      // pp_vgt_ip.inc:16-20
      const double du[3] = {vj0 - vi0, vj1 - vi1, vj2 - vi2};
      const double dx[3] = {d0, d1, d2};
      const double gfac = gw / ni * ir;
#pragma unroll
      for (int b = 0; b < DIM; ++b)
#pragma unroll
        for (int a = 0; a < DIM; ++a) L[a + DIM * b] = L[a + DIM * b] + du[a] * dx[b] * gfac;
      // rhs_force_ff_non_Newtonian_HuAdams.F90:105-131 with
      // Pi = (p + G) I - G c (particles_compute_pressure_tensor.F90:114-133)
      const double pj = S.p[j];
      const double ev[3] = {e0, e1, e2};
      double fc[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int jj = 0; jj < DIM; ++jj)
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          double ptj = -P.G * S.ct[(int64_t)(a + DIM * jj) * n + j];
          if (a == jj) ptj = ptj + pj + P.G;
          fc[jj] = fc[jj] + pti[a + DIM * jj] * ev[a] * bi + ptj * ev[a] * bj;
        }
      fi0 = 0.0 - fc[0] * gw * inv_mi;
      fi1 = 0.0 - fc[1] * gw * inv_mi;
      fi2 = DIM == 3 ? 0.0 - fc[2] * gw * inv_mi : 0.0;
      const double f_d = P.eta * (bi + bj) * gw * ir;
      fi0 = fi0 + f_d * u0 * inv_mi;
      fi1 = fi1 + f_d * u1 * inv_mi;
      if (DIM == 3) fi2 = fi2 + f_d * u2 * inv_mi;
      if (P.noise) {
        double w0, w1, w2;
        noise_vector<DIM, double>(P, pidi, (uint32_t)S.pid[j], e0, e1, e2, w0, w1, w2);
        const double f_r = sqrt(-4.0 * P.kt * f_d) * P.inv_sqrt_dt;
        fi0 = fi0 + f_r * w0 * inv_mi;
        fi1 = fi1 + f_r * w1 * inv_mi;
        if (DIM == 3) fi2 = fi2 + f_r * w2 * inv_mi;
      }
    } else if (FT == 3) {
      // rhs_force_ff_Newtonian_HuAdams.F90:140-237
      const double f_c = (ai + S.an[j]) * gw;
      fi0 = 0.0 - f_c * e0 * inv_mi;
      fi1 = 0.0 - f_c * e1 * inv_mi;
      fi2 = DIM == 3 ? 0.0 - f_c * e2 * inv_mi : 0.0;
      const double f_d = P.eta * (bi + bj) * gw * ir;
      fi0 = fi0 + f_d * u0 * inv_mi;
      fi1 = fi1 + f_d * u1 * inv_mi;
      if (DIM == 3) fi2 = fi2 + f_d * u2 * inv_mi;
      if (P.noise) {
        double w0, w1, w2;
        noise_vector<DIM, double>(P, pidi, (uint32_t)S.pid[j], e0, e1, e2, w0, w1, w2);
        const double f_r = sqrt(-4.0 * P.kt * f_d) * P.inv_sqrt_dt;
        fi0 = fi0 + f_r * w0 * inv_mi;
        fi1 = fi1 + f_r * w1 * inv_mi;
        if (DIM == 3) fi2 = fi2 + f_r * w2 * inv_mi;
      }
    } else {
      // rhs_force_ff_Newtonian_HuAdams_angular.F90:166-229
      double ev = e0 * u0 + e1 * u1;
      if (DIM == 3) ev = ev + e2 * u2;
      const double pij = (pi + S.p[j]) * 0.5;
      const double aij = (bi + bj) * gw;
      const double f_c = aij * pij;
      fi0 = 0.0 - f_c * e0 * inv_mi;
      fi1 = 0.0 - f_c * e1 * inv_mi;
      fi2 = DIM == 3 ? 0.0 - f_c * e2 * inv_mi : 0.0;
      const double f_d = P.eta_coef * P.eta * aij * ev * ir;
      fi0 = fi0 + f_d * e0 * inv_mi;
      fi1 = fi1 + f_d * e1 * inv_mi;
      if (DIM == 3) fi2 = fi2 + f_d * e2 * inv_mi;
      if (P.noise) {
        const uint32_t pj_ = (uint32_t)S.pid[j];
        float g[4];
        pair_normals(P.seed, pidi < pj_ ? pidi : pj_, pidi < pj_ ? pj_ : pidi, P.step, g);
        const double f_r = (double)g[0] * sqrt(2.0 * P.kt * P.eta) *
                           sqrt(-P.eta_coef * aij * ir) * P.inv_sqrt_dt;
        fi0 = fi0 + f_r * e0 * inv_mi;
        fi1 = fi1 + f_r * e1 * inv_mi;
        if (DIM == 3) fi2 = fi2 + f_r * e2 * inv_mi;
      }
    }
    f0 = f0 + fi0;
    f1 = f1 + fi1;
    if (DIM == 3) f2 = f2 + fi2;
  }
};

template <int DIM, int FT>
struct ForceMixed {
  const DevParams& P;
  DecideMixed<DIM> dec;
  const DeviceState& S;
  float vi0, vi1, vi2, ai, bi, pi;
  uint32_t pidi;
  KahanF f0, f1, f2;
  int cnt = 0;
  __device__ __forceinline__ bool test(int j, const CellInfo& ci) { return dec(j, ci); }
  __device__ __forceinline__ void body(int j, const CellInfo& ci) {
    float4 q = S.xr[j];
    float d0 = dec.a0 - (q.x + ci.o0);
    float d1 = dec.a1 - (q.y + ci.o1);
    float d2 = DIM == 3 ? dec.a2 - (q.z + ci.o2) : 0.f;
    float r2 = d0 * d0 + d1 * d1 + d2 * d2;
    float r = __builtin_sqrtf(r2);
    float ir = __builtin_amdgcn_rcpf(r);
    float w, gw;
    kernel_w_gradw_f32(P.kernel_type, r * P.inv_hf, P.coeff, P.coef_gradf, w, gw);
    float e0 = d0 * ir, e1 = d1 * ir, e2 = d2 * ir;
    float4 vj = S.vr[j];
    float2 pbj = S.pb[j];
    float u0 = vi0 - vj.x, u1 = vi1 - vj.y, u2 = vi2 - vj.z;
    float fi0, fi1, fi2;
    ++cnt;
    if (FT == 3) {
      const float f_c = (ai + vj.w) * gw;
      const float f_d = P.etaf * (bi + pbj.x) * gw * ir;
      fi0 = f_d * u0 - f_c * e0;
      fi1 = f_d * u1 - f_c * e1;
      fi2 = f_d * u2 - f_c * e2;
      if (P.noise) {
        float w0, w1, w2;
        noise_vector<DIM, float>(P, pidi, __float_as_uint(q.w), e0, e1, e2, w0, w1, w2);
        const float f_r = __builtin_sqrtf(-4.0f * P.ktf * f_d) * P.inv_sqrt_dtf;
        fi0 = fi0 + f_r * w0;
        fi1 = fi1 + f_r * w1;
        fi2 = fi2 + f_r * w2;
      }
    } else {
      const float ev = e0 * u0 + e1 * u1 + e2 * u2;
      const float pij = (pi + pbj.y) * 0.5f;
      const float aij = (bi + pbj.x) * gw;
      const float f_c = aij * pij;
      float f_d = P.eta_coeff * P.etaf * aij * ev * ir;
      float s = f_d - f_c;
      if (P.noise) {
        const uint32_t pj_ = __float_as_uint(q.w);
        float g[4];
        pair_normals(P.seed, pidi < pj_ ? pidi : pj_, pidi < pj_ ? pj_ : pidi, P.step, g);
        s = s + g[0] * P.sqrt_2kt_etaf * __builtin_sqrtf(-P.eta_coeff * aij * ir) * P.inv_sqrt_dtf;
      }
      fi0 = s * e0;
      fi1 = s * e1;
      fi2 = s * e2;
    }
    f0.add(fi0);
    f1.add(fi1);
    if (DIM == 3) f2.add(fi2);
  }
};

template <int DIM>
__device__ __forceinline__ void finish_force(const DevParams& P, const DeviceState& S, int i,
                                             double f0, double f1, double f2) {
  const int64_t n = P.n;
  if (P.body_force) {
    // particles_apply_body_force.F90:125-193
    if (fabs(P.bf[0]) > 4.440892098500626e-16) f0 = f0 + P.bf[0];
    if (fabs(P.bf[1]) > 4.440892098500626e-16) f1 = f1 + P.bf[1];
    if (DIM == 3 && fabs(P.bf[2]) > 4.440892098500626e-16) f2 = f2 + P.bf[2];
  }
  S.f[i] = f0;
  S.f[n + i] = f1;
  if (DIM == 3) S.f[2 * n + i] = f2;
}

template <int DIM, int FT, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_forces(DevParams P, DeviceState S) {
  __shared__ uint32_t sbuf[kWaves][kBuf * kWave];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool active = i < P.n;
  const int ii = active ? i : 0;
  const int64_t n = P.n;
  double xi0 = S.x[ii], xi1 = S.x[n + ii], xi2 = DIM == 3 ? S.x[2 * n + ii] : 0.0;
  CellCoord c = decode_key(P, S.key_sorted[ii]);
  uint32_t* buf = sbuf[threadIdx.x / kWave];
  const double inv_mi = 1.0 / S.m[ii];
  if (MIXED) {
    float4 a = S.xr[ii];
    float4 vi = S.vr[ii];
    float2 pbi = S.pb[ii];
    ForceMixed<DIM, FT> pol{P, DecideMixed<DIM>{P, S.x, S.xr, xi0, xi1, xi2, a.x, a.y, a.z, 0}, S,
                            vi.x, vi.y, vi.z, vi.w, pbi.x, pbi.y, __float_as_uint(a.w)};
    gather<DIM>(P, S.start, i, active, c, pol, buf);
    if (!active) return;
    finish_force<DIM>(P, S, i, pol.f0.value() * inv_mi, pol.f1.value() * inv_mi,
                      DIM == 3 ? pol.f2.value() * inv_mi : 0.0);
    S.nbr_count[i] = pol.cnt;
    S.recheck[i] = pol.dec.rechecks;
  } else {
    ForceF64<DIM, FT> pol{P, DecideF64<DIM>{P, S.x, xi0, xi1, xi2}, S};
    pol.vi0 = S.v[ii];
    pol.vi1 = S.v[n + ii];
    pol.vi2 = DIM == 3 ? S.v[2 * n + ii] : 0.0;
    pol.ai = S.an[ii];
    pol.bi = S.bn[ii];
    pol.pi = S.p[ii];
    pol.ni = S.rho[ii];
    pol.inv_mi = inv_mi;
    pol.pidi = (uint32_t)S.pid[ii];
#pragma unroll
    for (int q = 0; q < 9; ++q) pol.L[q] = 0.0;
    if (FT == 5) {
      // particles_compute_pressure_tensor.F90:114-133
#pragma unroll
      for (int b = 0; b < DIM; ++b)
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          double val = -P.G * S.ct[(int64_t)(a + DIM * b) * n + ii];
          if (a == b) val = val + pol.pi + P.G;
          pol.pti[a + DIM * b] = val;
        }
    }
    gather<DIM>(P, S.start, i, active, c, pol, buf);
    if (!active) return;
    finish_force<DIM>(P, S, i, pol.f0, pol.f1, pol.f2);
    S.nbr_count[i] = pol.cnt;
    S.recheck[i] = 0;
    if (FT == 5) {
      // particles_compute_act.F90:140-155 on this particle's own ct and vgt
      const int d2n = DIM * DIM;
      double cc[9];
#pragma unroll
      for (int q = 0; q < d2n; ++q) {
        cc[q] = S.ct[(int64_t)q * n + i];
        S.vgt[(int64_t)q * n + i] = pol.L[q];
      }
#pragma unroll
      for (int b = 0; b < DIM; ++b)
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          double t = 0.0;
#pragma unroll
          for (int k = 0; k < DIM; ++k)
            t = t + cc[a + DIM * k] * pol.L[b + DIM * k] + pol.L[a + DIM * k] * cc[k + DIM * b];
          double id = a == b ? 1.0 : 0.0;
          S.act[(int64_t)(a + DIM * b) * n + i] = t - (cc[a + DIM * b] - id) / P.tau;
        }
    }
  }
}

// ---------------------------------------------------------------------
// neighbour-set signature: per particle, the number of in-cutoff
// neighbours and the wrapping sum of a hash of (pid_i, pid_j)
// ---------------------------------------------------------------------
__device__ __forceinline__ uint32_t pair_hash(uint32_t a, uint32_t b) {
  uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return h;
}

template <class Dec>
struct SignaturePol {
  Dec dec;
  const int32_t* __restrict__ pid;
  uint32_t pidi;
  int cnt;
  uint32_t hs;
  __device__ __forceinline__ bool test(int j, const CellInfo& ci) { return dec(j, ci); }
  __device__ __forceinline__ void body(int j, const CellInfo&) {
    ++cnt;
    hs += pair_hash(pidi, (uint32_t)pid[j]);
  }
};

template <int DIM, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_signature(DevParams P, DeviceState S) {
  __shared__ uint32_t sbuf[kWaves][kBuf * kWave];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool active = i < P.n;
  const int ii = active ? i : 0;
  const int64_t n = P.n;
  double xi0 = S.x[ii], xi1 = S.x[n + ii], xi2 = DIM == 3 ? S.x[2 * n + ii] : 0.0;
  CellCoord c = decode_key(P, S.key_sorted[ii]);
  uint32_t* buf = sbuf[threadIdx.x / kWave];
  const uint32_t pidi = (uint32_t)S.pid[ii];
  int cnt;
  uint32_t hs;
  int rechecks = 0;
  if (MIXED) {
    float4 a = S.xr[ii];
    SignaturePol<DecideMixed<DIM>> pol{
        DecideMixed<DIM>{P, S.x, S.xr, xi0, xi1, xi2, a.x, a.y, a.z, 0}, S.pid, pidi, 0, 0u};
    gather<DIM>(P, S.start, i, active, c, pol, buf);
    cnt = pol.cnt;
    hs = pol.hs;
    rechecks = pol.dec.rechecks;
  } else {
    SignaturePol<DecideF64<DIM>> pol{DecideF64<DIM>{P, S.x, xi0, xi1, xi2}, S.pid, pidi, 0, 0u};
    gather<DIM>(P, S.start, i, active, c, pol, buf);
    cnt = pol.cnt;
    hs = pol.hs;
  }
  if (!active) return;
  S.nbr_count[i] = cnt;
  S.nbr_hash[i] = hs;
  S.recheck[i] = rechecks;
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
static inline int nblk(int n) { return (n + kBlock - 1) / kBlock; }

void launch_density(hipStream_t s, const DevParams& P, const DeviceState& S, bool mixed) {
  dim3 g(nblk(P.n)), b(kBlock);
  if (P.dim == 3) {
    if (mixed)
      hipLaunchKernelGGL((k_density<3, true>), g, b, 0, s, P, S);
    else
      hipLaunchKernelGGL((k_density<3, false>), g, b, 0, s, P, S);
  } else {
    if (mixed)
      hipLaunchKernelGGL((k_density<2, true>), g, b, 0, s, P, S);
    else
      hipLaunchKernelGGL((k_density<2, false>), g, b, 0, s, P, S);
  }
}

void launch_eos(hipStream_t s, const DevParams& P, const DeviceState& S) {
  hipLaunchKernelGGL(k_eos, dim3((P.n + 255) / 256), dim3(256), 0, s, P, S);
}

template <int DIM>
static void launch_forces_dim(hipStream_t s, const DevParams& P, const DeviceState& S,
                              bool mixed) {
  dim3 g(nblk(P.n)), b(kBlock);
  const int ft = P.newtonian ? P.force_type : 5;
  if (mixed) {
    if (ft == 3)
      hipLaunchKernelGGL((k_forces<DIM, 3, true>), g, b, 0, s, P, S);
    else
      hipLaunchKernelGGL((k_forces<DIM, 4, true>), g, b, 0, s, P, S);
  } else {
    if (ft == 3)
      hipLaunchKernelGGL((k_forces<DIM, 3, false>), g, b, 0, s, P, S);
    else if (ft == 4)
      hipLaunchKernelGGL((k_forces<DIM, 4, false>), g, b, 0, s, P, S);
    else
      hipLaunchKernelGGL((k_forces<DIM, 5, false>), g, b, 0, s, P, S);
  }
}

void launch_forces(hipStream_t s, const DevParams& P, const DeviceState& S, bool mixed) {
  if (P.dim == 3)
    launch_forces_dim<3>(s, P, S, mixed);
  else
    launch_forces_dim<2>(s, P, S, mixed);
}

void launch_signature(hipStream_t s, const DevParams& P, const DeviceState& S, bool mixed) {
  dim3 g(nblk(P.n)), b(kBlock);
  if (P.dim == 3) {
    if (mixed)
      hipLaunchKernelGGL((k_signature<3, true>), g, b, 0, s, P, S);
    else
      hipLaunchKernelGGL((k_signature<3, false>), g, b, 0, s, P, S);
  } else {
    if (mixed)
      hipLaunchKernelGGL((k_signature<2, true>), g, b, 0, s, P, S);
    else
      hipLaunchKernelGGL((k_signature<2, false>), g, b, 0, s, P, S);
  }
}

}  // namespace sdpd
