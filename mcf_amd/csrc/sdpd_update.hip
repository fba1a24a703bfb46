// sdpd_update.hip -- K5 velocity-Verlet / Oldroyd-B streaming updates and
// K6 compensated statistics (gfx950).
//
// Replaces particles_integrate_velocity (particles_integrate_velocity.F90:84-94),
// particles_integrate_position (particles_integrate_position.F90:82-123),
// particles_adjust_particles' periodic branch (particles_adjust_particles.F90:
// 154-161, 272-279), particles_integrate_ct (particles_integrate_ct.F90:66-68)
// and statistic_compute_statistic (statistic_compute_statistic.F90:230-283).
// Expressions keep the reference's evaluation order: v + (lambda*f)*dt,
// x + v*dt, ct + (act*dt)*lambda.
#include "sdpd_kernels.h"

namespace sdpd {

// Colloid boundary particles (sid > 0) are not integrated: they move with
// their colloid (particles_integrate_velocity.F90:84-94 and
// particles_integrate_position.F90:82-123 act on fluid particles only).
__device__ __forceinline__ bool is_fluid(const DevParams& P, const DeviceState& S, int i) {
  return P.ncoll == 0 || S.sid[i] == 0;
}

__global__ void k_kick(DevParams P, DeviceState S, double lambda) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  if (!is_fluid(P, S, i)) return;
  const int64_t n = P.n;
  for (int d = 0; d < P.dim; ++d)
    S.v[d * n + i] = S.v[d * n + i] + lambda * S.f[d * n + i] * P.dt;
}

// One dimension D of the wrap; true when the particle crossed a
// Lees-Edwards face, which ends its check.  Crossing that face moves x by
// the shear length, modulo its period, and wraps the third coordinate
// (compile-time dimensions throughout, so the coordinates stay in
// registers).
template <int DIM, int D, class Geo>
__device__ __forceinline__ bool adjust_dim(const DevParams& P, const Geo& g, double* xv,
                                           double* vv, double le_time) {
  const double lo = g.lo[D], L = g.L[D];
  int side = 0;  // -1: crossed the lower face, +1: the upper one
  if (xv[D] < lo) {
    xv[D] = xv[D] + L;
    side = -1;
  }
  // (a periodic dimension checks the upper face on the shifted value too)
  if (!(side < 0 && D == P.le_g) && xv[D] >= lo + L) {
    xv[D] = xv[D] - L;
    side = 1;
  }
  if (side == 0 || D != P.le_g) return false;
  // shear_length(x, lower) = (v_lower - v_upper) t, shear_length(x, upper)
  // = (v_upper - v_lower) t; the velocity jump has the opposite sign
  const double du = P.le_vhi - P.le_vlo;
  const double sl = side < 0 ? -du * le_time : du * le_time;
  xv[0] = f_modulo(xv[0] - sl, g.L[0]);
  if (DIM == 3) {
    constexpr int K = D == 1 ? 2 : 1;
    xv[K] = f_modulo(xv[K], g.L[K]);
  }
  vv[0] = side < 0 ? vv[0] + du : vv[0] - du;
  return true;
}

// Periodic and Lees-Edwards faces after a drift (particles_adjust_particles
// .F90:137-389): dimensions in order; crossing a Lees-Edwards face moves
// every other coordinate k by MODULO(x_k - shear_length_k, L_k) (the box
// starts at 0 then; the shear length is zero but along x, so for the third
// coordinate this is a plain periodic wrap), adds the face velocity jump,
// and ends the particle's check.
template <int DIM, class Geo>
__device__ __forceinline__ void adjust_particle(const DevParams& P, const Geo& g, double* xv,
                                                double* vv, double le_time) {
  if (adjust_dim<DIM, 0>(P, g, xv, vv, le_time)) return;
  if (adjust_dim<DIM, 1>(P, g, xv, vv, le_time)) return;
  if (DIM == 3) (void)adjust_dim<DIM, DIM == 3 ? 2 : 1>(P, g, xv, vv, le_time);
}

__global__ void k_drift_wrap(DevParams P, DeviceState S) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int64_t n = P.n;
  const RepGeom& g = S.geo[S.rep[i]];
  double xv[3] = {0.0, 0.0, 0.0}, vv[3] = {0.0, 0.0, 0.0};
  const bool fluid = is_fluid(P, S, i);
  for (int d = 0; d < P.dim; ++d) {
    vv[d] = S.v[d * n + i];
    xv[d] = fluid ? S.x[d * n + i] + vv[d] * P.dt : S.x[d * n + i];
  }
  if (P.dim == 3)
    adjust_particle<3>(P, g, xv, vv, S.time_dev[1]);
  else
    adjust_particle<2>(P, g, xv, vv, S.time_dev[1]);
  for (int d = 0; d < P.dim; ++d) {
    S.x[d * n + i] = xv[d];
    S.v[d * n + i] = vv[d];
  }
}

__global__ void k_conformation(DevParams P, DeviceState S, double lambda) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int64_t n = P.n;
  const int d2 = P.dim * P.dim;
  for (int q = 0; q < d2; ++q)
    S.ct[q * n + i] = S.ct[q * n + i] + S.act[q * n + i] * P.dt * lambda;
}

// Verlet lists: particle i's displacement since the lists were built and,
// in mixed precision, its FP32 position relative to the corner of the cell
// it was sorted into then.  Per dimension the displacement x - x_ref is
// taken to the nearest period, so a particle that crossed a periodic face
// keeps its build cell: its FP32 position goes on from there, unwrapped,
// and stays consistent with the stencil offsets stored in the lists.
// Returns the squared displacement.
//
// FP64 runs without sheared faces keep the FP32 positions of the build:
// their walks take distances from the FP64 positions, and the only readers
// of the FP32 ones between builds (the rescans of overflowed lists and the
// neighbour-set signature) screen at rc + skin.  Every particle has moved
// at most skin / 2 since the build, so the build-time positions put every
// pair that is within rc now inside that screen, and the exact FP64
// decision follows.  With sheared faces the scan's windows follow the
// current shear length and FP32 positions, which are kept current.
// A replica's corner and edges in registers.
struct BoxFrame {
  double lo[3], L[3];
};
template <int DIM>
__device__ __forceinline__ BoxFrame load_frame(const RepGeom& g) {
  BoxFrame f;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    f.lo[d] = d < DIM ? g.lo[d] : 0.0;
    f.L[d] = d < DIM ? g.L[d] : 1.0;
  }
  return f;
}

template <int DIM>
__device__ __forceinline__ double track_particle(const DevParams& P, const DeviceState& S,
                                                 const RepGeom& g, const BoxFrame& f, int i,
                                                 const double* xv) {
  const int64_t n = P.n;
  double w[3] = {0.0, 0.0, 0.0};
  double d2 = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double L = f.L[d];
    const double raw = xv[d] - S.x_ref[d * n + i];
    w[d] = raw > 0.5 * L ? -L : (raw < -0.5 * L ? L : 0.0);
    const double dd = raw + w[d];
    d2 = d2 + dd * dd;
  }
  if (P.mixed || P.le_g >= 0) {
    const int kl = S.key_sorted[i] - g.cell_base;
    const int cc[3] = {kl % g.nm[0], (kl / g.nm[0]) % g.nm[1], kl / (g.nm[0] * g.nm[1])};
    float a[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      a[d] = (float)((xv[d] - (g.lo[d] + (double)cc[d] * g.cs[d])) + w[d]);
    const float4 old = S.xa[i];
    S.xa[i] = make_float4(a[0], a[1], a[2], old.w);
  }
  return d2;
}

// The largest squared displacement of the workgroup (256 lanes) into disp.
// Non-negative doubles order like their bit patterns.  One lane per
// workgroup reads the running maximum first and only raises it: after the
// first workgroups nearly none of the 8k per launch needs the atomic, which
// one address would otherwise serialize.
__device__ __forceinline__ void block_max_disp(unsigned long long* disp, double d2) {
  __shared__ double wmax[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) d2 = fmax(d2, __shfl_xor(d2, off));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wmax[wave] = d2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmax[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, wmax[w]);
    const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
    if (m > 0.0 && bits > __atomic_load_n(disp, __ATOMIC_RELAXED)) atomicMax(disp, bits);
  }
}

// Step boundary: the closing half kick of the previous step, then the
// opening half kick, drift, Oldroyd-B half update and periodic wrap of the
// next one.  Then either the cell key and cell histogram of the next sort
// (lists rebuilt every step; cell counts must be zero on entry) or, with
// Verlet lists, the displacement since the lists' build.
template <int DIM, bool VERLET>
__global__ __launch_bounds__(256) void k_boundary(DevParams P, DeviceState S, int closing,
                                                  int opening) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i0 < P.n;
  const int i = act ? i0 : P.n - 1;  // lanes past the end compute but store nothing
  const int64_t n = P.n;
  constexpr int dim = DIM;
  // the replica's corner and edges in registers: their loads issue with
  // the state's (read through a reference into global memory, they were
  // issued one dimension after another inside the wrap)
  const RepGeom& g = S.geo[S.rep[i]];
  const BoxFrame fr = load_frame<DIM>(g);
  double xv[3] = {0.0, 0.0, 0.0}, vw[3] = {0.0, 0.0, 0.0};
  const bool fluid = is_fluid(P, S, i);
#pragma unroll
  for (int d = 0; d < dim; ++d) {
    double vv = S.v[d * n + i];
    double ff = S.f[d * n + i];
    if (closing && fluid) vv = vv + 0.5 * ff * P.dt;
    if (opening && fluid) vv = vv + 0.5 * ff * P.dt;
    vw[d] = vv;
    double x = S.x[d * n + i];
    if (opening && fluid) x = x + vv * P.dt;
    xv[d] = x;
  }
  if (opening) adjust_particle<DIM>(P, fr, xv, vw, S.time_dev[1]);
  if (act) {
#pragma unroll
    for (int d = 0; d < dim; ++d) {
      S.v[d * n + i] = vw[d];
      S.x[d * n + i] = xv[d];
    }
    if (!P.newtonian) {
      constexpr int d2 = dim * dim;
#pragma unroll
      for (int q = 0; q < d2; ++q) {
        double c = S.ct[q * n + i];
        double a = S.act[q * n + i];
        if (closing) c = c + a * P.dt * 0.5;
        if (opening) c = c + a * P.dt * 0.5;
        S.ct[q * n + i] = c;
      }
    }
  }
  if (VERLET) {
    const double d2 = act ? track_particle<DIM>(P, S, g, fr, i, xv) : 0.0;
    block_max_disp(S.disp, d2);
    return;
  }
  // cell histogram of the next sort; the particles are in the cell order
  // of the previous one and move little
  const int32_t k = cell_key_geo(g, dim, xv[0], xv[1], xv[2]);
  const int32_t sl = histogram_slot(S.count, k, act);
  if (act) {
    S.key[i] = k;
    S.slot[i] = sl;
  }
}

// Displacement tracking alone (Verlet lists, after a separate drift).
__global__ __launch_bounds__(256) void k_track(DevParams P, DeviceState S) {
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i0 < P.n;
  const int i = act ? i0 : P.n - 1;
  const int64_t n = P.n;
  double xv[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < P.dim; ++d) xv[d] = S.x[d * n + i];
  const RepGeom& g = S.geo[S.rep[i]];
  const double d2 = !act ? 0.0
                          : (P.dim == 3 ? track_particle<3>(P, S, g, load_frame<3>(g), i, xv)
                                        : track_particle<2>(P, S, g, load_frame<2>(g), i, xv));
  block_max_disp(S.disp, d2);
}

// The step's gate: rebuild the lists when a particle moved more than
// skin / 2 since their build (a pair then may have closed in by more than
// the skin); the displacement maximum starts again from zero.  gate[1]
// counts the builds; gate[2], set by the build, flags lists that
// overflowed their capacity.
__global__ void k_decide(DevParams P, DeviceState S) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const double d2 = __longlong_as_double((long long)S.disp[0]);
  double lim = P.half_skin2;
  int rebuild = 0;
  if (P.le_g >= 0) {
    // sheared faces: the copies across a face move by |v_upper - v_lower|
    // per unit time against the box, so a pair across it may have closed
    // in by that drift since the build on top of the particles' own
    // displacements: the lists hold while 2 max |dx| + drift < skin.  A
    // list that overflowed rebuilds every step (the rescans then always
    // see the positions of a build).
    const double drift = fabs(P.le_vhi - P.le_vlo) * fabs(S.time_dev[1] - S.time_dev[3]);
    const double room = 0.5 * (P.skin - drift);
    lim = room > 0.0 ? room * room : -1.0;
    rebuild = S.gate[2] != 0 ? 1 : 0;
  }
  rebuild = rebuild | (d2 > lim ? 1 : 0);
  // a build still pending (a sort whose lists no density pass has built
  // yet) stays pending
  S.gate[0] |= rebuild;
  S.gate[1] += rebuild;
  if (rebuild) S.gate[2] = 0;
  S.disp[0] = 0ull;
}

// After the density pass: the lists are current (no build pending).
__global__ void k_lists_current(DeviceState S) {
  if (blockIdx.x == 0 && threadIdx.x == 0) S.gate[0] = 0;
}

// A sort requested outside the step (upload, a cell-list call, a check):
// the next density pass builds the lists.
__global__ void k_force_build(DeviceState S) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  S.gate[0] = 1;
  S.gate[1] += 1;
  S.gate[2] = 0;
  S.disp[0] = 0ull;
}

// Neumaier-compensated sum helpers
struct KSum {
  double s, c;
};
__device__ __forceinline__ void kadd(KSum& k, double v) {
  double t = k.s + v;
  if (fabs(k.s) >= fabs(v))
    k.c += (k.s - t) + v;
  else
    k.c += (v - t) + k.s;
  k.s = t;
}

// Block partial sums of (E_k, P_x, P_y, P_z) over fluid particles; a
// second pass reduces the partials in a fixed order.
__global__ __launch_bounds__(256) void k_stats_partial(DevParams P, DeviceState S) {
  __shared__ double sh[4][256];
  const int64_t n = P.n;
  const int dim = P.dim;
  KSum acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    if (!is_fluid(P, S, (int)i)) continue;  // statistic_compute_statistic.F90:230-246
    double mi = S.m[i];
    double v2 = 0.0;
    for (int d = 0; d < dim; ++d) {
      double vv = S.v[d * n + i];
      v2 = v2 + vv * vv;
      kadd(acc[1 + d], mi * vv);
    }
    kadd(acc[0], 0.5 * mi * v2);
  }
  for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = acc[q].s + acc[q].c;
  __syncthreads();
  if (threadIdx.x < 4) {
    KSum t = {0, 0};
    for (int k = 0; k < 256; ++k) kadd(t, sh[threadIdx.x][k]);
    S.partials[(int64_t)blockIdx.x * 4 + threadIdx.x] = t.s + t.c;
  }
}

__global__ void k_stats_final(int nblocks, DeviceState S) {
  if (threadIdx.x < 4) {
    KSum t = {0, 0};
    for (int b = 0; b < nblocks; ++b) kadd(t, S.partials[(int64_t)b * 4 + threadIdx.x]);
    S.stats_out[threadIdx.x] = t.s + t.c;
  }
}

// Per-replica (E_k, P) with one workgroup per replica over its particles,
// which are contiguous in cell order: [start[cell_base_q], start[cell_base_q+1]).
__global__ __launch_bounds__(256) void k_replica_stats(DevParams P, DeviceState S) {
  __shared__ double sh[4][256];
  const int q = blockIdx.x;
  const int64_t n = P.n;
  const int dim = P.dim;
  const int c0 = S.geo[q].cell_base;
  const int c1 = q + 1 < P.nrep ? S.geo[q + 1].cell_base : P.ncell;
  const int p0 = S.start[c0], p1 = S.start[c1];
  KSum acc[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (int i = p0 + threadIdx.x; i < p1; i += 256) {
    double mi = S.m[i];
    double v2 = 0.0;
    for (int d = 0; d < dim; ++d) {
      double vv = S.v[d * n + i];
      v2 = v2 + vv * vv;
      kadd(acc[1 + d], mi * vv);
    }
    kadd(acc[0], 0.5 * mi * v2);
  }
  for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = acc[k].s + acc[k].c;
  __syncthreads();
  if (threadIdx.x < 4) {
    KSum t = {0, 0};
    for (int k = 0; k < 256; ++k) kadd(t, sh[threadIdx.x][k]);
    S.rep_stats[(int64_t)q * 4 + threadIdx.x] = t.s + t.c;
  }
}

__global__ void k_set_step(int64_t* __restrict__ step_dev, double* __restrict__ time_dev,
                           int64_t step, double time, double le_time, double le_prev) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    step_dev[0] = step;
    time_dev[0] = time;
    time_dev[1] = le_time;
    time_dev[2] = le_prev;
  }
}

// One marching step: step + 1, time_current + dt (marching_marching.F90:
// 1392-1393), and the Lees-Edwards shear length is evaluated at
// time_current + dt (marching_integrate_VV.F90:784-786).
__global__ void k_advance_step(int64_t* __restrict__ step_dev, double* __restrict__ time_dev,
                               double dt) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    step_dev[0] = step_dev[0] + 1;
    const double t = time_dev[0] + dt;
    time_dev[2] = time_dev[1];
    time_dev[0] = t;
    time_dev[1] = t + dt;
  }
}

__global__ void k_pair_normals(const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi,
                               int count, uint64_t seed, int64_t step, float* __restrict__ g) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float r[4];
  pair_normals(seed, lo[i], hi[i], step, r);
  for (int q = 0; q < 4; ++q) g[4 * (int64_t)i + q] = r[q];
}

static inline int nblk(int64_t n) { return (int)((n + 255) / 256); }


int stats_blocks(int n) {
  int b = nblk(n);
  return b > 1024 ? 1024 : (b < 1 ? 1 : b);
}

void launch_kick(hipStream_t s, const DevParams& P, const DeviceState& S, double lambda) {
  hipLaunchKernelGGL(k_kick, dim3(nblk(P.n)), dim3(256), 0, s, P, S, lambda);
}

void launch_drift_wrap(hipStream_t s, const DevParams& P, const DeviceState& S) {
  hipLaunchKernelGGL(k_drift_wrap, dim3(nblk(P.n)), dim3(256), 0, s, P, S);
}

void launch_conformation(hipStream_t s, const DevParams& P, const DeviceState& S, double lambda) {
  hipLaunchKernelGGL(k_conformation, dim3(nblk(P.n)), dim3(256), 0, s, P, S, lambda);
}

void launch_boundary(hipStream_t s, const DevParams& P, const DeviceState& S, int closing,
                     int opening) {
  const dim3 g(nblk(P.n)), b(256);
  if (P.verlet) {
    if (P.dim == 3)
      hipLaunchKernelGGL((k_boundary<3, true>), g, b, 0, s, P, S, closing, opening);
    else
      hipLaunchKernelGGL((k_boundary<2, true>), g, b, 0, s, P, S, closing, opening);
    return;
  }
  launch_fill_i32(s, S.count, 0, P.ncell);
  if (P.dim == 3)
    hipLaunchKernelGGL((k_boundary<3, false>), g, b, 0, s, P, S, closing, opening);
  else
    hipLaunchKernelGGL((k_boundary<2, false>), g, b, 0, s, P, S, closing, opening);
}

void launch_track(hipStream_t s, const DevParams& P, const DeviceState& S) {
  hipLaunchKernelGGL(k_track, dim3(nblk(P.n)), dim3(256), 0, s, P, S);
}

void launch_decide(hipStream_t s, const DevParams& P, const DeviceState& S) {
  hipLaunchKernelGGL(k_decide, dim3(1), dim3(64), 0, s, P, S);
}

void launch_force_build(hipStream_t s, const DeviceState& S) {
  hipLaunchKernelGGL(k_force_build, dim3(1), dim3(64), 0, s, S);
}

void launch_lists_current(hipStream_t s, const DeviceState& S) {
  hipLaunchKernelGGL(k_lists_current, dim3(1), dim3(64), 0, s, S);
}

void launch_stats(hipStream_t s, const DevParams& P, const DeviceState& S, int nblocks) {
  hipLaunchKernelGGL(k_stats_partial, dim3(nblocks), dim3(256), 0, s, P, S);
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, s, nblocks, S);
}

void launch_replica_stats(hipStream_t s, const DevParams& P, const DeviceState& S) {
  hipLaunchKernelGGL(k_replica_stats, dim3(P.nrep), dim3(256), 0, s, P, S);
}

void launch_set_step(hipStream_t s, const DeviceState& S, int64_t step, double time,
                     double le_time, double le_prev) {
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, s, S.step_dev, S.time_dev, step, time,
                     le_time, le_prev);
}

void launch_advance_step(hipStream_t s, const DevParams& P, const DeviceState& S) {
  hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(64), 0, s, S.step_dev, S.time_dev, P.dt);
}

void launch_pair_normals(hipStream_t s, const uint32_t* lo, const uint32_t* hi, int count,
                         uint64_t seed, int64_t step, float* g) {
  hipLaunchKernelGGL(k_pair_normals, dim3(nblk(count)), dim3(256), 0, s, lo, hi, count, seed,
                     step, g);
}

}  // namespace sdpd
