// sdpd_colloid.hip -- rigid colloids made of boundary particles (gfx950):
// drag and torque collection, the rigid-body block of a step and the
// colloids' share of the statistics.  Disks in 2D, spheres in 3D.
//
// Replaces particles_collect_colloid_interaction.F90:93-136 (+ the MPI sum
// of colloid_collect_particles_interaction.F90:112-125, one GPU here), the
// colloid block of marching_integrate_VV.F90:455-765 with the explicit
// integrators (colloid_compute_rotation_vector.F90, _rotation_matrix.F90,
// _accumulation_matrix.F90, particles_compute_colloid_relative_position.F90,
// colloid_integrate_rotate_velocity.F90, _translate_position.F90,
// _translate_velocity.F90, colloid_apply_body_force.F90,
// colloid_compute_translate_acceleration.F90, colloid_adjust_colloid.F90,
// colloid_compute_rotate_acceleration.F90,
// particles_compute_colloid_absolute_position.F90), the start of the
// acceleration histories in marching_marching.F90:1003-1080 and
// colloid_compute_statistic.F90.  Colloid-colloid and colloid-wall forces
// are not modelled (pp_interact_cc off, no walls).
//
// Work shapes: one 64-lane wave per colloid for the collection (lanes
// stride over the colloid's boundary particles, then a fixed butterfly
// reduction: deterministic), one thread per colloid for the rigid-body
// update (a few hundred flops), one thread per boundary particle to carry
// it along.
#include "sdpd_kernels.h"

namespace sdpd {

namespace {

constexpr int kCollWave = 64;

__global__ __launch_bounds__(kCollWave) void k_colloid_collect(DevParams P, DeviceState S,
                                                               int which) {
  const int k = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t n = P.n;
  const CollHot C = S.coll_hot[k];
  const double t = S.time_dev[which];
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int q0 = S.coll_start[k], q1 = S.coll_start[k + 1];
  for (int q = q0 + lane; q < q1; q += kCollWave) {
    const int sl = S.slot_of[S.coll_members[q]];
    const double mj = S.m[sl];
    double F[3] = {0.0, 0.0, 0.0}, xp[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < P.dim; ++d) {
      F[d] = S.f[d * n + sl] * mj;
      xp[d] = S.x[d * n + sl];
    }
    double rx[3], vim[3], T[3];
    if (P.dim == 3)
      coll_nearest<3>(P, C, xp, t, rx, vim);
    else
      coll_nearest<2>(P, C, xp, t, rx, vim);
    cross3(rx, F, T);
    for (int d = 0; d < 3; ++d) {
      acc[d] += F[d];
      acc[3 + d] += T[d];
    }
  }
#pragma unroll
  for (int off = kCollWave / 2; off >= 1; off >>= 1)
#pragma unroll
    for (int u = 0; u < 6; ++u) acc[u] += __shfl_xor(acc[u], off);
  if (lane == 0) {
    CollState& cs = S.coll[k];
    for (int d = 0; d < 3; ++d) {
      cs.drag[d] = acc[d];
      cs.torque[d] = acc[3 + d];
    }
  }
}

// history levels move one back (level h - 1 -> h, h = ab - 1 .. 1)
__device__ __forceinline__ void history_shift(double (*h)[3], int ab) {
  for (int l = ab - 1; l >= 1; --l)
    for (int d = 0; d < 3; ++d) h[l][d] = h[l - 1][d];
}

// Adams-Bashforth increment of the given order from a history, newest first
__device__ __forceinline__ void ab_increment(const double (*h)[3], int order, double dt,
                                             double* out) {
  for (int d = 0; d < 3; ++d) {
    switch (order) {
      case 1: out[d] = h[0][d] * dt; break;
      case 2: out[d] = (3.0 * h[0][d] - h[1][d]) * dt / 2.0; break;
      case 3: out[d] = (23.0 * h[0][d] - 16.0 * h[1][d] + 5.0 * h[2][d]) * dt / 12.0; break;
      case 4:
        out[d] = (55.0 * h[0][d] - 59.0 * h[1][d] + 37.0 * h[2][d] - 9.0 * h[3][d]) * dt / 24.0;
        break;
      default:
        out[d] = (1901.0 * h[0][d] - 2774.0 * h[1][d] + 2616.0 * h[2][d] - 1274.0 * h[3][d] +
                  251.0 * h[4][d]) * dt / 720.0;
        break;
    }
  }
}

// colloid_compute_translate_acceleration.F90 / _rotate_acceleration.F90
// (disks: alpha_z = torque_z / I_z; spheres: alpha = torque / I)
__device__ __forceinline__ void new_accelerations(const DevParams& P, CollState& c) {
  history_shift(c.f, P.coll_ab);
  for (int d = 0; d < P.dim; ++d) c.f[0][d] = P.coll_translate ? c.drag[d] / c.m : 0.0;
  history_shift(c.al, P.coll_ab);
  if (!P.coll_rotate) {
    c.al[0][0] = c.al[0][1] = c.al[0][2] = 0.0;
  } else if (P.dim == 2) {
    c.al[0][2] = c.torque[2] / c.mmi[2];
  } else {
    for (int d = 0; d < 3; ++d) c.al[0][d] = c.torque[d] / c.mmi[d];
  }
}

// the collected drag plus the colloid body force (colloid_apply_body_force.F90)
__device__ __forceinline__ void add_body_force(const DevParams& P, CollState& c) {
  if (P.coll_bf_type == 1)
    for (int d = 0; d < P.dim; ++d) c.drag[d] += P.coll_bf[d] * c.m;
}

__global__ void k_colloid_prologue(DevParams P, DeviceState S) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.ncoll) return;
  CollState& c = S.coll[k];
  add_body_force(P, c);
  new_accelerations(P, c);
}

// tool_rotation_matrix.F90: rotation by phi about the unit axis, column-major
__device__ __forceinline__ void rotation_matrix(const double* ax, double phi, double* rm) {
  const double cp = cos(phi), cp1 = 1.0 - cp, sp = sin(phi);
  rm[0] = ax[0] * ax[0] * cp1 + cp;
  rm[3] = ax[0] * ax[1] * cp1 - ax[2] * sp;
  rm[6] = ax[0] * ax[2] * cp1 + ax[1] * sp;
  rm[1] = ax[1] * ax[0] * cp1 + ax[2] * sp;
  rm[4] = ax[1] * ax[1] * cp1 + cp;
  rm[7] = ax[1] * ax[2] * cp1 - ax[0] * sp;
  rm[2] = ax[2] * ax[0] * cp1 - ax[1] * sp;
  rm[5] = ax[2] * ax[1] * cp1 + ax[0] * sp;
  rm[8] = ax[2] * ax[2] * cp1 + cp;
}

// colloid_adjust_colloid.F90: a centre leaving through a periodic face
// comes back by whole box lengths; through a Lees-Edwards face every other
// coordinate k becomes MODULO(x_k - shear_length_k, L_k) and the velocity
// takes the jump
__device__ __forceinline__ void le_wrap_others(const DevParams& P, CollState& c, int d,
                                               double sl0) {
  c.x[0] = f_modulo(c.x[0] - sl0, P.length[0]);
  if (P.dim == 3) {
    if (d == 1)
      c.x[2] = f_modulo(c.x[2], P.length[2]);
    else
      c.x[1] = f_modulo(c.x[1], P.length[1]);
  }
}

__device__ __forceinline__ void adjust_centre(const DevParams& P, CollState& c, double t) {
  for (int d = 0; d < P.dim; ++d) {
    const double lo = P.min_phys[d], L = P.length[d], hi = lo + L;
    if (c.x[d] < lo) {
      if (d == P.le_g) {
        c.x[d] = c.x[d] + L;
        le_wrap_others(P, c, d, (P.le_vlo - P.le_vhi) * t);
        c.v[0][0] = c.v[0][0] + (P.le_vhi - P.le_vlo);
        return;
      }
      const double cross = (hi - c.x[d]) / L;
      c.x[d] = c.x[d] + L * (double)(int)cross;
    }
    if (c.x[d] >= hi) {
      if (d == P.le_g) {
        c.x[d] = c.x[d] - L;
        le_wrap_others(P, c, d, (P.le_vhi - P.le_vlo) * t);
        c.v[0][0] = c.v[0][0] + (P.le_vlo - P.le_vhi);
        return;
      }
      const double cross = (c.x[d] - lo) / L;
      c.x[d] = c.x[d] - L * (double)(int)cross;
    }
  }
}

// The rigid-body update of one step, per colloid: the sub-steps of the
// colloid block with the drag and torque just collected.  Records the
// centre before each sub-step and after the last, and each sub-step's
// rotation, for k_colloid_move.
__global__ void k_colloid_update(DevParams P, DeviceState S) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.ncoll) return;
  CollState& c = S.coll[k];
  const double t = S.time_dev[2];
  const int64_t step = S.step_dev[0];
  const int ab = P.coll_ab, nsub = P.coll_sub;
  const double dts = P.dt / nsub;
  const double drag0[3] = {c.drag[0], c.drag[1], c.drag[2]};
  for (int d = 0; d < 3; ++d) c.xs[0][d] = c.x[d];
  for (int i = 1; i <= nsub; ++i) {
    // the AB order ramps up with the step since the start of the run; the
    // rotating velocity is passed step - 1 + i + step_start
    const int64_t st = step - 1 + i - P.coll_step0, sw = step - 1 + i + P.coll_step0;
    const int ord = (int)(st < ab ? (st < 1 ? 1 : st) : ab);
    const int ordw = (int)(sw < ab ? (sw < 1 ? 1 : sw) : ab);
    double* R = c.rot[i - 1];
    if (P.coll_rotate) {
      double axis[3];
      ab_increment(c.w, ord, dts, axis);
      double phi = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
      if (phi < 1e-14) {
        phi = 0.0;
        axis[0] = 1.0;
        axis[1] = axis[2] = 0.0;
      } else {
        for (int d = 0; d < 3; ++d) axis[d] = axis[d] / phi;
      }
      c.theta[2] += phi;
      rotation_matrix(axis, phi, R);
      double RA[9];
      for (int cc = 0; cc < 3; ++cc)
        for (int r = 0; r < 3; ++r)
          RA[r + 3 * cc] =
              R[r] * c.acc[3 * cc] + R[r + 3] * c.acc[1 + 3 * cc] + R[r + 6] * c.acc[2 + 3 * cc];
      for (int q = 0; q < 9; ++q) c.acc[q] = RA[q];
      history_shift(c.w, ab);
      double inc[3];
      ab_increment(c.al, ordw, dts, inc);
      for (int d = 0; d < 3; ++d) c.w[0][d] += inc[d];
    }
    if (P.coll_translate) {
      double inc[3];
      ab_increment(c.v, ord, dts, inc);
      for (int d = 0; d < P.dim; ++d) c.x[d] += inc[d];
      history_shift(c.v, ab);
      ab_increment(c.f, ord, dts, inc);
      for (int d = 0; d < P.dim; ++d) c.v[0][d] += inc[d];
    }
    for (int d = 0; d < 3; ++d) c.drag[d] = drag0[d];
    add_body_force(P, c);
    // translate acceleration (before the centre is wrapped), wrap, then the
    // rotate acceleration, as the reference orders them
    history_shift(c.f, ab);
    for (int d = 0; d < P.dim; ++d) c.f[0][d] = P.coll_translate ? c.drag[d] / c.m : 0.0;
    adjust_centre(P, c, t);
    history_shift(c.al, ab);
    if (!P.coll_rotate) {
      c.al[0][0] = c.al[0][1] = c.al[0][2] = 0.0;
    } else if (P.dim == 2) {
      c.al[0][2] = c.torque[2] / c.mmi[2];
    } else {
      for (int d = 0; d < 3; ++d) c.al[0][d] = c.torque[d] / c.mmi[d];
    }
    for (int d = 0; d < 3; ++d) c.xs[i][d] = c.x[d];
  }
  CollHot& hcol = S.coll_hot[k];
  for (int d = 0; d < 3; ++d) {
    hcol.x[d] = c.x[d];
    hcol.v[d] = c.v[0][d];
    hcol.w[d] = c.w[0][d];
  }
}

// Each boundary particle follows its colloid through the sub-steps: its
// position relative to the nearest image of the centre before the
// sub-step, rotated by the sub-step's rotation, added to the centre after
// it.
__global__ void k_colloid_move(DevParams P, DeviceState S, int nmembers) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nmembers) return;
  const int64_t n = P.n;
  const int sl = S.slot_of[S.coll_members[q]];
  const int k = S.sid[sl] - 1;
  const CollState& c = S.coll[k];
  const double t = S.time_dev[2];
  double xp[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < P.dim; ++d) xp[d] = S.x[d * n + sl];
  CollHot at = S.coll_hot[k];
  for (int i = 1; i <= P.coll_sub; ++i) {
    for (int d = 0; d < 3; ++d) at.x[d] = c.xs[i - 1][d];
    double rx[3], vim[3];
    if (P.dim == 3)
      coll_nearest<3>(P, at, xp, t, rx, vim);
    else
      coll_nearest<2>(P, at, xp, t, rx, vim);
    double rb[3] = {rx[0], rx[1], rx[2]};
    if (P.coll_rotate) {
      const double* R = c.rot[i - 1];
      for (int r = 0; r < P.dim; ++r) {
        double s = 0.0;
        for (int cc = 0; cc < P.dim; ++cc) s = s + R[r + 3 * cc] * rx[cc];
        rb[r] = s;
      }
    }
    for (int d = 0; d < P.dim; ++d) xp[d] = rb[d] + c.xs[i][d];
  }
  for (int d = 0; d < P.dim; ++d) S.x[d * n + sl] = xp[d];
}

// colloid_compute_statistic.F90: translational kinetic energy and momentum
// of the colloids, summed in colloid order into stats_out[4..7]
__global__ void k_colloid_stats(DevParams P, DeviceState S) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double ek = 0.0, mom[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < P.ncoll; ++k) {
    const CollState& c = S.coll[k];
    double v2 = 0.0, mk[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < P.dim; ++d) {
      v2 = v2 + c.v[0][d] * c.v[0][d];
      mk[d] = mk[d] + c.m * c.v[0][d];
    }
    ek = ek + (0.0 + 0.5 * c.m * v2);
    for (int d = 0; d < P.dim; ++d) mom[d] = mom[d] + mk[d];
  }
  S.stats_out[4] = ek;
  for (int d = 0; d < 3; ++d) S.stats_out[5 + d] = mom[d];
}

}  // namespace

void launch_colloid_collect(hipStream_t s, const DevParams& P, const DeviceState& S, int which) {
  hipLaunchKernelGGL(k_colloid_collect, dim3(P.ncoll), dim3(kCollWave), 0, s, P, S, which);
}

void launch_colloid_prologue(hipStream_t s, const DevParams& P, const DeviceState& S) {
  hipLaunchKernelGGL(k_colloid_prologue, dim3((P.ncoll + 63) / 64), dim3(64), 0, s, P, S);
}

void launch_colloid_block(hipStream_t s, const DevParams& P, const DeviceState& S,
                          int nmembers) {
  launch_colloid_collect(s, P, S, 2);
  if (!P.coll_translate && !P.coll_rotate) return;
  hipLaunchKernelGGL(k_colloid_update, dim3((P.ncoll + 63) / 64), dim3(64), 0, s, P, S);
  if (nmembers > 0)
    hipLaunchKernelGGL(k_colloid_move, dim3((nmembers + 255) / 256), dim3(256), 0, s, P, S,
                       nmembers);
}

void launch_colloid_stats(hipStream_t s, const DevParams& P, const DeviceState& S) {
  hipLaunchKernelGGL(k_colloid_stats, dim3(1), dim3(64), 0, s, P, S);
}

}  // namespace sdpd
