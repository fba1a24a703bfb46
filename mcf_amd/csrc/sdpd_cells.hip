// sdpd_cells.hip -- K1: cell list by a stable counting sort (gfx950).
//
// Replaces technique_build_list -> ppm_neighlist_clist -> ppm_util_rank3d
// (ppm.tgz:src/ppm_util_rank3d.F90:395-531).  The reference sorts particle
// indices into boxes with a stable counting sort: histogram, prefix sum,
// scatter in ascending particle index.  Here:
//   1. key + histogram   one thread per particle; the cell key is PPM's
//                        FLOOR((x - x0) * rdx) with its upper-face
//                        corrections; the slot inside the cell comes from an
//                        integer atomic (arrival order, not yet stable)
//   2. scan              exclusive prefix sum of the cell counts
//   3. scatter           unsorted[start[key] + slot] = i
//   4. place             each particle ranks itself among the members of its
//                        cell by index (the stable order the reference
//                        produces), then moves its state to that slot
// The result is identical run to run and equal to the reference's
// lhbx / lpdx for the same particle order.  Particles stay in cell order
// between steps, so the next sort is a near-identity permutation and the
// place pass writes almost coalesced.
#include "sdpd_kernels.h"

namespace sdpd {

__global__ void k_fill_i32(int32_t* __restrict__ a, int32_t v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) a[i] = v;
}

__global__ void k_fill_f64(double* __restrict__ a, double v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) a[i] = v;
}

__device__ __forceinline__ int cell_coord(double xv, double x0, double rd, double xmax, int nm) {
  int k = (int)floor((xv - x0) * rd);
  // ppm_util_rank3d.F90:433-459 with no lower ghost layer
  if (xv >= xmax && k < nm) k = nm;
  if (xv < xmax && k >= nm) k = nm - 1;
  // particles are wrapped into [min, max) before every sort; clamp so a
  // stray value can never address outside the cell arrays
  if (k < 0) k = 0;
  if (k >= nm) k = nm - 1;
  return k;
}

__device__ __forceinline__ int32_t cell_key_of(const DevParams& P, double x0, double x1,
                                               double x2) {
  int cx = cell_coord(x0, P.min_phys[0], P.rd[0], P.min_phys[0] + P.length[0], P.nm[0]);
  int cy = cell_coord(x1, P.min_phys[1], P.rd[1], P.min_phys[1] + P.length[1], P.nm[1]);
  int cz = P.dim == 3
               ? cell_coord(x2, P.min_phys[2], P.rd[2], P.min_phys[2] + P.length[2], P.nm[2])
               : 0;
  return cx + P.nm[0] * (cy + P.nm[1] * cz);
}

__global__ void k_cell_key(DevParams P, const double* __restrict__ x, int32_t* __restrict__ key,
                           int32_t* __restrict__ slot, int32_t* __restrict__ count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int n = P.n;
  double x2 = P.dim == 3 ? x[2 * (int64_t)n + i] : 0.0;
  int32_t k = cell_key_of(P, x[i], x[(int64_t)n + i], x2);
  key[i] = k;
  slot[i] = atomicAdd(&count[k], 1);
}

// Exclusive scan of count[0..ncell) into start[0..ncell], one workgroup of
// 1024 threads walking the array in tiles of 4096 with a running carry.
__global__ __launch_bounds__(1024) void k_scan_cells(const int32_t* __restrict__ count,
                                                     int32_t* __restrict__ start, int32_t ncell) {
  __shared__ int32_t wsum[16];
  __shared__ int32_t carry_s;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  if (tid == 0) carry_s = 0;
  __syncthreads();
  for (int64_t base = 0; base < ncell; base += 4096) {
    int32_t v[4];
    int32_t local = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int64_t idx = base + (int64_t)tid * 4 + q;
      v[q] = idx < ncell ? count[idx] : 0;
      local += v[q];
    }
    // inclusive wave scan of local
    int32_t incl = local;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      int32_t t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    if (wid == 0) {
      int32_t s = lane < 16 ? wsum[lane] : 0;
      int32_t si = s;
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        int32_t t = __shfl_up(si, off, 64);
        if (lane >= off) si += t;
      }
      if (lane < 16) wsum[lane] = si - s;  // exclusive wave offsets
    }
    __syncthreads();
    int32_t carry = carry_s;
    int32_t excl = carry + wsum[wid] + incl - local;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int64_t idx = base + (int64_t)tid * 4 + q;
      if (idx < ncell) start[idx] = excl;
      excl += v[q];
    }
    __syncthreads();
    if (tid == 1023) carry_s = excl;
    __syncthreads();
  }
  if (tid == 0) start[ncell] = carry_s;
}

__global__ void k_scatter(int n, const int32_t* __restrict__ key, const int32_t* __restrict__ slot,
                          const int32_t* __restrict__ start, int32_t* __restrict__ unsorted) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsorted[start[key[i]] + slot[i]] = i;
}

// Rank inside the cell by index (stable order) and move the state there.
__global__ void k_place(DevParams P, const int32_t* __restrict__ key,
                        const int32_t* __restrict__ start, const int32_t* __restrict__ unsorted,
                        const double* __restrict__ x, const double* __restrict__ v,
                        const double* __restrict__ m, const int32_t* __restrict__ pid,
                        const int32_t* __restrict__ perm, const double* __restrict__ ct,
                        double* __restrict__ x2, double* __restrict__ v2, double* __restrict__ m2,
                        int32_t* __restrict__ pid2, int32_t* __restrict__ perm2,
                        double* __restrict__ ct2, int32_t* __restrict__ key2,
                        float4* __restrict__ xr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = P.n;
  if (i >= n) return;
  int32_t k = key[i];
  int32_t s = start[k], e = start[k + 1];
  int32_t r = 0;
  for (int32_t q = s; q < e; ++q) r += unsorted[q] < i ? 1 : 0;
  int32_t p = s + r;
  const int dim = P.dim;
  for (int d = 0; d < dim; ++d) {
    x2[(int64_t)d * n + p] = x[(int64_t)d * n + i];
    v2[(int64_t)d * n + p] = v[(int64_t)d * n + i];
  }
  m2[p] = m[i];
  pid2[p] = pid[i];
  perm2[p] = perm[i];
  key2[p] = k;
  if (ct) {
    const int d2 = dim * dim;
    for (int q = 0; q < d2; ++q) ct2[(int64_t)q * n + p] = ct[(int64_t)q * n + i];
  }
  if (xr) {
    // FP32 position relative to the lower corner of the own cell
    int cx = k % P.nm[0];
    int cy = (k / P.nm[0]) % P.nm[1];
    int cz = k / (P.nm[0] * P.nm[1]);
    float a0 = (float)(x[i] - (P.min_phys[0] + (double)cx * P.cs[0]));
    float a1 = (float)(x[(int64_t)n + i] - (P.min_phys[1] + (double)cy * P.cs[1]));
    float a2 = dim == 3 ? (float)(x[2 * (int64_t)n + i] - (P.min_phys[2] + (double)cz * P.cs[2]))
                        : 0.f;
    xr[p] = make_float4(a0, a1, a2, __uint_as_float((uint32_t)pid[i]));
  }
}

// ---------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------
static inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

void launch_fill_i32(hipStream_t s, int32_t* a, int32_t v, int64_t n) {
  if (n <= 0) return;
  int nb = blocks_for(n, 256);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, s, a, v, n);
}

void launch_fill_f64(hipStream_t s, double* a, double v, int64_t n) {
  if (n <= 0) return;
  int nb = blocks_for(n, 256);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(256), 0, s, a, v, n);
}

void launch_cell_keys(hipStream_t s, const DevParams& P, const DeviceState& S) {
  launch_fill_i32(s, S.count, 0, P.ncell);
  hipLaunchKernelGGL(k_cell_key, dim3(blocks_for(P.n, 256)), dim3(256), 0, s, P, S.x, S.key,
                     S.slot, S.count);
}

void launch_cell_sort_after_keys(hipStream_t s, const DevParams& P, const DeviceState& S) {
  const int n = P.n;
  hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, s, S.count, S.start, P.ncell);
  hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, S.key, S.slot,
                     S.start, S.unsorted);
  hipLaunchKernelGGL(k_place, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, S.key, S.start,
                     S.unsorted, S.x, S.v, S.m, S.pid, S.perm, P.newtonian ? nullptr : S.ct,
                     S.x_alt, S.v_alt, S.m_alt, S.pid_alt, S.perm_alt,
                     P.newtonian ? nullptr : S.ct_alt, S.key_sorted, P.mixed ? S.xr : nullptr);
}

}  // namespace sdpd
