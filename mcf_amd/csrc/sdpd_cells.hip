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

// gate: a kernel of a sort that runs on some steps only returns at once
// unless gate[0] != 0 (a null gate always runs)
__device__ __forceinline__ bool gated_off(const int32_t* gate) { return gate && gate[0] == 0; }

__global__ void k_fill_i32(int32_t* __restrict__ a, int32_t v, int64_t n,
                           const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) a[i] = v;
}

__global__ void k_fill_f64(double* __restrict__ a, double v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) a[i] = v;
}

__global__ void k_cell_key(DevParams P, const double* __restrict__ x,
                           const int32_t* __restrict__ rep, const RepGeom* __restrict__ geo,
                           int32_t* __restrict__ key, int32_t* __restrict__ slot,
                           int32_t* __restrict__ count, const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i0 < P.n;
  const int i = act ? i0 : P.n - 1;  // every lane takes part in the histogram
  const int64_t n = P.n;
  const RepGeom& g = geo[rep[i]];
  const int32_t k = cell_key_geo(g, P.dim, x[i], x[n + i], P.dim == 3 ? x[2 * n + i] : 0.0);
  const int32_t sl = histogram_slot(count, k, act);
  if (act) {
    key[i] = k;
    slot[i] = sl;
  }
}

// Exclusive scan of count[0..m) into start[0..m] in three passes over
// tiles of 4096 entries: tile sums, a scan of the tile sums in one
// workgroup, and the tile scans with their offsets.
constexpr int kScanTile = 4096;

__device__ __forceinline__ int32_t block_exclusive_scan_1024(int32_t v, int32_t& total,
                                                             int32_t* wsum) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  int32_t incl = v;
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
    if (lane < 16) wsum[lane] = si - s;
    if (lane == 15) wsum[16] = si;
  }
  __syncthreads();
  total = wsum[16];
  int32_t r = wsum[wid] + incl - v;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(1024) void k_scan_tile_sums(const int32_t* __restrict__ count,
                                                         int32_t* __restrict__ sums, int32_t m,
                                                         const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  __shared__ int32_t wsum[17];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
  int32_t local = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (base + q < m) local += count[base + q];
  int32_t total;
  (void)block_exclusive_scan_1024(local, total, wsum);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_sums(int32_t* __restrict__ sums, int32_t ntile,
                                                    int32_t* __restrict__ out_total,
                                                    const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  __shared__ int32_t wsum[17];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int32_t b0 = 0; b0 < ntile; b0 += 1024) {
    const int32_t idx = b0 + threadIdx.x;
    const int32_t v = idx < ntile ? sums[idx] : 0;
    int32_t total;
    const int32_t ex = block_exclusive_scan_1024(v, total, wsum);
    const int32_t c = carry;
    if (idx < ntile) sums[idx] = c + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry = c + total;
    __syncthreads();
  }
  if (threadIdx.x == 0) *out_total = carry;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(const int32_t* __restrict__ count,
                                                     const int32_t* __restrict__ sums,
                                                     int32_t* __restrict__ start, int32_t m,
                                                     const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  __shared__ int32_t wsum[17];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * 4;
  int32_t v[4], local = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[q] = base + q < m ? count[base + q] : 0;
    local += v[q];
  }
  int32_t total;
  int32_t ex = block_exclusive_scan_1024(local, total, wsum) + sums[blockIdx.x];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (base + q < m) start[base + q] = ex;
    ex += v[q];
  }
}

// start[m] receives the total
void launch_exclusive_scan(hipStream_t s, const int32_t* count, int32_t* start, int32_t m,
                           int32_t* tile_sums, const int32_t* gate) {
  const int32_t ntile = (m + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(ntile), dim3(1024), 0, s, count, tile_sums, m, gate);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, tile_sums, ntile, start + m, gate);
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntile), dim3(1024), 0, s, count, tile_sums, start, m,
                     gate);
}

// Work map of the pair kernels: each (y, z) row of the sort grid is cut
// into chunks of at most 64 consecutive particles; chunk c belongs to row
// chunk_row[c] and is the (c - row_chunk_off[row])-th chunk of that row.
__global__ void k_row_chunks(DevParams P, const int32_t* __restrict__ start,
                             const int32_t* __restrict__ row_cell,
                             const int32_t* __restrict__ row_len,
                             int32_t* __restrict__ row_nchunk, const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P.nrows) return;
  const int c0 = row_cell[r];
  int cnt = start[c0 + row_len[r]] - start[c0];
  row_nchunk[r] = (cnt + 63) / 64;
}

__global__ void k_chunk_map(DevParams P, const int32_t* __restrict__ row_chunk_off,
                            int32_t* __restrict__ chunk_row, const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P.nrows) return;
  for (int c = row_chunk_off[r]; c < row_chunk_off[r + 1]; ++c) chunk_row[c] = r;
}

__global__ void k_scatter(int n, const int32_t* __restrict__ key, const int32_t* __restrict__ slot,
                          const int32_t* __restrict__ start, int32_t* __restrict__ unsorted,
                          const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsorted[start[key[i]] + slot[i]] = i;
}

// Rank inside the cell by index (stable order) and move the state there.
__global__ void k_place(DevParams P, const int32_t* __restrict__ key,
                        const int32_t* __restrict__ start, const int32_t* __restrict__ unsorted,
                        const double* __restrict__ x, const double* __restrict__ v,
                        const double* __restrict__ m, const int32_t* __restrict__ pid,
                        const int32_t* __restrict__ perm, const int32_t* __restrict__ rep,
                        const double* __restrict__ ct, const RepGeom* __restrict__ geo,
                        double* __restrict__ x2, double* __restrict__ v2, double* __restrict__ m2,
                        int32_t* __restrict__ pid2, int32_t* __restrict__ perm2,
                        int32_t* __restrict__ rep2, double* __restrict__ ct2,
                        int32_t* __restrict__ key2, float4* __restrict__ xa,
                        const int32_t* __restrict__ sid, int32_t* __restrict__ sid2,
                        int32_t* __restrict__ slot_of, const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
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
  if (sid) {
    // colloids: species id moves with the particle; the slot of each upload
    // index locates the boundary particles of a colloid
    sid2[p] = sid[i];
    slot_of[perm[i]] = p;
  }
  const int q = rep[i];
  rep2[p] = q;
  key2[p] = k;
  if (ct) {
    const int d2 = dim * dim;
    for (int q = 0; q < d2; ++q) ct2[(int64_t)q * n + p] = ct[(int64_t)q * n + i];
  }
  // FP32 position relative to the lower corner of the own cell, the cell's
  // x index in w (records of the neighbour scan)
  const RepGeom& g = geo[q];
  const int kl = k - g.cell_base;
  const int cx = kl % g.nm[0];
  const int cy = (kl / g.nm[0]) % g.nm[1];
  const int cz = kl / (g.nm[0] * g.nm[1]);
  xa[p] = make_float4((float)(x[i] - (g.lo[0] + (double)cx * g.cs[0])),
                      (float)(x[(int64_t)n + i] - (g.lo[1] + (double)cy * g.cs[1])),
                      dim == 3 ? (float)(x[2 * (int64_t)n + i] - (g.lo[2] + (double)cz * g.cs[2]))
                               : 0.f,
                      __int_as_float(cx));
}

// The sorted state back from the *_alt buffers into the primary ones, and
// the sorted positions into x_ref (the positions of the lists' build).
__global__ void k_copy_back(DevParams P, DeviceState S, const int32_t* __restrict__ gate) {
  if (gated_off(gate)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = P.n;
  if (i >= n) return;
  for (int d = 0; d < P.dim; ++d) {
    const double xv = S.x_alt[(int64_t)d * n + i];
    S.x[(int64_t)d * n + i] = xv;
    if (S.x_ref) S.x_ref[(int64_t)d * n + i] = xv;
    S.v[(int64_t)d * n + i] = S.v_alt[(int64_t)d * n + i];
  }
  S.m[i] = S.m_alt[i];
  S.pid[i] = S.pid_alt[i];
  S.perm[i] = S.perm_alt[i];
  S.rep[i] = S.rep_alt[i];
  if (!P.newtonian) {
    const int d2 = P.dim * P.dim;
    for (int q = 0; q < d2; ++q) S.ct[(int64_t)q * n + i] = S.ct_alt[(int64_t)q * n + i];
  }
  if (P.ncoll) S.sid[i] = S.sid_alt[i];
}

// ---------------------------------------------------------------------
// PPM's cell list on PPM's grid (cell edge >= 1.0001 cut_off), in upload
// indices: the order inside a cell is ascending upload index, which is
// what ppm_util_rank3d produces for particles stored in upload order.
// ---------------------------------------------------------------------
__global__ void k_ref_key(DevParams P, const double* __restrict__ x, int32_t* __restrict__ key,
                          int32_t* __restrict__ slot, int32_t* __restrict__ count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int64_t n = P.n;
  int cx = cell_coord(x[i], P.min_phys[0], P.rd_ref[0], P.min_phys[0] + P.length[0], P.nm_ref[0]);
  int cy = cell_coord(x[n + i], P.min_phys[1], P.rd_ref[1], P.min_phys[1] + P.length[1],
                      P.nm_ref[1]);
  int cz = P.dim == 3 ? cell_coord(x[2 * n + i], P.min_phys[2], P.rd_ref[2],
                                   P.min_phys[2] + P.length[2], P.nm_ref[2])
                      : 0;
  int32_t k = cx + P.nm_ref[0] * (cy + P.nm_ref[1] * cz);
  key[i] = k;
  slot[i] = atomicAdd(&count[k], 1);
}

__global__ void k_ref_rank(int n, const int32_t* __restrict__ key,
                           const int32_t* __restrict__ start,
                           const int32_t* __restrict__ unsorted,
                           const int32_t* __restrict__ perm, int32_t* __restrict__ lpdx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t k = key[i];
  int32_t s = start[k], e = start[k + 1];
  int32_t me = perm[i], r = 0;
  for (int32_t q = s; q < e; ++q) r += perm[unsorted[q]] < me ? 1 : 0;
  lpdx[s + r] = me;
}

// ---------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------
static inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

void launch_fill_i32(hipStream_t s, int32_t* a, int32_t v, int64_t n, const int32_t* gate) {
  if (n <= 0) return;
  int nb = blocks_for(n, 256);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, s, a, v, n, gate);
}

void launch_fill_f64(hipStream_t s, double* a, double v, int64_t n) {
  if (n <= 0) return;
  int nb = blocks_for(n, 256);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_fill_f64, dim3(nb), dim3(256), 0, s, a, v, n);
}

void launch_cell_keys(hipStream_t s, const DevParams& P, const DeviceState& S,
                      const int32_t* gate) {
  launch_fill_i32(s, S.count, 0, P.ncell, gate);
  hipLaunchKernelGGL(k_cell_key, dim3(blocks_for(P.n, 256)), dim3(256), 0, s, P, S.x, S.rep,
                     S.geo, S.key, S.slot, S.count, gate);
}

void launch_cell_sort_after_keys(hipStream_t s, const DevParams& P, const DeviceState& S,
                                 const int32_t* gate, bool copy_back) {
  const int n = P.n;
  launch_exclusive_scan(s, S.count, S.start, P.ncell, S.scan_tmp, gate);
  hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, S.key, S.slot,
                     S.start, S.unsorted, gate);
  hipLaunchKernelGGL(k_place, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, S.key, S.start,
                     S.unsorted, S.x, S.v, S.m, S.pid, S.perm, S.rep,
                     P.newtonian ? nullptr : S.ct, S.geo, S.x_alt, S.v_alt, S.m_alt, S.pid_alt,
                     S.perm_alt, S.rep_alt, P.newtonian ? nullptr : S.ct_alt, S.key_sorted, S.xa,
                     P.ncoll ? S.sid : nullptr, S.sid_alt, S.slot_of, gate);
  if (copy_back)
    hipLaunchKernelGGL(k_copy_back, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, S, gate);
  const int nrows = P.nrows;
  hipLaunchKernelGGL(k_row_chunks, dim3(blocks_for(nrows, 256)), dim3(256), 0, s, P, S.start,
                     S.row_cell, S.row_len, S.row_nchunk, gate);
  launch_exclusive_scan(s, S.row_nchunk, S.row_chunk_off, nrows, S.scan_tmp, gate);
  hipLaunchKernelGGL(k_chunk_map, dim3(blocks_for(nrows, 256)), dim3(256), 0, s, P,
                     S.row_chunk_off, S.chunk_row, gate);
}

void launch_reference_cells(hipStream_t s, const DevParams& P, const DeviceState& S) {
  const int n = P.n;
  launch_fill_i32(s, S.rcount, 0, P.ncell_ref);
  hipLaunchKernelGGL(k_ref_key, dim3(blocks_for(n, 256)), dim3(256), 0, s, P, S.x, S.rkey,
                     S.rslot, S.rcount);
  launch_exclusive_scan(s, S.rcount, S.rstart, P.ncell_ref, S.scan_tmp);
  hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, S.rkey, S.rslot,
                     S.rstart, S.runsorted, nullptr);
  hipLaunchKernelGGL(k_ref_rank, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, S.rkey, S.rstart,
                     S.runsorted, S.perm, S.rlpdx);
}

}  // namespace sdpd
