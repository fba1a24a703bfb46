// sdpd_kernels.h -- device buffers of one box and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sdpd_device.h"

namespace sdpd {

// Device buffers.  Per-particle arrays are structure-of-arrays in cell
// order: component d of particle i lives at a[d * n + i].
struct DeviceState {
  // particle state (cell order) and the ping-pong copies the sort writes
  double *x, *v, *m;
  int32_t *pid, *perm;   // perm: upload index of the particle in this slot
  double *x_alt, *v_alt, *m_alt;
  int32_t *pid_alt, *perm_alt;
  double *f;             // acceleration (dim, n)
  double *rho, *p;       // density (number density for type 2), pressure
  double *an, *bn;       // p / n^2 and 1 / n^2 of the last density pass
  // Oldroyd-B: conformation, its rate, velocity gradient (dim*dim, n)
  double *ct, *ct_alt, *act, *vgt;
  // cell list
  int32_t *key, *slot, *count, *start, *unsorted, *key_sorted;
  // mixed precision records (cell order)
  float4 *xa;            // FP32 position relative to the own cell corner, w = cell x index
  // work map of the pair kernels (chunks of <= 64 particles of one row)
  int32_t *row_nchunk, *row_chunk_off, *chunk_row;
  int32_t *scan_tmp;     // tile sums of the exclusive scans
  // FP64 records of the force pass: x, y, z, p/n^2, vx, vy, vz, 1/n^2;
  // Oldroyd-B also Pi / n^2 as (xx, yy, zz, xy, xz, yz) (2D: xx, yy, 0, xy)
  double *rec64;
  double *pt64;
  float4 *vr;            // velocity, w = 1 / n^2; Oldroyd-B: w = pid bits
  float4 *qr;            // (cell-relative x, p / n^2 or p); Oldroyd-B: (1/n^2, Pxx, Pyy, Pzz)
  float4 *qr2;           // Oldroyd-B: (Pxy, Pxz, Pyz, 0), P = Pi / n^2
  // reductions and counters
  double *partials;      // stats partial sums (blocks x 8)
  double *stats_out;     // 8 doubles
  int32_t *nbr_count;    // per-particle in-cutoff count of the last density pass
  uint32_t *nlist;       // in-cutoff neighbour lists of the last density pass,
                         // one block of groups of 8 entries per chunk
  int32_t *nbr_sig_count;  // neighbour counts of the last signature pass
  int32_t *recheck;      // per-particle FP64 re-decisions (mixed mode)
  uint32_t *nbr_hash;    // per-particle sum of pair hashes (checking)
  // replicas: geometry table, replica of each particle (cell order, with
  // its ping-pong copy), and per row of the sort grids its first cell,
  // its length (cells) and its replica
  const RepGeom* geo;
  int32_t *rep, *rep_alt;
  int32_t *row_cell, *row_len, *row_rep;
  double* rep_stats;     // per replica (E_k, P_x, P_y, P_z)
  int64_t* step_dev;     // marching step read by the force pass (noise key)
  // [0] time_current, [1] time of the Lees-Edwards shear length of this
  // step, [2] that of the previous step (read by the colloid block, which
  // runs before the step's boundary update)
  double* time_dev;
  // colloids: species id of each particle (cell order, with its ping-pong
  // copy), slot of each upload index, colloid records, and the boundary
  // particles of colloid k as upload indices coll_members[coll_start[k] ..
  // coll_start[k + 1])
  int32_t *sid, *sid_alt, *slot_of;
  CollHot* coll_hot;
  CollState* coll;
  int32_t *coll_start, *coll_members;
  // reference-grid cell list (built on request by sdpd_get_cells)
  int32_t *rkey, *rslot, *rcount, *rstart, *runsorted, *rlpdx;
  // Verlet lists: positions at the last build (cell order), the step's
  // gate (gate[0] = 1: the lists are rebuilt this step; gate[1] counts the
  // builds; gate[2] = 1: a chunk's lists overflowed at the last build, so
  // the rescan passes have work), the bits of the largest squared
  // displacement since the build
  // seen this step, and the in-cut-off neighbours of each particle in the
  // last density pass (every mode)
  double* x_ref;
  int32_t* gate;
  unsigned long long* disp;
  int32_t* nbr_in;
};

void launch_fill_i32(hipStream_t s, int32_t* a, int32_t v, int64_t n,
                     const int32_t* gate = nullptr);
void launch_fill_f64(hipStream_t s, double* a, double v, int64_t n);
// key + histogram (counts zeroed first), then scan, scatter and place into
// the *_alt buffers.  Without copy_back the caller swaps the ping-pong
// pointers afterwards; with it the sorted state is copied back into the
// primary buffers (and into x_ref), so the buffers never change, which lets
// a captured graph sort on some steps only.  A non-null gate makes every
// kernel of the sort return at once unless gate[0] != 0.
void launch_cell_keys(hipStream_t s, const DevParams& P, const DeviceState& S,
                      const int32_t* gate = nullptr);
void launch_exclusive_scan(hipStream_t s, const int32_t* count, int32_t* start, int32_t m,
                           int32_t* tile_sums, const int32_t* gate = nullptr);
void launch_cell_sort_after_keys(hipStream_t s, const DevParams& P, const DeviceState& S,
                                 const int32_t* gate = nullptr, bool copy_back = false);
// PPM's cell list on PPM's grid, particles ordered by upload index inside a
// cell (ppm_util_rank3d on the particles in upload order); fills rstart and
// rlpdx without moving the state
void launch_reference_cells(hipStream_t s, const DevParams& P, const DeviceState& S);

// scan_all (Verlet lists, the autotuner): every chunk scans its stencil,
// without storing lists, whatever the step's gate
void launch_density(hipStream_t s, const DevParams& P, const DeviceState& S, bool mixed,
                    bool scan_all = false);
void launch_eos(hipStream_t s, const DevParams& P, const DeviceState& S);
void launch_forces(hipStream_t s, const DevParams& P, const DeviceState& S, bool mixed);
void launch_signature(hipStream_t s, const DevParams& P, const DeviceState& S, bool mixed);
void launch_refresh_records(hipStream_t s, const DevParams& P, const DeviceState& S);

// Verlet lists: displacement since the last build and the FP32 cell-relative
// positions after a drift (every particle), then the step's gate: 1 when a
// particle moved more than skin / 2 (the displacement maximum is cleared)
void launch_track(hipStream_t s, const DevParams& P, const DeviceState& S);
void launch_decide(hipStream_t s, const DevParams& P, const DeviceState& S);
// after a sort outside the step: gate[0] = 1 (build at the next density pass)
void launch_force_build(hipStream_t s, const DeviceState& S);
// after a density pass: gate[0] = 0 (the lists are current)
void launch_lists_current(hipStream_t s, const DeviceState& S);
void launch_kick(hipStream_t s, const DevParams& P, const DeviceState& S, double lambda);
void launch_drift_wrap(hipStream_t s, const DevParams& P, const DeviceState& S);
void launch_conformation(hipStream_t s, const DevParams& P, const DeviceState& S, double lambda);
void launch_boundary(hipStream_t s, const DevParams& P, const DeviceState& S, int closing,
                     int opening);
void launch_stats(hipStream_t s, const DevParams& P, const DeviceState& S, int nblocks);
void launch_replica_stats(hipStream_t s, const DevParams& P, const DeviceState& S);
// the device copy of the marching step: set, or advance by one (a
// captured step graph advances it itself)
void launch_set_step(hipStream_t s, const DeviceState& S, int64_t step, double time,
                     double le_time, double le_prev);
void launch_advance_step(hipStream_t s, const DevParams& P, const DeviceState& S);
// colloids (single box): drag and torque of every colloid from its
// boundary particles' last forces, with relative positions at the shear
// length of time_dev[which]; the prologue's start of the acceleration
// histories; the colloid block of one step (rigid-body update of the
// colloids, then the boundary particles carried along)
void launch_colloid_collect(hipStream_t s, const DevParams& P, const DeviceState& S, int which);
void launch_colloid_prologue(hipStream_t s, const DevParams& P, const DeviceState& S);
void launch_colloid_block(hipStream_t s, const DevParams& P, const DeviceState& S,
                          int nmembers);
void launch_colloid_stats(hipStream_t s, const DevParams& P, const DeviceState& S);

void launch_pair_normals(hipStream_t s, const uint32_t* lo, const uint32_t* hi, int count,
                         uint64_t seed, int64_t step, float* g);

int stats_blocks(int n);

}  // namespace sdpd
