// sdpd_pair.hip -- K2 density + equation of state and K4 pair forces
// (gfx950), FP64 and mixed precision.
//
// Replaces the two hot loops of the reference: particles_compute_density
// (particles_compute_density.F90:578-867) and particles_compute_interaction
// (particles_compute_interaction.F90:618-911 with pp_interaction.inc,
// pp_interaction_ff.inc, pp_vgt_ip.inc and rhs_force_ff_*).  The reference
// visits each unordered pair once in a half stencil and adds the result to
// both particles and to periodic ghost copies that are folded back later.
// Here every particle gathers over the cells around its own, with the
// periodic image applied per stencil segment, and writes only its own
// sums: no atomics, no ghosts, identical results run to run.
//
// Work decomposition.  The sort grid has half of PPM's cell edge, so the
// neighbours of a particle lie in the 5 x 5 x 5 (2D: 5 x 5) block of cells
// around its own.  One wavefront takes a chunk of at most 64 consecutive
// particles of one (y, z) row of cells, so all its lanes share the 25
// (2D: 5) stencil rows.  For each stencil row the wave stages the union of
// its lanes' candidate cells into LDS with coalesced loads -- laid out in
// "virtual cell" order, where the cells seen across a periodic x face
// follow the row with their image shift -- and every lane scans the five
// cells around its own from LDS.
//
// Cut-off decision (both modes): r^2 in FP32 on positions relative to the
// cell corner; a candidate within a band around rc^2 wider than the FP32
// error bound is re-decided in FP64 exactly as the reference forms it
// (pp_interaction.inc:37-46) with the reference's periodic image
// convention.  Neighbour sets are therefore bit-exact in both modes.
//
// The density pass appends the accepted neighbours of every particle to a
// per-lane LDS list, evaluates the density terms whenever the lists fill
// up, and stores the lists of a chunk as one block of groups of 8 entries
// per lane: every store fills whole 32-byte sectors and the force pass of
// the same step (positions unchanged) reads 2 KB contiguous per group.
#include "sdpd_kernels.h"

namespace sdpd {

constexpr int kWave = 64;
// Defaults of the two tunable shapes of the candidate scan (the density
// kernel is instantiated for the autotuner's other choices too):
constexpr int kBuf = 16;          // LDS list entries per lane between flushes (>= 10)
constexpr int kTile = 224;        // staged candidate records per stencil row batch
constexpr int kCodeShift = 25;    // list entry = j | (stencil code << 25)
constexpr uint32_t kIndexMask = (1u << kCodeShift) - 1u;

// Stencil code of a neighbour: offsets o in -2..2 per dimension,
// code = (ox+2) + 5 (oy+2) + 25 (oz+2).
__device__ __forceinline__ uint32_t make_code(int ox, int oy, int oz) {
  return (uint32_t)((ox + 2) + 5 * (oy + 2) + 25 * (oz + 2));
}

struct Offsets {
  int ox, oy, oz;
};
__device__ __forceinline__ Offsets decode_code(uint32_t code) {
  Offsets o;
  o.ox = (int)(code % 5u) - 2;
  o.oy = (int)((code / 5u) % 5u) - 2;
  o.oz = (int)(code / 25u) - 2;
  return o;
}

// FP64 lists store the neighbour's periodic image instead of its stencil
// offset: bits 25-26, 27-28 and 29-30 hold, per dimension, the image
// shift of x_i - x_j in box lengths as a 2-bit signed field (+1 for a
// virtual cell below 0, -1 for one at or above nm, else 0); bit 31 flags
// a colloid boundary particle.  The FP32 path needs the offsets
// (cell-relative positions) and keeps the codes.
struct Images {
  int sx, sy, sz;
};
// Bit 31 of an FP64 entry: the neighbour is a colloid boundary particle
// (boxes with colloids), so the pair body needs no species gather for the
// fluid-fluid pairs.
constexpr uint32_t kBoundaryBit = 0x80000000u;
__device__ __forceinline__ int image_class(int v, int nm) { return v < 0 ? 1 : (v >= nm ? -1 : 0); }
__device__ __forceinline__ uint32_t image_field(int v, int nm, int dim) {
  return ((uint32_t)image_class(v, nm) & 3u) << (kCodeShift + 2 * dim);
}
__device__ __forceinline__ Images decode_images(uint32_t ent) {
  Images m;
  m.sx = (int)(ent << 5) >> 30;
  m.sy = (int)(ent << 3) >> 30;
  m.sz = (int)(ent << 1) >> 30;
  return m;
}

// The cell-relative position of an xa record without its cell index: a
// 12-byte gather instead of 16.
__device__ __forceinline__ float4 load_xyz(const float4* __restrict__ rec) {
  const float* f = reinterpret_cast<const float*>(rec);
  return make_float4(f[0], f[1], f[2], 0.f);
}

// Geometry of the replica a chunk belongs to (uniform across the wave).
struct Geo {
  int nm0, nm1, nm2;    // sort-grid cells per dimension
  int cell_base, row_base;
  double L0, L1, L2;    // edge lengths
  double csd0;          // FP64 cell edge along x
  float cs0, cs1, cs2;  // FP32 cell edges
};

// Lees-Edwards view of a lane (P.le_g >= 0).  Rows across the upper face
// (sigma 0) see the neighbours' upper copies, shifted along x by the shear
// length S_up = (v_upper - v_lower) t; rows across the lower face (sigma 1)
// by S_down = -S_up.  In the frame shifted by sp = MODULO(S, Lx) the lane
// sits in cell wc, dxf = (cx - wc) cs - sp is its FP32 offset from its own
// cell corner in that frame, and wmin / wmax bound wc over the wave.
// Per-lane Lees-Edwards frame of the two sheared faces (sigma 0: across
// the upper face, 1: across the lower face): x shift S' of the images, the
// window centre, the FP32 x offset of the sheared frame, and the wave's
// window range.  Named members with select accessors: an array indexed by
// the run-time sigma would be placed in scratch memory.
struct LeLane {
  double t;             // time of the shear length
  double sp0, sp1;
  int wc0, wc1;
  float dxf0, dxf1;
  int wmin0, wmin1, wmax0, wmax1;
  __device__ __forceinline__ double sp(int s) const { return s == 0 ? sp0 : sp1; }
  __device__ __forceinline__ int wc(int s) const { return s == 0 ? wc0 : wc1; }
  __device__ __forceinline__ float dxf(int s) const { return s == 0 ? dxf0 : dxf1; }
  __device__ __forceinline__ int wmin(int s) const { return s == 0 ? wmin0 : wmin1; }
  __device__ __forceinline__ int wmax(int s) const { return s == 0 ? wmax0 : wmax1; }
};

// Per-lane context of the home particle.
struct Home {
  int i;                // own index (valid for active lanes)
  int ii;               // a safe index for every lane
  bool active;
  int cx, cy, cz;       // own cell (replica-local)
  float4 a;             // FP32 position relative to the own cell corner
  double x0, x1, x2;    // FP64 position
  Geo G;
  LeLane le;
};

// Lees-Edwards sign of the stencil row at offsets (dy, dz): 0 across the
// upper face of the shearing dimension, 1 across the lower face, -1 none
__device__ __forceinline__ int le_sigma(const DevParams& P, const Home& h, int dy, int dz) {
  if (P.le_g == 1) {
    const int c = h.cy + dy;
    return c >= h.G.nm1 ? 0 : (c < 0 ? 1 : -1);
  }
  if (P.le_g == 2) {
    const int c = h.cz + dz;
    return c >= h.G.nm2 ? 0 : (c < 0 ? 1 : -1);
  }
  return -1;
}
// The same sign from a listed neighbour's images: the image below the
// upper face (shift -1) is the upper copy.
__device__ __forceinline__ int le_sigma(const DevParams& P, const Images& m) {
  const int c = P.le_g == 1 ? m.sy : (P.le_g == 2 ? m.sz : 0);
  return c < 0 ? 0 : (c > 0 ? 1 : -1);
}

// Image modes of a neighbour at offsets o from the home cell.
struct Modes {
  int m0, m1, m2;
};
// Exact FP64 difference and squared distance for a pair.
template <int DIM>
__device__ __forceinline__ double exact_d2(const DevParams& P, const double* __restrict__ x, int j,
                                           const Home& h, const Modes& m, double& d0, double& d1,
                                           double& d2) {
  const int64_t n = P.n;
  d0 = image_diff(h.x0, x[j], h.G.L0, m.m0);
  d1 = image_diff(h.x1, x[n + j], h.G.L1, m.m1);
  d2 = DIM == 3 ? image_diff(h.x2, x[2 * n + j], h.G.L2, m.m2) : 0.0;
  return DIM == 3 ? dist2_exact(d0, d1, d2) : dist2_exact2(d0, d1);
}

// Exact FP64 squared distance of a pair with the image chosen from the
// positions (image_diff_pos): the reference's formula and convention
// without the cell classes, so it holds between sorts (Verlet lists),
// after either particle crossed a periodic face.
template <int DIM>
__device__ __forceinline__ double exact_d2_pos(const DevParams& P, const Home& h, double xj0,
                                               double xj1, double xj2, double& d0, double& d1,
                                               double& d2) {
  d0 = image_diff_pos(h.x0, xj0, h.G.L0, P.symmetry);
  d1 = image_diff_pos(h.x1, xj1, h.G.L1, P.symmetry);
  d2 = DIM == 3 ? image_diff_pos(h.x2, xj2, h.G.L2, P.symmetry) : 0.0;
  return DIM == 3 ? dist2_exact(d0, d1, d2) : dist2_exact2(d0, d1);
}

// The Verlet lists' decision: the exact squared distance on the current
// positions with the image taken from them.  In a box with sheared faces a
// pair whose coordinates along the faces' normal differ by more than half
// the box meets across a face, and j's copy there is the one the candidate
// scan sees: sigma 0, the copy above the upper face, or 1, below the lower
// one (returned in sg; -1 when no sheared face lies between them), with the
// formula and roundings of exact_d2_le.
template <int DIM>
__device__ __forceinline__ double exact_d2_verlet(const DevParams& P, const Home& h, double xj0,
                                                  double xj1, double xj2, double& d0,
                                                  double& d1, double& d2, int& sg) {
  sg = -1;
  if (P.le_g >= 0) {
    const bool gy = P.le_g == 1;
    const double raw = gy ? h.x1 - xj1 : h.x2 - xj2;
    const double Lg = gy ? h.G.L1 : h.G.L2;
    sg = raw > 0.5 * Lg ? 0 : (raw < -0.5 * Lg ? 1 : -1);
  }
  if (sg < 0) return exact_d2_pos<DIM>(P, h, xj0, xj1, xj2, d0, d1, d2);
  const double L0 = h.G.L0;
  const double su = (P.le_vhi - P.le_vlo) * h.le.t, sd = (P.le_vlo - P.le_vhi) * h.le.t;
  double xi = h.x0, xj = xj0;
  if (sg == 0)
    xj = f_modulo(xj + su, L0);
  else if (P.symmetry)
    xi = f_modulo(xi + su, L0);
  else
    xj = f_modulo(xj + sd, L0);
  const double rx = xi - xj;
  const int mx = rx > 0.5 * L0 ? 1 : (rx < -0.5 * L0 ? (P.symmetry ? 2 : 3) : 0);
  const int mg = sg == 0 ? 1 : (P.symmetry ? 2 : 3);
  d0 = image_diff(xi, xj, L0, mx);
  if (P.le_g == 1) {
    d1 = image_diff(h.x1, xj1, h.G.L1, mg);
    d2 = DIM == 3 ? image_diff_pos(h.x2, xj2, h.G.L2, P.symmetry) : 0.0;
  } else {
    d1 = image_diff_pos(h.x1, xj1, h.G.L1, P.symmetry);
    d2 = image_diff(h.x2, xj2, h.G.L2, mg);
  }
  return DIM == 3 ? dist2_exact(d0, d1, d2) : dist2_exact2(d0, d1);
}

// FP32 squared distance, rounded as the candidate scan rounds it, so the
// in / out decisions of the density and force passes coincide
__device__ __forceinline__ float mixed_s2(float d0, float d1, float d2) {
  return fmaf(d2, d2, fmaf(d1, d1, d0 * d0));
}

// FP32 separation from cell-relative positions and the stencil offsets.
template <int DIM>
__device__ __forceinline__ void rel_diff(const DevParams& P, const Home& h, const float4& q,
                                         const Offsets& o, float& d0, float& d1, float& d2) {
  float ax = h.a.x;
  if (P.le_g >= 0) {
    // across a Lees-Edwards face ox counts from the sheared window centre
    const int sg = le_sigma(P, h, o.oy, o.oz);
    if (sg >= 0) ax = ax + h.le.dxf(sg);
  }
  d0 = ax - (q.x + (float)o.ox * h.G.cs0);
  d1 = h.a.y - (q.y + (float)o.oy * h.G.cs1);
  d2 = DIM == 3 ? h.a.z - (q.z + (float)o.oz * h.G.cs2) : 0.f;
}

// The chunk of the sort grid this wavefront works on; false when the wave
// has none (the launch covers an upper bound of chunks).
struct Chunk {
  int c;                // chunk index
  int row;              // global row; replica-local row = cy + nm1 * cz
  int p0, p1;           // particles [p0, p1)
  int xmin, xmax;       // x range of the home cells
  Geo G;
};

// Chunk slots of the pair launches: an upper bound of the chunks (one
// partial chunk per row plus n / 64 full ones), rounded up to a multiple of
// the eight XCDs.
__host__ __device__ inline int chunk_slots(const DevParams& P) {
  const int m = (P.n + kWave - 1) / kWave + P.nrows;
  return (m + 7) / 8 * 8;
}

// Workgroups are handed to the eight XCDs round-robin.  With xcd_order,
// slot b takes chunk (b % 8) ceil(chunks / 8) + b / 8, i.e. each XCD
// works through a contiguous eighth of the chunk sequence and its L2 serves
// a compact slab of space; otherwise chunk b.  The slots, a multiple of 8
// (chunk_slots), are at least the number of chunks; a workgroup takes slot
// blockIdx.x, or every gridDim.x-th slot on the reduced grid of a rescan.
__device__ __forceinline__ bool load_chunk(const DevParams& P, const DeviceState& S, Chunk& ch,
                                           int b) {
  const int total = S.row_chunk_off[P.nrows];
  int c = b;
  if (P.xcd_order) {
    const int per = (total + 7) >> 3;
    c = (b & 7) * per + (b >> 3);
    if ((b >> 3) >= per) return false;
  }
  if (c >= total) return false;
  ch.c = c;
  ch.row = S.chunk_row[c];
  const int k = c - S.row_chunk_off[ch.row];
  const RepGeom& g = S.geo[S.row_rep[ch.row]];
  ch.G.nm0 = g.nm[0];
  ch.G.nm1 = g.nm[1];
  ch.G.nm2 = g.nm[2];
  ch.G.cell_base = g.cell_base;
  ch.G.row_base = g.row_base;
  ch.G.L0 = g.L[0];
  ch.G.L1 = g.L[1];
  ch.G.L2 = g.L[2];
  ch.G.csd0 = g.cs[0];
  ch.G.cs0 = g.csf[0];
  ch.G.cs1 = g.csf[1];
  ch.G.cs2 = g.csf[2];
  const int c0 = S.row_cell[ch.row];
  const int rs = S.start[c0];
  const int re = S.start[c0 + ch.G.nm0];
  ch.p0 = rs + kWave * k;
  ch.p1 = min(ch.p0 + kWave, re);
  ch.xmin = (S.key_sorted[ch.p0] - ch.G.cell_base) % ch.G.nm0;
  ch.xmax = (S.key_sorted[ch.p1 - 1] - ch.G.cell_base) % ch.G.nm0;
  return true;
}

// Neighbour list block of a chunk: list_cap / 8 groups, each 64 lanes x 8
// entries; a lane's group g holds its entries 8g .. 8g+7.
__device__ __forceinline__ uint32_t* list_block(const DevParams& P, uint32_t* list, const Chunk& ch) {
  return list + (int64_t)ch.c * P.list_cap * kWave;
}
__device__ __forceinline__ int64_t group_offset(int q, int lane) {
  return ((int64_t)(q >> 3) * kWave + lane) * 8;
}

template <int DIM>
__device__ __forceinline__ Home load_home(const DevParams& P, const DeviceState& S,
                                          const Chunk& ch) {
  Home h;
  const int lane = threadIdx.x;
  h.i = ch.p0 + lane;
  h.active = h.i < ch.p1;
  h.ii = h.active ? h.i : ch.p0;
  const int64_t n = P.n;
  h.G = ch.G;
  h.cx = (S.key_sorted[h.ii] - ch.G.cell_base) % ch.G.nm0;
  const int lr = ch.row - ch.G.row_base;
  h.cy = lr % ch.G.nm1;
  h.cz = lr / ch.G.nm1;
  h.a = S.xa[h.ii];
  h.x0 = S.x[h.ii];
  h.x1 = S.x[n + h.ii];
  h.x2 = DIM == 3 ? S.x[2 * n + h.ii] : 0.0;
  if (P.le_g >= 0) {
    // boxes with Lees-Edwards faces start at 0 (the reference's MODULO)
    h.le.t = S.time_dev[1];
    const double cs = ch.G.csd0;
#pragma unroll
    for (int sg = 0; sg < 2; ++sg) {
      const double sl = sg == 0 ? (P.le_vhi - P.le_vlo) * h.le.t : (P.le_vlo - P.le_vhi) * h.le.t;
      const double sp = f_modulo(sl, ch.G.L0);
      // Verlet lists keep the build's cells, and the FP32 position goes on
      // from the build cell unwrapped: the window centre is taken in that
      // frame, so a lane that crossed an x face since stays with its wave
      const double xw = P.verlet ? (double)h.cx * cs + (double)h.a.x : h.x0;
      const int wc = (int)floor((xw - sp) / cs);
      const float dxf = (float)((double)(h.cx - wc) * cs - sp);
      int mn = h.active ? wc : INT_MAX, mx = h.active ? wc : INT_MIN;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, __shfl_xor(mn, off));
        mx = max(mx, __shfl_xor(mx, off));
      }
      if (sg == 0) {
        h.le.sp0 = sp;
        h.le.wc0 = wc;
        h.le.dxf0 = dxf;
        h.le.wmin0 = mn;
        h.le.wmax0 = mx;
      } else {
        h.le.sp1 = sp;
        h.le.wc1 = wc;
        h.le.dxf1 = dxf;
        h.le.wmin1 = mn;
        h.le.wmax1 = mx;
      }
    }
  }
  return h;
}

// ---------------------------------------------------------------------
// stencil rows of a chunk
// ---------------------------------------------------------------------
// One stencil row (dy, dz) of a chunk covers the virtual cells
// [vlo, vhi] of the target row of cells: virtual cell v is physical cell
// v - m nm0 seen through m periodic x faces.  Lanes look at the five cells
// around their own (home cell, or the Lees-Edwards window centre), so the
// range is [xmin - 2, xmax + 2] (or the window centres' extremes +-2) and
// splits into at most three pieces of one period each, contiguous in
// memory: piece k spans virtual [a_k, a_k+1) of period m0 + k.  Sequence
// index t runs over the pieces in that order.
struct RowPieces {
  int rb;                 // first cell of the target row
  int my, mz;             // image modes of the row in y and z
  int sigma;              // Lees-Edwards sign of the row (-1 none)
  int vlo, a1, a2, aend;  // piece bounds in virtual cells
  int m0;                 // period of piece 0
  int g0, len0, g1, len1, g2, len2;
  int tot;
};

__device__ __forceinline__ int floor_div(int a, int b) {
  return a >= 0 ? a / b : -((-a + b - 1) / b);
}

template <int DIM>
__device__ __forceinline__ RowPieces row_pieces(const DevParams& P, const DeviceState& S,
                                                const Chunk& ch, const Home& h, int row) {
  RowPieces rp;
  const int nm0 = h.G.nm0;
  const int32_t* __restrict__ start = S.start;
  const int dy = row % 5 - 2, dz = DIM == 3 ? row / 5 - 2 : 0;
  rp.mz = 0;
  const int ny = wrap_cell(h.cy + dy, h.G.nm1, P.symmetry, rp.my);
  const int nz = DIM == 3 ? wrap_cell(h.cz + dz, h.G.nm2, P.symmetry, rp.mz) : 0;
  rp.rb = h.G.cell_base + nm0 * (ny + h.G.nm1 * nz);
  rp.sigma = P.le_g >= 0 ? le_sigma(P, h, dy, dz) : -1;
  int vhi;
  if (rp.sigma < 0) {
    rp.vlo = ch.xmin - 2;
    vhi = ch.xmax + 2;
  } else {
    rp.vlo = h.le.wmin(rp.sigma) - 2;
    vhi = h.le.wmax(rp.sigma) + 2;
  }
  rp.aend = vhi + 1;
  rp.m0 = floor_div(rp.vlo, nm0);
  rp.a1 = min(rp.aend, (rp.m0 + 1) * nm0);
  rp.a2 = min(rp.aend, (rp.m0 + 2) * nm0);
  const int b0 = rp.m0 * nm0, b1 = b0 + nm0, b2 = b1 + nm0;
  rp.g0 = start[rp.rb + rp.vlo - b0];
  rp.len0 = start[rp.rb + rp.a1 - b0] - rp.g0;
  rp.g1 = rp.a2 > rp.a1 ? start[rp.rb + rp.a1 - b1] : rp.g0;
  rp.len1 = rp.a2 > rp.a1 ? start[rp.rb + rp.a2 - b1] - rp.g1 : 0;
  rp.g2 = rp.aend > rp.a2 ? start[rp.rb + rp.a2 - b2] : rp.g0;
  rp.len2 = rp.aend > rp.a2 ? start[rp.rb + rp.aend - b2] - rp.g2 : 0;
  rp.tot = rp.len0 + rp.len1 + rp.len2;
  return rp;
}

// sequence index of the first particle of virtual cell v (v in [vlo, aend])
__device__ __forceinline__ int seq_of(const RowPieces& rp, const int32_t* __restrict__ start,
                                      int nm0, int v) {
  if (v >= rp.aend) return rp.tot;
  const int b0 = rp.m0 * nm0;
  if (v < rp.a1) return start[rp.rb + v - b0] - rp.g0;
  if (v < rp.a2) return rp.len0 + (start[rp.rb + v - b0 - nm0] - rp.g1);
  return rp.len0 + rp.len1 + (start[rp.rb + v - b0 - 2 * nm0] - rp.g2);
}

// particle index of sequence position t
__device__ __forceinline__ int seq_particle(const RowPieces& rp, int t) {
  return t < rp.len0 ? rp.g0 + t
                     : (t < rp.len0 + rp.len1 ? rp.g1 + (t - rp.len0) : rp.g2 + (t - rp.len0 - rp.len1));
}

// Exact FP64 squared distance of a pair across a Lees-Edwards face, as the
// reference forms it from its images (ppm_bcdef_le_y.inc, then the
// periodic cascade): the shifted copy is MODULO(x + shear_length, Lx)
// along x (j's upper copy for sigma 0; for sigma 1 i's upper copy with
// symmetry, j's lower copy without), whose periodic image nearest the
// other particle is the one inside the cut-off.
template <int DIM>
__device__ __forceinline__ double exact_d2_le(const DevParams& P, const double* __restrict__ x,
                                              int j, const Home& h, int sigma, int my, int mz) {
  const int64_t n = P.n;
  const double L0 = h.G.L0;
  const double su = (P.le_vhi - P.le_vlo) * h.le.t, sd = (P.le_vlo - P.le_vhi) * h.le.t;
  double xi = h.x0, xj = x[j];
  if (sigma == 0)
    xj = f_modulo(xj + su, L0);
  else if (P.symmetry)
    xi = f_modulo(xi + su, L0);
  else
    xj = f_modulo(xj + sd, L0);
  const double raw = xi - xj;
  const int mx = raw > 0.5 * L0 ? 1 : (raw < -0.5 * L0 ? (P.symmetry ? 2 : 3) : 0);
  const int mg = sigma == 0 ? 1 : (P.symmetry ? 2 : 3);
  const double d0 = image_diff(xi, xj, L0, mx);
  const double d1 = image_diff(h.x1, x[n + j], h.G.L1, P.le_g == 1 ? mg : my);
  if (DIM == 2) return dist2_exact2(d0, d1);
  const double d2 = image_diff(h.x2, x[2 * n + j], h.G.L2, P.le_g == 2 ? mg : mz);
  return dist2_exact(d0, d1, d2);
}

// ---------------------------------------------------------------------
// the staged scan
// ---------------------------------------------------------------------
// Candidate scan over the stencil rows of a chunk; returns the number of
// accepted neighbours.
// SKIN (Verlet lists): accept every candidate within rc + skin (FP32, with
// the band's margin); the policy's body takes the in / out decision.
template <int DIM, class Pol, int TILE = kTile, int BUF = kBuf, bool SKIN = false>
__device__ __forceinline__ int staged_scan(const DevParams& P, const DeviceState& S,
                                           const Chunk& ch, const Home& h, Pol& pol,
                                           float4* __restrict__ tile, uint32_t* __restrict__ buf,
                                           int& rechecks, uint32_t* __restrict__ list) {
  const int lane = threadIdx.x;
  const int nm0 = h.G.nm0;
  const int32_t* __restrict__ start = S.start;
  const int low_mode = P.symmetry ? 2 : 3;
  // the boundary flag of the lane's own record (FP64 entries with colloids)
  const uint32_t self_flag = (Pol::kImages && P.ncoll && S.sid[h.ii] != 0) ? kBoundaryBit : 0u;
  int cnt = 0, stored = 0;
  // buf holds cnt entries of this lane, the first ev of them evaluated
  // already; entries reach the list block in whole groups of 8
  int ev = 0;
  uint32_t* __restrict__ blk = list ? list_block(P, list, ch) : nullptr;
  auto flush = [&](bool final) {
    if constexpr (Pol::kBatched) {
      // Pol::kBatch entries at a time: the policy issues all their gathers
      // before it evaluates the first pair (slots past cnt read the lane's
      // own particle and add nothing)
      constexpr int NB = Pol::kBatch;
      for (int q = ev; __ballot(q < cnt) != 0ull; q += NB) {
        uint32_t e[NB];
        bool ok[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
          ok[k] = q + k < cnt;
          e[k] = ok[k] ? buf[(q + k) * kWave + lane] : (uint32_t)h.ii;
        }
        pol.body_batch(e, ok);
      }
    } else {
      for (int q = ev; __ballot(q < cnt) != 0ull; ++q) {
        if (q < cnt) {
          pol.body(buf[q * kWave + lane]);
        }
      }
    }
    const int w = final ? cnt : (cnt & ~7);
    if (blk) {
      for (int q = 0; __ballot(q < w) != 0ull; q += 8) {
        if (q < w && stored + q + 8 <= P.list_cap) {
          uint4 lo, hi;
          lo.x = buf[(q + 0) * kWave + lane];
          lo.y = buf[(q + 1) * kWave + lane];
          lo.z = buf[(q + 2) * kWave + lane];
          lo.w = buf[(q + 3) * kWave + lane];
          hi.x = buf[(q + 4) * kWave + lane];
          hi.y = buf[(q + 5) * kWave + lane];
          hi.z = buf[(q + 6) * kWave + lane];
          hi.w = buf[(q + 7) * kWave + lane];
          uint4* dst = reinterpret_cast<uint4*>(blk + group_offset(stored + q, lane));
          dst[0] = lo;
          dst[1] = hi;
        }
      }
    }
    stored += w;
    // keep the unwritten tail (fewer than 8 entries) at the front
    const int rem = cnt - w;
    for (int r = 0; r < rem; ++r) buf[r * kWave + lane] = buf[(w + r) * kWave + lane];
    cnt = rem;
    ev = rem;
  };
  const int nrow = DIM == 3 ? 25 : 5;
#pragma unroll 1
  for (int row = 0; row < nrow; ++row) {
    const int dy = row % 5 - 2, dz = DIM == 3 ? row / 5 - 2 : 0;
    const RowPieces rp = row_pieces<DIM>(P, S, ch, h, row);
    const int my = rp.my, mz = rp.mz;
    const int tot = rp.tot;
    const int v_lo = rp.vlo;
    // Positions in the tile are relative to the corner of virtual cell
    // v_lo in x (and to the candidate's own cell in y, z).  With stencil
    // codes the tile's .w holds the list entry with the virtual cell (mod
    // 128) in place of the code, so that entry + kadd is the entry with
    // the code (the offset from the lane's window centre); with images it
    // holds the entry with the x image, and the row adds the y and z ones.
    // Across a Lees-Edwards face the window centre and the lane's x come
    // from the sheared frame.
    const bool le_row = rp.sigma >= 0;
    const int wc = le_row ? h.le.wc(rp.sigma) : h.cx;
    const int crel = wc - v_lo;
    const float ax = (le_row ? h.a.x + h.le.dxf(rp.sigma) : h.a.x) + (float)crel * h.G.cs0;
    // the lane's own copy: same particle, own virtual cell (x image 0)
    const uint32_t self_w =
        le_row ? 0xFFFFFFFFu
               : (Pol::kImages ? ((uint32_t)h.ii | self_flag)
                               : ((uint32_t)h.ii | ((uint32_t)crel << kCodeShift)));
    // this lane's window: virtual cells wc-2 .. wc+2 as a range of the
    // row sequence
    int tb = 0, te = 0;
    if (h.active) {
      tb = seq_of(rp, start, nm0, wc - 2);
      te = seq_of(rp, start, nm0, wc + 3);
    }
    const float ay = h.a.y - (float)dy * h.G.cs1;
    const float az = DIM == 3 ? h.a.z - (float)dz * h.G.cs2 : 0.f;
    const uint32_t kadd =
        Pol::kImages
            ? image_field(h.cy + dy, h.G.nm1, 1) | (DIM == 3 ? image_field(h.cz + dz, h.G.nm2, 2) : 0u)
            : (uint32_t)(2 - crel + 5 * (dy + 2) + 25 * (dz + 2)) << kCodeShift;
    // the row in batches of at most TILE records (one batch in practice)
    for (int base = 0; base < tot; base += TILE) {
      const int bend = min(tot, base + TILE);
      for (int t = base + lane; t < bend; t += kWave) {
        const int piece = t < rp.len0 ? 0 : (t < rp.len0 + rp.len1 ? 1 : 2);
        const int g = seq_particle(rp, t);
        const float4 q = S.xa[g];
        const int v = __float_as_int(q.w) + (rp.m0 + piece) * nm0;
        const int vrel = v - v_lo;
        uint32_t code = Pol::kImages ? image_field(v, nm0, 0) : (uint32_t)vrel << kCodeShift;
        if (Pol::kImages && P.ncoll && S.sid[g] != 0) code |= kBoundaryBit;
        tile[t - base] = make_float4(fmaf((float)vrel, h.G.cs0, q.x), q.y, q.z,
                                     __uint_as_float((uint32_t)g | code));
      }
      __syncthreads();
      const int lb = max(tb, base) - base, le = min(te, bend) - base;
      for (int t = lb; __ballot(t < le) != 0ull; t += 2) {
        if (__ballot(cnt > BUF - 2) != 0ull) flush(false);
        // records t and t + 1 from one clamped base (the tile holds TILE + 1
        // records; a clamped base only serves lanes past their window)
        const float4* __restrict__ tp = tile + min(t, TILE - 1);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int tt = t + u;
          const bool live = tt < le;
          const float4 q = tp[u];
          const uint32_t wv = __float_as_uint(q.w);
          const float d0 = ax - q.x;
          const float d1 = ay - q.y;
          float s2 = fmaf(d1, d1, d0 * d0);
          if (DIM == 3) {
            const float d2 = az - q.z;
            s2 = fmaf(d2, d2, s2);
          }
          const bool ok = live && wv != self_w;
          const uint32_t ent = wv + kadd;
          const bool below_lo = SKIN ? s2 < P.skin2_hi_f32 : s2 < P.cut2_lo_f32;
          const bool below_hi = SKIN ? below_lo : s2 < P.cut2_hi_f32;
          bool in = below_lo;
          if (!SKIN && ok && below_hi && !below_lo) {
            // FP64 re-decision with the reference's image convention
            ++rechecks;
            const int j = (int)(ent & kIndexMask);
            if (le_row) {
              in = exact_d2_le<DIM>(P, S.x, j, h, rp.sigma, my, mz) < P.cut2;
            } else {
              const int cls = Pol::kImages ? decode_images(ent).sx
                                           : image_class(h.cx + decode_code(ent >> kCodeShift).ox, nm0);
              Modes m;
              m.m0 = cls > 0 ? low_mode : (cls < 0 ? 1 : 0);
              m.m1 = my;
              m.m2 = mz;
              double e0, e1, e2;
              in = exact_d2<DIM>(P, S.x, j, h, m, e0, e1, e2) < P.cut2;
            }
          }
          if (ok && in) {
            buf[cnt * kWave + lane] = ent;
            ++cnt;
          }
        }
      }
      __syncthreads();
    }
  }
  flush(true);
  return stored;
}

// FP32 terms summed in FP64 (each add rounds once, at FP64 precision)
struct WideF {
  double s = 0.0;
  __device__ __forceinline__ void add(float v) { s = s + (double)v; }
  __device__ __forceinline__ double value() const { return s; }
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

// ---------------------------------------------------------------------
// K2: density + EOS
// ---------------------------------------------------------------------
// Velocity of a neighbour's Lees-Edwards copy relative to its own: + the
// face velocity jump across the upper face, - across the lower one
// (particles_set_boundary_ghost_velocity.F90:175-240).
__device__ __forceinline__ double le_dv(const DevParams& P, const Images& m) {
  if (P.le_g < 0) return 0.0;
  const int sg = le_sigma(P, m);
  return sg == 0 ? P.le_u : (sg == 1 ? -P.le_u : 0.0);
}
__device__ __forceinline__ float le_dvf(const DevParams& P, const Home& h, const Offsets& o) {
  if (P.le_g < 0) return 0.f;
  const int sg = le_sigma(P, h, o.oy, o.oz);
  return sg == 0 ? P.le_uf : (sg == 1 ? -P.le_uf : 0.f);
}

// FP64 separation of a listed pair: x_i - x_j plus the periodic image
// shift of the neighbour's virtual cell, s L with s in {-1, 0, 1} (the fma
// rounds once, as the sum with the shift does).  The in/out decision was
// taken with the reference's own formula; the value here differs from it
// by at most an ulp of the box length.
template <int DIM>
__device__ __forceinline__ double pair_d2(const DevParams& P, const Home& h, const Images& m,
                                          double xj0, double xj1, double xj2, double& d0,
                                          double& d1, double& d2) {
  d0 = fma((double)m.sx, h.G.L0, h.x0 - xj0);
  if (P.le_g >= 0) {
    // across a Lees-Edwards face: the sheared copy, nearest periodic image
    const int sg = le_sigma(P, m);
    if (sg >= 0) {
      // x_i - x_j - S' lies in (-2 Lx, Lx): nearest period
      const double raw = (h.x0 - xj0) - h.le.sp(sg);
      d0 = raw - h.G.L0 * rint(raw / h.G.L0);
    }
  }
  d1 = fma((double)m.sy, h.G.L1, h.x1 - xj1);
  d2 = DIM == 3 ? fma((double)m.sz, h.G.L2, h.x2 - xj2) : 0.0;
  return DIM == 3 ? fma(d2, d2, fma(d1, d1, d0 * d0)) : fma(d1, d1, d0 * d0);
}

// DECIDE (Verlet lists): the entries are candidates within rc + skin; the
// body takes the exact in / out decision on the current positions and
// counts the in-cut-off neighbours (nin) and the FP64 re-decisions.
// Verlet lists evaluate in batches of eight: the positions of the batch's
// neighbours are gathered before the first pair is evaluated (three 8-byte
// gathers each, whose latency a one-at-a-time walk exposes); the sums keep
// the entries' order.  On the 128^3 box the walk takes 0.89 ms in batches
// of four, 0.72 in eight, 0.74 in sixteen.  The per-step lists go one at a
// time (the batch's registers cost the candidate scan occupancy every
// step).
template <int DIM, bool DECIDE = false>
struct DensityF64 {
  const DevParams& P;
  const DeviceState& S;
  const Home& h;
  double acc;
  int nin = 0, rechecks = 0;
  static constexpr bool kImages = true;
  static constexpr bool kBatched = DECIDE;
  static constexpr int kBatch = 8;
  __device__ __forceinline__ void body(uint32_t ent) {
    const int j = (int)(ent & kIndexMask);
    const int64_t n = P.n;
    double d0, d1, d2;
    const double r2 = pair_d2<DIM>(P, h, decode_images(ent), S.x[j], S.x[n + j],
                                   DIM == 3 ? S.x[2 * n + j] : 0.0, d0, d1, d2);
    double w, gw;
    kernel_eval(P, sqrt_nr(r2) * P.inv_h, w, gw);
    acc = acc + (P.density_type == 1 ? S.m[j] * w : w);
  }
  template <int NB>
  __device__ __forceinline__ void body_batch(const uint32_t (&e)[NB], const bool (&ok)[NB]) {
    const int64_t n = P.n;
    double x0[NB], x1[NB], x2[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const int j = (int)(e[k] & kIndexMask);
      x0[k] = S.x[j];
      x1[k] = S.x[n + j];
      x2[k] = DIM == 3 ? S.x[2 * n + j] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      double d0, d1, d2, r2;
      bool use = ok[k];
      if (DECIDE) {
        // the reference's decision (pp_interaction.inc:37-46) on the
        // current positions
        int sg;
        r2 = exact_d2_verlet<DIM>(P, h, x0[k], x1[k], x2[k], d0, d1, d2, sg);
        use = use && r2 < P.cut2;
        nin += use ? 1 : 0;
      } else {
        r2 = pair_d2<DIM>(P, h, decode_images(e[k]), x0[k], x1[k], x2[k], d0, d1, d2);
      }
      // beyond the cut-off the kernel is 0: no branch on the decision
      double w, gw;
      kernel_eval(P, sqrt_nr(r2) * P.inv_h, w, gw);
      if (use) acc = acc + (P.density_type == 1 ? S.m[(int)(e[k] & kIndexMask)] * w : w);
    }
  }
};

// The in / out decision of a candidate from its FP32 squared distance: in
// below the band, out above it, inside it the reference's FP64 formula.
template <int DIM>
__device__ __forceinline__ bool mixed_inside(const DevParams& P, const DeviceState& S,
                                             const Home& h, int j, float s2, int& rechecks) {
  if (s2 < P.cut2_lo_f32) return true;
  if (!(s2 < P.cut2_hi_f32)) return false;
  ++rechecks;
  const int64_t n = P.n;
  double e0, e1, e2;
  return exact_d2_pos<DIM>(P, h, S.x[j], S.x[n + j], DIM == 3 ? S.x[2 * n + j] : 0.0, e0, e1,
                           e2) < P.cut2;
}

template <int DIM, bool DECIDE = false>
struct DensityMixed {
  const DevParams& P;
  const DeviceState& S;
  const Home& h;
  KahanF acc;
  int nin = 0, rechecks = 0;
  static constexpr bool kImages = false;
  // in batches of eight (the 128^3 Tait box in mixed precision: 0.59 ms in
  // fours, 0.51 in eights, 0.55 in sixteens)
  static constexpr bool kBatched = true;
  static constexpr int kBatch = 8;
  template <int NB>
  __device__ __forceinline__ void body_batch(const uint32_t (&e)[NB], const bool (&ok)[NB]) {
    float4 q[NB];
    float mj[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const int j = (int)(e[k] & kIndexMask);
      q[k] = load_xyz(S.xa + j);
      mj[k] = P.density_type == 1 ? (float)S.m[j] : 1.f;
    }
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      // the lane's own particle in a padding slot: code 62 is offset 0
      const Offsets o = decode_code(ok[k] ? e[k] >> kCodeShift : 62u);
      float d0, d1, d2;
      rel_diff<DIM>(P, h, q[k], o, d0, d1, d2);
      const float s2 = mixed_s2(d0, d1, d2);
      bool use = ok[k];
      if (DECIDE && use) {
        use = mixed_inside<DIM>(P, S, h, (int)(e[k] & kIndexMask), s2, rechecks);
        nin += use ? 1 : 0;
      }
      const float r = __builtin_amdgcn_sqrtf(s2);
      float w, gw;
      kernel_w_gradw_f32(P.kernel_type, r * P.inv_hf, P.coeff, P.coef_gradf, w, gw);
      if (use) acc.add(P.density_type == 1 ? mj[k] * w : w);
    }
  }
};

// Records of the force pass, written once the density is known.
//   FP64: (x, y, z, p/n^2, vx, vy, vz, 1/n^2).
//   FP32 Newtonian: (cell-relative x, y, z, p/n^2 for Hu-Adams or p for the
//   angular form) and (vx, vy, vz, 1/n^2); the pair body gathers these two
//   and the particle id.
//   FP32 Oldroyd-B: (v, pid), (1/n^2, Pxx, Pyy, Pzz) and (Pxy, Pxz, Pyz, 0)
//   with P = Pi / n^2, Pi = (p + G) I - G c (particles_compute_pressure_
//   tensor.F90:114-133; c is symmetric); the position comes from xa.
template <int DIM>
__device__ __forceinline__ void write_force_records(const DevParams& P, const DeviceState& S, int i,
                                                    double an, double bn, double press) {
  const int64_t n = P.n;
  if (P.mixed) {
    const float v0 = (float)S.v[i], v1 = (float)S.v[n + i], v2 = DIM == 3 ? (float)S.v[2 * n + i] : 0.f;
    if (P.newtonian) {
      const float4 xa = S.xa[i];
      S.qr[i] = make_float4(xa.x, xa.y, xa.z, (float)(P.force_type == 4 ? press : an));
      S.vr[i] = make_float4(v0, v1, v2, (float)bn);
    } else {
      S.vr[i] = make_float4(v0, v1, v2, __uint_as_float((uint32_t)S.pid[i]));
      auto pi = [&](int a, int b) {
        const double c = S.ct[(int64_t)(a + DIM * b) * n + i];
        return ((a == b ? press + P.G : 0.0) - P.G * c) * bn;
      };
      if (DIM == 3) {
        S.qr[i] = make_float4((float)bn, (float)pi(0, 0), (float)pi(1, 1), (float)pi(2, 2));
        S.qr2[i] = make_float4((float)pi(0, 1), (float)pi(0, 2), (float)pi(1, 2), 0.f);
      } else {
        S.qr[i] = make_float4((float)bn, (float)pi(0, 0), (float)pi(1, 1), 0.f);
        S.qr2[i] = make_float4((float)pi(0, 1), 0.f, 0.f, 0.f);
      }
    }
  } else {
    double4* r = reinterpret_cast<double4*>(S.rec64 + 8 * (int64_t)i);
    r[0] = make_double4(S.x[i], S.x[n + i], DIM == 3 ? S.x[2 * n + i] : 0.0, an);
    r[1] = make_double4(S.v[i], S.v[n + i], DIM == 3 ? S.v[2 * n + i] : 0.0, bn);
    if (!P.newtonian) {
      // Oldroyd-B: Pi / n^2 as (xx, yy, zz, xy, xz, yz), Pi = (p + G) I - G c
      // (particles_compute_pressure_tensor.F90:114-133; c is symmetric)
      auto pib = [&](int a, int b) {
        const double c = S.ct[(int64_t)(a + DIM * b) * n + i];
        return ((a == b ? press + P.G : 0.0) - P.G * c) * bn;
      };
      double2* q = reinterpret_cast<double2*>(S.pt64 + 6 * (int64_t)i);
      if (DIM == 3) {
        q[0] = make_double2(pib(0, 0), pib(1, 1));
        q[1] = make_double2(pib(2, 2), pib(0, 1));
        q[2] = make_double2(pib(0, 2), pib(1, 2));
      } else {
        q[0] = make_double2(pib(0, 0), pib(1, 1));
        q[1] = make_double2(0.0, pib(0, 1));
        q[2] = make_double2(0.0, 0.0);
      }
    }
  }
}

// A chunk one of whose particles has more in-cutoff neighbours than the
// stored lists hold: its wave scans the stencil again in the force pass.
__device__ __forceinline__ bool list_overflow(const DevParams& P, const DeviceState& S,
                                              const Chunk& ch) {
  const int i = ch.p0 + (int)threadIdx.x;
  const int cnt = i < ch.p1 ? S.nbr_count[i] : 0;
  return __ballot(cnt > P.list_cap) != 0ull;
}

// Walk the neighbours stored in a chunk's lists (waves without list
// overflow): the in-cut-off pairs of the step's density pass, or the Verlet
// lists' candidates.  Policies that evaluate in batches get Pol::kBatch
// entries at a time from every lane (slots past a lane's count carry its
// own particle and are flagged off).
template <int DIM, class Pol>
__device__ __forceinline__ void for_each_listed(const DevParams& P, const DeviceState& S,
                                                const Chunk& ch, const Home& h, Pol& pol,
                                                uint32_t* __restrict__ buf) {
  const int cnt = h.active ? S.nbr_count[h.i] : 0;
  const uint32_t* __restrict__ blk = list_block(P, S.nlist, ch);
  const int lane = threadIdx.x;
  if constexpr (Pol::kBatched) {
    (void)buf;
    // G groups of eight entries per iteration (the batch may span two);
    // a group past the lane's count is not read
    constexpr int NB = Pol::kBatch;
    constexpr int G = (NB + 7) / 8;
    for (int q = 0; __ballot(q < cnt) != 0ull; q += 8 * G) {
      uint32_t ee[8 * G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        uint4 lo = make_uint4(0u, 0u, 0u, 0u), hi = lo;
        if (q + 8 * g < cnt) {
          const uint4* src = reinterpret_cast<const uint4*>(blk + group_offset(q + 8 * g, lane));
          lo = src[0];
          hi = src[1];
        }
        ee[8 * g + 0] = lo.x;
        ee[8 * g + 1] = lo.y;
        ee[8 * g + 2] = lo.z;
        ee[8 * g + 3] = lo.w;
        ee[8 * g + 4] = hi.x;
        ee[8 * g + 5] = hi.y;
        ee[8 * g + 6] = hi.z;
        ee[8 * g + 7] = hi.w;
      }
#pragma unroll
      for (int b0 = 0; b0 < 8 * G; b0 += NB) {
        uint32_t e[NB];
        bool ok[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
          ok[k] = q + b0 + k < cnt;
          e[k] = ok[k] ? ee[b0 + k] : (uint32_t)h.ii;
        }
        pol.body_batch(e, ok);
      }
    }
  } else {
    for (int q = 0; __ballot(q < cnt) != 0ull; q += 8) {
      if (q < cnt) {
        // eight entries per pair of 16-byte loads; entries past cnt are
        // padding inside the row and are not evaluated.  The group goes to
        // LDS (entry u of a lane at buf[u * 64 + lane], conflict-free), and
        // the loop reads one entry per pair, the next one ahead of the body.
        const uint4* src = reinterpret_cast<const uint4*>(blk + group_offset(q, lane));
        const uint4 lo = src[0], hi = src[1];
        buf[0 * kWave + lane] = lo.x;
        buf[1 * kWave + lane] = lo.y;
        buf[2 * kWave + lane] = lo.z;
        buf[3 * kWave + lane] = lo.w;
        buf[4 * kWave + lane] = hi.x;
        buf[5 * kWave + lane] = hi.y;
        buf[6 * kWave + lane] = hi.z;
        buf[7 * kWave + lane] = hi.w;
        const int m = min(8, cnt - q);
        uint32_t ent = buf[lane];
#pragma unroll 1
        for (int u = 0; u < m; ++u) {
          const uint32_t cur = ent;
          ent = buf[((u + 1) & 7) * kWave + lane];
          pol.body(cur);
        }
      }
    }
  }
}

// Whether a chunk's row of cells lies within `cells` cells of a face of
// dimension le_g (2: its stencil rows reach across the face).
__device__ __forceinline__ bool chunk_meets_le_face(int le_g, const Chunk& ch, int cells) {
  const int lr = ch.row - ch.G.row_base;
  const int c = le_g == 1 ? lr % ch.G.nm1 : lr / ch.G.nm1;
  const int nm = le_g == 1 ? ch.G.nm1 : ch.G.nm2;
  return c < cells || c >= nm - cells;
}

// LE = false compiles the Lees-Edwards paths out: le_g is then the
// constant -1 in every inlined helper.  A box with sheared faces runs two
// launches over the same grid, LE = true for the chunks next to a face and
// LE = false for the rest, whose rows never cross a face (the sheared
// frame's state would cost the whole pass registers and occupancy).
//
// Modes.  kStore: the in-cut-off pairs found by the candidate scan are
// stored; lists rebuilt every step.  Verlet lists decide in / out of the
// cut-off per candidate on the current positions: kBuild (steps whose gate
// is set) scans the stencil for every candidate within rc + skin and stores
// them; kWalk (the other steps) walks the stored candidates of the chunks
// whose lists hold them all; kRescan (those steps) scans the stencil again,
// without storing, for the chunks with an overflowed list; kScanAll does
// that for every chunk whatever the gate (the autotuner).  The walk and the
// scans visit the candidates in the same order, so they give the same sums.
enum DensityMode { kStore = 0, kBuild = 1, kWalk = 2, kRescan = 3, kScanAll = 4 };

// One chunk slot of the density pass.
template <int DIM, bool MIXED, int TILE, int BUF, bool LE, int MODE>
__device__ __forceinline__ void density_chunk(const DevParams& P, const DeviceState& S, int slot,
                                              int le_box, float4* __restrict__ tile,
                                              uint32_t* __restrict__ buf) {
  constexpr bool kDecide = MODE != kStore;
  constexpr bool kScan = MODE != kWalk;
  Chunk ch;
  if (!load_chunk(P, S, ch, slot)) return;
  if (le_box >= 0 && chunk_meets_le_face(le_box, ch, 2) != LE) return;
  if ((MODE == kWalk || MODE == kRescan) && list_overflow(P, S, ch) != (MODE == kRescan)) return;
  const Home h = load_home<DIM>(P, S, ch);
  const double mi = S.m[h.ii];
  uint32_t* list = (MODE == kStore || MODE == kBuild) ? S.nlist : nullptr;
  double rho;
  int rechecks = 0, cnt = 0, nin;
  if (MIXED) {
    using Pol = DensityMixed<DIM, kDecide>;
    Pol pol{P, S, h, KahanF{}};
    pol.acc.add(P.density_type == 1 ? (float)(P.w0 * mi) : (float)P.w0);
    if constexpr (kScan)
      cnt = staged_scan<DIM, Pol, TILE, BUF, kDecide>(P, S, ch, h, pol, tile, buf, rechecks, list);
    else
      for_each_listed<DIM>(P, S, ch, h, pol, buf);
    rho = pol.acc.value();
    nin = kDecide ? pol.nin : cnt;
    rechecks += pol.rechecks;
  } else {
    // self term first (particles_compute_density.F90:369-558)
    using Pol = DensityF64<DIM, kDecide>;
    Pol pol{P, S, h, P.density_type == 1 ? P.w0 * mi : P.w0};
    if constexpr (kScan)
      cnt = staged_scan<DIM, Pol, TILE, BUF, kDecide>(P, S, ch, h, pol, tile, buf, rechecks, list);
    else
      for_each_listed<DIM>(P, S, ch, h, pol, buf);
    rho = pol.acc;
    nin = kDecide ? pol.nin : cnt;
    rechecks += pol.rechecks;
  }
  if (!h.active) return;
  const int i = h.i;
  // colloid boundary particles of constant density take the initial
  // density rho (number density rho / m) instead of the sum
  // (particles_compute_density.F90:442-473, 526-556)
  if (P.ncoll && P.coll_rho_type == 0 && S.sid[i] != 0)
    rho = P.density_type == 1 ? P.rho0 : P.rho0 / mi;
  const double press = eos_pressure(P, P.density_type == 2 ? rho * mi : rho);
  S.rho[i] = rho;
  S.p[i] = press;
  // per-particle factors of the pair forces: p/n^2 and 1/n^2 (Newtonian
  // Morris and Espanol, whose viscous terms divide by n_i n_j: 1/n)
  const double an = press / (rho * rho);
  const double bn = (P.newtonian && P.force_type <= 2) ? 1.0 / rho : 1.0 / (rho * rho);
  S.an[i] = an;
  S.bn[i] = bn;
  if (MODE == kStore || MODE == kBuild) S.nbr_count[i] = cnt;
  if (MODE == kBuild && cnt > P.list_cap) S.gate[2] = 1;  // every writer stores 1
  // the shear time of the build (the copies across sheared faces drift
  // from their build positions as it advances)
  if (MODE == kBuild && i == 0) S.time_dev[3] = S.time_dev[1];
  S.nbr_in[i] = nin;
  S.recheck[i] = rechecks;
  write_force_records<DIM>(P, S, i, an, bn, press);
}

// The rescan (kRescan) has work only after a list overflowed at the build;
// it is launched on a reduced grid, whose workgroups take every gridDim.x-th
// chunk slot, so that on the other steps its launch costs little.
template <int DIM, bool MIXED, int TILE, int BUF, bool LE, int MODE>
__global__ __launch_bounds__(kWave) void k_density(DevParams P, DeviceState S) {
  if (MODE == kBuild && S.gate[0] == 0) return;
  if ((MODE == kWalk || MODE == kRescan) && S.gate[0] != 0) return;
  // no list overflowed at the build: nothing to rescan
  if (MODE == kRescan && S.gate[2] == 0) return;
  constexpr bool kScan = MODE != kWalk;
  const int le_box = P.le_g;
  if (!LE) P.le_g = -1;
  __shared__ float4 tile[kScan ? TILE + 1 : 1];
  // whole groups of 8 rows: the final flush stores the last group of a
  // lane's entries as 8 rows, of which those past its count are padding
  __shared__ uint32_t buf[(BUF + 7) / 8 * 8 * kWave];
  if (MODE == kRescan) {
    const int slots = chunk_slots(P);
    for (int slot = blockIdx.x; slot < slots; slot += gridDim.x) {
      density_chunk<DIM, MIXED, TILE, BUF, LE, MODE>(P, S, slot, le_box, tile, buf);
      __syncthreads();
    }
  } else {
    density_chunk<DIM, MIXED, TILE, BUF, LE, MODE>(P, S, blockIdx.x, le_box, tile, buf);
  }
}

__global__ void k_eos(DevParams P, DeviceState S) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const double r = S.rho[i];
  const double press = eos_pressure(P, P.density_type == 2 ? r * S.m[i] : r);
  S.p[i] = press;
  S.an[i] = press / (r * r);
  S.bn[i] = 1.0 / (r * r);
}

// Force-pass records after the velocities or the pressure changed between
// the density and the force pass (separate kick / EOS calls).
__global__ void k_refresh_records(DevParams P, DeviceState S) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  if (P.dim == 3)
    write_force_records<3>(P, S, i, S.an[i], S.bn[i], S.p[i]);
  else
    write_force_records<2>(P, S, i, S.an[i], S.bn[i], S.p[i]);
}

// ---------------------------------------------------------------------
// K4: pair forces
// ---------------------------------------------------------------------

// Random force direction seen by particle i for neighbour j: a Gaussian
// vector with the covariance of the reference's symmetric traceless
// dW . e_ij (I/2 + e e^T/6 in 3D, I/2 in 2D), oriented by the lower pid.
// g3 (Espanol's trace draw): the block's fourth normal, shared by the pair.
template <int DIM, typename T>
__device__ __forceinline__ void noise_vector(const DevParams& P, int64_t step, uint32_t pidi,
                                             uint32_t pidj, T e0, T e1, T e2, T& w0, T& w1,
                                             T& w2, T* g3 = nullptr) {
  uint32_t lo = pidi < pidj ? pidi : pidj, hi = pidi < pidj ? pidj : pidi;
  float g[4];
  pair_normals(P.seed, lo, hi, step, g);
  if (g3) *g3 = (T)g[3];
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

// force types: 3 Hu-Adams, 4 Hu-Adams angular, 5 Oldroyd-B (non-Newtonian
// Hu-Adams).  The bodies sum m_i times the acceleration; finish_force
// divides by m_i once.  DECIDE (Verlet lists): the entries are candidates
// within rc + skin and the body skips those outside the cut-off, deciding
// exactly as the density pass of the same step did.
template <int DIM, int FT, bool COLL = false, bool DECIDE = false>
struct ForceF64 {
  const DevParams& P;
  const DeviceState& S;
  const Home& h;
  double vi0, vi1, vi2, ai, bi, pi, inv_ni;
  uint32_t pidi;
  int64_t step;            // noise step (the context's marching step)
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  double L[9];
  double pti[9];
  // colloids: own species id; a boundary particle's position relative to
  // the nearest image of its colloid's centre and that image's velocity;
  // time of the pass's shear length
  int sidi = 0;
  double rxc_i[3] = {0.0, 0.0, 0.0}, vcoll_i[3] = {0.0, 0.0, 0.0};
  double le_t = 0.0;
  static constexpr bool kImages = true;
  // Verlet lists: the walk takes eight entries at a time (two with the
  // Oldroyd-B body's registers) and gathers their records and pids before
  // it evaluates the first pair, so a wave keeps more loads in flight.
  // Measured on the 128^3 box: forces 1.824 ms one entry at a time, 1.819
  // two, 1.761 four, 1.65 eight (193 VGPRs, two waves per SIMD); the
  // Oldroyd-B batch of 256 boxes 6.62 -> 6.23 ms with two.
  // With colloids the Verlet walk takes twos (512 spheres of radius 5 in
  // the sheared 128^3 box: forces 3.66 ms one at a time, 3.31 in twos, 3.69
  // in fours; the substitution branch's registers leave no room for more).
  static constexpr bool kBatched = DECIDE;
  static constexpr int kBatch = COLL ? 2 : (FT >= 5 ? 2 : 8);
  __device__ __forceinline__ void body(uint32_t ent) {
    const int j = (int)(ent & kIndexMask);
    if (COLL) {
      // two boundary particles do not interact (pp_interaction.inc: same
      // colloid CYCLE, colloid-colloid pair forces off)
      if (sidi != 0 && (ent & kBoundaryBit) != 0u) return;
    }
    // both halves of the record into registers up front (a reference into
    // global memory would leave each field's load at its first use)
    const double4* rj = reinterpret_cast<const double4*>(S.rec64 + 8 * (int64_t)j);
    const double4 ra = rj[0], rb = rj[1];
    eval<false>(ent, ra, rb, 0u);
  }
  template <int NB>
  __device__ __forceinline__ void body_batch(const uint32_t (&e)[NB], const bool (&ok)[NB]) {
    double4 ra[NB], rb[NB];
    uint32_t pj[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const int j = (int)(e[k] & kIndexMask);
      const double4* rj = reinterpret_cast<const double4*>(S.rec64 + 8 * (int64_t)j);
      ra[k] = rj[0];
      rb[k] = rj[1];
      pj[k] = P.noise ? (uint32_t)S.pid[j] : 0u;
    }
#pragma unroll
    for (int k = 0; k < NB; ++k)
      if (ok[k]) eval<true>(e[k], ra[k], rb[k], pj[k]);
  }
  // One pair from j's record; PID: j's pid given (else gathered for the
  // noise term).
  template <bool PID>
  __device__ __forceinline__ void eval(uint32_t ent, const double4& ra, const double4& rb,
                                       uint32_t pid_in) {
    const int j = (int)(ent & kIndexMask);
    const Images o = decode_images(ent);
    const bool jb = COLL && (ent & kBoundaryBit) != 0u;
    // two boundary particles do not interact
    if (COLL && sidi != 0 && jb) return;
    double d0, d1, d2, r2;
    double jump;  // x velocity of j's image minus j's own
    if (DECIDE) {
      // the image and the sheared face crossed come from the positions
      int sg;
      r2 = exact_d2_verlet<DIM>(P, h, ra.x, ra.y, ra.z, d0, d1, d2, sg);
      if (!(r2 < P.cut2)) return;
      jump = sg == 0 ? P.le_u : (sg == 1 ? -P.le_u : 0.0);
    } else {
      r2 = pair_d2<DIM>(P, h, o, ra.x, ra.y, ra.z, d0, d1, d2);
      jump = le_dv(P, o);
    }
    const double ir = rsqrt_nr(r2);
    const double r = r2 * ir;
    double w, gw;
    kernel_eval(P, r * P.inv_h, w, gw);
    // (without sheared faces the jump is the constant 0 and no add is made)
    double u0 = vi0 - (P.le_g >= 0 ? rb.x + jump : rb.x), u1 = vi1 - rb.y,
           u2 = DIM == 3 ? vi2 - rb.z : 0.0;
    const double bj = rb.w;
    // pressure terms: p_i / n_i^2 + p_j / n_j^2 (type 3), (p_i + p_j) / 2 (type 4)
    double psum = 0.0, pij = 0.0;
    if (FT <= 3) psum = ai + ra.w;
    if (FT == 4) pij = (pi + S.p[j]) * 0.5;
    // with a colloid boundary particle in an Oldroyd-B flow both sides take
    // the fluid particle's 1/n^2 and Pi/n^2
    // (pp_interaction_fc.inc:80-113, pp_interaction_cf.inc:80-110)
    double bsum = bi + bj;
    const bool cpair = COLL && (sidi != 0 || jb);
    if (cpair) {
      // a fluid particle and a colloid boundary particle: the boundary
      // particle's velocity changes, and its pressure (Newtonian) or its
      // number density and pressure tensor (Oldroyd-B)
      double pb;
      colloid_substitute(j, jb ? S.sid[j] : 0, ra, rb, d0, d1, d2, jump, u0, u1, u2, pb);
      if (FT == 3) psum = sidi == 0 ? ai + pb * bj : pb * bi + ra.w;
      // Morris and Espanol: bi, bj = 1/rho
      if (FT <= 2) psum = sidi == 0 ? ai + pb * (bj * bj) : pb * (bi * bi) + ra.w;
      if (FT == 4) pij = pb;
      if (FT == 5) bsum = sidi == 0 ? bi + bi : bj + bj;
    }
    double a0, a1, a2;
    if (FT == 1) {
      // rhs_force_ff_Newtonian_Morris.F90:112-136: accelerations, each term
      // carrying the partner's mass (the caller skips the division by m_i);
      // bi, bj = 1/rho
      const double mj = S.m[j];
      const double fcm = mj * psum * gw * ir;
      const double fdm = mj * 2.0 * P.eta * gw * bi * bj * ir;
      a0 = fdm * u0 - fcm * d0;
      a1 = fdm * u1 - fcm * d1;
      a2 = fdm * u2 - fcm * d2;
    } else if (FT == 2) {
      // rhs_force_ff_Newtonian_Espanol.F90:152-251; bi, bj = 1/n
      const double gnn = gw * bi * bj * ir;
      const double ev = (DIM == 3 ? fma(d2, u2, fma(d1, u1, d0 * u0)) : fma(d1, u1, d0 * u0)) * ir;
      const double fd1 = (2.0 - 1.0 / DIM) * P.eta * gnn;
      const double fd2 = P.eta * (DIM + 2.0) / DIM * gnn * ev;
      // pressure and the dissipative term along e, both along d / r
      const double se = (psum * gw * -1.0 + fd2) * ir;
      a0 = fd1 * u0 + se * d0;
      a1 = fd1 * u1 + se * d1;
      a2 = fd1 * u2 + se * d2;
      if (P.noise) espanol_noise(PID ? pid_in : (uint32_t)S.pid[j], d0, d1, d2, ir, gnn, a0, a1, a2);
    } else if (FT == 4) {
      // rhs_force_ff_Newtonian_HuAdams_angular.F90:166-229: everything
      // along e = d / r
      const double du = DIM == 3 ? fma(d2, u2, fma(d1, u1, d0 * u0)) : fma(d1, u1, d0 * u0);
      const double aij = (bi + bj) * gw;
      double sc = P.eta_coef * P.eta * aij * du * (ir * ir) - aij * pij;
      if (P.noise) {
        const uint32_t pj_ = PID ? pid_in : (uint32_t)S.pid[j];
        float g[4];
        pair_normals(P.seed, pidi < pj_ ? pidi : pj_, pidi < pj_ ? pj_ : pidi, step, g);
        sc = sc + (double)g[0] * P.sqrt_2kt_eta * sqrt_nr(-P.eta_coef * aij * ir) * P.inv_sqrt_dt;
      }
      sc = sc * ir;
      a0 = sc * d0;
      a1 = sc * d1;
      a2 = sc * d2;
    } else {
      const double gir = gw * ir;
      const double fd = P.eta * (FT >= 5 ? bsum : bi + bj) * gir;
      if (FT == 3) {
        // rhs_force_ff_Newtonian_HuAdams.F90:140-237
        const double fcr = psum * gir;
        a0 = fd * u0 - fcr * d0;
        a1 = fd * u1 - fcr * d1;
        a2 = fd * u2 - fcr * d2;
      } else {
        // rhs_force_ff_non_Newtonian_HuAdams.F90:105-131 (FT 5; Espanol's,
        // FT 6, differs only in the noise; the angular form, FT 7, below) with
        // Pi = (p + G) I - G c (particles_compute_pressure_tensor.F90:114-133),
        // and the velocity gradient of pp_vgt_ip.inc:16-20
        const double du[3] = {-u0, -u1, -u2};
        const double dx[3] = {d0, d1, d2};
        // a boundary particle's gradient divides by its fluid partner's n
        const double gfac = gw * ((cpair && sidi != 0) ? 1.0 / S.rho[j] : inv_ni) * ir;
#pragma unroll
        for (int b = 0; b < DIM; ++b)
#pragma unroll
          for (int a = 0; a < DIM; ++a) L[a + DIM * b] = fma(du[a] * dx[b], gfac, L[a + DIM * b]);
        // j's Pi / n^2 from its 48-byte record (c is symmetric)
        const double2* __restrict__ rp = reinterpret_cast<const double2*>(S.pt64 + 6 * (int64_t)j);
        const double2 q0 = rp[0], q1 = rp[1], q2 = rp[2];
        const double pjb[9] = {q0.x, q1.y, q2.x, q1.y, q0.y, q2.y, q2.x, q2.y, q1.x};
        double fc[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int jj = 0; jj < DIM; ++jj)
#pragma unroll
          for (int a = 0; a < DIM; ++a) {
            const double ptjb = DIM == 3 ? pjb[a + 3 * jj] : pjb[(a == jj ? (a == 0 ? 0 : 4) : 1)];
            // colloid pairs: the fluid particle's Pi / n^2 on both sides
            double t = cpair ? (sidi == 0 ? 2.0 * (pti[a + DIM * jj] * bi) : ptjb + ptjb)
                             : pti[a + DIM * jj] * bi + ptjb;
            // rhs_force_ff_non_Newtonian_HuAdams_angular.F90: the diagonal
            // averaged, (Pi_i + Pi_j) / 2 (1/n_i^2 + 1/n_j^2)
            if (FT == 7 && a == jj) t = 0.5 * (pti[a + DIM * jj] + ptjb / bj) * (bi + bj);
            fc[jj] = fc[jj] + t * dx[a];
          }
        if (FT == 7) {
          // its dissipative term acts along e: eta_coef eta (1/n_i^2 +
          // 1/n_j^2) W' (e . v_ij) / r
          const double dv = DIM == 3 ? fma(d2, u2, fma(d1, u1, d0 * u0)) : fma(d1, u1, d0 * u0);
          const double fde = P.eta_coef * P.eta * (bi + bj) * gw * dv * (ir * ir * ir);
          a0 = fde * d0 - fc[0] * gir;
          a1 = fde * d1 - fc[1] * gir;
          a2 = DIM == 3 ? fde * d2 - fc[2] * gir : 0.0;
        } else {
          a0 = fd * u0 - fc[0] * gir;
          a1 = fd * u1 - fc[1] * gir;
          a2 = DIM == 3 ? fd * u2 - fc[2] * gir : 0.0;
        }
      }
      if (FT == 6 && P.noise) {
        espanol_noise(PID ? pid_in : (uint32_t)S.pid[j], d0, d1, d2, ir,
                      gw * ir * inv_ni / S.rho[j], a0, a1, a2);
      } else if (FT != 7 && P.noise) {
        // f_r W e with W e = s (ca g + cb (g.e) e), e = d / r, s = +1 when
        // pid_i < pid_j (noise_vector), written on d to save operations
        const uint32_t pidj = PID ? pid_in : (uint32_t)S.pid[j];
        float g[4];
        pair_normals(P.seed, pidi < pidj ? pidi : pidj, pidi < pidj ? pidj : pidi, step, g);
        // sqrt(-4 kT f_d) / sqrt(dt) as one root (the factor is uniform)
        const double fr = sqrt_nr((-4.0 * P.kt / P.dt) * fd);
        const double sfr = pidi < pidj ? fr : -fr;
        const double ca = 0.70710678118654752440;
        const double g0 = (double)g[0], g1 = (double)g[1];
        if (DIM == 3) {
          const double cb = 0.81649658092772603273 - 0.70710678118654752440;
          const double g2 = (double)g[2];
          const double k = cb * fma(g2, d2, fma(g1, d1, g0 * d0)) * (ir * ir);
          a0 = fma(sfr, fma(k, d0, ca * g0), a0);
          a1 = fma(sfr, fma(k, d1, ca * g1), a1);
          a2 = fma(sfr, fma(k, d2, ca * g2), a2);
        } else {
          a0 = fma(sfr, ca * g0, a0);
          a1 = fma(sfr, ca * g1, a1);
        }
      }
    }
    f0 = f0 + a0;
    f1 = f1 + a1;
    if (DIM == 3) f2 = f2 + a2;
  }

  // Espanol's thermal force (rhs_force_ff_Newtonian_Espanol.F90:198-249,
  // rhs_force_ff_non_Newtonian_Espanol.F90): (A S e + B tr e) / sqrt(dt)
  // with S the symmetric traceless part of the pair's Gaussian dW and tr
  // its trace / num_dim, Z = 4 kT W' / (r n_i n_j) (gnn = W' / (r n_i n_j)).
  // S and tr are independent: S e is drawn as in the Hu-Adams body, oriented
  // by the lower pid, and tr is the pair block's fourth normal over
  // sqrt(num_dim), shared by both sides.
  __device__ __forceinline__ void espanol_noise(uint32_t pidj, double d0, double d1, double d2,
                                                double ir, double gnn, double& a0, double& a1,
                                                double& a2) const {
    float g[4];
    pair_normals(P.seed, pidi < pidj ? pidi : pidj, pidi < pidj ? pidj : pidi, step, g);
    const double zij = 4.0 * P.kt * gnn;
    const double ea = 5.0 * P.eta / 3.0, eb = (DIM + 2.0) * P.eta / 3.0;
    const double aij = sqrt_nr(-zij * ea) * P.inv_sqrt_dt;
    const double bij = sqrt_nr(-zij * DIM / 2.0 * (eb + ea * (2.0 / DIM - 1.0))) * P.inv_sqrt_dt;
    const double saij = pidi < pidj ? aij : -aij;
    const double ca = 0.70710678118654752440;
    const double g0 = (double)g[0], g1 = (double)g[1];
    const double btr = bij * (double)g[3] * (DIM == 3 ? 0.57735026918962576451 : ca) * ir;
    if (DIM == 3) {
      const double cb = 0.81649658092772603273 - 0.70710678118654752440;
      const double g2 = (double)g[2];
      const double k = cb * fma(g2, d2, fma(g1, d1, g0 * d0)) * (ir * ir);
      a0 = fma(saij, fma(k, d0, ca * g0), fma(btr, d0, a0));
      a1 = fma(saij, fma(k, d1, ca * g1), fma(btr, d1, a1));
      a2 = fma(saij, fma(k, d2, ca * g2), fma(btr, d2, a2));
    } else {
      a0 = fma(saij, ca * g0, fma(btr, d0, a0));
      a1 = fma(saij, ca * g1, fma(btr, d1, a1));
    }
  }

  // A fluid particle and a colloid boundary particle
  // (pp_interaction_fc.inc, pp_interaction_cf.inc): the boundary particle
  // moves at the no-slip velocity seen from the fluid particle and carries
  // the fluid particle's pressure (pb); its own number density stays.  The
  // no-slip velocity is formed in the boundary particle's frame, where the
  // fluid particle sits at r_xc + (x_f - x_c) and moves at its image
  // velocity, so both sides of the pair see the same relative velocity
  // u = v_i - v_j.
  __device__ __forceinline__ void colloid_substitute(int j, int sj, const double4& ra,
                                                     const double4& rb, double d0, double d1,
                                                     double d2, double jump, double& u0,
                                                     double& u1, double& u2, double& pb) {
    double vc[3];
    if (sidi == 0) {
      const CollHot C = S.coll_hot[sj - 1];
      const double xj[3] = {ra.x, ra.y, DIM == 3 ? ra.z : 0.0};
      double rxc[3], vcoll[3];
      coll_nearest<DIM>(P, C, xj, le_t, rxc, vcoll);
      const double rxf[3] = {rxc[0] + d0, rxc[1] + d1, DIM == 3 ? rxc[2] + d2 : 0.0};
      const double vf[3] = {vi0 - jump, vi1, DIM == 3 ? vi2 : 0.0};
      coll_noslip<DIM>(P, C, rxc, vcoll, rxf, vf, vc);
      u0 = vf[0] - vc[0];
      u1 = vf[1] - vc[1];
      u2 = DIM == 3 ? vf[2] - vc[2] : 0.0;
      pb = pi;
    } else {
      const CollHot C = S.coll_hot[sidi - 1];
      const double rxf[3] = {rxc_i[0] - d0, rxc_i[1] - d1, DIM == 3 ? rxc_i[2] - d2 : 0.0};
      const double vf[3] = {P.le_g >= 0 ? rb.x + jump : rb.x, rb.y, DIM == 3 ? rb.z : 0.0};
      coll_noslip<DIM>(P, C, rxc_i, vcoll_i, rxf, vf, vc);
      u0 = vc[0] - vf[0];
      u1 = vc[1] - vf[1];
      u2 = DIM == 3 ? vc[2] - vf[2] : 0.0;
      pb = S.p[j];
    }
  }
};

template <int DIM, int FT, bool DECIDE = false>
struct ForceMixed {
  const DevParams& P;
  const DeviceState& S;
  const Home& h;
  float vi0, vi1, vi2, ai, bi, pi;
  uint32_t pidi;
  int64_t step;            // noise step (the context's marching step)
  KahanF f0, f1, f2;
  // Oldroyd-B: own Pi / n^2 (xx, yy, zz, xy, xz, yz), 1 / n, velocity gradient
  float pb[6];
  float inv_ni;
  WideF L[DIM * DIM];
  static constexpr bool kImages = false;
  // Verlet lists: entries four at a time, the records of all four
  // gathered before the first pair is evaluated (the 128^3 Tait box: forces
  // 1.27 ms one at a time, 1.16 in fours, 1.19 in eights; 256 Oldroyd-B
  // boxes of 2k-60k particles: 3.94 in twos, 3.87 in fours)
  static constexpr bool kBatched = DECIDE;
  static constexpr int kBatch = 4;
  // j's records: FP32 position (w: p/n^2 or p, Newtonian), velocity (w:
  // 1/n^2, or the pid bits for Oldroyd-B), pid, and the Oldroyd-B
  // (1/n^2, Pxx, Pyy, Pzz) and (Pxy, Pxz, Pyz, 0)
  struct Rec {
    float4 x, v, q, q2;
    uint32_t pid;
  };
  __device__ __forceinline__ Rec load_rec(int j) const {
    Rec r;
    if (FT >= 5) {
      r.v = S.vr[j];
      r.x = load_xyz(S.xa + j);
      r.q = S.qr[j];
      r.q2 = S.qr2[j];
      r.pid = __float_as_uint(r.v.w);
    } else {
      r.x = S.qr[j];
      r.v = S.vr[j];
      r.q = r.q2 = make_float4(0.f, 0.f, 0.f, 0.f);
      r.pid = P.noise ? (uint32_t)S.pid[j] : 0u;
    }
    return r;
  }
  __device__ __forceinline__ void body(uint32_t ent) {
    const int j = (int)(ent & kIndexMask);
    const Rec r = load_rec(j);
    body_rec(j, r.x, r.v, r.pid, r.q, r.q2, decode_code(ent >> kCodeShift));
  }
  template <int NB>
  __device__ __forceinline__ void body_batch(const uint32_t (&e)[NB], const bool (&ok)[NB]) {
    Rec r[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) r[k] = load_rec((int)(e[k] & kIndexMask));
#pragma unroll
    for (int k = 0; k < NB; ++k)
      if (ok[k])
        body_rec((int)(e[k] & kIndexMask), r[k].x, r[k].v, r[k].pid, r[k].q, r[k].q2,
                 decode_code(e[k] >> kCodeShift));
  }
  // xj: cell-relative position (w: p/n^2 for Hu-Adams, p for the angular
  // form); vj: velocity (w: 1/n^2, Newtonian); Oldroyd-B qj: (1/n^2, Pxx,
  // Pyy, Pzz), q2: (Pxy, Pxz, Pyz, 0)
  __device__ __forceinline__ void body_rec(int jidx, const float4& xj, const float4& vj,
                                           uint32_t pidj, const float4& qj, const float4& q2,
                                           const Offsets& o) {
    float d0, d1, d2;
    rel_diff<DIM>(P, h, xj, o, d0, d1, d2);
    const float r2 = mixed_s2(d0, d1, d2);
    // Verlet candidates: the pair is evaluated either way (outside the
    // cut-off the kernel gradient is 0) and only added when inside.  The
    // record's w is pinned to a register ahead of the decision (an empty
    // asm), so the record arrives in one 16-byte gather with the position
    // instead of a second 4-byte one after the decision's branch.
    bool use = true;
    float xjw = xj.w;
    if (DECIDE) {
      asm volatile("" : "+v"(xjw));
      int rechecks = 0;
      use = mixed_inside<DIM>(P, S, h, jidx, r2, rechecks);
    }
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * ir;
    float w, gw;
    kernel_w_gradw_f32(P.kernel_type, r * P.inv_hf, P.coeff, P.coef_gradf, w, gw);
    const float u0 = vi0 - (vj.x + le_dvf(P, h, o)), u1 = vi1 - vj.y, u2 = vi2 - vj.z;
    float fi0, fi1, fi2;
    if (FT >= 5) {
      // rhs_force_ff_non_Newtonian_HuAdams.F90:105-131 with the records'
      // Pi / n^2, and the velocity gradient of pp_vgt_ip.inc:16-20 (FT 6,
      // Espanol: its thermal force below; FT 7, angular: the diagonal
      // averaged, (Pi_i + Pi_j) / 2 (1/n_i^2 + 1/n_j^2), and dissipation
      // along e)
      float sxx = pb[0] + qj.y, syy = pb[1] + qj.z, szz = pb[2] + qj.w;
      if (FT == 7) {
        const float bsum = bi + qj.x;
        sxx = 0.5f * (pb[0] / bi + qj.y / qj.x) * bsum;
        syy = 0.5f * (pb[1] / bi + qj.z / qj.x) * bsum;
        szz = 0.5f * (pb[2] / bi + qj.w / qj.x) * bsum;
      }
      const float sxy = pb[3] + q2.x, sxz = pb[4] + q2.y, syz = pb[5] + q2.z;
      float c0, c1, c2;
      if (DIM == 3) {
        c0 = sxx * d0 + sxy * d1 + sxz * d2;
        c1 = sxy * d0 + syy * d1 + syz * d2;
        c2 = sxz * d0 + syz * d1 + szz * d2;
      } else {
        c0 = sxx * d0 + sxy * d1;
        c1 = sxy * d0 + syy * d1;
        c2 = 0.f;
      }
      const float gir = gw * ir;
      const float f_d = P.etaf * (bi + qj.x) * gir;
      if (FT == 7) {
        const float fde = P.eta_coeff * P.etaf * (bi + qj.x) * gw * (d0 * u0 + d1 * u1 + d2 * u2) *
                          (ir * ir * ir);
        fi0 = fde * d0 - c0 * gir;
        fi1 = fde * d1 - c1 * gir;
        fi2 = fde * d2 - c2 * gir;
      } else {
        fi0 = f_d * u0 - c0 * gir;
        fi1 = f_d * u1 - c1 * gir;
        fi2 = f_d * u2 - c2 * gir;
      }
      const float gfac = gir * inv_ni;
      const float du[3] = {-u0, -u1, -u2};
      const float dx[3] = {d0, d1, d2};
      if (use) {
#pragma unroll
        for (int b = 0; b < DIM; ++b)
#pragma unroll
          for (int a = 0; a < DIM; ++a) L[a + DIM * b].add(du[a] * dx[b] * gfac);
      }
      if (FT == 6 && P.noise) {
        // Espanol's (A S e + B tr e) / sqrt(dt), Z = 4 kT W' / (r n_i n_j)
        float w0, w1, w2, g3;
        noise_vector<DIM, float>(P, step, pidi, pidj, d0 * ir, d1 * ir, d2 * ir, w0, w1, w2, &g3);
        const float zij = 4.0f * P.ktf * gir * inv_ni * __builtin_amdgcn_sqrtf(qj.x);
        const float ea = 5.0f * P.etaf / 3.0f, eb = (DIM + 2.0f) * P.etaf / 3.0f;
        const float aij = __builtin_amdgcn_sqrtf(-zij * ea) * P.inv_sqrt_dtf;
        const float bij = __builtin_amdgcn_sqrtf(-zij * DIM / 2.0f * (eb + ea * (2.0f / DIM - 1.0f))) *
                          P.inv_sqrt_dtf;
        const float btr = bij * g3 * (DIM == 3 ? 0.57735026918962576f : 0.70710678118654752f) * ir;
        fi0 = fi0 + aij * w0 + btr * d0;
        fi1 = fi1 + aij * w1 + btr * d1;
        fi2 = fi2 + aij * w2 + btr * d2;
      } else if (FT == 5 && P.noise) {
        float w0, w1, w2;
        noise_vector<DIM, float>(P, step, pidi, pidj, d0 * ir, d1 * ir, d2 * ir, w0, w1, w2);
        const float f_r = __builtin_amdgcn_sqrtf(-4.0f * P.ktf * f_d) * P.inv_sqrt_dtf;
        fi0 = fi0 + f_r * w0;
        fi1 = fi1 + f_r * w1;
        fi2 = fi2 + f_r * w2;
      }
    } else if (FT == 1) {
      // rhs_force_ff_Newtonian_Morris.F90:112-136 (accelerations carrying
      // the partner's mass; bi, vj.w = 1/rho)
      const float mj = (float)S.m[jidx];
      const float fcm = mj * (ai + xjw) * gw * ir;
      const float fdm = mj * 2.0f * P.etaf * gw * bi * vj.w * ir;
      fi0 = fdm * u0 - fcm * d0;
      fi1 = fdm * u1 - fcm * d1;
      fi2 = fdm * u2 - fcm * d2;
    } else if (FT == 2) {
      // rhs_force_ff_Newtonian_Espanol.F90:152-251 (bi, vj.w = 1/n)
      const float gnn = gw * bi * vj.w * ir;
      const float ev = (d0 * u0 + d1 * u1 + d2 * u2) * ir;
      const float fd1 = (2.0f - 1.0f / DIM) * P.etaf * gnn;
      const float fd2 = P.etaf * (DIM + 2.0f) / DIM * gnn * ev;
      const float se = (fd2 - (ai + xjw) * gw) * ir;
      fi0 = fd1 * u0 + se * d0;
      fi1 = fd1 * u1 + se * d1;
      fi2 = fd1 * u2 + se * d2;
      if (P.noise) {
        // (A S e + B tr e) / sqrt(dt), tr = g[3] / sqrt(num_dim)
        float w0, w1, w2, g3;
        noise_vector<DIM, float>(P, step, pidi, pidj, d0 * ir, d1 * ir, d2 * ir, w0, w1, w2, &g3);
        const float zij = 4.0f * P.ktf * gnn;
        const float ea = 5.0f * P.etaf / 3.0f, eb = (DIM + 2.0f) * P.etaf / 3.0f;
        const float aij = __builtin_amdgcn_sqrtf(-zij * ea) * P.inv_sqrt_dtf;
        const float bij = __builtin_amdgcn_sqrtf(-zij * DIM / 2.0f * (eb + ea * (2.0f / DIM - 1.0f))) *
                          P.inv_sqrt_dtf;
        const float btr = bij * g3 * (DIM == 3 ? 0.57735026918962576f : 0.70710678118654752f) * ir;
        fi0 = fi0 + aij * w0 + btr * d0;
        fi1 = fi1 + aij * w1 + btr * d1;
        fi2 = fi2 + aij * w2 + btr * d2;
      }
    } else if (FT == 3) {
      const float fcr = (ai + xjw) * gw * ir;
      const float f_d = P.etaf * (bi + vj.w) * gw * ir;
      fi0 = f_d * u0 - fcr * d0;
      fi1 = f_d * u1 - fcr * d1;
      fi2 = f_d * u2 - fcr * d2;
      if (P.noise) {
        float w0, w1, w2;
        noise_vector<DIM, float>(P, step, pidi, pidj, d0 * ir, d1 * ir, d2 * ir, w0, w1, w2);
        const float f_r = __builtin_amdgcn_sqrtf(-4.0f * P.ktf * f_d) * P.inv_sqrt_dtf;
        fi0 = fi0 + f_r * w0;
        fi1 = fi1 + f_r * w1;
        fi2 = fi2 + f_r * w2;
      }
    } else {
      const float ev = (d0 * u0 + d1 * u1 + d2 * u2) * ir;
      const float pij = (pi + xjw) * 0.5f;
      const float aij = (bi + vj.w) * gw;
      const float f_c = aij * pij;
      const float f_d = P.eta_coeff * P.etaf * aij * ev * ir;
      float sc = f_d - f_c;
      if (P.noise) {
        float g[4];
        pair_normals(P.seed, pidi < pidj ? pidi : pidj, pidi < pidj ? pidj : pidi, step, g);
        sc = sc + g[0] * P.sqrt_2kt_etaf * __builtin_amdgcn_sqrtf(-P.eta_coeff * aij * ir) *
                      P.inv_sqrt_dtf;
      }
      sc = sc * ir;
      fi0 = sc * d0;
      fi1 = sc * d1;
      fi2 = sc * d2;
    }
    if (!use) return;
    f0.add(fi0);
    f1.add(fi1);
    if (DIM == 3) f2.add(fi2);
  }
};

template <int DIM>
__device__ __forceinline__ void finish_force(const DevParams& P, const DeviceState& S, int i,
                                             double f0, double f1, double f2,
                                             bool fluid = true) {
  const int64_t n = P.n;
  // fluid particles only (particles_apply_body_force.F90:173); the body
  // forces of types 2 and 3 follow in k_body_force
  if (P.body_force == 1 && fluid) {
    // particles_apply_body_force.F90:125-193
    if (fabs(P.bf[0]) > 1e-14) f0 = f0 + P.bf[0];
    if (fabs(P.bf[1]) > 1e-14) f1 = f1 + P.bf[1];
    if (DIM == 3 && fabs(P.bf[2]) > 1e-14) f2 = f2 + P.bf[2];
  }
  S.f[i] = f0;
  S.f[n + i] = f1;
  if (DIM == 3) S.f[2 * n + i] = f2;
}

// particles_compute_act.F90:140-155 on particle i's own ct and velocity
// gradient L: stores vgt = L and act = c L^T + L c - (c - I) / tau.
template <int DIM>
__device__ __forceinline__ void conformation_rate(const DevParams& P, const DeviceState& S, int i,
                                                  const double* L) {
  const int64_t n = P.n;
  const int d2n = DIM * DIM;
  double cc[9];
#pragma unroll
  for (int q = 0; q < d2n; ++q) {
    cc[q] = S.ct[(int64_t)q * n + i];
    S.vgt[(int64_t)q * n + i] = L[q];
  }
#pragma unroll
  for (int b = 0; b < DIM; ++b)
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < DIM; ++k)
        t = t + cc[a + DIM * k] * L[b + DIM * k] + L[a + DIM * k] * cc[k + DIM * b];
      double id = a == b ? 1.0 : 0.0;
      S.act[(int64_t)(a + DIM * b) * n + i] = t - (cc[a + DIM * b] - id) / P.tau;
    }
}

// The neighbours of a chunk: its stored lists, or (SCAN) the stencil
// scanned again.
template <int DIM, bool SCAN, bool SKIN, class Pol>
__device__ __forceinline__ void visit_neighbours(const DevParams& P, const DeviceState& S,
                                                 const Chunk& ch, const Home& h, Pol& pol,
                                                 float4* __restrict__ tile,
                                                 uint32_t* __restrict__ buf) {
  if constexpr (SCAN) {
    int rechecks = 0;
    (void)staged_scan<DIM, Pol, kTile, kBuf, SKIN>(P, S, ch, h, pol, tile, buf, rechecks, nullptr);
  } else {
    (void)tile;
    for_each_listed<DIM>(P, S, ch, h, pol, buf);
  }
}

// The FP64 walk of per-step lists (without colloids) takes 8 KB of LDS per
// workgroup, its entry buffer plus a reserve.  That holds it to five waves
// per SIMD.  Its registers alone would allow six, but on the 128^3 box six
// waves' gathers ran 6 % slower (1.99 against 1.88 ms per pass).  The
// Verlet walk evaluates in batches from registers and takes no LDS.
constexpr int kForceLdsWordsF64 = 2048;

// Two launches over the same grid: SCAN = false takes the chunks whose
// lists hold every neighbour and walks them, SCAN = true takes the others
// and scans their stencil again.  Every chunk is done by exactly one of
// them, and the list walk carries none of the scan's state (it needs 40 %
// fewer registers, so more waves share a SIMD).  With colloids a box with
// sheared faces splits the chunks once more: the Lees-Edwards
// instantiation takes the rows within coll_le_cells of a face, the plain
// one (a quarter fewer registers) the interior.  VERLET: the lists hold
// the Verlet candidates and the bodies decide in / out of the cut-off.
// One chunk slot of the force pass.
template <int DIM, int FT, bool MIXED, bool COLL, bool LE, bool SCAN, bool VERLET>
__device__ __forceinline__ void forces_chunk(const DevParams& P, const DeviceState& S, int slot,
                                             int le_box, float4* __restrict__ tile,
                                             uint32_t* __restrict__ buf) {
  Chunk ch;
  if (!load_chunk(P, S, ch, slot)) return;
  if (list_overflow(P, S, ch) != SCAN) return;
  if (COLL && le_box >= 0 && chunk_meets_le_face(le_box, ch, P.coll_le_cells) != LE) return;
  const Home h = load_home<DIM>(P, S, ch);
  const int64_t n = P.n;
  const int ii = h.ii;
  // the Morris body sums accelerations already
  const double inv_mi = FT == 1 ? 1.0 : 1.0 / S.m[ii];
  if (MIXED) {
    const float4 vi = S.vr[ii];
    const float4 qi = S.qr[ii];
    // Newtonian records: qi.w = p/n^2 (Hu-Adams) or p (angular), vi.w = 1/n^2
    ForceMixed<DIM, FT, VERLET> pol{P, S, h, vi.x, vi.y, vi.z, FT <= 3 ? qi.w : 0.f, vi.w,
                            FT == 4 ? qi.w : 0.f,
                            FT >= 5 ? __float_as_uint(vi.w) : (uint32_t)S.pid[ii],
                            S.step_dev[0]};
    if (FT >= 5) {
      const float4 q2 = S.qr2[ii];
      pol.bi = qi.x;
      pol.pb[0] = qi.y;
      pol.pb[1] = qi.z;
      pol.pb[2] = qi.w;
      pol.pb[3] = q2.x;
      pol.pb[4] = q2.y;
      pol.pb[5] = q2.z;
      pol.inv_ni = (float)(1.0 / S.rho[ii]);
    }
    visit_neighbours<DIM, SCAN, VERLET>(P, S, ch, h, pol, tile, buf);
    if (!h.active) return;
    finish_force<DIM>(P, S, h.i, pol.f0.value() * inv_mi, pol.f1.value() * inv_mi,
                      DIM == 3 ? pol.f2.value() * inv_mi : 0.0);
    if (FT >= 5) {
      double Ld[DIM * DIM];
#pragma unroll
      for (int q = 0; q < DIM * DIM; ++q) Ld[q] = pol.L[q].value();
      conformation_rate<DIM>(P, S, h.i, Ld);
    }
  } else {
    ForceF64<DIM, FT, COLL, VERLET> pol{P, S, h};
    if (COLL) {
      pol.sidi = S.sid[ii];
      pol.le_t = S.time_dev[1];
      if (pol.sidi != 0) {
        const double xi[3] = {h.x0, h.x1, DIM == 3 ? h.x2 : 0.0};
        coll_nearest<DIM>(P, S.coll_hot[pol.sidi - 1], xi, pol.le_t, pol.rxc_i, pol.vcoll_i);
      }
    }
    pol.vi0 = S.v[ii];
    pol.vi1 = S.v[n + ii];
    pol.vi2 = DIM == 3 ? S.v[2 * n + ii] : 0.0;
    pol.ai = S.an[ii];
    pol.bi = S.bn[ii];
    pol.pi = S.p[ii];
    pol.inv_ni = 1.0 / S.rho[ii];
    pol.pidi = (uint32_t)S.pid[ii];
    pol.step = S.step_dev[0];
#pragma unroll
    for (int q = 0; q < 9; ++q) pol.L[q] = 0.0;
    if (FT >= 5) {
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
    visit_neighbours<DIM, SCAN, VERLET>(P, S, ch, h, pol, tile, buf);
    if (!h.active) return;
    const int i = h.i;
    finish_force<DIM>(P, S, i, pol.f0 * inv_mi, pol.f1 * inv_mi, pol.f2 * inv_mi,
                      !COLL || pol.sidi == 0);
    if (FT >= 5) conformation_rate<DIM>(P, S, i, pol.L);
  }
}

// With Verlet lists the rescan (SCAN) has work only if a list overflowed at
// the build; it is launched on a reduced grid, whose workgroups take every
// gridDim.x-th chunk slot, so that on the other steps its launch costs
// little.
template <int DIM, int FT, bool MIXED, bool COLL, bool LE, bool SCAN, bool VERLET = false>
__global__ __launch_bounds__(kWave) void k_forces(DevParams P, DeviceState S) {
  if (SCAN && VERLET && S.gate[2] == 0) return;
  const int le_box = P.le_g;
  if (!LE) P.le_g = -1;
  __shared__ float4 tile[SCAN ? kTile + 1 : 1];
  __shared__ uint32_t buf[(!MIXED && !COLL && !SCAN) ? kForceLdsWordsF64 : kBuf * kWave];
  if (SCAN && VERLET) {
    const int slots = chunk_slots(P);
    for (int slot = blockIdx.x; slot < slots; slot += gridDim.x) {
      forces_chunk<DIM, FT, MIXED, COLL, LE, SCAN, VERLET>(P, S, slot, le_box, tile, buf);
      __syncthreads();
    }
  } else {
    forces_chunk<DIM, FT, MIXED, COLL, LE, SCAN, VERLET>(P, S, blockIdx.x, le_box, tile, buf);
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

// DECIDE (Verlet lists, between sorts): the scan takes every candidate
// within rc + skin and the body keeps the ones inside the cut-off on the
// current positions.
template <int DIM, bool DECIDE>
struct SignaturePol {
  const DevParams& P;
  const DeviceState& S;
  const Home& h;
  uint32_t pidi;
  uint32_t hs;
  int nin = 0;
  static constexpr bool kImages = true;
  static constexpr bool kBatched = false;
  __device__ __forceinline__ void body(uint32_t ent) {
    const int j = (int)(ent & kIndexMask);
    if (DECIDE) {
      const int64_t n = P.n;
      double d0, d1, d2;
      int sg;
      if (!(exact_d2_verlet<DIM>(P, h, S.x[j], S.x[n + j], DIM == 3 ? S.x[2 * n + j] : 0.0, d0,
                                 d1, d2, sg) < P.cut2))
        return;
    }
    ++nin;
    hs += pair_hash(pidi, (uint32_t)S.pid[j]);
  }
};

template <int DIM, bool LE, bool VERLET>
__global__ __launch_bounds__(kWave) void k_signature(DevParams P, DeviceState S) {
  if (!LE) P.le_g = -1;
  __shared__ float4 tile[kTile + 1];
  __shared__ uint32_t buf[kBuf * kWave];
  Chunk ch;
  if (!load_chunk(P, S, ch, blockIdx.x)) return;
  const Home h = load_home<DIM>(P, S, ch);
  SignaturePol<DIM, VERLET> pol{P, S, h, (uint32_t)S.pid[h.ii], 0u};
  int rechecks = 0;
  (void)staged_scan<DIM, SignaturePol<DIM, VERLET>, kTile, kBuf, VERLET>(P, S, ch, h, pol, tile,
                                                                         buf, rechecks, nullptr);
  if (!h.active) return;
  S.nbr_hash[h.i] = pol.hs;
  S.nbr_sig_count[h.i] = pol.nin;
  S.recheck[h.i] = rechecks;
}

// ---------------------------------------------------------------------
// launchers: one 64-lane workgroup per chunk; the grid covers the upper
// bound of chunks (one partial chunk per row plus n / 64 full ones)
// ---------------------------------------------------------------------
static inline int nchunks_max(const DevParams& P) { return chunk_slots(P); }

// Workgroups of a rescan launch (a multiple of 8, at most the slots): few
// enough that the launch is cheap on the steps without an overflowed list
// (on the 128^3 box, whose 33k slots each took a workgroup before, the two
// rescans now cost the step 15 us less), and enough that a step with
// overflowed lists checks its slots 16 at a time.
static inline dim3 rescan_grid(const DevParams& P) { return dim3(std::min(chunk_slots(P), 2048)); }

template <int DIM, bool MIXED, bool LE, int MODE>
static void launch_density_shape(hipStream_t s, const DevParams& P, const DeviceState& S) {
  dim3 g(nchunks_max(P)), b(kWave);
  switch (P.density_shape) {
    case 1:
      hipLaunchKernelGGL((k_density<DIM, MIXED, 128, 16, LE, MODE>), g, b, 0, s, P, S);
      break;
    case 2:
      hipLaunchKernelGGL((k_density<DIM, MIXED, 224, 12, LE, MODE>), g, b, 0, s, P, S);
      break;
    case 3:
      hipLaunchKernelGGL((k_density<DIM, MIXED, 128, 12, LE, MODE>), g, b, 0, s, P, S);
      break;
    default:
      hipLaunchKernelGGL((k_density<DIM, MIXED, kTile, kBuf, LE, MODE>), g, b, 0, s, P, S);
  }
}

template <int DIM, bool MIXED>
static void launch_density_dim(hipStream_t s, const DevParams& P, const DeviceState& S,
                               bool scan_all) {
  if (!P.verlet) {
    if (P.le_g >= 0) launch_density_shape<DIM, MIXED, true, kStore>(s, P, S);
    launch_density_shape<DIM, MIXED, false, kStore>(s, P, S);
    return;
  }
  if (scan_all) {
    launch_density_shape<DIM, MIXED, false, kScanAll>(s, P, S);
    return;
  }
  // Verlet lists: each kernel returns at once unless the step's gate
  // selects it (build, or walk + rescan of overflowed chunks); a box with
  // sheared faces splits the chunks between the Lees-Edwards and the plain
  // instantiations, as its per-step lists do
  dim3 g(nchunks_max(P)), b(kWave);
  if constexpr (!MIXED) {  // mixed precision keeps per-step lists in sheared boxes
    if (P.le_g >= 0) {
      launch_density_shape<DIM, MIXED, true, kBuild>(s, P, S);
      hipLaunchKernelGGL((k_density<DIM, MIXED, kTile, kBuf, true, kWalk>), g, b, 0, s, P, S);
      hipLaunchKernelGGL((k_density<DIM, MIXED, kTile, kBuf, true, kRescan>), rescan_grid(P), b,
                         0, s, P, S);
    }
  }
  launch_density_shape<DIM, MIXED, false, kBuild>(s, P, S);
  hipLaunchKernelGGL((k_density<DIM, MIXED, kTile, kBuf, false, kWalk>), g, b, 0, s, P, S);
  hipLaunchKernelGGL((k_density<DIM, MIXED, kTile, kBuf, false, kRescan>), rescan_grid(P), b, 0,
                     s, P, S);
  launch_lists_current(s, S);
}

void launch_density(hipStream_t s, const DevParams& P, const DeviceState& S, bool mixed,
                    bool scan_all) {
  if (P.dim == 3) {
    if (mixed)
      launch_density_dim<3, true>(s, P, S, scan_all);
    else
      launch_density_dim<3, false>(s, P, S, scan_all);
  } else {
    if (mixed)
      launch_density_dim<2, true>(s, P, S, scan_all);
    else
      launch_density_dim<2, false>(s, P, S, scan_all);
  }
}

void launch_eos(hipStream_t s, const DevParams& P, const DeviceState& S) {
  hipLaunchKernelGGL(k_eos, dim3((P.n + 255) / 256), dim3(256), 0, s, P, S);
}

void launch_refresh_records(hipStream_t s, const DevParams& P, const DeviceState& S) {
  hipLaunchKernelGGL(k_refresh_records, dim3((P.n + 255) / 256), dim3(256), 0, s, P, S);
}

template <int DIM, int FT, bool MIXED, bool COLL>
static void launch_forces_le(hipStream_t s, const DevParams& P, const DeviceState& S) {
  dim3 g(nchunks_max(P)), b(kWave);
  if (P.verlet) {
    // as the lists rebuilt every step: with colloids the chunks near a
    // sheared face take the Lees-Edwards instantiation and the others the
    // plain one; without, a sheared box runs the Lees-Edwards one alone
    if constexpr (!MIXED) {  // mixed precision keeps per-step lists in sheared boxes
      if (P.le_g >= 0) {
        hipLaunchKernelGGL((k_forces<DIM, FT, MIXED, COLL, true, false, true>), g, b, 0, s, P, S);
        hipLaunchKernelGGL((k_forces<DIM, FT, MIXED, COLL, true, true, true>), rescan_grid(P), b,
                           0, s, P, S);
      }
    }
    if (P.le_g < 0 || COLL) {
      hipLaunchKernelGGL((k_forces<DIM, FT, MIXED, COLL, false, false, true>), g, b, 0, s, P, S);
      hipLaunchKernelGGL((k_forces<DIM, FT, MIXED, COLL, false, true, true>), rescan_grid(P), b,
                         0, s, P, S);
    }
    return;
  }
  if (P.le_g >= 0) {
    hipLaunchKernelGGL((k_forces<DIM, FT, MIXED, COLL, true, false>), g, b, 0, s, P, S);
    hipLaunchKernelGGL((k_forces<DIM, FT, MIXED, COLL, true, true>), g, b, 0, s, P, S);
  }
  if (P.le_g < 0 || COLL) {
    hipLaunchKernelGGL((k_forces<DIM, FT, MIXED, COLL, false, false>), g, b, 0, s, P, S);
    hipLaunchKernelGGL((k_forces<DIM, FT, MIXED, COLL, false, true>), g, b, 0, s, P, S);
  }
}

// Body forces that depend on the position (particles_apply_body_force.F90:
// 196-353), added to the accelerations of fluid particles after the force
// pass: type 2, +g_x below half of max_phys along the last dimension and
// -g_x above it (nothing on the line itself); type 3, g_x sin(k y) with
// k = 2 pi / (max_phys(2) - min_phys(2)).  Kept out of the pair kernels,
// whose registers the sine would otherwise take.
__global__ void k_body_force(DevParams P, DeviceState S) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  if (P.ncoll && S.sid[i] != 0) return;  // :246, :252, :330
  const int64_t n = P.n;
  double f0 = S.f[i];
  if (P.body_force == 2) {
    const double xl = S.x[(int64_t)(P.dim - 1) * n + i];
    if (xl < P.bf_split)
      f0 = f0 + P.bf[0];
    else if (xl > P.bf_split)
      f0 = f0 - P.bf[0];
  } else {
    f0 = f0 + P.bf[0] * sin(P.bf_k * S.x[n + i]);
  }
  S.f[i] = f0;
}

template <int DIM>
static void launch_forces_dim(hipStream_t s, const DevParams& P, const DeviceState& S,
                              bool mixed) {
  // Oldroyd-B: 5 Hu-Adams, 6 Espanol, 7 angular (rhs_force_ff_non_Newtonian.F90)
  const int ft = P.newtonian ? P.force_type : (P.force_type == 3 ? 5 : (P.force_type == 2 ? 6 : 7));
  if (mixed) {
    if (ft == 3)
      launch_forces_le<DIM, 3, true, false>(s, P, S);
    else if (ft == 4)
      launch_forces_le<DIM, 4, true, false>(s, P, S);
    else if (ft == 1)
      launch_forces_le<DIM, 1, true, false>(s, P, S);
    else if (ft == 2)
      launch_forces_le<DIM, 2, true, false>(s, P, S);
    else if (ft == 6)
      launch_forces_le<DIM, 6, true, false>(s, P, S);
    else if (ft == 7)
      launch_forces_le<DIM, 7, true, false>(s, P, S);
    else
      launch_forces_le<DIM, 5, true, false>(s, P, S);
  } else if (P.ncoll) {
    // colloids: FP64 only (checked by the API)
    if (ft == 3)
      launch_forces_le<DIM, 3, false, true>(s, P, S);
    else if (ft == 4)
      launch_forces_le<DIM, 4, false, true>(s, P, S);
    else if (ft == 1)
      launch_forces_le<DIM, 1, false, true>(s, P, S);
    else if (ft == 2)
      launch_forces_le<DIM, 2, false, true>(s, P, S);
    else
      launch_forces_le<DIM, 5, false, true>(s, P, S);
  } else {
    if (ft == 3)
      launch_forces_le<DIM, 3, false, false>(s, P, S);
    else if (ft == 4)
      launch_forces_le<DIM, 4, false, false>(s, P, S);
    else if (ft == 1)
      launch_forces_le<DIM, 1, false, false>(s, P, S);  // Morris (FP64 only)
    else if (ft == 2)
      launch_forces_le<DIM, 2, false, false>(s, P, S);  // Espanol (FP64 only)
    else if (ft == 6)
      launch_forces_le<DIM, 6, false, false>(s, P, S);  // Oldroyd-B Espanol (FP64 only)
    else if (ft == 7)
      launch_forces_le<DIM, 7, false, false>(s, P, S);  // Oldroyd-B angular (FP64 only)
    else
      launch_forces_le<DIM, 5, false, false>(s, P, S);
  }
}

void launch_forces(hipStream_t s, const DevParams& P, const DeviceState& S, bool mixed) {
  if (P.dim == 3)
    launch_forces_dim<3>(s, P, S, mixed);
  else
    launch_forces_dim<2>(s, P, S, mixed);
  if (P.body_force >= 2)
    hipLaunchKernelGGL(k_body_force, dim3((P.n + 255) / 256), dim3(256), 0, s, P, S);
}

void launch_signature(hipStream_t s, const DevParams& P, const DeviceState& S, bool mixed) {
  (void)mixed;  // both modes share the decision
  dim3 g(nchunks_max(P)), b(kWave);
  const bool le = P.le_g >= 0;
  if (P.dim == 3) {
    if (P.verlet && le)
      hipLaunchKernelGGL((k_signature<3, true, true>), g, b, 0, s, P, S);
    else if (P.verlet)
      hipLaunchKernelGGL((k_signature<3, false, true>), g, b, 0, s, P, S);
    else if (le)
      hipLaunchKernelGGL((k_signature<3, true, false>), g, b, 0, s, P, S);
    else
      hipLaunchKernelGGL((k_signature<3, false, false>), g, b, 0, s, P, S);
  } else {
    if (P.verlet && le)
      hipLaunchKernelGGL((k_signature<2, true, true>), g, b, 0, s, P, S);
    else if (P.verlet)
      hipLaunchKernelGGL((k_signature<2, false, true>), g, b, 0, s, P, S);
    else if (le)
      hipLaunchKernelGGL((k_signature<2, true, false>), g, b, 0, s, P, S);
    else
      hipLaunchKernelGGL((k_signature<2, false, false>), g, b, 0, s, P, S);
  }
}

}  // namespace sdpd
