// sdpd_noalloc.hip -- density, equation of state and pair forces of
// particles held in device memory, without any allocation: the working
// buffers of the cell sort, the density pass and the force pass (sorted
// copies, cell tables, pair records, neighbour lists) are carved from one
// static device array.  For callers that cannot allocate device memory
// (code under HIP graph capture, serving loops); the context API allocates
// its buffers at sdpd_upload instead.
//
// It runs the kernels of sdpd_build_cells, sdpd_density and sdpd_forces on
// a single periodic box of Newtonian fluid, FP64 or mixed precision, as a
// context's first evaluation runs them (the sort, the build of the Verlet
// lists with the density, the list walk of the forces), with the context
// path's launch shapes, so its results equal that path's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <utility>

#include "../../include/sdpd.h"
#include "sdpd_host.h"
#include "sdpd_kernels.h"

namespace {

// The static array: every call on a device uses that device's copy and
// leaves its contents behind; calls that overlap on one device race.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

using sdpd::DeviceState;
using sdpd::DevParams;
using sdpd::RepGeom;

// Device address of the array on each device, looked up once per device
// with that device current (the array's name is only a host-side stand-in).
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxDevices] = {};

int scratch_address(unsigned char** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return SDPD_ERR_DEVICE;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      std::fprintf(stderr, "sdpd_evaluate_noalloc : hipGetSymbolAddress: %s\n",
                   hipGetErrorString(e));
      return SDPD_ERR_DEVICE;
    }
    g_scratch_dev[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_dev[dev];
  return SDPD_OK;
}

// Bump carving of 256-byte aligned pieces; with base == nullptr it only
// counts the bytes.
struct Carver {
  unsigned char* base;
  size_t used = 0;
  template <typename T>
  T* take(size_t count) {
    const size_t bytes = ((count ? count : 1) * sizeof(T) + 255) & ~(size_t)255;
    T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
    used += bytes;
    return p;
  }
};

// The buffers the three passes touch for one Newtonian box (the context
// path's alloc_state without the Oldroyd-B, colloid, replica-statistics and
// checking arrays).  Returns the bytes used.
size_t carve_state(unsigned char* base, const DevParams& P, DeviceState& S) {
  Carver c{base};
  const size_t dim = P.dim, N = (size_t)P.n, nrows = (size_t)P.nrows;
  S.x = c.take<double>(dim * N);
  S.v = c.take<double>(dim * N);
  S.x_alt = c.take<double>(dim * N);
  S.v_alt = c.take<double>(dim * N);
  S.m = c.take<double>(N);
  S.m_alt = c.take<double>(N);
  S.pid = c.take<int32_t>(N);
  S.pid_alt = c.take<int32_t>(N);
  S.perm = c.take<int32_t>(N);
  S.perm_alt = c.take<int32_t>(N);
  S.rep = c.take<int32_t>(N);
  S.rep_alt = c.take<int32_t>(N);
  S.f = c.take<double>(dim * N);
  S.rho = c.take<double>(N);
  S.p = c.take<double>(N);
  S.an = c.take<double>(N);
  S.bn = c.take<double>(N);
  S.key = c.take<int32_t>(N);
  S.slot = c.take<int32_t>(N);
  S.count = c.take<int32_t>((size_t)P.ncell);
  S.start = c.take<int32_t>((size_t)P.ncell + 1);
  S.unsorted = c.take<int32_t>(N);
  S.key_sorted = c.take<int32_t>(N);
  S.xa = c.take<float4>(N);
  if (P.mixed) {
    S.vr = c.take<float4>(N);
    S.qr = c.take<float4>(N);
  } else {
    S.rec64 = c.take<double>(8 * N);
  }
  S.geo = c.take<RepGeom>(1);
  S.row_cell = c.take<int32_t>(nrows);
  S.row_len = c.take<int32_t>(nrows);
  S.row_rep = c.take<int32_t>(nrows);
  S.step_dev = c.take<int64_t>(1);
  S.time_dev = c.take<double>(4);
  const size_t nchunk = (N + 63) / 64 + nrows;
  const size_t ntile = (std::max((size_t)P.ncell, nrows) + 4095) / 4096 + 1;
  S.row_nchunk = c.take<int32_t>(nrows);
  S.row_chunk_off = c.take<int32_t>(nrows + 1);
  S.chunk_row = c.take<int32_t>(nchunk);
  S.scan_tmp = c.take<int32_t>(ntile);
  S.nbr_count = c.take<int32_t>(N);
  S.nbr_in = c.take<int32_t>(N);
  S.recheck = c.take<int32_t>(N);
  S.nlist = c.take<uint32_t>((size_t)P.list_cap * 64 * ((nchunk + 7) / 8 * 8));
  S.gate = c.take<int32_t>(4);
  S.disp = c.take<unsigned long long>(1);
  if (P.verlet) S.x_ref = c.take<double>(dim * N);
  return c.used;
}

// Geometry tables of the single box: its RepGeom and, per row of the sort
// grid, first cell, length and replica (0).
__global__ void k_na_geometry(RepGeom g, int32_t nrows, RepGeom* __restrict__ geo,
                              int32_t* __restrict__ row_cell, int32_t* __restrict__ row_len,
                              int32_t* __restrict__ row_rep) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r == 0) *geo = g;
  if (r >= nrows) return;
  row_cell[r] = r * g.nm[0];
  row_len[r] = g.nm[0];
  row_rep[r] = 0;
}

// The caller's column-major x(dim, n), v(dim, n), m(n), id(2, n) into the
// structure-of-arrays state in upload order.
__global__ void k_na_ingest(int32_t n, int32_t dim, const double* __restrict__ x,
                            const double* __restrict__ v, const double* __restrict__ m,
                            const int32_t* __restrict__ id, DeviceState S) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int d = 0; d < dim; ++d) {
    S.x[(int64_t)d * n + i] = x[(int64_t)dim * i + d];
    S.v[(int64_t)d * n + i] = v[(int64_t)dim * i + d];
  }
  S.m[i] = m[i];
  S.pid[i] = id[2 * (int64_t)i];
  S.perm[i] = i;
  S.rep[i] = 0;
}

// Results back in upload order: rho(n), p(n), f(dim, n) column-major.
__global__ void k_na_egress(int32_t n, int32_t dim, DeviceState S, double* __restrict__ rho,
                            double* __restrict__ p, double* __restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int u = S.perm[i];
  if (rho) rho[u] = S.rho[i];
  if (p) p[u] = S.p[i];
  for (int d = 0; d < dim; ++d) f[(int64_t)dim * u + d] = S.f[(int64_t)d * n + i];
}

int reject(int code, const char* msg) {
  std::fprintf(stderr, "sdpd_evaluate_noalloc : %s\n", msg);
  return code;
}

}  // namespace

extern "C" int sdpd_evaluate_noalloc(const sdpd_params* params, int64_t step, const double* x,
                                     const double* v, const double* m, const int32_t* id,
                                     int32_t n, double* rho, double* p, double* f, void* stream) {
  if (!params || !x || !v || !m || !id || !f) return reject(SDPD_ERR_ARG, "null argument");
  if (n < 1 || n > SDPD_NOALLOC_MAX_PARTICLES)
    return reject(SDPD_ERR_ARG, "n must be 1 .. SDPD_NOALLOC_MAX_PARTICLES (65536)");
  DevParams P{};
  const char* err = "";
  int rc = sdpd::derive_params(*params, P, err);
  if (rc != SDPD_OK) return reject(rc, err);
  if (!P.newtonian) return reject(SDPD_ERR_ARG, "Newtonian flow only (no conformation input)");

  // one box: the sort grid of derive_params
  RepGeom g{};
  double mmax = 0.0;
  for (int d = 0; d < 3; ++d) {
    g.nm[d] = P.nm[d];
    g.lo[d] = P.min_phys[d];
    g.L[d] = P.length[d];
    g.rd[d] = P.rd[d];
    g.cs[d] = P.cs[d];
    g.csf[d] = (float)P.cs[d];
    if (d < P.dim) mmax = std::max(mmax, P.length[d] + 5.0 * P.cs[d]);
  }
  P.nrep = 1;
  P.nrows = P.nm[1] * P.nm[2];
  // the lists a context builds for this box: Verlet candidates within the
  // largest skin of the grid (or in-cut-off pairs when it is too thin)
  const double skin = sdpd::grid_skin(P, &g, 1);
  if (sdpd::skin_usable(P, skin)) {
    P.verlet = 1;
    P.skin = skin;
    P.half_skin2 = 0.25 * skin * skin;
  }
  sdpd::set_cut_band(P, mmax);
  P.n = n;
  P.step = step;
  double volume = 1.0;
  for (int d = 0; d < P.dim; ++d) volume *= P.length[d];

  // the automatic list capacity, smaller if the array cannot hold it (any
  // capacity gives the same results: chunks whose lists overflow are
  // scanned again by the force pass)
  DeviceState S{};
  P.list_cap = sdpd::auto_list_capacity(P, volume, n);
  while (carve_state(nullptr, P, S) > kScratchBytes && P.list_cap > 8)
    P.list_cap = std::max(8, (P.list_cap / 2 + 7) / 8 * 8);
  if (carve_state(nullptr, P, S) > kScratchBytes)
    return reject(SDPD_ERR_ARG,
                  "the box needs more than the 32 MiB scratch array (too many particles or cells)");

  unsigned char* base = nullptr;
  if ((rc = scratch_address(&base)) != SDPD_OK) return reject(rc, "no scratch array on this device");
  carve_state(base, P, S);

  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_na_geometry, dim3((P.nrows + 255) / 256), dim3(256), 0, s, g, P.nrows,
                     const_cast<RepGeom*>(S.geo), S.row_cell, S.row_len, S.row_rep);
  // noise step; the times only serve sheared faces, which this entry lacks
  sdpd::launch_set_step(s, S, step, 0.0, 0.0, 0.0);
  hipLaunchKernelGGL(k_na_ingest, dim3(nb), dim3(256), 0, s, n, P.dim, x, v, m, id, S);
  // technique_build_list: the sort writes the *_alt copies, which are
  // copied back (Verlet lists, whose build the gate then asks for) or swap
  sdpd::launch_cell_keys(s, P, S);
  sdpd::launch_cell_sort_after_keys(s, P, S, nullptr, P.verlet != 0);
  if (P.verlet) {
    sdpd::launch_force_build(s, S);
  } else {
    std::swap(S.x, S.x_alt);
    std::swap(S.v, S.v_alt);
    std::swap(S.m, S.m_alt);
    std::swap(S.pid, S.pid_alt);
    std::swap(S.perm, S.perm_alt);
    std::swap(S.rep, S.rep_alt);
  }
  sdpd::launch_density(s, P, S, P.mixed != 0);
  sdpd::launch_forces(s, P, S, P.mixed != 0);
  hipLaunchKernelGGL(k_na_egress, dim3(nb), dim3(256), 0, s, n, P.dim, S, rho, p, f);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    std::fprintf(stderr, "sdpd_evaluate_noalloc : launch: %s\n", hipGetErrorString(e));
    return SDPD_ERR_DEVICE;
  }
  return SDPD_OK;
}
