// sdpd_api.hip -- the C-ABI declared in include/sdpd.h: context, device
// buffers, host <-> device copies and the velocity-Verlet step driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sdpd.h"
#include "sdpd_host.h"
#include "sdpd_kernels.h"

using sdpd::DevParams;
using sdpd::DeviceState;

struct sdpd_ctx {
  sdpd_params prm{};
  DevParams P{};
  DeviceState S{};
  hipStream_t stream = nullptr;
  int device = 0;
  int n = 0;
  int stats_nb = 1;
  int64_t step = 0;
  bool cells_valid = false;  // particles sorted for the current positions
  bool rho_valid = false;    // rho / p current for the current order
  bool f_valid = false;      // f (and act) current for the current order
  bool vr_stale = false;     // mixed: FP32 velocity records older than v
  // phase timing: pairs of events recorded around each phase's launches
  bool timing = false;
  struct Mark {
    int phase;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> event_pool;
  std::string err;
  // boxes: one (params' min_phys / max_phys) or a batch of replicas
  std::vector<double> rep_lo, rep_hi;  // (3, nrep) when batched
  std::vector<sdpd::RepGeom> geo;
  std::vector<int32_t> row_cell, row_len, row_rep;
  bool sorted_once = false;
  // captured two-step graphs of sdpd_step, one per parity of the sort's
  // ping-pong buffers (kernel arguments hold the buffer pointers)
  bool use_graphs = true;
  int list_cap_req = 0;      // stored neighbour entries per particle (0: automatic)
  double skin_req = -1.0;    // Verlet skin asked for (< 0: automatic, 0: lists every step)
  double mmax = 0.0;         // largest box extent plus five cells (the FP32 band)
  int parity = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  // colloids: parameters and initial state until sdpd_upload puts them on
  // the device; boundary particles of all colloids; whether the initial
  // evaluation has started the colloids' acceleration histories
  sdpd_colloid_params cprm{};
  std::vector<double> c_radius, c_m, c_mmi, c_x, c_v, c_w;
  int nmembers = 0;
  bool coll_primed = false;
};

namespace {

constexpr double kPi = 3.141592653589793238;
// mcf_machine_zero as the reference sets it at start-up:
// 10^(INT(LOG10(EPSILON(1d0))) + 1) = 1e-14 (technique_new.F90:361-363)
constexpr double kMachineZero = 1e-14;

int fail(sdpd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIP_TRY(ctx, call)                                                         \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail((ctx), SDPD_ERR_DEVICE,                                          \
                  std::string(#call) + ": " + hipGetErrorString(e_));              \
  } while (0)

// time_current after `step` marching steps, accumulated by repeated
// addition of dt as marching_marching.F90:1393 does, the time at which
// that step evaluates the Lees-Edwards shear length (time_current + dt,
// marching_integrate_VV.F90:784-786; the initial evaluation uses
// time_current, marching_marching.F90:437), and that time of the step
// before (the shear length the step's colloid block still sees)
void marching_time(int64_t step, double dt, double* t, double* t_le, double* t_prev = nullptr) {
  double tt = 0.0, prev = 0.0;
  for (int64_t s = 0; s < step; ++s) {
    prev = tt;
    tt = tt + dt;
  }
  *t = tt;
  *t_le = step > 0 ? tt + dt : tt;
  if (t_prev) *t_prev = step > 1 ? prev + dt : (step == 1 ? prev : *t_le);
}

int check_launch(sdpd_ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(c, SDPD_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  return SDPD_OK;
}

}  // namespace

namespace sdpd {

// Derived constants, with the reference's start-up rules; err names the
// failed check.
int derive_params(const sdpd_params& p, DevParams& P, const char*& err) {
  std::memset(&P, 0, sizeof(P));
  if (p.num_dim != 2 && p.num_dim != 3) return err = "num_dim must be 2 or 3", SDPD_ERR_ARG;
  if (!(p.cut_off > 0)) return err = "cut_off must be positive", SDPD_ERR_ARG;
  if (p.kernel_type != 1 && p.kernel_type != 2)
    return err = "kernel_type must be 1 (quintic spline) or 2 (Lucy)", SDPD_ERR_ARG;
  if (p.rhs_density_type != 1 && p.rhs_density_type != 2)
    return err = "rhs_density_type must be 1 or 2", SDPD_ERR_ARG;
  if (p.rhs_force_type < 1 || p.rhs_force_type > 4)
    return err = "rhs_force_type must be 1 (Morris), 2 (Espanol), 3 (Hu-Adams) or 4 (angular)",
           SDPD_ERR_ARG;
  // rhs_force_ff_Newtonian_Morris.F90 has no thermal noise (ctrl.mcf:
  // Brownian needs rhs_force_type 2, 3 or 4)
  if (p.brownian && p.rhs_force_type == 1)
    return err = "Brownian motion requires rhs_force_type 2, 3 or 4", SDPD_ERR_ARG;
  if (p.state_equation_type != 1 && p.state_equation_type != 2)
    return err = "stateEquation_type must be 1 or 2", SDPD_ERR_ARG;
  if (p.state_equation_type == 1 && p.c == 0.0)
    return err = "stateEquation_init_Morris: sound speed should not be zero", SDPD_ERR_ARG;
  if (p.state_equation_type == 2 && (p.rho_ref == 0.0 || p.gamma == 0.0))
    return err = "stateEquation_init_Batchelor: neither rho_ref nor gamma can be zero",
           SDPD_ERR_ARG;
  // rhs_force_ff_non_Newtonian.F90: Espanol (2), Hu-Adams (3), angular (4)
  if (!p.newtonian && p.rhs_force_type == 1)
    return err = "non-Newtonian flow needs rhs_force_type 2, 3 or 4", SDPD_ERR_ARG;
  if (!p.newtonian && p.rhs_force_type == 4 && p.brownian)
    return err = "rhs_force_ff_non_Newtonian_HuAdams_angular: Brownian force is not available",
           SDPD_ERR_ARG;
  if (!p.newtonian && std::fabs(p.tau) < kMachineZero)
    return err = "particles_compute_act: tau should not be zero", SDPD_ERR_ARG;
  if (p.brownian && p.rhs_density_type != 2)
    return err = "Brownian motion requires rhs_density_type = 2", SDPD_ERR_ARG;
  if (p.eta < 0) return err = "dissipative force should be non-positive (eta >= 0)", SDPD_ERR_ARG;
  if (!(p.dt > 0)) return err = "dt must be positive", SDPD_ERR_ARG;
  if (p.precision != SDPD_PRECISION_FP64 && p.precision != SDPD_PRECISION_MIXED)
    return err = "precision must be 0 (FP64) or 1 (mixed)", SDPD_ERR_ARG;

  P.dim = p.num_dim;
  P.symmetry = p.symmetry ? 1 : 0;
  P.kernel_type = p.kernel_type;
  P.density_type = p.rhs_density_type;
  P.force_type = p.rhs_force_type;
  P.eos_type = p.state_equation_type;
  P.newtonian = p.newtonian ? 1 : 0;
  const double kt = p.brownian ? p.kt : 0.0;  // physics_adjust_parameters.F90:415-417
  P.noise = (p.brownian && kt > kMachineZero) ? 1 : 0;
  // particles_apply_body_force.F90:96-108: 1 constant, 2 +-x on the two
  // halves of the last dimension, 3 sine along y; type 4 (the stochastic
  // model, a Langevin force from the reference's sequential random stream,
  // with its own tau_sm / k_sm keys) is not provided
  if (p.body_force_type < 0 || p.body_force_type > 3)
    return err = "body_force_type must be 0..3 (4, the stochastic model, is not supported)",
           SDPD_ERR_ARG;
  P.body_force = 0;
  if (p.body_force_type == 1)
    for (int d = 0; d < p.num_dim; ++d)
      if (std::fabs(p.body_force[d]) > kMachineZero) P.body_force = 1;
  // types 2 and 3 act along x only, when that component is not zero
  if ((p.body_force_type == 2 || p.body_force_type == 3) &&
      std::fabs(p.body_force[0]) > kMachineZero)
    P.body_force = p.body_force_type;
  P.bf_split = p.max_phys[p.num_dim - 1] * 0.5;                // :245-252
  P.bf_k = 2.0 * kPi / (p.max_phys[1] - p.min_phys[1]);        // :323-324
  const double cell = p.cut_off * 1.0001;  // mcf.F90:617
  int64_t ncell = 1, ncell_ref = 1;
  double lmax = 0.0;
  for (int d = 0; d < 3; ++d) {
    if (d < p.num_dim) {
      double L = p.max_phys[d] - p.min_phys[d];
      if (!(L > 0)) return err = "max_phys must exceed min_phys", SDPD_ERR_ARG;
      // PPM's grid (ppm_neighlist_clist.F90:370-376) and the sort grid with
      // half its cell edge
      int nm = (int)(L / cell);
      if (nm < 1) nm = 1;
      if (nm < 3)
        return err = "the box must span at least 3 cells of 1.0001*cut_off per dimension",
               SDPD_ERR_ARG;
      P.nm_ref[d] = nm;
      P.nm[d] = 2 * nm;
      P.min_phys[d] = p.min_phys[d];
      P.length[d] = L;
      P.rd_ref[d] = (double)nm / L;
      P.rd[d] = (double)(2 * nm) / L;
      P.cs[d] = L / (double)(2 * nm);
      lmax = std::max(lmax, L);
    } else {
      P.nm[d] = P.nm_ref[d] = 1;
      P.min_phys[d] = 0.0;
      P.length[d] = 1.0;
      P.rd[d] = P.rd_ref[d] = 1.0;
      P.cs[d] = 1.0;
    }
    ncell *= P.nm[d];
    ncell_ref *= P.nm_ref[d];
  }
  if (ncell > (int64_t)1 << 30) return err = "too many cells", SDPD_ERR_ARG;
  P.ncell = (int32_t)ncell;
  P.ncell_ref = (int32_t)ncell_ref;
  for (int d = 0; d < 3; ++d) {
    P.lengthf[d] = (float)P.length[d];
    P.csf[d] = (float)P.cs[d];
  }
  (void)lmax;
  P.cut_off = p.cut_off;
  P.cut2 = p.cut_off * p.cut_off;
  // kernel_new.F90:47-93
  if (p.kernel_type == 1) {
    P.h = p.cut_off / 3.0;
    P.coef = p.num_dim == 2 ? 7.0 / (478.0 * kPi * P.h * P.h)
                            : 3.0 / (359.0 * kPi * P.h * P.h * P.h);
  } else {
    P.h = p.cut_off;
    P.coef = p.num_dim == 2 ? 5.0 / (kPi * P.h * P.h)
                            : 105.0 / (16.0 * kPi * P.h * P.h * P.h);
  }
  P.coef_grad = -P.coef / P.h;
  P.inv_h = 1.0 / P.h;
  P.w0 = p.kernel_type == 1 ? P.coef * (243.0 + -192.0 + 15.0) : P.coef;
  P.c2 = p.c * p.c;
  P.rho0 = p.rho;
  P.rho_ref = p.rho_ref;
  P.gamma = p.gamma;
  P.p0 = p.state_equation_type == 2
             ? p.c * p.c * p.rho * std::pow(p.rho_ref / p.rho, p.gamma) / p.gamma
             : 0.0;  // stateEquation_new.F90:102
  P.eta = p.eta;
  P.eta_coef = p.eta_coef;
  P.kt = kt;
  P.dt = p.dt;
  P.sqrt_dt = std::sqrt(p.dt);
  P.inv_sqrt_dt = 1.0 / P.sqrt_dt;
  P.sqrt_2kt_eta = std::sqrt(2.0 * P.kt * P.eta);
  P.tau = p.tau;
  P.G = p.n_p * p.kt_p;  // particles_compute_pressure_tensor.F90:94
  for (int d = 0; d < 3; ++d) P.bf[d] = p.body_force[d];
  P.seed = p.seed;
  P.step = 0;
  P.mixed = p.precision == SDPD_PRECISION_MIXED ? 1 : 0;
  P.density_shape = 0;
  P.xcd_order = 1;
  P.le_g = -1;
  P.inv_hf = (float)(1.0 / P.h);
  P.coeff = (float)P.coef;
  P.coef_gradf = (float)P.coef_grad;
  P.etaf = (float)P.eta;
  P.eta_coeff = (float)P.eta_coef;
  P.ktf = (float)P.kt;
  P.inv_sqrt_dtf = (float)P.inv_sqrt_dt;
  P.sqrt_2kt_etaf = (float)std::sqrt(2.0 * P.kt * P.eta);
  return SDPD_OK;
}

// Band of the FP32 candidate test.  The scan compares positions relative
// to the corner of a chunk's first candidate cell in x (at most L + 5 cs
// in size, mmax over the boxes) and to a neighbouring cell corner in y, z;
// with six FP32 roundings per difference each difference is off by at most
// dd = 2^-21 mmax, and r^2 by 2 sqrt(3) r dd + 4 2^-24 r^2.  The band is
// twice that bound relative to rc^2, and at least 2e-5.
void set_cut_band(DevParams& P, double mmax) {
  const double dd = std::ldexp(mmax, -21);
  double rel = 2.0 * (2.0 * std::sqrt(3.0) * dd / P.cut_off + 4.0 * std::ldexp(1.0, -24));
  rel = std::max(rel, 2e-5);
  const double c2 = P.cut_off * P.cut_off;
  P.cut2_lo_f32 = (float)(c2 * std::max(0.0, 1.0 - rel)) * (1.0f - 1e-6f);
  P.cut2_hi_f32 = (float)(c2 * (1.0 + rel)) * (1.0f + 1e-6f);
  // the Verlet build scan keeps every candidate whose exact distance may be
  // below rc + skin
  const double rs = P.cut_off + P.skin;
  P.skin2_hi_f32 = (float)(rs * rs * (1.0 + rel)) * (1.0f + 1e-6f);
}

// The largest Verlet skin of boxes with these sort grids: the candidate
// scan reaches two cells each way, so a build finds every pair within
// 2 cs - rc beyond the cut-off; less a margin for the rounding of cell keys.
double grid_skin(const DevParams& P, const RepGeom* geo, int nrep) {
  double smax = 1e300, lmax = 0.0;
  for (int q = 0; q < nrep; ++q)
    for (int d = 0; d < P.dim; ++d) {
      smax = std::min(smax, 2.0 * geo[q].cs[d] - P.cut_off);
      lmax = std::max(lmax, geo[q].L[d]);
    }
  return smax - 1e-9 * lmax;
}

// Whether a skin serves Verlet lists: 1e-6 rc or more.
bool skin_usable(const DevParams& P, double skin) { return skin > 1e-6 * P.cut_off; }

// Capacity of the stored neighbour lists: 1.6 times the neighbours a
// uniform fluid of n particles in `volume` has inside the cut-off (plus the
// Verlet skin), plus 32, rounded up to 16 (at most 4096).  A particle with more is handled by
// scanning again in the force pass.
int auto_list_capacity(const DevParams& P, double volume, int64_t n) {
  const double kPiL = 3.141592653589793238;
  const double rc = P.cut_off + P.skin;
  const double ball = P.dim == 3 ? 4.0 / 3.0 * kPiL * rc * rc * rc : kPiL * rc * rc;
  const double expect = (double)n / volume * ball;
  int cap = (int)std::ceil(1.6 * expect) + 32;
  cap = (cap + 15) / 16 * 16;
  return std::min(cap, 4096);
}

}  // namespace sdpd

namespace {

int derive(const sdpd_params& p, DevParams& P, std::string& err) {
  const char* e = "";
  const int rc = sdpd::derive_params(p, P, e);
  if (rc != SDPD_OK) err = e;
  return rc;
}

// Geometry tables of the context's boxes: the single box of the params,
// or the replicas of sdpd_set_replicas.  Sets P.nrep, P.ncell, P.nrows and
// the FP32 cut-off band, which is sized for the largest box.
int build_geometry(sdpd_ctx* c) {
  DevParams& P = c->P;
  const int dim = P.dim;
  const bool batched = !c->rep_lo.empty();
  const int nrep = batched ? (int)(c->rep_lo.size() / 3) : 1;
  const double cell = c->prm.cut_off * 1.0001;  // mcf.F90:617
  c->geo.assign(nrep, sdpd::RepGeom{});
  c->row_cell.clear();
  c->row_len.clear();
  c->row_rep.clear();
  int64_t cells = 0, rows = 0;
  double mmax = 0.0;
  for (int q = 0; q < nrep; ++q) {
    sdpd::RepGeom& g = c->geo[q];
    for (int d = 0; d < 3; ++d) {
      if (d < dim) {
        const double lo = batched ? c->rep_lo[3 * (size_t)q + d] : P.min_phys[d];
        const double L = batched ? c->rep_hi[3 * (size_t)q + d] - lo : P.length[d];
        if (!(L > 0)) return fail(c, SDPD_ERR_ARG, "replica box with non-positive edge");
        const int nm_ref = (int)(L / cell);
        if (nm_ref < 3)
          return fail(c, SDPD_ERR_ARG,
                      "every box must span at least 3 cells of 1.0001*cut_off per dimension");
        g.nm[d] = 2 * nm_ref;
        g.lo[d] = lo;
        g.L[d] = L;
        g.rd[d] = (double)g.nm[d] / L;
        g.cs[d] = L / (double)g.nm[d];
        mmax = std::max(mmax, L + 5.0 * g.cs[d]);
      } else {
        g.nm[d] = 1;
        g.lo[d] = 0.0;
        g.L[d] = 1.0;
        g.rd[d] = 1.0;
        g.cs[d] = 1.0;
      }
      g.csf[d] = (float)g.cs[d];
    }
    g.cell_base = (int32_t)cells;
    g.row_base = (int32_t)rows;
    const int nr = g.nm[1] * g.nm[2];
    for (int r = 0; r < nr; ++r) {
      c->row_cell.push_back((int32_t)(cells + (int64_t)r * g.nm[0]));
      c->row_len.push_back(g.nm[0]);
      c->row_rep.push_back(q);
    }
    cells += (int64_t)g.nm[0] * g.nm[1] * g.nm[2];
    rows += nr;
    if (cells > ((int64_t)1 << 30)) return fail(c, SDPD_ERR_ARG, "too many cells");
  }
  P.nrep = nrep;
  P.ncell = (int32_t)cells;
  P.nrows = (int32_t)rows;
  c->mmax = mmax;
  sdpd::set_cut_band(P, mmax);
  return SDPD_OK;
}

template <typename T>
int dalloc(sdpd_ctx* c, T** p, size_t count) {
  if (count == 0) count = 1;
  HIP_TRY(c, hipMalloc((void**)p, count * sizeof(T)));
  return SDPD_OK;
}

void drop_graphs(sdpd_ctx* c) {
  for (hipGraphExec_t& g : c->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

void free_state(DeviceState& S) {
  void* ptrs[] = {S.x,        S.v,        S.m,         S.pid,     S.perm,     S.x_alt,
                  S.v_alt,    S.m_alt,    S.pid_alt,   S.perm_alt, S.f,       S.rho,
                  S.p,        S.an,       S.bn,        S.ct,      S.ct_alt,   S.act,
                  S.vgt,      S.key,      S.slot,      S.count,   S.start,    S.unsorted,
                  S.key_sorted, S.xa,     S.pt64,      S.vr,        S.qr,      S.partials, S.stats_out,
                  S.nbr_count, S.recheck, S.nbr_hash, S.rkey,   S.rslot,    S.rcount,
                  S.rstart,   S.runsorted, S.rlpdx,  S.nlist,   S.nbr_sig_count,
                  S.row_nchunk, S.row_chunk_off, S.chunk_row, S.scan_tmp, S.rec64,
                  (void*)S.geo, S.rep,      S.rep_alt,  S.row_cell, S.row_len,  S.row_rep,
                  S.rep_stats, S.qr2,     S.step_dev, S.time_dev, S.sid,
                  S.sid_alt,  S.slot_of,  (void*)S.coll_hot, (void*)S.coll, S.coll_start,
                  S.coll_members, S.x_ref, S.gate, (void*)S.disp, S.nbr_in};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  std::memset(&S, 0, sizeof(S));
}

int alloc_state(sdpd_ctx* c, int n) {
  DeviceState& S = c->S;
  const DevParams& P = c->P;
  const size_t dim = P.dim, d2 = dim * dim, N = (size_t)n;
  int rc;
#define ALLOC(ptr, cnt) \
  if ((rc = dalloc(c, &S.ptr, (cnt))) != SDPD_OK) return rc;
  ALLOC(x, dim * N);
  ALLOC(v, dim * N);
  ALLOC(m, N);
  ALLOC(pid, N);
  ALLOC(perm, N);
  ALLOC(x_alt, dim * N);
  ALLOC(v_alt, dim * N);
  ALLOC(m_alt, N);
  ALLOC(pid_alt, N);
  ALLOC(perm_alt, N);
  ALLOC(f, dim * N);
  ALLOC(rho, N);
  ALLOC(p, N);
  ALLOC(an, N);
  ALLOC(bn, N);
  if (!P.newtonian) {
    ALLOC(ct, d2 * N);
    ALLOC(ct_alt, d2 * N);
    ALLOC(act, d2 * N);
    ALLOC(vgt, d2 * N);
  }
  ALLOC(key, N);
  ALLOC(slot, N);
  ALLOC(count, (size_t)P.ncell);
  ALLOC(start, (size_t)P.ncell + 1);
  ALLOC(unsorted, N);
  ALLOC(key_sorted, N);
  ALLOC(xa, N);
  if (c->prm.precision == SDPD_PRECISION_MIXED) {
    ALLOC(vr, N);
    ALLOC(qr, N);
    if (!P.newtonian) ALLOC(qr2, N);
  } else {
    ALLOC(rec64, 8 * N);
    if (!P.newtonian) ALLOC(pt64, 6 * N);
  }
  {
    // replica tables (static for the context's lifetime)
    sdpd::RepGeom* geo = nullptr;
    if ((rc = dalloc(c, &geo, c->geo.size())) != SDPD_OK) return rc;
    S.geo = geo;
    HIP_TRY(c, hipMemcpy(geo, c->geo.data(), c->geo.size() * sizeof(sdpd::RepGeom),
                         hipMemcpyHostToDevice));
    ALLOC(rep, N);
    ALLOC(rep_alt, N);
    ALLOC(row_cell, c->row_cell.size());
    ALLOC(row_len, c->row_len.size());
    ALLOC(row_rep, c->row_rep.size());
    HIP_TRY(c, hipMemcpy(S.row_cell, c->row_cell.data(), c->row_cell.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(S.row_len, c->row_len.data(), c->row_len.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(S.row_rep, c->row_rep.data(), c->row_rep.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
    ALLOC(rep_stats, 4 * (size_t)P.nrep);
    ALLOC(step_dev, 1);
    ALLOC(time_dev, 4);
  }
  {
    const size_t nrows = (size_t)P.nrows;
    const size_t nchunk = (N + 63) / 64 + nrows;
    const size_t ntile = ((size_t)std::max(P.ncell, (int32_t)nrows) + 4095) / 4096 + 1;
    ALLOC(row_nchunk, nrows);
    ALLOC(row_chunk_off, nrows + 1);
    ALLOC(chunk_row, nchunk);
    ALLOC(scan_tmp, std::max(ntile, (size_t)((P.ncell_ref + 4095) / 4096 + 1)));
  }
  c->stats_nb = sdpd::stats_blocks(n);
  ALLOC(partials, (size_t)c->stats_nb * 4);
  ALLOC(stats_out, 8);
  ALLOC(nbr_count, N);
  ALLOC(nbr_in, N);
  ALLOC(nbr_sig_count, N);
  ALLOC(gate, 4);
  ALLOC(disp, 1);
  if (P.verlet) ALLOC(x_ref, dim * N);
  // neighbour lists: one block of list_cap x 64 entries per chunk of the
  // sort grid (at most N / 64 + rows chunks, rounded up to 8)
  ALLOC(nlist, (size_t)P.list_cap * 64 * (((N + 63) / 64 + (size_t)P.nrows + 7) / 8 * 8));
  ALLOC(recheck, N);
  ALLOC(nbr_hash, N);
  if (P.ncoll > 0) {
    ALLOC(sid, N);
    ALLOC(sid_alt, N);
    ALLOC(slot_of, N);
    ALLOC(coll_hot, (size_t)P.ncoll);
    ALLOC(coll, (size_t)P.ncoll);
    ALLOC(coll_start, (size_t)P.ncoll + 1);
    ALLOC(coll_members, N);
  }
#undef ALLOC
  return SDPD_OK;
}

void swap_sorted(sdpd_ctx* c) {
  DeviceState& S = c->S;
  c->parity ^= 1;
  std::swap(S.x, S.x_alt);
  std::swap(S.v, S.v_alt);
  std::swap(S.m, S.m_alt);
  std::swap(S.pid, S.pid_alt);
  std::swap(S.perm, S.perm_alt);
  std::swap(S.rep, S.rep_alt);
  std::swap(S.ct, S.ct_alt);
  std::swap(S.sid, S.sid_alt);
}

hipEvent_t take_event(sdpd_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// RAII bracket of one phase on the context stream (no-op unless timing)
struct PhaseMark {
  sdpd_ctx* c;
  int phase;
  hipEvent_t a = nullptr;
  PhaseMark(sdpd_ctx* ctx, int ph) : c(ctx), phase(ph) {
    if (c->timing && (a = take_event(c))) (void)hipEventRecord(a, c->stream);
  }
  ~PhaseMark() {
    if (!a) return;
    hipEvent_t b = take_event(c);
    if (!b) {
      c->event_pool.push_back(a);
      return;
    }
    (void)hipEventRecord(b, c->stream);
    c->marks.push_back({phase, a, b});
  }
};

void release_events(sdpd_ctx* c) {
  for (auto& m : c->marks) {
    c->event_pool.push_back(m.a);
    c->event_pool.push_back(m.b);
  }
  c->marks.clear();
}

int need_particles(sdpd_ctx* c) {
  if (!c) return SDPD_ERR_ARG;
  if (c->n <= 0) return fail(c, SDPD_ERR_STATE, "no particles: call sdpd_upload first");
  return SDPD_OK;
}

// K1 for the current positions (keys, histogram, scan, scatter, place).
// Lists rebuilt every step: the sort writes the ping-pong copies, which
// then swap.  Verlet lists: the sorted state is copied back (the buffers
// stay put), the positions of the sort become the lists' reference, and
// the next density pass builds the lists.
int run_sort(sdpd_ctx* c, bool keys_done) {
  c->P.step = c->step;
  {
    PhaseMark pm(c, 1);
    if (c->P.verlet) {
      sdpd::launch_cell_keys(c->stream, c->P, c->S);
      sdpd::launch_cell_sort_after_keys(c->stream, c->P, c->S, nullptr, true);
      sdpd::launch_force_build(c->stream, c->S);
    } else {
      if (!keys_done) sdpd::launch_cell_keys(c->stream, c->P, c->S);
      sdpd::launch_cell_sort_after_keys(c->stream, c->P, c->S);
    }
  }
  int rc = check_launch(c, "cell sort");
  if (rc) return rc;
  if (!c->P.verlet) swap_sorted(c);
  c->cells_valid = true;
  c->sorted_once = true;
  c->rho_valid = false;
  c->f_valid = false;
  return SDPD_OK;
}

// Verlet lists after a drift whose displacements have been tracked: the
// step's gate (a particle moved more than skin / 2 since the lists' build),
// then the sort, every kernel of which returns at once unless the gate is
// set.  Nothing here depends on the host, so captured steps replay it.
int run_list_update(sdpd_ctx* c) {
  c->P.step = c->step;
  {
    PhaseMark pm(c, 1);
    sdpd::launch_decide(c->stream, c->P, c->S);
    sdpd::launch_cell_keys(c->stream, c->P, c->S, c->S.gate);
    sdpd::launch_cell_sort_after_keys(c->stream, c->P, c->S, c->S.gate, true);
  }
  int rc = check_launch(c, "list update");
  if (rc) return rc;
  c->cells_valid = true;
  c->sorted_once = true;
  c->rho_valid = false;
  c->f_valid = false;
  return SDPD_OK;
}

int run_density(sdpd_ctx* c, bool scan_all = false) {
  c->P.step = c->step;
  {
    PhaseMark pm(c, 2);
    sdpd::launch_density(c->stream, c->P, c->S, c->P.mixed != 0, scan_all);
  }
  int rc = check_launch(c, "density");
  if (rc) return rc;
  c->rho_valid = true;
  c->vr_stale = false;
  return SDPD_OK;
}

int run_forces(sdpd_ctx* c) {
  c->P.step = c->step;
  {
    PhaseMark pm(c, 3);
    if (c->vr_stale) {
      sdpd::launch_refresh_records(c->stream, c->P, c->S);
      c->vr_stale = false;
    }
    sdpd::launch_forces(c->stream, c->P, c->S, c->P.mixed != 0);
  }
  int rc = check_launch(c, "forces");
  if (rc) return rc;
  c->f_valid = true;
  return SDPD_OK;
}

// After the initial evaluation: the colloids' drag and torque of the
// initial forces start their acceleration histories (marching_marching.F90:
// 1003-1080); the run's step_start is the current step.
int prime_colloids(sdpd_ctx* c) {
  if (!c->P.ncoll || c->coll_primed) return SDPD_OK;
  drop_graphs(c);  // coll_step0 is a kernel argument
  c->P.coll_step0 = c->step;
  sdpd::launch_colloid_collect(c->stream, c->P, c->S, 1);
  sdpd::launch_colloid_prologue(c->stream, c->P, c->S);
  int rc = check_launch(c, "colloid prologue");
  if (rc) return rc;
  c->coll_primed = true;
  return SDPD_OK;
}

// One velocity-Verlet step after the prologue: advance the marching step
// (host and device copies), step boundary (closing kick of the previous
// step if asked, opening kick, drift, ct half update, wrap, cell keys),
// sort, density + EOS, forces.
int step_body(sdpd_ctx* c, int closing) {
  c->step += 1;
  c->P.step = c->step;
  sdpd::launch_advance_step(c->stream, c->P, c->S);
  {
    PhaseMark pm(c, 0);
    // the colloid block moves only boundary particles, the drift only fluid
    // ones: the block may run first
    if (c->P.ncoll) sdpd::launch_colloid_block(c->stream, c->P, c->S, c->nmembers);
    sdpd::launch_boundary(c->stream, c->P, c->S, closing, 1);
  }
  int rc = check_launch(c, "step boundary");
  if (rc) return rc;
  c->cells_valid = false;
  if ((rc = c->P.verlet ? run_list_update(c) : run_sort(c, true))) return rc;
  if ((rc = run_density(c))) return rc;
  return run_forces(c);
}

// Records two closing-kick steps on the context stream into a graph.
// Nothing runs during capture: the host step counter is restored, and the
// two sorts leave the ping-pong parity where it was.
int capture_two_steps(sdpd_ctx* c, hipGraphExec_t* out) {
  const int64_t step0 = c->step;
  HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = step_body(c, 1);
  if (!rc) rc = step_body(c, 1);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &graph);
  c->step = step0;
  c->P.step = step0;
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(c, SDPD_ERR_DEVICE, std::string("step graph capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    *out = nullptr;
    return fail(c, SDPD_ERR_DEVICE, std::string("step graph instantiate: ") + hipGetErrorString(e));
  }
  return SDPD_OK;
}

// column-major host (dim, n) in upload order  <->  device SoA in cell order
int download_vec(sdpd_ctx* c, const double* dev, int comps, double* host) {
  const int n = c->n;
  std::vector<double> tmp((size_t)comps * n);
  std::vector<int32_t> perm(n);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemcpy(tmp.data(), dev, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(perm.data(), c->S.perm, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) {
    int32_t u = perm[i];
    if (u < 0 || u >= n) return fail(c, SDPD_ERR_STATE, "corrupt permutation");
    for (int q = 0; q < comps; ++q) host[(size_t)q + (size_t)comps * u] = tmp[(size_t)q * n + i];
  }
  return SDPD_OK;
}

}  // namespace

extern "C" {

int sdpd_create(const sdpd_params* params, sdpd_ctx** out) {
  if (!params || !out) return SDPD_ERR_ARG;
  *out = nullptr;
  sdpd_ctx* c = new sdpd_ctx();
  c->prm = *params;
  std::string err;
  int rc = derive(c->prm, c->P, err);
  if (rc != SDPD_OK) {
    std::fprintf(stderr, "sdpd_create : %s\n", err.c_str());
    delete c;
    return rc;
  }
  if ((rc = build_geometry(c)) != SDPD_OK) {
    std::fprintf(stderr, "sdpd_create : %s\n", c->err.c_str());
    delete c;
    return rc;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::fprintf(stderr, "sdpd_create : no HIP device available\n");
    delete c;
    return SDPD_ERR_DEVICE;
  }
  if (hipGetDevice(&c->device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    std::fprintf(stderr, "sdpd_create : stream creation failed\n");
    delete c;
    return SDPD_ERR_DEVICE;
  }
  *out = c;
  return SDPD_OK;
}

int sdpd_destroy(sdpd_ctx* c) {
  if (!c) return SDPD_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  release_events(c);
  for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
  c->event_pool.clear();
  drop_graphs(c);
  free_state(c->S);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return SDPD_OK;
}

const char* sdpd_last_error(const sdpd_ctx* c) { return c ? c->err.c_str() : "null context"; }

}  // extern "C"

namespace {

// Verlet lists or lists rebuilt every step.  The candidate scan reaches two
// cells of the sort grid each way, so a build finds every pair within
// 2 cs - rc of the cut-off: that is the largest skin (less a margin for the
// rounding of cell keys).  Mixed precision with sheared faces, a request
// of 0 or a skin thinner than 1e-6 rc rebuild the lists every step.
//
// The automatic skin trades the candidates past the cut-off, which every
// step's passes walk (about 3 skin / rc of the neighbours), against the
// rebuilds: it is what a particle at speed v_est covers in kSkinSteps
// steps, twice over (lists are rebuilt after skin / 2), between 1/20 of
// the grid's largest and that largest.  v_est is the fastest upload speed,
// or five thermal standard deviations sqrt(kT / m) of the lightest
// particle when those are faster.  The neighbour sets do not depend on
// the skin.
constexpr double kSkinSteps = 256.0;

double auto_skin(const DevParams& P, double smax, double v_est) {
  const double want = 2.0 * v_est * P.dt * kSkinSteps;
  return std::min(smax, std::max(0.05 * smax, want));
}

int choose_skin(sdpd_ctx* c, double v_est) {
  DevParams& P = c->P;
  const double smax = sdpd::grid_skin(P, c->geo.data(), (int)c->geo.size());
  if (c->skin_req > 0.0 && c->skin_req > smax)
    return fail(c, SDPD_ERR_ARG,
                "sdpd_set_skin: the skin exceeds what the sort grid allows (2 cell edges - cut_off = " +
                    std::to_string(smax) + ")");
  const double skin = c->skin_req < 0.0 ? auto_skin(P, smax, v_est) : c->skin_req;
  // (mixed precision in a sheared box keeps lists rebuilt every step)
  const bool on = !(P.le_g >= 0 && P.mixed) && sdpd::skin_usable(P, skin);
  P.verlet = on ? 1 : 0;
  P.skin = on ? skin : 0.0;
  P.half_skin2 = 0.25 * P.skin * P.skin;
  sdpd::set_cut_band(P, c->mmax);
  return SDPD_OK;
}

// Species ids, the boundary particles of each colloid (upload indices in
// upload order) and the colloids' initial state (histories zero, no
// accumulated rotation) onto the device.
int upload_colloids(sdpd_ctx* c, const int32_t* id, int32_t n) {
  const int ncoll = c->P.ncoll;
  {
    double rmax = 0.0;
    for (int k = 0; k < ncoll; ++k) rmax = std::max(rmax, c->c_radius[k]);
    int cells = 2;
    if (c->P.le_g >= 0) {
      const double reach = rmax + c->P.coll_dout + c->P.cut_off;
      cells = std::max(2, (int)std::ceil(reach / c->P.cs[c->P.le_g]) + 1);
    }
    c->P.coll_le_cells = cells;
  }
  std::vector<int32_t> sid(n), start(ncoll + 1, 0), slot(n);
  for (int32_t i = 0; i < n; ++i) {
    sid[i] = id[2 * (size_t)i + 1];
    slot[i] = i;
    if (sid[i] > 0) ++start[sid[i]];
  }
  for (int k = 0; k < ncoll; ++k) start[k + 1] += start[k];
  std::vector<int32_t> members(std::max(start[ncoll], 1)), fill(start.begin(), start.end() - 1);
  for (int32_t i = 0; i < n; ++i)
    if (sid[i] > 0) members[fill[sid[i] - 1]++] = i;
  c->nmembers = start[ncoll];
  std::vector<sdpd::CollState> cs(ncoll);
  std::vector<sdpd::CollHot> hot(ncoll);
  for (int k = 0; k < ncoll; ++k) {
    sdpd::CollState& a = cs[k];
    std::memset(&a, 0, sizeof(a));
    for (int d = 0; d < 3; ++d) {
      a.x[d] = c->c_x[3 * (size_t)k + d];
      a.v[0][d] = c->c_v[3 * (size_t)k + d];
      a.w[0][d] = c->c_w[3 * (size_t)k + d];
      a.mmi[d] = c->c_mmi[3 * (size_t)k + d];
      a.acc[4 * d] = 1.0;
    }
    a.radius = c->c_radius[k];
    a.m = c->c_m[k];
    sdpd::CollHot& h = hot[k];
    std::memset(&h, 0, sizeof(h));
    for (int d = 0; d < 3; ++d) {
      h.x[d] = a.x[d];
      h.v[d] = a.v[0][d];
      h.w[d] = a.w[0][d];
    }
    h.radius = a.radius;
  }
  const DeviceState& S = c->S;
  HIP_TRY(c, hipMemcpy(S.sid, sid.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(S.slot_of, slot.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(S.coll_start, start.data(), start.size() * sizeof(int32_t),
                       hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(S.coll_members, members.data(), (size_t)c->nmembers * sizeof(int32_t),
                       hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(S.coll, cs.data(), cs.size() * sizeof(sdpd::CollState), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(S.coll_hot, hot.data(), hot.size() * sizeof(sdpd::CollHot),
                       hipMemcpyHostToDevice));
  return SDPD_OK;
}

int upload_impl(sdpd_ctx* c, const double* x, const double* v, const double* m,
                const int32_t* id, const int32_t* rep, int32_t n, const char* who) {
  const std::string w(who);
  if (n <= 0 || !x || !v || !m || !id) return fail(c, SDPD_ERR_ARG, w + ": bad arguments");
  if (n >= (1 << 25))
    return fail(c, SDPD_ERR_ARG, w + ": at most 2^25 - 1 particles per context");
  const int dim = c->P.dim;
  const int nrep = c->P.nrep;
  const int ncoll = c->P.ncoll;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t sid = id[2 * (size_t)i + 1];
    if (sid < 0 || sid > ncoll)
      return fail(c, SDPD_ERR_ARG,
                  w + (ncoll ? ": species id must be 0 (fluid) or a colloid 1..num"
                             : ": only fluid particles (sid = 0) without sdpd_set_colloids"));
    if (!(m[i] > kMachineZero))
      return fail(c, SDPD_ERR_PHYSICS, w + ": particle mass must be positive");
    const int q = rep ? rep[i] : 0;
    if (q < 0 || q >= nrep) return fail(c, SDPD_ERR_ARG, w + ": replica index out of range");
    const sdpd::RepGeom& g = c->geo[q];
    for (int d = 0; d < dim; ++d) {
      double xv = x[(size_t)d + (size_t)dim * i];
      if (!(xv >= g.lo[d] && xv < g.lo[d] + g.L[d]))
        return fail(c, SDPD_ERR_ARG, w + ": particle outside its box [min, max)");
    }
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  free_state(c->S);
  c->parity = 0;
  c->n = 0;
  c->P.n = n;
  {
    // capacity of the stored neighbour lists: 1.6 times the neighbours a
    // uniform fluid of this mean density has inside the cut-off, plus 32;
    // a particle with more is handled by scanning again in the force pass
    double vol = 0.0;
    for (const sdpd::RepGeom& g : c->geo) {
      double vq = 1.0;
      for (int d = 0; d < dim; ++d) vq *= g.L[d];
      vol += vq;
    }
    // the speed that sizes the automatic Verlet skin
    double vmax2 = 0.0, mmin = 1e300;
    for (int32_t i = 0; i < n; ++i) {
      double v2 = 0.0;
      for (int d = 0; d < dim; ++d) v2 += v[(size_t)d + (size_t)dim * i] * v[(size_t)d + (size_t)dim * i];
      vmax2 = std::max(vmax2, v2);
      mmin = std::min(mmin, m[i]);
    }
    const double v_th = c->P.noise ? 5.0 * std::sqrt(c->P.kt / mmin) : 0.0;
    int rc = choose_skin(c, std::max(std::sqrt(vmax2), v_th));
    if (rc) return rc;
    c->P.list_cap = c->list_cap_req > 0 ? c->list_cap_req : sdpd::auto_list_capacity(c->P, vol, n);
  }
  int rc = alloc_state(c, n);
  if (rc) {
    free_state(c->S);
    return rc;
  }
  std::vector<double> xs((size_t)dim * n), vs((size_t)dim * n);
  std::vector<int32_t> pid(n), perm(n);
  for (int32_t i = 0; i < n; ++i) {
    for (int d = 0; d < dim; ++d) {
      xs[(size_t)d * n + i] = x[(size_t)d + (size_t)dim * i];
      vs[(size_t)d * n + i] = v[(size_t)d + (size_t)dim * i];
    }
    pid[i] = id[2 * (size_t)i];
    perm[i] = i;
  }
  HIP_TRY(c, hipMemcpy(c->S.x, xs.data(), xs.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->S.v, vs.data(), vs.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->S.m, m, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->S.pid, pid.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->S.perm, perm.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
  if (ncoll > 0 && (rc = upload_colloids(c, id, n))) return rc;
  if (rep)
    HIP_TRY(c, hipMemcpy(c->S.rep, rep, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
  else
    sdpd::launch_fill_i32(c->stream, c->S.rep, 0, n);
  {
    double t, tle;
    double tprev;
    marching_time(c->step, c->P.dt, &t, &tle, &tprev);
    sdpd::launch_set_step(c->stream, c->S, c->step, t, tle, tprev);
  }
  sdpd::launch_fill_i32(c->stream, c->S.gate, 0, 4);
  sdpd::launch_fill_i32(c->stream, reinterpret_cast<int32_t*>(c->S.disp), 0, 2);
  sdpd::launch_fill_f64(c->stream, c->S.f, 0.0, (int64_t)dim * n);
  sdpd::launch_fill_f64(c->stream, c->S.rho, 0.0, n);
  sdpd::launch_fill_f64(c->stream, c->S.p, 0.0, n);
  if (!c->P.newtonian) {
    // ct = I (particles_init_partial_inter.F90:145-149), act = vgt = 0
    const int d2 = dim * dim;
    for (int q = 0; q < d2; ++q) {
      double val = (q % (dim + 1) == 0) ? 1.0 : 0.0;
      sdpd::launch_fill_f64(c->stream, c->S.ct + (size_t)q * n, val, n);
    }
    sdpd::launch_fill_f64(c->stream, c->S.act, 0.0, (int64_t)d2 * n);
    sdpd::launch_fill_f64(c->stream, c->S.vgt, 0.0, (int64_t)d2 * n);
  }
  rc = check_launch(c, "sdpd_upload");
  if (rc) return rc;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->n = n;
  c->cells_valid = c->rho_valid = c->f_valid = false;
  c->sorted_once = false;
  c->coll_primed = false;
  return SDPD_OK;
}

}  // namespace

extern "C" {

int sdpd_upload(sdpd_ctx* c, const double* x, const double* v, const double* m,
                const int32_t* id, int32_t n) {
  if (!c) return SDPD_ERR_ARG;
  if (c->P.nrep != 1 || !c->rep_lo.empty())
    return fail(c, SDPD_ERR_STATE, "sdpd_upload: batched context, use sdpd_upload_replicas");
  return upload_impl(c, x, v, m, id, nullptr, n, "sdpd_upload");
}

int sdpd_set_replicas(sdpd_ctx* c, int32_t nrep, const double* box_lo, const double* box_hi) {
  if (!c) return SDPD_ERR_ARG;
  if (nrep < 1 || !box_lo || !box_hi)
    return fail(c, SDPD_ERR_ARG, "sdpd_set_replicas: bad arguments");
  if (c->P.le_g >= 0)
    return fail(c, SDPD_ERR_STATE, "sdpd_set_replicas: Lees-Edwards faces are for one box only");
  if (c->P.ncoll > 0)
    return fail(c, SDPD_ERR_STATE, "sdpd_set_replicas: colloids are for one box only");
  if (c->P.body_force >= 2)
    return fail(c, SDPD_ERR_STATE,
                "sdpd_set_replicas: body_force_type 2 and 3 depend on the box bounds (one box only)");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  free_state(c->S);
  c->parity = 0;
  c->n = 0;
  c->cells_valid = c->rho_valid = c->f_valid = false;
  c->rep_lo.assign(box_lo, box_lo + 3 * (size_t)nrep);
  c->rep_hi.assign(box_hi, box_hi + 3 * (size_t)nrep);
  int rc = build_geometry(c);
  if (rc) {
    c->rep_lo.clear();
    c->rep_hi.clear();
    (void)build_geometry(c);
  }
  return rc;
}

int sdpd_upload_replicas(sdpd_ctx* c, const double* x, const double* v, const double* m,
                         const int32_t* id, const int32_t* rep, int32_t n) {
  if (!c) return SDPD_ERR_ARG;
  if (!rep) return fail(c, SDPD_ERR_ARG, "sdpd_upload_replicas: null replica array");
  return upload_impl(c, x, v, m, id, rep, n, "sdpd_upload_replicas");
}

int sdpd_stats_replicas(sdpd_ctx* c, double* mom, double* ekin) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (!mom || !ekin) return fail(c, SDPD_ERR_ARG, "sdpd_stats_replicas: null output");
  // the replicas' particle ranges come from the last sort
  if (!c->sorted_once && (rc = run_sort(c, false))) return rc;
  {
    PhaseMark pm(c, 4);
    sdpd::launch_replica_stats(c->stream, c->P, c->S);
  }
  rc = check_launch(c, "replica stats");
  if (rc) return rc;
  const int nrep = c->P.nrep;
  std::vector<double> out(4 * (size_t)nrep);
  HIP_TRY(c, hipMemcpyAsync(out.data(), c->S.rep_stats, out.size() * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int q = 0; q < nrep; ++q) {
    ekin[q] = out[4 * (size_t)q];
    for (int d = 0; d < 3; ++d) mom[3 * (size_t)q + d] = d < c->P.dim ? out[4 * (size_t)q + 1 + d] : 0.0;
  }
  return SDPD_OK;
}

int sdpd_upload_ct(sdpd_ctx* c, const double* ct) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (c->P.newtonian) return fail(c, SDPD_ERR_STATE, "sdpd_upload_ct: Newtonian context");
  if (!ct) return fail(c, SDPD_ERR_ARG, "sdpd_upload_ct: null array");
  const int n = c->n, d2 = c->P.dim * c->P.dim;
  std::vector<int32_t> perm(n);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemcpy(perm.data(), c->S.perm, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  std::vector<double> tmp((size_t)d2 * n);
  for (int i = 0; i < n; ++i)
    for (int q = 0; q < d2; ++q) tmp[(size_t)q * n + i] = ct[(size_t)q + (size_t)d2 * perm[i]];
  HIP_TRY(c, hipMemcpy(c->S.ct, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
  c->vr_stale = true;  // mixed records carry Pi / n^2
  return SDPD_OK;
}

int sdpd_build_cells(sdpd_ctx* c) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (!c->P.verlet || !c->sorted_once) return run_sort(c, false);
  // Verlet lists: the displacements since the lists' build decide whether
  // they (and the cell order) are rebuilt at the next density pass
  sdpd::launch_track(c->stream, c->P, c->S);
  return run_list_update(c);
}

int sdpd_density(sdpd_ctx* c) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (!c->cells_valid) return fail(c, SDPD_ERR_STATE, "sdpd_density: build the cell list first");
  return run_density(c);
}

int sdpd_eos(sdpd_ctx* c) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (!c->rho_valid) return fail(c, SDPD_ERR_STATE, "sdpd_eos: density is not current");
  sdpd::launch_eos(c->stream, c->P, c->S);
  c->vr_stale = true;
  return check_launch(c, "eos");
}

int sdpd_forces(sdpd_ctx* c) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (!c->cells_valid || !c->rho_valid)
    return fail(c, SDPD_ERR_STATE, "sdpd_forces: density is not current");
  if ((rc = run_forces(c))) return rc;
  return prime_colloids(c);
}

int sdpd_kick(sdpd_ctx* c, double lambda) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (!c->f_valid) return fail(c, SDPD_ERR_STATE, "sdpd_kick: forces are not current");
  sdpd::launch_kick(c->stream, c->P, c->S, lambda);
  c->vr_stale = true;
  return check_launch(c, "kick");
}

int sdpd_drift_wrap(sdpd_ctx* c) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (c->P.ncoll) {
    if (!c->coll_primed)
      return fail(c, SDPD_ERR_STATE, "sdpd_drift_wrap: evaluate the initial forces first (colloids)");
    sdpd::launch_colloid_block(c->stream, c->P, c->S, c->nmembers);
  }
  sdpd::launch_drift_wrap(c->stream, c->P, c->S);
  c->cells_valid = false;
  return check_launch(c, "drift_wrap");
}

int sdpd_conformation(sdpd_ctx* c, double lambda) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (c->P.newtonian) return fail(c, SDPD_ERR_STATE, "sdpd_conformation: Newtonian context");
  if (!c->f_valid) return fail(c, SDPD_ERR_STATE, "sdpd_conformation: act is not current");
  sdpd::launch_conformation(c->stream, c->P, c->S, lambda);
  c->vr_stale = true;  // mixed records carry Pi / n^2
  return check_launch(c, "conformation");
}

int sdpd_stats(sdpd_ctx* c, double* mom, double* ekin) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (!mom || !ekin) return fail(c, SDPD_ERR_ARG, "sdpd_stats: null output");
  {
    PhaseMark pm(c, 4);
    sdpd::launch_stats(c->stream, c->P, c->S, c->stats_nb);
    if (c->P.ncoll) sdpd::launch_colloid_stats(c->stream, c->P, c->S);
  }
  rc = check_launch(c, "stats");
  if (rc) return rc;
  double out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIP_TRY(c, hipMemcpyAsync(out, c->S.stats_out, (c->P.ncoll ? 8 : 4) * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  // statistic_compute_statistic.F90:465-475 adds the colloids' share
  *ekin = out[0] + out[4];
  for (int d = 0; d < c->P.dim; ++d) mom[d] = out[1 + d] + out[5 + d];
  for (int d = c->P.dim; d < 3; ++d) mom[d] = 0.0;
  return SDPD_OK;
}

int sdpd_download(sdpd_ctx* c, double* x, double* v, double* rho, double* p, double* f) {
  int rc = need_particles(c);
  if (rc) return rc;
  const int dim = c->P.dim;
  if ((rho || p) && !c->rho_valid)
    return fail(c, SDPD_ERR_STATE, "sdpd_download: density is not current");
  if (f && !c->f_valid) return fail(c, SDPD_ERR_STATE, "sdpd_download: forces are not current");
  if (x && (rc = download_vec(c, c->S.x, dim, x))) return rc;
  if (v && (rc = download_vec(c, c->S.v, dim, v))) return rc;
  if (rho && (rc = download_vec(c, c->S.rho, 1, rho))) return rc;
  if (p && (rc = download_vec(c, c->S.p, 1, p))) return rc;
  if (f && (rc = download_vec(c, c->S.f, dim, f))) return rc;
  return SDPD_OK;
}

int sdpd_download_tensors(sdpd_ctx* c, double* ct, double* vgt) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (c->P.newtonian) return fail(c, SDPD_ERR_STATE, "sdpd_download_tensors: Newtonian context");
  const int d2 = c->P.dim * c->P.dim;
  if (ct && (rc = download_vec(c, c->S.ct, d2, ct))) return rc;
  if (vgt) {
    if (!c->f_valid) return fail(c, SDPD_ERR_STATE, "sdpd_download_tensors: vgt not current");
    if ((rc = download_vec(c, c->S.vgt, d2, vgt))) return rc;
  }
  return SDPD_OK;
}

int sdpd_step(sdpd_ctx* c, int32_t nsteps) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (nsteps < 0) return fail(c, SDPD_ERR_ARG, "sdpd_step: nsteps < 0");
  if (!c->f_valid) {
    // initial evaluation (marching_marching.F90:528-1123)
    if ((rc = run_sort(c, false))) return rc;
    if ((rc = run_density(c))) return rc;
    if ((rc = run_forces(c))) return rc;
  }
  if ((rc = prime_colloids(c))) return rc;
  if (nsteps == 0) return SDPD_OK;
  // the first step opens without a closing kick (the previous call ended
  // with it); every later one starts with the closing kick of the one
  // before, two at a time through a captured graph when possible
  if ((rc = step_body(c, 0))) return rc;
  int32_t left = nsteps - 1;
  if (c->use_graphs && !c->timing && left >= 2) {
    hipGraphExec_t& g = c->graph[c->parity];
    if (!g && (rc = capture_two_steps(c, &g))) return rc;
    while (left >= 2) {
      HIP_TRY(c, hipGraphLaunch(g, c->stream));
      c->step += 2;
      left -= 2;
    }
  }
  for (; left > 0; --left)
    if ((rc = step_body(c, 1))) return rc;
  {
    // closing half kick (and conformation half update) of the last step
    PhaseMark pm(c, 0);
    if (!c->P.newtonian) sdpd::launch_conformation(c->stream, c->P, c->S, 0.5);
    sdpd::launch_kick(c->stream, c->P, c->S, 0.5);
    if ((rc = check_launch(c, "closing kick"))) return rc;
  }
  // the pair records still hold the velocities (and Pi) of the force pass
  c->vr_stale = true;
  return SDPD_OK;
}

int sdpd_set_lees_edwards(sdpd_ctx* c, int32_t le_dim, const double* v_lower,
                          const double* v_upper) {
  if (!c) return SDPD_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);  // kernel arguments change
  if (le_dim == 0) {
    c->P.le_g = -1;
    return SDPD_OK;
  }
  if (!v_lower || !v_upper) return fail(c, SDPD_ERR_ARG, "sdpd_set_lees_edwards: null velocity");
  const int dim = c->P.dim;
  if (le_dim < 2 || le_dim > dim)
    return fail(c, SDPD_ERR_ARG, "sdpd_set_lees_edwards: le_dim must be 2 (y) or 3 (z)");
  if (c->P.nrep != 1 || !c->rep_lo.empty())
    return fail(c, SDPD_ERR_STATE, "sdpd_set_lees_edwards: one box only (batched context)");
  // the colloid force pass sizes its band of sheared rows at upload, and
  // Verlet lists (chosen at upload) serve boxes without sheared faces only
  if ((c->P.ncoll > 0 || c->P.verlet) && c->n > 0)
    return fail(c, SDPD_ERR_STATE,
                "sdpd_set_lees_edwards: with colloids or Verlet lists, set the faces before "
                "sdpd_upload");
  for (int d = 0; d < dim; ++d)
    if (c->P.min_phys[d] != 0.0)
      return fail(c, SDPD_ERR_ARG, "sdpd_set_lees_edwards: min_phys must be 0 (MODULO wraps into [0, L))");
  for (int k = 1; k < dim; ++k)
    if (k != le_dim - 1 && v_upper[k] - v_lower[k] != 0.0)
      return fail(c, SDPD_ERR_ARG,
                  "sdpd_set_lees_edwards: the faces may only shear along x (rows of cells)");
  c->P.le_g = le_dim - 1;
  c->P.le_vlo = v_lower[0];
  c->P.le_vhi = v_upper[0];
  c->P.le_u = v_upper[0] - v_lower[0];
  c->P.le_uf = (float)c->P.le_u;
  return SDPD_OK;
}

int sdpd_set_colloids(sdpd_ctx* c, const sdpd_colloid_params* prm, int32_t num,
                      const double* radius, const double* m, const double* mmi, const double* x,
                      const double* v, const double* omega) {
  if (!c) return SDPD_ERR_ARG;
  if (c->n > 0) return fail(c, SDPD_ERR_STATE, "sdpd_set_colloids: call before sdpd_upload");
  if (num == 0) {
    c->P.ncoll = 0;
    return SDPD_OK;
  }
  if (num < 0 || !prm || !radius || !m || !mmi || !x || !v || !omega)
    return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: bad arguments");
  if (c->P.nrep != 1 || !c->rep_lo.empty())
    return fail(c, SDPD_ERR_STATE, "sdpd_set_colloids: one box only (batched context)");
  if (c->P.mixed) return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: FP64 only");
  if (!c->P.newtonian && c->P.force_type != 3)
    return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: rhs_force_type 3 in Oldroyd-B flows");
  if (prm->noslip != 1 && prm->noslip != 2)
    return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: noslip must be 1 (frozen) or 2 (Morris)");
  if (prm->rho_type != 0 && prm->rho_type != 1)
    return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: rho_type must be 0 or 1");
  if (prm->integrate_ab < 1 || prm->integrate_ab > sdpd::kCollHist)
    return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: integrate_ab must be 1..5");
  if (prm->sub_steps < 1 || prm->sub_steps > sdpd::kCollSub)
    return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: sub_steps must be 1..8");
  if (prm->body_force_type != 0 && prm->body_force_type != 1)
    return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: body_force_type must be 0 or 1");
  if (!(prm->dout >= 0.0)) return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: dout < 0");
  const int dim = c->P.dim;
  for (int k = 0; k < num; ++k) {
    if (!(radius[k] > 0.0) || !(m[k] > 0.0))
      return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: radius and mass must be positive");
    if (prm->rotate)
      for (int d = (dim == 2 ? 2 : 0); d < 3; ++d)
        if (!(mmi[3 * (size_t)k + d] > 0.0))
          return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: moment of inertia must be positive");
    for (int d = 0; d < dim; ++d) {
      const double xv = x[3 * (size_t)k + d];
      if (!(xv >= c->P.min_phys[d] && xv < c->P.min_phys[d] + c->P.length[d]))
        return fail(c, SDPD_ERR_ARG, "sdpd_set_colloids: centre outside the box [min, max)");
    }
  }
  c->cprm = *prm;
  c->c_radius.assign(radius, radius + num);
  c->c_m.assign(m, m + num);
  c->c_mmi.assign(mmi, mmi + 3 * (size_t)num);
  c->c_x.assign(x, x + 3 * (size_t)num);
  c->c_v.assign(v, v + 3 * (size_t)num);
  c->c_w.assign(omega, omega + 3 * (size_t)num);
  if (dim == 2)
    for (int k = 0; k < num; ++k) {
      // disks move in the plane and spin about z
      c->c_x[3 * (size_t)k + 2] = c->c_v[3 * (size_t)k + 2] = 0.0;
      c->c_w[3 * (size_t)k] = c->c_w[3 * (size_t)k + 1] = 0.0;
    }
  DevParams& P = c->P;
  P.ncoll = num;
  P.coll_noslip = prm->noslip;
  P.coll_rho_type = prm->rho_type;
  P.coll_translate = prm->translate ? 1 : 0;
  P.coll_rotate = prm->rotate ? 1 : 0;
  P.coll_ab = prm->integrate_ab;
  P.coll_sub = prm->sub_steps;
  P.coll_bf_type = prm->body_force_type;
  for (int d = 0; d < 3; ++d) P.coll_bf[d] = d < dim ? prm->body_force[d] : 0.0;
  P.coll_dout = prm->dout;
  return SDPD_OK;
}

int sdpd_get_colloids(sdpd_ctx* c, double* x, double* v, double* omega, double* drag,
                      double* torque) {
  int rc = need_particles(c);
  if (rc) return rc;
  const int num = c->P.ncoll;
  if (num == 0) return fail(c, SDPD_ERR_STATE, "sdpd_get_colloids: no colloids");
  std::vector<sdpd::CollState> cs(num);
  HIP_TRY(c, hipMemcpyAsync(cs.data(), c->S.coll, cs.size() * sizeof(sdpd::CollState),
                            hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < num; ++k)
    for (int d = 0; d < 3; ++d) {
      const size_t o = 3 * (size_t)k + d;
      if (x) x[o] = cs[k].x[d];
      if (v) v[o] = cs[k].v[0][d];
      if (omega) omega[o] = cs[k].w[0][d];
      if (drag) drag[o] = cs[k].drag[d];
      if (torque) torque[o] = cs[k].torque[d];
    }
  return SDPD_OK;
}

int sdpd_set_graphs(sdpd_ctx* c, int32_t enable) {
  if (!c) return SDPD_ERR_ARG;
  c->use_graphs = enable != 0;
  if (!c->use_graphs) {
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    drop_graphs(c);
  }
  return SDPD_OK;
}

int sdpd_set_list_capacity(sdpd_ctx* c, int32_t cap) {
  if (!c) return SDPD_ERR_ARG;
  if (c->n > 0) return fail(c, SDPD_ERR_STATE, "sdpd_set_list_capacity: call before sdpd_upload");
  if (cap != 0 && (cap < 8 || cap > 4096 || cap % 8 != 0))
    return fail(c, SDPD_ERR_ARG, "sdpd_set_list_capacity: 0, or a multiple of 8 in 8..4096");
  c->list_cap_req = cap;
  return SDPD_OK;
}

int sdpd_get_list_capacity(sdpd_ctx* c, int32_t* cap) {
  if (!c || !cap) return SDPD_ERR_ARG;
  *cap = c->n > 0 ? c->P.list_cap : c->list_cap_req;
  return SDPD_OK;
}

int sdpd_set_skin(sdpd_ctx* c, double skin) {
  if (!c) return SDPD_ERR_ARG;
  if (c->n > 0) return fail(c, SDPD_ERR_STATE, "sdpd_set_skin: call before sdpd_upload");
  if (std::isnan(skin)) return fail(c, SDPD_ERR_ARG, "sdpd_set_skin: skin is NaN");
  c->skin_req = skin < 0.0 ? -1.0 : skin;
  return SDPD_OK;
}

int sdpd_get_skin(sdpd_ctx* c, double* skin) {
  if (!c || !skin) return SDPD_ERR_ARG;
  *skin = c->n > 0 ? c->P.skin : c->skin_req;
  return SDPD_OK;
}

int sdpd_get_list_builds(sdpd_ctx* c, int64_t* builds) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (!builds) return fail(c, SDPD_ERR_ARG, "sdpd_get_list_builds: null output");
  int32_t g[4] = {0, 0, 0, 0};
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemcpy(g, c->S.gate, sizeof(g), hipMemcpyDeviceToHost));
  // lists rebuilt every step: one build per density pass; the counter then
  // stays 0
  *builds = c->P.verlet ? (int64_t)g[1] : -1;
  return SDPD_OK;
}

int sdpd_set_tuning(sdpd_ctx* c, const sdpd_tuning* t) {
  if (!c || !t) return SDPD_ERR_ARG;
  if (t->density_shape < 0 || t->density_shape > 3 || (t->chunk_order != 0 && t->chunk_order != 1))
    return fail(c, SDPD_ERR_ARG, "sdpd_set_tuning: density_shape 0..3, chunk_order 0 or 1");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);  // kernel arguments change
  c->P.density_shape = t->density_shape;
  c->P.xcd_order = t->chunk_order;
  return SDPD_OK;
}

int sdpd_get_tuning(sdpd_ctx* c, sdpd_tuning* t) {
  if (!c || !t) return SDPD_ERR_ARG;
  t->density_shape = c->P.density_shape;
  t->chunk_order = c->P.xcd_order;
  return SDPD_OK;
}

int sdpd_autotune(sdpd_ctx* c, int32_t reps, sdpd_tuning* best, double* ms_best) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (reps < 1) return fail(c, SDPD_ERR_ARG, "sdpd_autotune: reps < 1");
  if (!c->cells_valid && (rc = run_sort(c, false))) return rc;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  // The passes below recompute f (and act, vgt) from the current velocities.
  // After sdpd_step those are v(t + dt), while the next step must open with
  // the forces of the last step's evaluation (half-step velocities): keep a
  // copy and put it back at the end.
  struct Saved {
    double* dev;
    double* keep;
    size_t count;
  };
  std::vector<Saved> saved;
  if (c->f_valid) {
    const size_t n = (size_t)c->n, dim = c->P.dim;
    saved.push_back({c->S.f, nullptr, dim * n});
    if (!c->P.newtonian) {
      saved.push_back({c->S.act, nullptr, dim * dim * n});
      saved.push_back({c->S.vgt, nullptr, dim * dim * n});
    }
  }
  auto release_saved = [&]() {
    for (Saved& s : saved)
      if (s.keep) (void)hipFree(s.keep);
  };
  for (Saved& s : saved) {
    hipError_t e = hipMalloc((void**)&s.keep, s.count * sizeof(double));
    if (e == hipSuccess)
      e = hipMemcpyAsync(s.keep, s.dev, s.count * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) {
      release_saved();
      return fail(c, SDPD_ERR_DEVICE, std::string("sdpd_autotune: ") + hipGetErrorString(e));
    }
  }
  const bool timing = c->timing;
  c->timing = false;
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    if (a) (void)hipEventDestroy(a);
    c->timing = timing;
    release_saved();
    return fail(c, SDPD_ERR_DEVICE, "sdpd_autotune: event creation failed");
  }
  // reps timed passes after an untimed one, at the current positions (the
  // state is recomputed identically each time): SCAN times the candidate
  // scan of every chunk (the density pass of a Verlet build), otherwise a
  // density + force pass pair as a step runs it.  Returns the time in ms,
  // or -1 after an error (rc set).
  auto time_passes = [&](bool scan) -> float {
    auto pass = [&]() -> int {
      if (scan) return run_density(c, true);
      int r = run_density(c);
      if (!r) r = run_forces(c);
      return r;
    };
    if ((rc = pass())) return -1.f;
    (void)hipEventRecord(a, c->stream);
    for (int r = 0; r < reps && rc == SDPD_OK; ++r) rc = pass();
    (void)hipEventRecord(b, c->stream);
    if (rc) return -1.f;
    if (hipEventSynchronize(b) != hipSuccess) {
      rc = fail(c, SDPD_ERR_DEVICE, "sdpd_autotune: event wait failed");
      return -1.f;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
  };
  // Each choice is timed twice, in a sequence and then backwards, and keeps
  // its faster time: the GPU's clock still settles during the first passes,
  // which would otherwise favour whatever is timed last.  The linear chunk
  // order is taken only when it is faster by more than kOrderMargin: the
  // two orders often time within noise of each other, and the
  // XCD-contiguous one reads less through the fabric (each XCD's L2 serves
  // a compact slab).
  constexpr float kOrderMargin = 0.99f;
  auto keep_min = [](float& t, float ms) {
    if (ms >= 0.f && (t < 0.f || ms < t)) t = ms;
  };
  sdpd_tuning pick{0, 1};
  float tbest = -1.f;
  if (c->P.verlet) {
    // Verlet lists: a step walks the lists, whose cost depends on the chunk
    // order alone; the density shape serves the candidate scan of the
    // builds.  The order is chosen on walk passes, then the shape on scans.
    float t_order[2] = {-1.f, -1.f};
    for (int k = 0; k < 4 && rc == SDPD_OK; ++k) {
      const int order = k == 0 || k == 3 ? 1 : 0;
      c->P.xcd_order = order;
      keep_min(t_order[order], time_passes(false));
    }
    if (rc == SDPD_OK) {
      pick.chunk_order = t_order[0] < kOrderMargin * t_order[1] ? 0 : 1;
      tbest = t_order[pick.chunk_order];
    }
    c->P.xcd_order = pick.chunk_order;
    float t_shape[4] = {-1.f, -1.f, -1.f, -1.f};
    for (int k = 0; k < 8 && rc == SDPD_OK; ++k) {
      const int shape = k < 4 ? k : 7 - k;
      c->P.density_shape = shape;
      keep_min(t_shape[shape], time_passes(true));
    }
    for (int shape = 1; shape < 4; ++shape)
      if (t_shape[shape] < t_shape[pick.density_shape]) pick.density_shape = shape;
  } else {
    // lists rebuilt every step: the density pass is the scan itself
    float t[2][4] = {{-1.f, -1.f, -1.f, -1.f}, {-1.f, -1.f, -1.f, -1.f}};
    for (int k = 0; k < 16 && rc == SDPD_OK; ++k) {
      const int m = k < 8 ? k : 15 - k;  // the sequence, then backwards
      const int order = 1 - m / 4, shape = m % 4;
      c->P.density_shape = shape;
      c->P.xcd_order = order;
      keep_min(t[order][shape], time_passes(false));
    }
    if (rc == SDPD_OK) {
      int s_order[2] = {0, 0};
      for (int order = 0; order < 2; ++order)
        for (int shape = 1; shape < 4; ++shape)
          if (t[order][shape] < t[order][s_order[order]]) s_order[order] = shape;
      const float t0 = t[0][s_order[0]], t1 = t[1][s_order[1]];
      pick.chunk_order = t0 < kOrderMargin * t1 ? 0 : 1;
      pick.density_shape = s_order[pick.chunk_order];
      tbest = pick.chunk_order ? t1 : t0;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  c->timing = timing;
  if (!rc) {
    c->P.density_shape = pick.density_shape;
    c->P.xcd_order = pick.chunk_order;
    if (!(rc = run_density(c))) rc = run_forces(c);
  }
  for (Saved& s : saved) {
    hipError_t e = hipMemcpyAsync(s.dev, s.keep, s.count * sizeof(double), hipMemcpyDeviceToDevice,
                                  c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess && !rc)
      rc = fail(c, SDPD_ERR_DEVICE, std::string("sdpd_autotune: ") + hipGetErrorString(e));
  }
  release_saved();
  if (rc) return rc;
  if (best) *best = pick;
  if (ms_best) *ms_best = (double)tbest / reps;
  return SDPD_OK;
}

int sdpd_set_step(sdpd_ctx* c, int64_t step) {
  if (!c) return SDPD_ERR_ARG;
  c->step = step;
  if (c->S.step_dev) {
    double t, tle;
    double tprev;
    marching_time(step, c->P.dt, &t, &tle, &tprev);
    sdpd::launch_set_step(c->stream, c->S, step, t, tle, tprev);
    return check_launch(c, "sdpd_set_step");
  }
  return SDPD_OK;
}

int sdpd_get_step(sdpd_ctx* c, int64_t* step) {
  if (!c || !step) return SDPD_ERR_ARG;
  *step = c->step;
  return SDPD_OK;
}

int sdpd_get_cells(sdpd_ctx* c, int32_t* nm, int32_t* lhbx, int32_t* lpdx) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (c->P.nrep != 1 || !c->rep_lo.empty())
    return fail(c, SDPD_ERR_STATE, "sdpd_get_cells: one box only (batched context)");
  if (nm)
    for (int d = 0; d < 3; ++d) nm[d] = c->P.nm_ref[d];
  if (!lhbx && !lpdx) return SDPD_OK;
  DeviceState& S = c->S;
  const size_t N = (size_t)c->n;
  if (!S.rkey) {
#define RALLOC(ptr, cnt) \
  if ((rc = dalloc(c, &S.ptr, (cnt))) != SDPD_OK) return rc;
    RALLOC(rkey, N);
    RALLOC(rslot, N);
    RALLOC(rcount, (size_t)c->P.ncell_ref);
    RALLOC(rstart, (size_t)c->P.ncell_ref + 1);
    RALLOC(runsorted, N);
    RALLOC(rlpdx, N);
#undef RALLOC
  }
  sdpd::launch_reference_cells(c->stream, c->P, c->S);
  if ((rc = check_launch(c, "reference cells"))) return rc;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (lhbx)
    HIP_TRY(c, hipMemcpy(lhbx, S.rstart, ((size_t)c->P.ncell_ref + 1) * sizeof(int32_t),
                         hipMemcpyDeviceToHost));
  if (lpdx) HIP_TRY(c, hipMemcpy(lpdx, S.rlpdx, N * sizeof(int32_t), hipMemcpyDeviceToHost));
  return SDPD_OK;
}

int sdpd_get_counters(sdpd_ctx* c, int64_t* pairs, int64_t* rechecks) {
  int rc = need_particles(c);
  if (rc) return rc;
  std::vector<int32_t> a(c->n), b(c->n);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemcpy(a.data(), c->S.nbr_in, c->n * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(b.data(), c->S.recheck, c->n * sizeof(int32_t), hipMemcpyDeviceToHost));
  int64_t sa = 0, sb = 0;
  for (int i = 0; i < c->n; ++i) {
    sa += a[i];
    sb += b[i];
  }
  if (pairs) *pairs = sa;
  if (rechecks) *rechecks = sb;
  return SDPD_OK;
}

int sdpd_neighbor_signature(sdpd_ctx* c, int32_t mixed_decision, int32_t* count,
                            uint32_t* hash) {
  int rc = need_particles(c);
  if (rc) return rc;
  if (!c->cells_valid)
    return fail(c, SDPD_ERR_STATE, "sdpd_neighbor_signature: cells are not current");
  sdpd::launch_signature(c->stream, c->P, c->S, mixed_decision != 0);
  if ((rc = check_launch(c, "signature"))) return rc;
  const int n = c->n;
  std::vector<int32_t> cnt(n), perm(n);
  std::vector<uint32_t> hs(n);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemcpy(cnt.data(), c->S.nbr_sig_count, n * sizeof(int32_t),
                       hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(hs.data(), c->S.nbr_hash, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  HIP_TRY(c, hipMemcpy(perm.data(), c->S.perm, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) {
    if (count) count[perm[i]] = cnt[i];
    if (hash) hash[perm[i]] = hs[i];
  }
  return SDPD_OK;
}

int sdpd_pair_normals(const uint32_t* lo, const uint32_t* hi, int32_t count, uint64_t seed,
                      int64_t step, float* g) {
  if (count <= 0 || !lo || !hi || !g) return SDPD_ERR_ARG;
  uint32_t *dlo = nullptr, *dhi = nullptr;
  float* dg = nullptr;
  hipError_t e = hipMalloc((void**)&dlo, count * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&dhi, count * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&dg, (size_t)count * 4 * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(dlo, lo, count * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dhi, hi, count * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    sdpd::launch_pair_normals(nullptr, dlo, dhi, count, seed, step, dg);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(g, dg, (size_t)count * 4 * sizeof(float), hipMemcpyDeviceToHost);
  if (dlo) (void)hipFree(dlo);
  if (dhi) (void)hipFree(dhi);
  if (dg) (void)hipFree(dg);
  return e == hipSuccess ? SDPD_OK : SDPD_ERR_DEVICE;
}

int sdpd_set_timing(sdpd_ctx* c, int32_t enable) {
  if (!c) return SDPD_ERR_ARG;
  c->timing = enable != 0;
  return SDPD_OK;
}

int sdpd_get_phase_times(sdpd_ctx* c, double* ms, int64_t* launches) {
  if (!c) return SDPD_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  double acc[SDPD_NUM_PHASES] = {0};
  int64_t cnt[SDPD_NUM_PHASES] = {0};
  for (auto& m : c->marks) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, m.a, m.b) == hipSuccess && m.phase >= 0 &&
        m.phase < SDPD_NUM_PHASES) {
      acc[m.phase] += t;
      cnt[m.phase] += 1;
    }
  }
  release_events(c);
  for (int q = 0; q < SDPD_NUM_PHASES; ++q) {
    if (ms) ms[q] = acc[q];
    if (launches) launches[q] = cnt[q];
  }
  return SDPD_OK;
}

int sdpd_sync(sdpd_ctx* c) {
  if (!c) return SDPD_ERR_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return SDPD_OK;
}

}  // extern "C"
