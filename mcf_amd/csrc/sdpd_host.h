// sdpd_host.h -- host-side set-up shared by the C-ABI's translation units.
#pragma once

#include <stdint.h>

#include "../../include/sdpd.h"
#include "sdpd_kernels.h"

namespace sdpd {

// Derived constants of a parameter record, with the reference's start-up
// checks (mcf.F90:616-649, physics_adjust_parameters.F90, kernel_new.F90,
// stateEquation_new.F90); on failure err names the check (a literal).
int derive_params(const sdpd_params& p, DevParams& P, const char*& err);

// FP32 cut-off band of the candidate test for boxes whose x extent plus
// five sort-grid cells is at most mmax.
void set_cut_band(DevParams& P, double mmax);

// Automatic capacity of the stored neighbour lists for n particles in the
// given volume (the cut-off sphere plus the Verlet skin).
int auto_list_capacity(const DevParams& P, double volume, int64_t n);

// The largest Verlet skin the sort grids of nrep boxes allow, and whether
// a skin is thick enough to serve Verlet lists.
double grid_skin(const DevParams& P, const RepGeom* geo, int nrep);
bool skin_usable(const DevParams& P, double skin);

}  // namespace sdpd
