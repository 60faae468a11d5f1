// Shared host-side helpers of libpoe2d: status/error reporting and the set-up routines the solver
// handle uses (C-PML profiles, wavelets).
#pragma once

#include "../../include/poe2d.h"

#include <string>
#include <vector>

namespace poe2d {

// Records `msg` as this thread's last error and returns `code`.
poe2d_status fail(poe2d_status code, const std::string &msg);

// Source-time function sample at time t (no amplitude, no cut).
double wavelet_shape(int kind, double f0, double t0, double t);

// Largest P velocity implied by the stress law, max sqrt((lambda_c + 2 mu + alpha^2 M) / rho).
double cp_max_from_coeffs(const poe2d_grid &g, const poe2d_coeffs &c);

}  // namespace poe2d
