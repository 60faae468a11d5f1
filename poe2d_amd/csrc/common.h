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

// The lambda the stress update uses at a point: lambda_c as written, lambda_c - alpha^2 M under the
// drained law.
inline double effective_lambda(const poe2d_grid &g, double lambdac, double alpha, double M)
{
    return g.stress_law == POE2D_STRESS_DRAINED ? lambdac - alpha * alpha * M : lambdac;
}

// Largest P velocity implied by the stress law, max sqrt((lambda + 2 mu + alpha^2 M) / rho) with
// lambda = effective_lambda.
double cp_max_from_coeffs(const poe2d_grid &g, const poe2d_coeffs &c);

}  // namespace poe2d
