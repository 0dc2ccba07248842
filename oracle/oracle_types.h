// oracle_types.h — types shared by the oracle's translation units.
// TEST INFRASTRUCTURE ONLY (see dic_oracle.cpp).
#pragma once

#include <cstdint>

extern "C" {

// SolveResult (subset.rs:146-191) plus the reference quantities of
// Subset::new (subset.rs:885-899); layout mirrors dic_subset_result.
struct OrcResult {
    double p[12];
    double c_zncc, c_znssd;
    double norm;       // last convergence norm (NaN if no iteration)
    double eta_u, eta_v;
    double f_m, delta_f, sssig, sigma;
    int32_t iterations, converged, quality_ok, status;
    int32_t n_px, order;
};

int orc_inverse(const double* M, int m, double* out);
int orc_subset_solve(const double* fcoef, const double* gcoef, int h, int w, int border, const double* coord,
                     const double* offsets, int64_t n, int order, const double* p0, int p0_len, int method,
                     double tolerance, double max_norm, int max_iterations, OrcResult* R, double* history,
                     double* f_out, double* grad_out, double* f_coords_out);
}
