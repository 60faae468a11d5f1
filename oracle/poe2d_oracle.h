/*
 * poe2d_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's poroelastic C-PML time loop, the answer compared with by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The product library never links it.
 * It shares the public header's structs (types only) so a test can drive both behind one
 * interface.  Functions return 0 on success, a poe2d_status value otherwise.
 */
#ifndef POE2D_ORACLE_H
#define POE2D_ORACLE_H

#include "../include/poe2d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct poe2d_oracle poe2d_oracle;

int poe2d_oracle_create(const poe2d_grid *grid, const poe2d_coeffs *coeffs, poe2d_oracle **out);
int poe2d_oracle_set_source(poe2d_oracle *o, const poe2d_source *src);
int poe2d_oracle_set_wavelet(poe2d_oracle *o, const double *w, int n);
int poe2d_oracle_set_receivers(poe2d_oracle *o, const poe2d_receivers *rec);
int poe2d_oracle_set_threads(poe2d_oracle *o, int nthreads);
/* Reference-structured extras for the CPU baseline: the three full-grid receiver scans of
 * seismic_CPML_2d_poroelastic.cpp:657-739 run every step (panj VSP traces at x = is, panj2
 * horizontal traces at z = 10, four point traces at a0/a1/50/js). */
int poe2d_oracle_set_reference_scans(poe2d_oracle *o, int panj, int panj2, int a0_pos, int a1_pos);
/* Samples 0..n-1 recorded by those scans: which = 0 VSP v_z, 1 VSP v_x (index = line position
 * 0..panj-1), 2 horizontal v_z, 3 horizontal v_x (0..panj2-1), 4..7 the point traces at rows 50,
 * a0_pos, a1_pos, js (index 0). */
int poe2d_oracle_reference_scan(poe2d_oracle *o, int which, int index, double *out, int n);
int poe2d_oracle_run(poe2d_oracle *o, int nsteps);
int poe2d_oracle_step_count(poe2d_oracle *o, int *nsteps);
int poe2d_oracle_seismograms(poe2d_oracle *o, void *out, size_t out_bytes);
int poe2d_oracle_snapshot(poe2d_oracle *o, int field, void *out, size_t out_bytes);
/* Overwrite one state field (VX, VY, WX, WY, EXX, EYY, ZETA, SXY) -- for impulse and
 * polynomial tests.  ZETA sets the x part and zeroes the z part. */
int poe2d_oracle_set_field(poe2d_oracle *o, int field, const void *in, size_t in_bytes);
/* 1-D C-PML tables in use: which = 0..5 -> a, b, K, a_half, b_half, K_half; axis 0 = x, 1 = z. */
int poe2d_oracle_cpml_table(poe2d_oracle *o, int axis, int which, double *out, int n);
void poe2d_oracle_destroy(poe2d_oracle *o);
const char *poe2d_oracle_last_error(void);

/* Set-up restatements (independent of the product's). */
int poe2d_oracle_cpml_profile(int axis, int n, double delta, double dt, int npml, double npower,
                              double rcoef, double kmax, double alpha_max, double cp_max,
                              int precision, double *a, double *b, double *K, double *ah,
                              double *bh, double *Kh);
int poe2d_oracle_material_biot(size_t n, const double *rho_s, const double *rho_f,
                               const double *phi, const double *kappa, const double *eta,
                               const double *tort, const double *K_s, const double *K_f,
                               const double *K_fr, const double *mu_fr, double *rho, double *rhof,
                               double *m, double *mu, double *lambdac, double *M, double *alpha,
                               double *b);
int poe2d_oracle_material_reference(size_t n, const float *cp, const float *cs,
                                    const float *rho_sat, const float *medium, int nmed,
                                    const float *kf_tab, const float *eta_tab,
                                    const float *kappa_tab, const float *phi_tab,
                                    const float *a_tab, const float *rhof_tab, float *rho,
                                    float *rhof, float *m, float *mu, float *lambdac, float *M,
                                    float *alpha, float *b);
int poe2d_oracle_wavelet(int kind, double f0, double t0, double amp, double t_cut, double dt,
                         int n, double *out);

#ifdef __cplusplus
}
#endif

#endif
