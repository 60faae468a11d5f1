/*
 * poe2d.h -- C ABI of the MI355X (gfx950) 2-D poroelastic (Biot) C-PML solver.
 *
 * The reference (thejatmik/poe2d) has no callable interface: its time loop, device
 * arrays and scalars all live in one main() (seismic_CPML_2D_poroelastic_withCUDA.cu:470-1922,
 * seismic_CPML_2d_poroelastic.cpp:43-929).  Each entry point below names the piece of that
 * main() it replaces, so a maintainer can swap the reference loop for these calls.
 *
 * Conventions
 *   - Every function returns a poe2d_status and never exits the process; the message of the
 *     last failure on the calling thread is poe2d_last_error().
 *   - The caller owns every host pointer.  Inputs are copied during the call; outputs go into
 *     caller-allocated buffers whose size the call checks.
 *   - One solver handle = one GPU device + one HIP stream.  A handle is not re-entrant; distinct
 *     handles may be used from distinct threads.
 *   - Grids follow the reference's indexing: NX, NZ are the largest point indices, so a grid has
 *     (NX+1) x (NZ+1) points, stored row-major with x fastest; z (the reference's "z", the
 *     configs' "y") increases with depth.
 *   - FP32 and FP64 are chosen per handle (poe2d_grid.precision), not per library.
 */
#ifndef POE2D_H
#define POE2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define POE2D_ABI_VERSION 1

typedef enum {
    POE2D_OK = 0,
    POE2D_E_INVALID = 1,  /* bad argument or out-of-range geometry                        */
    POE2D_E_CFL = 2,      /* Courant number above 6/7/sqrt(2): the reference refuses to run  */
    POE2D_E_NOMEM = 3,    /* device or host allocation failed                                */
    POE2D_E_HIP = 4,      /* HIP runtime error (no device, launch failure, ...)              */
    POE2D_E_IO = 5,       /* file could not be read or written                               */
    POE2D_E_UNSTABLE = 6, /* non-finite or runaway wavefield detected                        */
    POE2D_E_STATE = 7     /* call out of order (e.g. run before set_source)                  */
} poe2d_status;

enum { POE2D_F32 = 32, POE2D_F64 = 64 };

/* Grid, time step and C-PML parameters.
 * Replaces the hard-coded constants of seismic_CPML_2D_poroelastic_withCUDA.cu:486-510, 840-845
 * (NX, NZ, DELTAX, DELTAZ, DELTAT, NPOINTS_PML, NPOWER, K_MAX_PML, ALPHA_MAX_PML, Rcoef at :992). */
typedef struct {
    int nx, nz;        /* NX, NZ: point indices run 0..nx (x) and 0..nz (z)                       */
    double dx, dz, dt; /* grid spacing (m) and time step (s)                                      */
    int npml;          /* C-PML thickness in grid points (NPOINTS_PML)                             */
    double npower;     /* profile power (NPOWER), reference 2                                      */
    double rcoef;      /* target reflection coefficient (Rcoef), reference 1e-3                    */
    double kmax;       /* K_MAX_PML, reference 1                                                   */
    double alpha_max;  /* ALPHA_MAX_PML; <= 0 selects the reference's pi*f0                        */
    double f0;         /* dominant frequency (Hz): sets alpha_max and the default wavelet          */
    double cp_max;     /* P velocity for d0 and the Courant check; <= 0: derived from the
                          coefficients as max sqrt((lambdac + 2 mu + alpha^2 M) / rho)             */
    int precision;     /* POE2D_F32 or POE2D_F64                                                  */
    int check_cfl;     /* nonzero: refuse Courant > 6/7/sqrt(2) like the reference (:1211)         */
    int stress_law;    /* POE2D_STRESS_AS_WRITTEN (0) or POE2D_STRESS_DRAINED                      */
    int staggered_material; /* POE2D_MATERIAL_AS_WRITTEN (0) or POE2D_MATERIAL_AVERAGED            */
} poe2d_grid;

/* Stress law.  AS_WRITTEN is the reference's: sigma_xx = (lambda_c + 2 mu) e_xx + lambda_c e_yy
 * - alpha p with the undrained lambda_c = lambda_o + alpha^2 M and p = -alpha M (e_xx + e_yy) + M
 * zeta (seismic_CPML_2D_poroelastic_withCUDA.cu:241-242), which counts alpha^2 M twice in the
 * P-wave modulus (lambda_c + 2 mu + alpha^2 M).  DRAINED uses lambda_o = lambda_c - alpha^2 M in
 * place of lambda_c (the coefficient array still holds lambda_c), which is Biot's law: the
 * low-frequency P-wave modulus becomes Gassmann's lambda_c + 2 mu.  The P velocity that sets d0
 * and the Courant check follows the law. */
enum { POE2D_STRESS_AS_WRITTEN = 0, POE2D_STRESS_DRAINED = 1 };

/* Material at the staggered points.  AS_WRITTEN is the reference's: sigma_xy at (x, z+1/2) uses
 * mu of point (x, z+1) (the harmonic mean computed at seismic_CPML_2D_poroelastic_withCUDA.cu:217
 * is overwritten at :218) and the v_y / w_y update at (x, z+1/2) uses rho, rho_f, m, b of point
 * (x, z+1) (:330-333).  AVERAGED uses the harmonic mean 2 / (1/mu(x,z) + 1/mu(x,z+1)) for sigma_xy
 * and the arithmetic means of rho, rho_f, m and b over points (x, z) and (x, z+1) for v_y / w_y.
 * Handles advanced together in a batch must share this setting. */
enum { POE2D_MATERIAL_AS_WRITTEN = 0, POE2D_MATERIAL_AVERAGED = 1 };

/* The eight per-point coefficient arrays the time loop reads (the reference's rho, rhof, rsm,
 * rmu, rlambdac, rbM, alpha, etaokappa uploaded at seismic_CPML_2D_poroelastic_withCUDA.cu:1433-1461).
 * Each array has (nz+1) rows of (nx+1) values, x fastest, rows `pitch` elements apart. */
typedef struct {
    int dtype;         /* POE2D_F32 or POE2D_F64: element type of the arrays                       */
    const void *rho;   /* bulk density rho = (1-phi) rho_s + phi rho_f                             */
    const void *rhof;  /* fluid density rho_f                                                      */
    const void *m;     /* m = a rho_f / phi (reference rsm)                                        */
    const void *mu;    /* shear modulus (reference rmu)                                            */
    const void *lambdac; /* lambda_c = lambda_o + alpha^2 M (reference rlambdac)                  */
    const void *M;     /* Biot modulus (reference rbM)                                             */
    const void *alpha; /* Biot-Willis coefficient                                                  */
    const void *b;     /* eta / kappa (reference etaokappa)                                        */
    size_t pitch;      /* elements between rows; 0 means nx+1                                      */
} poe2d_coeffs;

/* Source kinds.
 *   PRESSURE_DGAUSS: the reference's source, amp * 2a(t-t0) exp(-a (t-t0)^2), a = pi^2 f0^2,
 *     added to the pore pressure p at (ix, iz) after the stress update of each step, only while
 *     t < t_cut (seismic_CPML_2D_poroelastic_withCUDA.cu:90-111; amp = faktor = 1e10 at :45).
 *   FORCE_Y_RICKER: vertical body force density amp * (1 - 2a(t-t0)^2) exp(-a (t-t0)^2) added
 *     to d(sigma_xy)/dx + d(sigma_yy)/dz in the v_y / w_y update at (ix, iz); no counterpart in
 *     the reference (its v_y source is commented out at :357-383). */
enum { POE2D_SRC_PRESSURE_DGAUSS = 0, POE2D_SRC_FORCE_Y_RICKER = 1 };

typedef struct {
    int kind;      /* POE2D_SRC_*                                                                 */
    int ix, iz;    /* source point (ISOURCE, JSOURCE)                                             */
    double amp;    /* amplitude                                                                   */
    double t0;     /* wavelet delay (s); <= 0 selects the reference's 1/f0                        */
    double t_cut;  /* the wavelet is zero for t >= t_cut; <= 0 means no cut (reference 0.4 s)     */
} poe2d_source;

/* Receiver components, recorded at each receiver's own staggered position (no interpolation). */
enum { POE2D_VX = 1, POE2D_VY = 2, POE2D_P = 4 };

/* Receivers.  Slot k of a trace holds the wavefield at the start of step k*decimate: v after
 * that many steps and p of that step's stress update (the reference's save1_kernel slot
 * convention, seismic_CPML_2D_poroelastic_withCUDA.cu:114-201 with i = k+1 at :1627). */
typedef struct {
    int n;                 /* number of receivers                                                 */
    const int *ix, *iz;    /* receiver points                                                     */
    unsigned comp_mask;    /* POE2D_VX | POE2D_VY | POE2D_P                                       */
    int decimate;          /* record every decimate-th step (reference samprek)                   */
    int nsamp;             /* samples per trace to keep (capacity)                                */
} poe2d_receivers;

/* Fields for poe2d_snapshot.  P, SXX and SYY are recomputed from the state exactly as the
 * stress update computes them. */
enum {
    POE2D_FIELD_VX = 0, POE2D_FIELD_VY = 1, POE2D_FIELD_WX = 2, POE2D_FIELD_WY = 3,
    POE2D_FIELD_EXX = 4, POE2D_FIELD_EYY = 5, POE2D_FIELD_ZETA = 6, POE2D_FIELD_SXY = 7,
    POE2D_FIELD_P = 8, POE2D_FIELD_SXX = 9, POE2D_FIELD_SYY = 10
};

typedef struct poe2d_solver poe2d_solver;
typedef struct poe2d_batch poe2d_batch;

/* Library version / device probe.  Returns POE2D_OK when a gfx950 device is usable. */
int poe2d_abi_version(void);
poe2d_status poe2d_device_info(int *device_count, char *arch, size_t arch_len);

/* Create a solver: build the C-PML profiles, check stability, upload the coefficients and
 * allocate the zeroed state.  Replaces seismic_CPML_2D_poroelastic_withCUDA.cu:985-1212 (C-PML
 * profiles, Courant check) and :1251-1592 (cudaMalloc / H2D of every array). */
poe2d_status poe2d_create(const poe2d_grid *grid, const poe2d_coeffs *coeffs, poe2d_solver **out);

/* Set the source and compute its wavelet table on the host in FP64 for steps 0..nsteps_max-1.
 * Replaces ISOURCE/JSOURCE/f0 (:757-758, :830-836) and the device-side source term (:90-111). */
poe2d_status poe2d_set_source(poe2d_solver *s, const poe2d_source *src);

/* Optional: replace the wavelet with caller samples w[0..n-1] (w[k] is used at step k; steps
 * beyond n see 0).  amp and t_cut of the source still apply. */
poe2d_status poe2d_set_wavelet(poe2d_solver *s, const double *w, int n);

/* Set the receivers and allocate the device trace buffer.  Replaces panj/panj2/a0_pos/a1_pos/
 * samprek (:496-502, :759-762) and the save1_kernel gather (:114-201). */
poe2d_status poe2d_set_receivers(poe2d_solver *s, const poe2d_receivers *rec);

/* Choose how a time step is executed on this handle (no counterpart in the reference, whose
 * six launches per step are fixed at seismic_CPML_2D_poroelastic_withCUDA.cu:1628-1634).
 *   schedule 0 (default of FP64 handles): one fused launch per step; tiles of block_w - 8 columns
 *     (block_w = 128, 256 or 512 threads, default 256) by tile_rows rows (default 14; see
 *     poe2d_autotune).  When the grid holds at most 256 distinct coefficient tuples (see
 *     poe2d_material_palette) the coefficients are read through a one-byte material index per
 *     point.
 *   schedule 1: a stress launch and a velocity launch per step, intermediates through memory.
 *   schedule 2: as schedule 0, always reading the per-point coefficient arrays.
 *   schedule 3 (default of FP32 handles, block_w 256, tile_rows 24): two time steps per launch
 *     (tiles of block_w - 12 columns, block_w 64, 128 or 256, by tile_rows rows, multiples of 6
 *     best), the intermediate state kept on chip; the coefficients through a palette of at most 28
 *     entries, or from the per-point coefficient groups on grids without one; an odd last step of
 *     a run takes one schedule-0 launch.  FP64 handles run schedule 0 (block_w 64 as 128), and
 *     poe2d_autotune then tunes that kernel.
 * Tile heights do not change a bit; schedules and block widths agree to rounding.  The choice
 * only changes speed. */
poe2d_status poe2d_set_schedule(poe2d_solver *s, int schedule, int block_w, int tile_rows);

/* Number of distinct coefficient tuples of the grid when it has at most 256 (the palette the
 * fused step reads through a one-byte material index), else 0.  No counterpart in the
 * reference, which reads every coefficient array per point (seismic_CPML_2D_poroelastic_withCUDA.cu:
 * 60-111, 289-340). */
poe2d_status poe2d_material_palette(poe2d_solver *s, int *entries);

/* Advance nsteps time steps on the handle's stream and wait for them.  Each step performs the
 * reference's stress1 -> save1 -> stress2 -> stress3 -> vx -> vz sequence
 * (seismic_CPML_2D_poroelastic_withCUDA.cu:1625-1636). */
poe2d_status poe2d_run(poe2d_solver *s, int nsteps);

/* Device time (ms, hipEvent on the handle's stream) of the last poe2d_run, and how many
 * steps it covered. */
poe2d_status poe2d_last_run_ms(poe2d_solver *s, double *ms, int *nsteps);

/* Number of steps taken since create/reset. */
poe2d_status poe2d_step_count(poe2d_solver *s, int *nsteps);

/* Copy the recorded traces to the host as out[comp][rec][sample] (components in the order
 * VX, VY, P restricted to comp_mask), in the handle's precision.  out_bytes must be at least
 * ncomp*nrec*nsamp*sizeof(real).  Replaces the seismogram D2H copies (:1737-1836). */
poe2d_status poe2d_seismograms(poe2d_solver *s, void *out, size_t out_bytes);

/* Copy one field to the host as (nz+1) rows of (nx+1) values in the handle's precision.
 * Replaces the per-chunk D2H of vz/vx (:1642-1645) behind the .vzo/.vxo snapshots. */
poe2d_status poe2d_snapshot(poe2d_solver *s, int field, void *out, size_t out_bytes);

/* Overwrite one state field (VX .. SXY; ZETA is the total fluid content) of the current state
 * from (nz+1) rows of (nx+1) values in the handle's precision: an initial condition other than
 * the reference's zero state (zero_kernel, :387-422).  Points the reference's updates never
 * reach keep what is set here, as in the reference's arrays; fields derived from the strains (p,
 * sxx, syy) are refused.  No counterpart in the reference's interface. */
poe2d_status poe2d_set_field(poe2d_solver *s, int field, const void *in, size_t in_bytes);

/* Full state (wavefield, C-PML memory, step counter, recorded traces) for bitwise
 * checkpoint/resume.  No counterpart in the reference. */
poe2d_status poe2d_checkpoint_size(poe2d_solver *s, size_t *bytes);
poe2d_status poe2d_checkpoint(poe2d_solver *s, void *blob, size_t bytes);
poe2d_status poe2d_restore(poe2d_solver *s, const void *blob, size_t bytes);

/* Zero the wavefield, C-PML memory, traces and step counter.  Replaces zero_kernel
 * (:387-422, called at :1736). */
poe2d_status poe2d_reset(poe2d_solver *s);

/* Pick the fastest tile shape of the handle's schedule for its grid: times nsteps steps (after
 * one warm-up step) for each block width in {128, 256, 512} and tile height in {8, 14, 20, 26,
 * 32, 44, 62, 98} (two-step schedule: widths {64, 128, 256}, heights {6, 12, ..., 60}), times the
 * three fastest again, twice each over 3 nsteps steps, keeps the best and resets the handle.  The result is cached for the process per grid class (device,
 * precision, schedule, grid size, palette): a later call for the same class applies the cached
 * shape without timing.  nsteps = 0 means 10; nsteps < 0 times -nsteps steps and ignores the
 * cache.  Needs step count 0 (POE2D_E_STATE otherwise).  The shape changes speed only, never a
 * result.  No counterpart in the reference, whose launch shape is fixed
 * (seismic_CPML_2D_poroelastic_withCUDA.cu:1620-1634). */
poe2d_status poe2d_autotune(poe2d_solver *s, int nsteps, int *block_w, int *tile_rows);

/* Launch mode of the single-pass schedule: with graphs on (the default) a run captures 32 steps
 * (the step kernel and the receiver gather, which also advances the device step counter) in a
 * HIP graph once and replays it; off,
 * every step is launched directly.  Results are identical either way.  No counterpart in the
 * reference, which launches six kernels per step (.cu:1628-1634). */
poe2d_status poe2d_set_graphs(poe2d_solver *s, int enable);

/* A new handle on the same grid and coefficients as s (shared, read-only), with its own
 * wavefield, memory variables, source and receivers (none set), starting at step 0 with s's
 * schedule.  For shots over one model: the reference re-runs its whole program per shot position
 * (.cu:1615-1736 inside the shot loop), re-reading and re-uploading the model. */
poe2d_status poe2d_clone(poe2d_solver *s, poe2d_solver **out);

/* Replace the coefficients of a handle's grid (same grid, source and receivers) and reset it, so
 * a handle serves one model after another without device allocations: the new P velocity's
 * Courant check and C-PML profiles, the coefficient groups and the material palette, then the
 * zero state of poe2d_reset.  The reference re-reads its model files and re-derives the
 * coefficients per run (seismic_CPML_2D_poroelastic_withCUDA.cu:604-713) and keeps its device
 * arrays across shots (zero_kernel, :1736).  POE2D_E_STATE when the model is shared with clones;
 * a refused set (POE2D_E_CFL, POE2D_E_INVALID) leaves the handle as it was. */
poe2d_status poe2d_set_coeffs(poe2d_solver *s, const poe2d_coeffs *coeffs);

/* ---- Batches: several handles advanced in lock-step by one launch per time step ----
 * A batch is a view over existing handles (distinct handles, same device and precision,
 * single-pass schedules, ragged grids allowed); they keep their own results, read with the
 * handle calls above.  One step of the batch is one launch over a work list of (handle, tile)
 * items plus one gather launch; with graphs on, 32 steps are captured once and replayed.  Each
 * member's results equal a run of that handle alone.  No counterpart in the reference, which runs
 * one model and one shot per process.  batch_run needs every member at the same step. */
poe2d_status poe2d_batch_create(int n, poe2d_solver *const *members, poe2d_batch **out);
poe2d_status poe2d_batch_set_schedule(poe2d_batch *b, int block_w, int tile_rows);
poe2d_status poe2d_batch_set_graphs(poe2d_batch *b, int enable);
poe2d_status poe2d_batch_run(poe2d_batch *b, int nsteps);
poe2d_status poe2d_batch_last_run_ms(poe2d_batch *b, double *ms, int *nsteps);
/* As poe2d_autotune, over the whole batch (every member at step 0; all are reset after), cached
 * per batch class (member grid sizes, schedules, precision). */
poe2d_status poe2d_batch_autotune(poe2d_batch *b, int nsteps, int *block_w, int *tile_rows);
void poe2d_batch_destroy(poe2d_batch *b);

void poe2d_destroy(poe2d_solver *s);

/* Message of the last failure on this thread.  Replaces HANDLE_ERROR's print-and-exit
 * (book.h, included at :37). */
const char *poe2d_last_error(void);

/* ---------------------------------------------------------------------------------------
 * Host-side set-up helpers (no GPU).  They are the same routines poe2d_create uses.
 * --------------------------------------------------------------------------------------- */

/* C-PML 1-D profiles for one axis (seismic_CPML_2D_poroelastic_withCUDA.cu:985-1159).
 * axis 0 = x (n = nx+1 points, origins left npml*dx+3dx and right (nx-2)dx-npml*dx), axis 1 = z
 * (origins top npml*dz, whose alpha carries an extra 0.1*alpha_max, and bottom (nz-2)dz-npml*dz).
 * precision POE2D_F32 evaluates in the reference's float arithmetic (float variables, float
 * libm overloads, double literals) and widens the results to double exactly; POE2D_F64
 * evaluates in double.  Each output has n values. */
poe2d_status poe2d_cpml_profile(int axis, int n, double delta, double dt, int npml, double npower,
                                double rcoef, double kmax, double alpha_max, double cp_max,
                                int precision, double *out_a, double *out_b, double *out_K,
                                double *out_a_half, double *out_b_half, double *out_K_half);

/* Biot front end: per-point (rho_s, rho_f, phi, kappa, eta, tortuosity a, K_s, K_f, K_fr, mu_fr)
 * to the eight coefficient arrays, with K_b := K_fr and rho = (1-phi) rho_s + phi rho_f, and
 * the remaining algebra of seismic_CPML_2d_poroelastic.cpp:194-205.  n points, FP64 in/out. */
poe2d_status poe2d_material_biot(size_t n, const double *rho_s, const double *rho_f,
                                 const double *phi, const double *kappa, const double *eta,
                                 const double *tort, const double *K_s, const double *K_f,
                                 const double *K_fr, const double *mu_fr,
                                 double *rho, double *rhof, double *m, double *mu,
                                 double *lambdac, double *M, double *alpha, double *b);

/* The reference's own front end (seismic_CPML_2d_poroelastic.cpp:162-211): per-point c_p, c_s,
 * saturated density and medium index, plus per-medium tables of K_f, eta, kappa, phi, a and
 * rho_f (nmed entries), to the eight float coefficient arrays, in the reference's float
 * arithmetic.  Points whose medium index matches no table entry keep zero fluid properties
 * (the reference leaves them uninitialised); the call reports them with POE2D_E_INVALID. */
poe2d_status poe2d_material_reference(size_t n, const float *cp, const float *cs,
                                      const float *rho_sat, const float *medium, int nmed,
                                      const float *kf_tab, const float *eta_tab,
                                      const float *kappa_tab, const float *phi_tab,
                                      const float *a_tab, const float *rhof_tab,
                                      float *rho, float *rhof, float *m, float *mu,
                                      float *lambdac, float *M, float *alpha, float *b);

/* Wavelet samples w[k] = f(k*dt) of a source kind (amp applied, t_cut applied), FP64. */
poe2d_status poe2d_wavelet(int kind, double f0, double t0, double amp, double t_cut, double dt,
                           int n, double *out);

/* Seeded random patchy-saturation mask on (nz+1) x (nx+1) points (x fastest): disks with
 * uniform random centres over [0, nx dx] x [0, nz dz] and log-uniform radii in [rmin, rmax]
 * metres are painted (value 1) until at least gas_fraction of the points are covered.  No
 * counterpart in the reference, whose patchymodel.m only draws fixed lattices (below). */
poe2d_status poe2d_patchy_mask(int nx, int nz, double dx, double dz, double gas_fraction,
                               uint64_t seed, double rmin, double rmax, uint8_t *mask,
                               double *realised);

/* poe2d_patchy_mask of nmodels models (model k: gas_fraction gas[k], rmin = rmax = radius[k];
 * all zero when gas[k] is 0) read at flat indices idx[0 .. nidx) of the mask: out[k * nidx + j].
 * threads host threads share the models (0: one per hardware thread, at most 16).  No
 * counterpart in the reference; it forms the surrogate's layout inputs. */
poe2d_status poe2d_patchy_samples(int nx, int nz, double dx, double dz, uint64_t seed, int nmodels,
                                  const double *gas, const double *radius, int nidx, const int64_t *idx,
                                  uint8_t *out, int threads);

/* One of the six disk lattices of patchymodel.m:13-42 and the per-type blocks after it
 * (radius r points, rows x cols layer, 1 inside a disk).  type = 1..6. */
poe2d_status poe2d_patchy_lattice(int type, int rows, int cols, int r, uint8_t *layer);

/* ---- Surrogate network (no counterpart in the reference) ----
 * An ensemble of fully connected networks sizes[0] -> ... -> sizes[nlayers] with GELU (erf form)
 * between layers, bf16 weights and activations, FP32 accumulation on the matrix cores, FP32 master
 * weights and Adam (beta 0.9 / 0.999, eps 1e-8), mean-squared-error loss.  It maps patchy-model
 * parameters to (decimated, normalised) receiver gathers produced by the solver; the ensemble's
 * spread decides which predictions the solver re-checks.  Batches are at most max_batch rows
 * (a multiple of 32); training batches are multiples of 32.  Arrays are host, row-major FP32;
 * weights are [out][in]. */
typedef struct poe2d_mlp poe2d_mlp;
poe2d_status poe2d_mlp_create(int nlayers, const int *sizes, int members, int max_batch, uint64_t seed,
                              poe2d_mlp **out);
poe2d_status poe2d_mlp_get_params(poe2d_mlp *m, int member, int layer, float *w, float *b);
poe2d_status poe2d_mlp_set_params(poe2d_mlp *m, int member, int layer, const float *w, const float *b);
/* The bf16 copies of a layer's weights that the products read: w16[out][in] (forward) and its
 * transpose w16t[in][out] (the backward G W product); either may be null. */
poe2d_status poe2d_mlp_get_bf16(poe2d_mlp *m, int member, int layer, uint16_t *w16, uint16_t *w16t);
/* y[n][sizes[nlayers]] = network `member` applied to x[n][sizes[0]]. */
poe2d_status poe2d_mlp_forward(poe2d_mlp *m, int member, int n, const float *x, float *y);
/* Loss and gradients (gw[l]: [out][in], gb[l]: [out]; null entries skipped) of member on a batch
 * of n rows (n a multiple of 32), without updating the weights. */
poe2d_status poe2d_mlp_gradients(poe2d_mlp *m, int member, int n, const float *x, const float *t, float *loss,
                                 float *const *gw, float *const *gb);
/* Upload a training set x[n][sizes[0]], t[n][sizes[nlayers]] (kept on the device). */
poe2d_status poe2d_mlp_set_dataset(poe2d_mlp *m, int n, const float *x, const float *t);
/* nsteps Adam steps of member on batches of the dataset rows indices[step * batch + i];
 * losses[step] (may be null) gets each batch's loss before its update. */
poe2d_status poe2d_mlp_train(poe2d_mlp *m, int member, int nsteps, int batch, const int *indices, float lr,
                             float *losses);
/* nsteps Adam steps of each of `count` distinct members (members[j]) on the dataset rows
 * indices[(j * nsteps + step) * batch + i]; losses[j * nsteps + step] (may be null).  The members'
 * steps run on two streams with a buffer set each, so one member's latency-bound launches overlap
 * another's bandwidth-bound ones; every member's weights come out as poe2d_mlp_train leaves them,
 * bit for bit. */
poe2d_status poe2d_mlp_train_many(poe2d_mlp *m, int count, const int *members, int nsteps, int batch,
                                  const int *indices, float lr, float *losses);
/* Ensemble prediction of n rows: mean[n][sizes[nlayers]] (may be null) and spread[n] (may be
 * null) = ||standard deviation over members|| / ||mean|| per row.  Either may be host memory or
 * device memory (poe2d_device_alloc): a consumer on the GPU keeps the predictions in HBM. */
poe2d_status poe2d_mlp_predict(poe2d_mlp *m, int n, const float *x, float *mean, float *spread);
/* Per-output factors scale[sizes[nlayers]] applied on the device to every predicted mean (the
 * outputs' units when the network was trained on normalised targets); the spread stays that of
 * the unscaled outputs.  A null scale removes them. */
poe2d_status poe2d_mlp_set_output_scale(poe2d_mlp *m, const float *scale);
/* Device time of one forward pass and of one training step (forward, backward, Adam) of member
 * 0 on a batch of `batch` rows, averaged over iters (a measurement helper). */
poe2d_status poe2d_mlp_time(poe2d_mlp *m, int batch, int iters, double *forward_ms, double *train_ms);
/* C[M][N] = A[M][K] B[N][K]^T with A and B rounded to bf16 (round to nearest even) and FP32
 * accumulation on the matrix cores, ksplit >= 1 K partitions (a test and measurement helper). */
poe2d_status poe2d_gemm_bf16(int M, int N, int K, const float *A, const float *B, float *C, int ksplit);
/* The same product without allocating: no counterpart in the reference; for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid hipMalloc).  Operand staging,
 * the product and the split-K partials live in a fixed 32 MiB device scratch that each GPU holds
 * from library load until the process ends; operands needing more (4 M K + 4 N K + 2 (M + N) Kp
 * + 4 M N + 4 ksplit M N bytes, Kp = K rounded up to 32, each part rounded up to 256) are refused
 * with POE2D_E_INVALID.  Results equal poe2d_gemm_bf16's bit for bit.  Every call on a GPU shares
 * that GPU's scratch: calls that overlap on one GPU (from several host threads) race. */
poe2d_status poe2d_gemm_bf16_noalloc(int M, int N, int K, const float *A, const float *B, float *C, int ksplit);
void poe2d_mlp_destroy(poe2d_mlp *m);
/* Page-locked host memory (no counterpart in the reference), for large results copied from the
 * device such as a screen's predicted gathers: copies into it run at the link's DMA rate and its
 * pages are resident, so no first-touch faults slow a copy down.  Freed by poe2d_host_free. */
poe2d_status poe2d_host_alloc(size_t bytes, void **out);
void poe2d_host_free(void *p);
/* Device memory on the calling thread's current device (no counterpart in the reference), for
 * results that stay in HBM for a consumer on the GPU, such as a screen's gathers; freed by
 * poe2d_device_free. */
poe2d_status poe2d_device_alloc(size_t bytes, void **out);
void poe2d_device_free(void *p);
/* Copy bytes from src to dst, each host or device memory (the direction comes from the
 * pointers); returns when the copy is done. */
poe2d_status poe2d_copy(void *dst, const void *src, size_t bytes);

/* ---- Reference-format input and output (host only; no device needed) ---- */

/* Layout of a float32 model grid file: x fastest (arr2bin.m:4-9 applied to a z-by-x matrix, the
 * order the reference's fread loops expect, seismic_CPML_2D_poroelastic_withCUDA.cu:603-640), or
 * z fastest (patchymodel.m:48-50, which reshapes the matrix column-major before arr2bin). */
#define POE2D_X_FASTEST 0
#define POE2D_Z_FASTEST 1

/* input1.txt: NSTEP, ISOURCE, JSOURCE, a0_pos, a1_pos as whitespace-separated integers, read in
 * that order (seismic_CPML_2D_poroelastic_withCUDA.cu:472-477, 525, 752-762).  POE2D_E_IO when
 * the file is missing or short. */
poe2d_status poe2d_read_input1(const char *path, int *nstep, int *isource, int *jsource,
                               int *a0_pos, int *a1_pos);

/* input2.txt: the source frequency f0, read as a float (:479-484, 827-836). */
poe2d_status poe2d_read_input2(const char *path, double *f0);

/* An (nx1 = NX+1) x (nz1 = NZ+1) float32 little-endian grid into out (x fastest).  The file must
 * hold exactly nx1 * nz1 values (the reference freads that many without checking, :603-640). */
poe2d_status poe2d_read_grid_f32(const char *path, int nx1, int nz1, int layout, float *out);

/* A full-grid float32 snapshot, x fastest: the vz%05i.vzo / vx%05i.vxo files (:1648-1661,
 * 1700-1713). */
poe2d_status poe2d_write_grid_f32(const char *path, int nx1, int nz1, const float *data);

/* ASCII traces as the reference writes them with default stream formatting (6 significant
 * digits), each value followed by a space: a table of nrows lines of ncols values each ending in
 * a newline (.vsp / .hor, :1749-1805), or with one_line a single line without newline (.ats /
 * .bwh / .src, :1846-1860). */
poe2d_status poe2d_write_traces(const char *path, int nrows, int ncols, const float *data, int one_line);

/* The reference's receiver layout (save1_kernel, :114-201): the VSP line at column ISOURCE, rows
 * y = m (NZ+1)/panj - 1, m = 1..panj (n_vsp positions, ending at the grid's last row); the
 * horizontal line at row 10, columns x = m (NX+1)/panj2 - 1, m = 1..panj2 (n_hor); and the point
 * traces at column (NX+1)/2 - 1, rows a0_pos, a1_pos, JSOURCE (n_pts = 3), written to ix/iz in
 * that order.  cap = 0 only returns the counts.  The reference records v_z and v_x on both lines
 * and v_z at the points, sample k holding the state at the start of step k. */
poe2d_status poe2d_reference_receivers(int nx, int nz, int isource, int jsource, int a0_pos,
                                       int a1_pos, int panj, int panj2, int cap, int *ix, int *iz,
                                       int *n_vsp, int *n_hor, int *n_pts);

#ifdef __cplusplus
}
#endif

#endif /* POE2D_H */
