// Device-side building blocks shared by the time-step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/poe2d.h"

namespace poe2d {

// C-PML memory variables kept in the x strips (columns with a nonzero x profile), one array each,
// indexed [z][strip column].  The coefficient set each one uses is noted.
enum {
    PX_DVX = 0,   // d vx / dx in the strain update          (half-grid x coefficients)
    PX_DWX = 1,   // d wx / dx in the fluid-content update   (half-grid x)
    PX_DVY = 2,   // d vy / dx in the shear-stress update    (grid x)
    PX_DSXX = 3,  // d sxx / dx in the vx update             (grid x)
    PX_DPX = 4,   // d p / dx in the wx update               (grid x)
    PX_DSXY = 5,  // d sxy / dx in the vy update             (half-grid x)
    NPSI_X = 6
};
// ... and in the z strips (rows with a nonzero z profile), indexed [strip row][x].
enum {
    PZ_DVZ = 0,   // d vy / dz in the strain update          (grid z)
    PZ_DWZ = 1,   // d wy / dz in the fluid-content update   (grid z)
    PZ_DVX = 2,   // d vx / dz in the shear-stress update    (half-grid z)
    PZ_DSXY = 3,  // d sxy / dz in the vx update             (grid z)
    PZ_DSYY = 4,  // d syy / dz in the vy update             (half-grid z)
    PZ_DPZ = 5,   // d p / dz in the wy update               (half-grid z)
    NPSI_Z = 6
};

template <class T>
struct StepArgs {
    int nx, nz;        // NX, NZ (largest point indices)
    int pitch;         // elements between rows (>= nx + 2, pad columns zero)
    int wxs;           // number of x-strip columns
    T rdx, rdz;        // 1 / (24 dx), 1 / (24 dz)
    T dt, inv_dt;
    // state
    T *vx, *vz, *wx, *wz, *exx, *eyy, *zeta, *sxy;
    // intermediates of the two-pass schedule
    T *p, *sxx, *syy;
    // material
    const T *rho, *rhof, *m, *mu, *lc, *M, *al, *b;
    // C-PML memory variables and 1-D profiles
    T *psx[NPSI_X];
    T *psz[NPSI_Z];
    const T *ax, *bx, *Kx, *axh, *bxh, *Kxh;
    const T *az, *bz, *Kz, *azh, *bzh, *Kzh;
    const int *xmap, *zmap;
    // source of this step
    int src_kind, src_x, src_z;
    T src_val;
};

// Fourth-order staggered difference (27 (f1 - f0) - (f2 - fm)) / (24 h) between f0 and f1, with
// f2 the far point beyond f1 and fm the far point before f0; rh = 1 / (24 h).
template <class T>
__device__ __forceinline__ T d4(T f1, T f0, T f2, T fm, T rh)
{
    return (T(27) * (f1 - f0) - (f2 - fm)) * rh;
}

// C-PML correction of a derivative d inside a strip: psi <- b psi + a d, returns d / K + psi.
template <class T>
__device__ __forceinline__ T cpml(T d, T *psi, T a, T b, T K)
{
    const T s = b * *psi + a * d;
    *psi = s;
    return d / K + s;
}

// Biot momentum update of one velocity point: solid velocity v and relative fluid velocity w
// from the stress divergence `div` and the pressure gradient `grad`,
//   [rho rho_f; rho_f m] d/dt [v; w] = [div; -grad - b w],
// with the viscous term b w taken Crank-Nicolson ((w_old + w_new) / 2), as the reference's vx and
// vz kernels do (seismic_CPML_2D_poroelastic_withCUDA.cu:293-311).
template <class T>
__device__ __forceinline__ void biot_velocity(T &v, T &w, T div, T grad, T rho, T rhof, T m, T b,
                                              T inv_dt)
{
    const T co = (rho * m - rhof * rhof) * inv_dt;
    const T half_rb = rho * b * T(0.5);
    const T c1 = co + half_rb;
    const T c2 = co - half_rb;
    const T wold = w;
    const T wnew = (c2 * wold + (-rhof * div - rho * grad)) / c1;
    const T wavg = (wold + wnew) * T(0.5);
    v = v + (m * div + rhof * grad + rhof * b * wavg) / co;
    w = wnew;
}

// Pore pressure from strains and fluid content, p = -alpha M (e_xx + e_yy) + M zeta
// (seismic_CPML_2D_poroelastic_withCUDA.cu:86 with zeta = xi_1 + xi_2).  Shared by the step
// kernels, the receiver gather and snapshots so every consumer rounds p the same way.
template <class T>
__device__ __forceinline__ T pore_pressure(T exx, T eyy, T zeta, T al, T M)
{
    return -al * M * (exx + eyy) + M * zeta;
}

// Arguments of the single-pass kernel: state at step n is read from the `*_o` arrays and the
// state at n+1 written to the `*_n` arrays (double buffering lets neighbouring tiles recompute
// halo stresses from unmodified inputs).  Field pointers address point (0, 0); rows are
// `pitch` apart and the arrays carry zero pad rows/columns around the grid.
template <class T>
struct FusedArgs {
    int nx, nz, pitch, wxs;
    int th;          // output rows per tile
    int ntx;         // tiles across x
    int ntiles;      // tiles in total
    int per_xcd;     // tiles per XCD slot (grid = 8 * per_xcd blocks)
    T rdx, rdz, dt, inv_dt;
    const T *vx_o, *vz_o, *wx_o, *wz_o, *exx_o, *eyy_o, *zeta_o, *sxy_o;
    T *vx_n, *vz_n, *wx_n, *wz_n, *exx_n, *eyy_n, *zeta_n, *sxy_n;
    const T *rho, *rhof, *m, *mu, *lc, *M, *al, *b;
    const T *psx_o[NPSI_X];
    T *psx_n[NPSI_X];
    const T *psz_o[NPSI_Z];
    T *psz_n[NPSI_Z];
    const T *ax, *bx, *Kx, *axh, *bxh, *Kxh;
    const T *az, *bz, *Kz, *azh, *bzh, *Kzh;
    const int *xmap, *zmap;
    int src_kind, src_x, src_z;
    T src_val;
};

template <class T>
void launch_twopass(const StepArgs<T> &A, hipStream_t stream, int phase);

// Launches the single-pass step with block width bw (128, 256 or 512 threads).
template <class T>
hipError_t launch_fused(const FusedArgs<T> &A, int bw, hipStream_t stream);

// Output columns of one tile for block width bw (halo of 4 columns on each side).
inline int fused_tile_width(int bw) { return bw - 8; }

}  // namespace poe2d
