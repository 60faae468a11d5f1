// Two-pass time step: a stress pass (strain / fluid-content / shear-stress integration with their
// C-PML memory variables, pore pressure and normal stresses) and a velocity pass (solid and
// relative-fluid velocities with Biot inertial coupling and Crank-Nicolson viscous term).
//
// One thread per grid point; every field is SoA with a padded row pitch whose pad columns are
// zero and never written.  C-PML memory variables live only in compact strip arrays: columns
// with a nonzero x profile (xmap >= 0) and rows with a nonzero z profile (zmap >= 0).  Outside
// the strips a = 0, b = 1 and K = 1, so the memory variable stays zero and the C-PML derivative
// d/K + psi equals d exactly; the interior therefore skips the memory update and the division.
#include "step_common.h"

namespace poe2d {

// Stress pass of step n.  Index ranges are those of the reference kernels
// (seismic_CPML_2D_poroelastic_withCUDA.cu): stress1 x in [1,NX), z in [2,NZ-1) (:60-62);
// stress2 x in [2,NX), z in [1,NZ-1) (:212-214); stress3 x in [1,NX-1), z in [2,NZ) (:237-239).
template <class T>
__global__ void __launch_bounds__(256) stress_pass(StepArgs<T> A)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y * blockDim.y + threadIdx.y;
    if (x > A.nx || z > A.nz) return;
    const size_t P = A.pitch;
    const size_t o = (size_t)z * P + x;
    const int sx = A.xmap[x];
    const int sz = A.zmap[z];

    const bool r1 = x >= 1 && x < A.nx && z >= 2 && z < A.nz - 1;
    if (r1) {
        T dvx = d4(A.vx[o + 1], A.vx[o], A.vx[o + 2], A.vx[o - 1], A.rdx);
        T dwx = d4(A.wx[o + 1], A.wx[o], A.wx[o + 2], A.wx[o - 1], A.rdx);
        T dvz = d4(A.vz[o], A.vz[o - P], A.vz[o + P], A.vz[o - 2 * P], A.rdz);
        T dwz = d4(A.wz[o], A.wz[o - P], A.wz[o + P], A.wz[o - 2 * P], A.rdz);
        if (sx >= 0) {
            const size_t k = (size_t)z * A.wxs + sx;
            dvx = cpml(dvx, A.psx[PX_DVX] + k, A.axh[x], A.bxh[x], A.Kxh[x]);
            dwx = cpml(dwx, A.psx[PX_DWX] + k, A.axh[x], A.bxh[x], A.Kxh[x]);
        }
        if (sz >= 0) {
            const size_t k = (size_t)sz * P + x;
            dvz = cpml(dvz, A.psz[PZ_DVZ] + k, A.az[z], A.bz[z], A.Kz[z]);
            dwz = cpml(dwz, A.psz[PZ_DWZ] + k, A.az[z], A.bz[z], A.Kz[z]);
        }
        const T exx = A.exx[o] + A.dt * dvx;
        const T eyy = A.eyy[o] + A.dt * dvz;
        const T zeta = (A.zeta[o] - dwx * A.dt) - dwz * A.dt;
        A.exx[o] = exx;
        A.eyy[o] = eyy;
        A.zeta[o] = zeta;
        const T al = A.al[o], M = A.M[o];
        T p = -al * M * (exx + eyy) + M * zeta;
        if (A.src_kind == POE2D_SRC_PRESSURE_DGAUSS && x == A.src_x && z == A.src_z) p += A.src_val;
        A.p[o] = p;
        if (x < A.nx - 1) {  // stress3 range minus its row NZ-1, where strains and p stay zero
            const T lc = A.lc[o], mu = A.mu[o];
            const T l2m = lc + T(2) * mu;
            A.sxx[o] = l2m * exx + lc * eyy - al * p;
            A.syy[o] = lc * exx + l2m * eyy - al * p;
        }
    }
    const bool r2 = x >= 2 && x < A.nx && z >= 1 && z < A.nz - 1;
    if (r2) {
        T ddx = d4(A.vz[o], A.vz[o - 1], A.vz[o + 1], A.vz[o - 2], A.rdx);
        T ddz = d4(A.vx[o + P], A.vx[o], A.vx[o + 2 * P], A.vx[o - P], A.rdz);
        if (sx >= 0) {
            const size_t k = (size_t)z * A.wxs + sx;
            ddx = cpml(ddx, A.psx[PX_DVY] + k, A.ax[x], A.bx[x], A.Kx[x]);
        }
        if (sz >= 0) {
            // memory variable on the half grid, division by the full-grid K_z as written in
            // the reference (seismic_CPML_2D_poroelastic_withCUDA.cu:224-226)
            const size_t k = (size_t)sz * P + x;
            ddz = cpml(ddz, A.psz[PZ_DVX] + k, A.azh[z], A.bzh[z], A.Kz[z]);
        }
        A.sxy[o] = A.sxy[o] + A.mu[o + P] * (ddx + ddz) * A.dt;
    }
}

// Velocity pass.  vx x in [2,NX), z in [2,NZ) (:289-291); vz x in [1,NX-1), z in [1,NZ-1)
// (:326-328) with the material of point (x, z+1) (:330-333).
template <class T>
__global__ void __launch_bounds__(256) velocity_pass(StepArgs<T> A)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y * blockDim.y + threadIdx.y;
    if (x > A.nx || z > A.nz) return;
    const size_t P = A.pitch;
    const size_t o = (size_t)z * P + x;
    const int sx = A.xmap[x];
    const int sz = A.zmap[z];

    if (x >= 2 && x < A.nx && z >= 2 && z < A.nz) {
        T e1 = d4(A.sxx[o], A.sxx[o - 1], A.sxx[o + 1], A.sxx[o - 2], A.rdx);
        T e2 = d4(A.p[o], A.p[o - 1], A.p[o + 1], A.p[o - 2], A.rdx);
        T e3 = d4(A.sxy[o], A.sxy[o - P], A.sxy[o + P], A.sxy[o - 2 * P], A.rdz);
        if (sx >= 0) {
            const size_t k = (size_t)z * A.wxs + sx;
            e1 = cpml(e1, A.psx[PX_DSXX] + k, A.ax[x], A.bx[x], A.Kx[x]);
            e2 = cpml(e2, A.psx[PX_DPX] + k, A.ax[x], A.bx[x], A.Kx[x]);
        }
        if (sz >= 0) {
            const size_t k = (size_t)sz * P + x;
            e3 = cpml(e3, A.psz[PZ_DSXY] + k, A.az[z], A.bz[z], A.Kz[z]);
        }
        biot_velocity(A.vx[o], A.wx[o], e1 + e3, e2, A.rho[o], A.rhof[o], A.m[o], A.b[o], A.inv_dt);
    }
    if (x >= 1 && x < A.nx - 1 && z >= 1 && z < A.nz - 1) {
        const size_t t = o + P;
        T e1 = d4(A.sxy[o + 1], A.sxy[o], A.sxy[o + 2], A.sxy[o - 1], A.rdx);
        T e2 = d4(A.syy[o + P], A.syy[o], A.syy[o + 2 * P], A.syy[o - P], A.rdz);
        T e3 = d4(A.p[o + P], A.p[o], A.p[o + 2 * P], A.p[o - P], A.rdz);
        if (sx >= 0) {
            const size_t k = (size_t)z * A.wxs + sx;
            e1 = cpml(e1, A.psx[PX_DSXY] + k, A.axh[x], A.bxh[x], A.Kxh[x]);
        }
        if (sz >= 0) {
            const size_t k = (size_t)sz * P + x;
            e2 = cpml(e2, A.psz[PZ_DSYY] + k, A.azh[z], A.bzh[z], A.Kzh[z]);
            e3 = cpml(e3, A.psz[PZ_DPZ] + k, A.azh[z], A.bzh[z], A.Kzh[z]);
        }
        T div = e1 + e2;
        if (A.src_kind == POE2D_SRC_FORCE_Y_RICKER && x == A.src_x && z == A.src_z) div += A.src_val;
        biot_velocity(A.vz[o], A.wz[o], div, e3, A.rho[t], A.rhof[t], A.m[t], A.b[t], A.inv_dt);
    }
}

template __global__ void stress_pass<float>(StepArgs<float>);
template __global__ void stress_pass<double>(StepArgs<double>);
template __global__ void velocity_pass<float>(StepArgs<float>);
template __global__ void velocity_pass<double>(StepArgs<double>);

template <class T>
void launch_twopass(const StepArgs<T> &A, hipStream_t stream, int phase)
{
    dim3 block(64, 4);
    dim3 grid((A.nx + 1 + 63) / 64, (A.nz + 1 + 3) / 4);
    if (phase == 0)
        stress_pass<T><<<grid, block, 0, stream>>>(A);
    else
        velocity_pass<T><<<grid, block, 0, stream>>>(A);
}

template void launch_twopass<float>(const StepArgs<float> &, hipStream_t, int);
template void launch_twopass<double>(const StepArgs<double> &, hipStream_t, int);

}  // namespace poe2d
