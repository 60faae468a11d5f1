// Two-pass time step: a stress pass (strain / fluid-content / shear-stress integration with their
// C-PML memory variables, pore pressure and normal stresses, written to the intermediate group I)
// and a velocity pass (solid and relative-fluid velocities with Biot inertial coupling and the
// Crank-Nicolson viscous term).  Both work in place on the current state buffer.  This schedule is
// kept beside the single-pass kernel as an independent GPU evaluation of the same step.
//
// One thread per grid point.  C-PML memory variables live only in compact strip arrays: columns
// with a nonzero x profile (xmap >= 0) and rows with a nonzero z profile (zmap >= 0).  Outside the
// strips a = 0, b = 1 and K = 1, so the memory variable stays zero and the C-PML derivative
// d/K + psi equals d exactly (the tables hold 1/K); the interior therefore skips the memory update and the division.
#include "step_common.h"

namespace poe2d {

template <class T, bool AVG>
__global__ void __launch_bounds__(256) stress_pass(StepArgs<T> A)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y * blockDim.y + threadIdx.y;
    if (x > A.nx || z > A.nz) return;
    const long long P = A.pitch;
    const long long o = (long long)z * P + x;
    const Vec4<T> *V = state_group(A, group_v(A.in));
    Vec4<T> *S = state_group(A, group_s(A.in));
    Vec4<T> *I = state_group(A, G_I);
    const Vec4<T> *MS = mat_group(A, G_MS);
    const int sx = A.xmap[x];
    const int sz = A.zmap[z];
    const int b = A.in;
    const T *tab = A.tab;
    const int tn = A.tn;
    Vec4<T> st = S[o];

    if (x >= 1 && x < A.nx && z >= 2 && z < A.nz - 1) {
        T dvx = d4(V[o + 1].x, V[o].x, V[o + 2].x, V[o - 1].x, A.rdx);
        T dwx = d4(V[o + 1].z, V[o].z, V[o + 2].z, V[o - 1].z, A.rdx);
        T dvz = d4(V[o].y, V[o - P].y, V[o + P].y, V[o - 2 * P].y, A.rdz);
        T dwz = d4(V[o].w, V[o - P].w, V[o + P].w, V[o - 2 * P].w, A.rdz);
        if (sx >= 0) {
            const long long k = (long long)z * A.wxs + sx;
            dvx = cpml(dvx, psi_x(A, b, k) + PX_DVX, psi_x(A, b, k) + PX_DVX, true, tab[TB_AXH * tn + x], tab[TB_BXH * tn + x], tab[TB_RKXH * tn + x]);
            dwx = cpml(dwx, psi_x(A, b, k) + PX_DWX, psi_x(A, b, k) + PX_DWX, true, tab[TB_AXH * tn + x], tab[TB_BXH * tn + x], tab[TB_RKXH * tn + x]);
        }
        if (sz >= 0) {
            const long long k = (long long)sz * P + x;
            dvz = cpml(dvz, psi_z(A, b, k) + PZ_DVZ, psi_z(A, b, k) + PZ_DVZ, true, tab[TB_AZ * tn + z], tab[TB_BZ * tn + z], tab[TB_RKZ * tn + z]);
            dwz = cpml(dwz, psi_z(A, b, k) + PZ_DWZ, psi_z(A, b, k) + PZ_DWZ, true, tab[TB_AZ * tn + z], tab[TB_BZ * tn + z], tab[TB_RKZ * tn + z]);
        }
        st.x = st.x + A.dt * dvx;
        st.y = st.y + A.dt * dvz;
        st.z = (st.z - dwx * A.dt) - dwz * A.dt;
        const Vec4<T> m = MS[o];
        T p = pore_pressure(st.x, st.y, st.z, m.x, m.y);
        if (A.src_kind == POE2D_SRC_PRESSURE_DGAUSS && x == A.src_x && z == A.src_z) p += A.src_val;
        T sxx = 0, syy = 0;
        if (x < A.nx - 1) {  // stress3 range minus its row NZ-1, where strains and p stay zero
            const T l2m = m.z + T(2) * m.w;
            sxx = l2m * st.x + m.z * st.y - m.x * p;
            syy = m.z * st.x + l2m * st.y - m.x * p;
        }
        I[o] = Vec4<T>{p, sxx, syy, T(0)};
    }
    if (x >= 2 && x < A.nx && z >= 1 && z < A.nz - 1) {
        T ddx = d4(V[o].y, V[o - 1].y, V[o + 1].y, V[o - 2].y, A.rdx);
        T ddz = d4(V[o + P].x, V[o].x, V[o + 2 * P].x, V[o - P].x, A.rdz);
        if (sx >= 0) {
            const long long k = (long long)z * A.wxs + sx;
            ddx = cpml(ddx, psi_x(A, b, k) + PX_DVY, psi_x(A, b, k) + PX_DVY, true, tab[TB_AX * tn + x], tab[TB_BX * tn + x], tab[TB_RKX * tn + x]);
        }
        if (sz >= 0) {
            // memory variable on the half grid, division by the full-grid K_z as written in the
            // reference (seismic_CPML_2D_poroelastic_withCUDA.cu:224-226)
            const long long k = (long long)sz * P + x;
            ddz = cpml(ddz, psi_z(A, b, k) + PZ_DVX, psi_z(A, b, k) + PZ_DVX, true, tab[TB_AZH * tn + z], tab[TB_BZH * tn + z], tab[TB_RKZ * tn + z]);
        }
        const T mu_xy = AVG ? harmonic_mu(MS[o].w, MS[o + P].w) : MS[o + P].w;
        st.w = st.w + mu_xy * (ddx + ddz) * A.dt;
    }
    S[o] = st;
}

template <class T, bool AVG>
__global__ void __launch_bounds__(256) velocity_pass(StepArgs<T> A)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int z = blockIdx.y * blockDim.y + threadIdx.y;
    if (x > A.nx || z > A.nz) return;
    const long long P = A.pitch;
    const long long o = (long long)z * P + x;
    Vec4<T> *V = state_group(A, group_v(A.in));
    const Vec4<T> *S = state_group(A, group_s(A.in));
    const Vec4<T> *I = state_group(A, G_I);
    const Vec4<T> *MV = mat_group(A, G_MV);
    const int sx = A.xmap[x];
    const int sz = A.zmap[z];
    const int b = A.in;
    const T *tab = A.tab;
    const int tn = A.tn;
    Vec4<T> v = V[o];

    if (x >= 2 && x < A.nx && z >= 2 && z < A.nz) {
        T e1 = d4(I[o].y, I[o - 1].y, I[o + 1].y, I[o - 2].y, A.rdx);
        T e2 = d4(I[o].x, I[o - 1].x, I[o + 1].x, I[o - 2].x, A.rdx);
        T e3 = d4(S[o].w, S[o - P].w, S[o + P].w, S[o - 2 * P].w, A.rdz);
        if (sx >= 0) {
            const long long k = (long long)z * A.wxs + sx;
            e1 = cpml(e1, psi_x(A, b, k) + PX_DSXX, psi_x(A, b, k) + PX_DSXX, true, tab[TB_AX * tn + x], tab[TB_BX * tn + x], tab[TB_RKX * tn + x]);
            e2 = cpml(e2, psi_x(A, b, k) + PX_DPX, psi_x(A, b, k) + PX_DPX, true, tab[TB_AX * tn + x], tab[TB_BX * tn + x], tab[TB_RKX * tn + x]);
        }
        if (sz >= 0) {
            const long long k = (long long)sz * P + x;
            e3 = cpml(e3, psi_z(A, b, k) + PZ_DSXY, psi_z(A, b, k) + PZ_DSXY, true, tab[TB_AZ * tn + z], tab[TB_BZ * tn + z], tab[TB_RKZ * tn + z]);
        }
        const Vec4<T> m = MV[o];
        biot_velocity(v.x, v.z, e1 + e3, e2, m.x, m.y, m.z, m.w, A.inv_dt);
    }
    if (x >= 1 && x < A.nx - 1 && z >= 1 && z < A.nz - 1) {
        T e1 = d4(S[o + 1].w, S[o].w, S[o + 2].w, S[o - 1].w, A.rdx);
        T e2 = d4(I[o + P].z, I[o].z, I[o + 2 * P].z, I[o - P].z, A.rdz);
        T e3 = d4(I[o + P].x, I[o].x, I[o + 2 * P].x, I[o - P].x, A.rdz);
        if (sx >= 0) {
            const long long k = (long long)z * A.wxs + sx;
            e1 = cpml(e1, psi_x(A, b, k) + PX_DSXY, psi_x(A, b, k) + PX_DSXY, true, tab[TB_AXH * tn + x], tab[TB_BXH * tn + x], tab[TB_RKXH * tn + x]);
        }
        if (sz >= 0) {
            const long long k = (long long)sz * P + x;
            e2 = cpml(e2, psi_z(A, b, k) + PZ_DSYY, psi_z(A, b, k) + PZ_DSYY, true, tab[TB_AZH * tn + z], tab[TB_BZH * tn + z], tab[TB_RKZH * tn + z]);
            e3 = cpml(e3, psi_z(A, b, k) + PZ_DPZ, psi_z(A, b, k) + PZ_DPZ, true, tab[TB_AZH * tn + z], tab[TB_BZH * tn + z], tab[TB_RKZH * tn + z]);
        }
        T div = e1 + e2;
        if (A.src_kind == POE2D_SRC_FORCE_Y_RICKER && x == A.src_x && z == A.src_z) div += A.src_val;
        // material of point (x, z+1) (reference :330-333), or the mean of (x, z) and (x, z+1)
        const Vec4<T> m = AVG ? mean_rows(MV[o], MV[o + P]) : MV[o + P];
        biot_velocity(v.y, v.w, div, e3, m.x, m.y, m.z, m.w, A.inv_dt);
    }
    V[o] = v;
}

template <class T>
void launch_twopass(const StepArgs<T> &A, hipStream_t stream, int phase, bool avg)
{
    dim3 block(64, 4);
    dim3 grid((A.nx + 1 + 63) / 64, (A.nz + 1 + 3) / 4);
    if (phase == 0) {
        if (avg) stress_pass<T, true><<<grid, block, 0, stream>>>(A);
        else stress_pass<T, false><<<grid, block, 0, stream>>>(A);
    } else {
        if (avg) velocity_pass<T, true><<<grid, block, 0, stream>>>(A);
        else velocity_pass<T, false><<<grid, block, 0, stream>>>(A);
    }
}

template void launch_twopass<float>(const StepArgs<float> &, hipStream_t, int, bool);
template void launch_twopass<double>(const StepArgs<double> &, hipStream_t, int, bool);

}  // namespace poe2d
