// Single-pass time step: one launch per step computes the stress update and the velocity update
// of a tile without writing the intermediate stresses to memory.
//
// Work decomposition.  A workgroup of BW threads owns a column strip of W = BW - 8 output columns
// and th output rows.  Thread `lane` holds column c = x0 - 4 + lane, so the block spans the tile
// plus a 4-column halo on each side.  The block marches down its rows; at step s it
//   1. loads row s+2 of the old velocities (vx, vy, wx, wy) into a per-thread 5-row window,
//   2. forms the stresses of row s for every held column (strains, fluid content, shear stress,
//      pore pressure and normal stresses, with their C-PML memory variables), taking x
//      neighbours through LDS and z neighbours from the register window,
//   3. forms the velocities of row j = s-2 from a per-thread window of stress rows s-4..s and the
//      x neighbours of row j through LDS.
// Stresses of the 2 halo columns and 2 halo rows on each side are recomputed from the old state
// (double buffering keeps it intact) instead of being exchanged, so each point's 8 state fields
// and 8 coefficients cross HBM once per step.  Tiles are mapped so that consecutive tiles of a
// row band land on the same XCD (blocks are dispatched round-robin over the 8 XCDs), keeping halo
// re-reads inside one L2.
#include "step_common.h"

namespace poe2d {

namespace {

template <class T>
__device__ __forceinline__ T ld(const T *p, ptrdiff_t i, bool ok)
{
    return ok ? p[i] : T(0);
}

// C-PML derivative with double-buffered memory variable: psi_new = b psi_old + a d,
// returns d / K + psi_new; psi_new is stored only by the tile that owns the point.
template <class T>
__device__ __forceinline__ T cpml_db(T d, const T *psi_o, T *psi_n, ptrdiff_t k, bool store, T a, T b, T K)
{
    const T s = b * psi_o[k] + a * d;
    if (store) psi_n[k] = s;
    return d / K + s;
}

}  // namespace

template <class T, int BW>
__global__ void __launch_bounds__(BW) fused_step(FusedArgs<T> A)
{
    constexpr int W = BW - 8;
    __shared__ T xv[3][BW + 4];  // row s: vx, wx, vy of every held column (entry lane + 2)
    __shared__ T xs[3][BW + 4];  // row j: sxx, p, sxy

    const int tile = (int)(blockIdx.x & 7) * A.per_xcd + (int)(blockIdx.x >> 3);
    if (tile >= A.ntiles) return;
    const int tx = tile % A.ntx;
    const int tz = tile / A.ntx;
    const int lane = threadIdx.x;
    const int nx = A.nx, nz = A.nz;
    const int c = tx * W - 4 + lane;
    const int zA = tz * A.th;
    const int zB = min(zA + A.th, nz + 1);
    const ptrdiff_t P = A.pitch;
    const bool in_mem = c >= 0 && c <= nx + 1;
    const bool own = lane >= 4 && lane < BW - 4 && c <= nx;
    const T rdx = A.rdx, rdz = A.rdz, dt = A.dt;

    // per-column C-PML profile (x strips only)
    const int sx = (c >= 0 && c <= nx) ? A.xmap[c] : -1;
    T ax = 0, bx = 1, Kx = 1, axh = 0, bxh = 1, Kxh = 1;
    if (sx >= 0) {
        ax = A.ax[c]; bx = A.bx[c]; Kx = A.Kx[c];
        axh = A.axh[c]; bxh = A.bxh[c]; Kxh = A.Kxh[c];
    }
    // column ranges of the reference kernels (seismic_CPML_2D_poroelastic_withCUDA.cu)
    const bool xr1 = c >= 1 && c < nx;      // stress1 (:60)
    const bool xr2 = c >= 2 && c < nx;      // stress2 (:212)
    const bool xr3 = c >= 1 && c < nx - 1;  // stress3 (:237)
    const bool xr4 = c >= 2 && c < nx;      // vx      (:289)
    const bool xr5 = c >= 1 && c < nx - 1;  // vz      (:326)
    const bool src_p = A.src_kind == POE2D_SRC_PRESSURE_DGAUSS && c == A.src_x;
    const bool src_f = A.src_kind == POE2D_SRC_FORCE_Y_RICKER && c == A.src_x;

    // velocity window: slot k holds row s-2+k
    T vxw[5], vzw[5], wxw[5], wzw[5];
    vxw[0] = vzw[0] = wxw[0] = wzw[0] = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const ptrdiff_t o = (ptrdiff_t)(zA - 4 + k) * P + c;
        vxw[k + 1] = ld(A.vx_o, o, in_mem);
        vzw[k + 1] = ld(A.vz_o, o, in_mem);
        wxw[k + 1] = ld(A.wx_o, o, in_mem);
        wzw[k + 1] = ld(A.wz_o, o, in_mem);
    }
    // stress windows: sxy rows s-4..s, syy and p rows s-3..s, sxx rows s-2..s
    T sxyw[5] = {0, 0, 0, 0, 0}, syyw[4] = {0, 0, 0, 0}, pw[4] = {0, 0, 0, 0}, sxxw[3] = {0, 0, 0};
    T mu_cur = ld(A.mu, (ptrdiff_t)(zA - 2) * P + c, in_mem);
    T rho_c = 0, rhof_c = 0, m_c = 0, b_c = 0;  // coefficients of row j   (vx update)
    T rho_n = 0, rhof_n = 0, m_n = 0, b_n = 0;  // coefficients of row j+1 (vy update)

    for (int s = zA - 2; s < zB + 2; ++s) {
        // ---- 1. slide the velocity window down one row
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            vxw[k] = vxw[k + 1]; vzw[k] = vzw[k + 1];
            wxw[k] = wxw[k + 1]; wzw[k] = wzw[k + 1];
        }
        {
            const ptrdiff_t o = (ptrdiff_t)(s + 2) * P + c;
            vxw[4] = ld(A.vx_o, o, in_mem);
            vzw[4] = ld(A.vz_o, o, in_mem);
            wxw[4] = ld(A.wx_o, o, in_mem);
            wzw[4] = ld(A.wz_o, o, in_mem);
        }

        // ---- 2. stresses of row s
        const ptrdiff_t os = (ptrdiff_t)s * P + c;
        const bool srow = s >= zA && s < zB;
        const bool zr1 = s >= 2 && s < nz - 1;
        const bool zr2 = s >= 1 && s < nz - 1;
        const int szs = (s >= 0 && s <= nz) ? A.zmap[s] : -1;
        const T exx_o = ld(A.exx_o, os, in_mem), eyy_o = ld(A.eyy_o, os, in_mem);
        const T zeta_o = ld(A.zeta_o, os, in_mem), sxy_o = ld(A.sxy_o, os, in_mem);
        const T al = ld(A.al, os, in_mem), Mb = ld(A.M, os, in_mem), lc = ld(A.lc, os, in_mem);
        const T mu_next = ld(A.mu, os + P, in_mem);

        xv[0][lane + 2] = vxw[2];
        xv[1][lane + 2] = wxw[2];
        xv[2][lane + 2] = vzw[2];
        __syncthreads();

        T exx_n = exx_o, eyy_n = eyy_o, zeta_n = zeta_o, sxy_n = sxy_o;
        T p = 0, sxx = 0, syy = 0;
        const bool store_s = own && srow;
        if (xr1 && zr1) {
            T dvx = d4(xv[0][lane + 3], vxw[2], xv[0][lane + 4], xv[0][lane + 1], rdx);
            T dwx = d4(xv[1][lane + 3], wxw[2], xv[1][lane + 4], xv[1][lane + 1], rdx);
            T dvz = d4(vzw[2], vzw[1], vzw[3], vzw[0], rdz);
            T dwz = d4(wzw[2], wzw[1], wzw[3], wzw[0], rdz);
            if (sx >= 0) {
                const ptrdiff_t k = (ptrdiff_t)s * A.wxs + sx;
                dvx = cpml_db(dvx, A.psx_o[PX_DVX], A.psx_n[PX_DVX], k, store_s, axh, bxh, Kxh);
                dwx = cpml_db(dwx, A.psx_o[PX_DWX], A.psx_n[PX_DWX], k, store_s, axh, bxh, Kxh);
            }
            if (szs >= 0) {
                const ptrdiff_t k = (ptrdiff_t)szs * P + c;
                const T az = A.az[s], bz = A.bz[s], Kz = A.Kz[s];
                dvz = cpml_db(dvz, A.psz_o[PZ_DVZ], A.psz_n[PZ_DVZ], k, store_s, az, bz, Kz);
                dwz = cpml_db(dwz, A.psz_o[PZ_DWZ], A.psz_n[PZ_DWZ], k, store_s, az, bz, Kz);
            }
            exx_n = exx_o + dt * dvx;
            eyy_n = eyy_o + dt * dvz;
            zeta_n = (zeta_o - dwx * dt) - dwz * dt;
            p = pore_pressure(exx_n, eyy_n, zeta_n, al, Mb);
            if (src_p && s == A.src_z) p += A.src_val;
            if (xr3) {
                const T l2m = lc + T(2) * mu_cur;
                sxx = l2m * exx_n + lc * eyy_n - al * p;
                syy = lc * exx_n + l2m * eyy_n - al * p;
            }
        }
        if (xr2 && zr2) {
            T ddx = d4(vzw[2], xv[2][lane + 1], xv[2][lane + 3], xv[2][lane], rdx);
            T ddz = d4(vxw[3], vxw[2], vxw[4], vxw[1], rdz);
            if (sx >= 0) {
                const ptrdiff_t k = (ptrdiff_t)s * A.wxs + sx;
                ddx = cpml_db(ddx, A.psx_o[PX_DVY], A.psx_n[PX_DVY], k, store_s, ax, bx, Kx);
            }
            if (szs >= 0) {
                // half-grid memory coefficients, full-grid K_z divisor (reference :224-226)
                const ptrdiff_t k = (ptrdiff_t)szs * P + c;
                ddz = cpml_db(ddz, A.psz_o[PZ_DVX], A.psz_n[PZ_DVX], k, store_s, A.azh[s], A.bzh[s], A.Kz[s]);
            }
            sxy_n = sxy_o + mu_next * (ddx + ddz) * dt;
        }
        if (store_s) {
            A.exx_n[os] = exx_n;
            A.eyy_n[os] = eyy_n;
            A.zeta_n[os] = zeta_n;
            A.sxy_n[os] = sxy_n;
        }
        mu_cur = mu_next;
#pragma unroll
        for (int k = 0; k < 4; ++k) sxyw[k] = sxyw[k + 1];
        sxyw[4] = sxy_n;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            syyw[k] = syyw[k + 1];
            pw[k] = pw[k + 1];
        }
        syyw[3] = syy;
        pw[3] = p;
        sxxw[0] = sxxw[1];
        sxxw[1] = sxxw[2];
        sxxw[2] = sxx;

        // ---- 3. velocities of row j = s - 2
        const int j = s - 2;
        const bool vrow = j >= zA && j < zB;
        xs[0][lane + 2] = sxxw[0];
        xs[1][lane + 2] = pw[1];
        xs[2][lane + 2] = sxyw[2];
        rho_c = rho_n; rhof_c = rhof_n; m_c = m_n; b_c = b_n;
        {
            const ptrdiff_t o1 = (ptrdiff_t)(s - 1) * P + c;
            rho_n = ld(A.rho, o1, own);
            rhof_n = ld(A.rhof, o1, own);
            m_n = ld(A.m, o1, own);
            b_n = ld(A.b, o1, own);
        }
        __syncthreads();
        if (vrow) {
            const int szj = A.zmap[j];
            const ptrdiff_t oj = (ptrdiff_t)j * P + c;
            T vxn = vxw[0], wxn = wxw[0], vzn = vzw[0], wzn = wzw[0];
            if (own && xr4 && j >= 2 && j < nz) {
                T e1 = d4(xs[0][lane + 2], xs[0][lane + 1], xs[0][lane + 3], xs[0][lane], rdx);
                T e2 = d4(xs[1][lane + 2], xs[1][lane + 1], xs[1][lane + 3], xs[1][lane], rdx);
                T e3 = d4(sxyw[2], sxyw[1], sxyw[3], sxyw[0], rdz);
                if (sx >= 0) {
                    const ptrdiff_t k = (ptrdiff_t)j * A.wxs + sx;
                    e1 = cpml_db(e1, A.psx_o[PX_DSXX], A.psx_n[PX_DSXX], k, true, ax, bx, Kx);
                    e2 = cpml_db(e2, A.psx_o[PX_DPX], A.psx_n[PX_DPX], k, true, ax, bx, Kx);
                }
                if (szj >= 0) {
                    const ptrdiff_t k = (ptrdiff_t)szj * P + c;
                    e3 = cpml_db(e3, A.psz_o[PZ_DSXY], A.psz_n[PZ_DSXY], k, true, A.az[j], A.bz[j], A.Kz[j]);
                }
                biot_velocity(vxn, wxn, e1 + e3, e2, rho_c, rhof_c, m_c, b_c, A.inv_dt);
            }
            if (own && xr5 && j >= 1 && j < nz - 1) {
                T e1 = d4(xs[2][lane + 3], xs[2][lane + 2], xs[2][lane + 4], xs[2][lane + 1], rdx);
                T e2 = d4(syyw[2], syyw[1], syyw[3], syyw[0], rdz);
                T e3 = d4(pw[2], pw[1], pw[3], pw[0], rdz);
                if (sx >= 0) {
                    const ptrdiff_t k = (ptrdiff_t)j * A.wxs + sx;
                    e1 = cpml_db(e1, A.psx_o[PX_DSXY], A.psx_n[PX_DSXY], k, true, axh, bxh, Kxh);
                }
                if (szj >= 0) {
                    const ptrdiff_t k = (ptrdiff_t)szj * P + c;
                    const T azh = A.azh[j], bzh = A.bzh[j], Kzh = A.Kzh[j];
                    e2 = cpml_db(e2, A.psz_o[PZ_DSYY], A.psz_n[PZ_DSYY], k, true, azh, bzh, Kzh);
                    e3 = cpml_db(e3, A.psz_o[PZ_DPZ], A.psz_n[PZ_DPZ], k, true, azh, bzh, Kzh);
                }
                T div = e1 + e2;
                if (src_f && j == A.src_z) div += A.src_val;
                biot_velocity(vzn, wzn, div, e3, rho_n, rhof_n, m_n, b_n, A.inv_dt);
            }
            if (own) {
                A.vx_n[oj] = vxn;
                A.wx_n[oj] = wxn;
                A.vz_n[oj] = vzn;
                A.wz_n[oj] = wzn;
            }
        }
    }
}

template <class T>
hipError_t launch_fused(const FusedArgs<T> &A, int bw, hipStream_t stream)
{
    const dim3 grid(8 * A.per_xcd);
    switch (bw) {
    case 128: fused_step<T, 128><<<grid, 128, 0, stream>>>(A); break;
    case 256: fused_step<T, 256><<<grid, 256, 0, stream>>>(A); break;
    case 512: fused_step<T, 512><<<grid, 512, 0, stream>>>(A); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_fused<float>(const FusedArgs<float> &, int, hipStream_t);
template hipError_t launch_fused<double>(const FusedArgs<double> &, int, hipStream_t);

}  // namespace poe2d
