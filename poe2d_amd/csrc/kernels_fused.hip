// Single-pass time step: one launch per step computes the stress update and the velocity update
// of a tile without writing the intermediate stresses to memory.
//
// Work decomposition.  A workgroup of BW threads owns a column strip of W = BW - 8 output columns
// and th output rows.  Thread `lane` holds column c = x0 - 4 + lane, so the block spans the tile
// plus a 4-column halo on each side.  The block marches down its rows; at row step s it
//   1. slides a per-thread 5-row register window of the old velocities (rows s-2 .. s+2),
//   2. forms the stresses of row s for every held column (strains, fluid content, shear stress,
//      pore pressure and normal stresses, with their C-PML memory variables), taking x
//      neighbours through LDS and z neighbours from the register window,
//   3. forms the velocities of row j = s-2 from per-thread windows of stress rows and the x
//      neighbours of stress row j, which were put in an LDS ring two row steps earlier.
// Each point's state vectors (V, S) are read with one 16-byte load each, prefetched one row step
// ahead, and written with one 16-byte store each; every access is a uniform group base plus one
// 32-bit per-lane offset.  The coefficients come either from the MS and MV groups (one 16-byte
// load each) or, in palette mode, from a one-byte material index per point that selects an
// (MS, MV) entry of a palette held in LDS: 1 byte per point instead of 32.  Stresses of the 2 halo columns
// and 2 halo rows on each side are recomputed from the unmodified old state (double buffering)
// instead of being exchanged.  One barrier per row step.  Tiles are mapped so that consecutive
// tiles of a row band land on the same XCD (blocks are dispatched round-robin over the 8 XCDs),
// keeping halo re-reads inside one L2.
#include <cstdint>

#include "step_common.h"

namespace poe2d {

template <class T, int BW, bool PAL>
__global__ void __launch_bounds__(BW) fused_step(StepArgs<T> A)
{
    constexpr int W = BW - 8;
    constexpr uint32_t VB = sizeof(Vec4<T>);
    __shared__ T lv[2][3][BW + 4];  // ring of 2 velocity rows: vx, wx, vy   (entry lane + 2)
    __shared__ T ls[3][3][BW + 4];  // ring of 3 stress rows:   sxx, p, sxy
    __shared__ Vec4<T> lpal[PAL ? 2 * kMaxPalette : 1];  // MS entries, then MV entries

    const int tile = (int)(blockIdx.x & 7) * A.per_xcd + (int)(blockIdx.x >> 3);
    if (tile >= A.ntiles) return;
    const int tx = tile % A.ntx;
    const int tz = tile / A.ntx;
    const int lane = threadIdx.x;
    const int nx = A.nx, nz = A.nz;
    const int c = tx * W - 4 + lane;
    const int zA = tz * A.th;
    const int zB = min(zA + A.th, nz + 1);
    const int P = A.pitch;
    const bool own = lane >= 4 && lane < BW - 4 && c <= nx;
    const T rdx = A.rdx, rdz = A.rdz, dt = A.dt;

    // Group bases (start of each group's allocation) and byte offsets: point (x, z) of a group sits
    // at ((z + kPadRows) * P + x + kPadCols) * VB.  Columns left of the grid (c >= -4) are zero pads
    // in every group, and so is column nx+1; lanes right of it read column nx+1, so every load is
    // unconditional.
    const char *const gb = (const char *)A.grp;
    const long long gsb = A.group_stride * (long long)VB;
    const char *MS = gb + G_MS * gsb;
    const char *MV = gb + G_MV * gsb;
    const char *Vo = gb + group_v(A.in) * gsb;
    const char *So = gb + group_s(A.in) * gsb;
    char *Vn = (char *)gb + group_v(A.out) * gsb;
    char *Sn = (char *)gb + group_s(A.out) * gsb;
    const uint32_t colb = (uint32_t)(min(c, nx + 1) + kPadCols) * VB;
    const uint32_t rowb = (uint32_t)P * VB;
    auto off = [&](int row) -> uint32_t { return colb + (uint32_t)(row + kPadRows) * rowb; };
    auto ld = [&](const char *g, int row) -> Vec4<T> { return *(const Vec4<T> *)(g + off(row)); };
    const Vec4<T> zero4 = {T(0), T(0), T(0), T(0)};
    // palette mode: material index of a row (same pads as the groups, one byte per point)
    const uint32_t colm = (uint32_t)(min(c, nx + 1) + kPadCols);
    auto ldm = [&](int row) -> int { return A.mid[colm + (uint32_t)(row + kPadRows) * (uint32_t)P]; };
    if constexpr (PAL) {
        for (int i = lane; i < 2 * kMaxPalette; i += BW)
            if ((i & (kMaxPalette - 1)) < A.npal) lpal[i] = A.pal[i];
        __syncthreads();
    }
    auto ms_at = [&](int id) -> Vec4<T> { return lpal[id]; };
    auto mv_at = [&](int id) -> Vec4<T> { return lpal[kMaxPalette + id]; };

    // per-column C-PML profile (x strips only)
    const int sx = (c >= 0 && c <= nx) ? A.xmap[c] : -1;
    T ax = 0, bx = 1, rkx = 1, axh = 0, bxh = 1, rkxh = 1;
    if (sx >= 0) {
        ax = A.tab[TB_AX * A.tn + c]; bx = A.tab[TB_BX * A.tn + c]; rkx = A.tab[TB_RKX * A.tn + c];
        axh = A.tab[TB_AXH * A.tn + c]; bxh = A.tab[TB_BXH * A.tn + c]; rkxh = A.tab[TB_RKXH * A.tn + c];
    }
    const bool xr1 = c >= 1 && c < nx;      // stress1
    const bool xr2 = c >= 2 && c < nx;      // stress2
    const bool xr3 = c >= 1 && c < nx - 1;  // stress3
    const bool xr4 = c >= 2 && c < nx;      // vx
    const bool xr5 = c >= 1 && c < nx - 1;  // vz
    const bool src_p = A.src_kind == POE2D_SRC_PRESSURE_DGAUSS && c == A.src_x;
    const bool src_f = A.src_kind == POE2D_SRC_FORCE_Y_RICKER && c == A.src_x;

    // velocity window: slot k holds row s-2+k as (vx, vy, wx, wy)
    Vec4<T> vw[5];
    vw[0] = zero4;
#pragma unroll
    for (int k = 0; k < 4; ++k) vw[k + 1] = ld(Vo, zA - 4 + k);
    // stress windows: sxy rows s-4..s, syy and p rows s-3..s
    T sxyw[5] = {0, 0, 0, 0, 0}, syyw[4] = {0, 0, 0, 0}, pw[4] = {0, 0, 0, 0};
    Vec4<T> ms_cur, mv_cur = zero4;  // (alpha, M, lambda_c, mu) of row s, (rho, rho_f, m, b) of row j = s-2
    // prefetched inputs of the next row step
    Vec4<T> v_pf = ld(Vo, zA);
    Vec4<T> s_pf = ld(So, zA - 2);
    Vec4<T> ms_pf, mv_pf;
    int id_m1 = 0, id_0 = 0, id_p1 = 0, id_pf = 0;  // palette mode: material index of rows s-1, s, s+1
    if constexpr (PAL) {
        id_0 = ldm(zA - 3);
        id_p1 = ldm(zA - 2);
        id_pf = ldm(zA - 1);
        ms_cur = ms_at(id_p1);
    } else {
        ms_cur = ld(MS, zA - 2);
        ms_pf = ld(MS, zA - 1);
        mv_pf = ld(MV, zA - 3);
    }

    for (int s = zA - 2; s < zB + 2; ++s) {
        // ---- inputs of this row step, then prefetch the next one
#pragma unroll
        for (int k = 0; k < 4; ++k) vw[k] = vw[k + 1];
        vw[4] = v_pf;
        const Vec4<T> so = s_pf;
        Vec4<T> ms_nxt, mv_nxt;  // coefficients of row s+1 (stress side) and row s-1 = j+1 (velocity side)
        if constexpr (PAL) {
            id_m1 = id_0;
            id_0 = id_p1;
            id_p1 = id_pf;
            ms_nxt = ms_at(id_p1);
            mv_nxt = mv_at(id_m1);
        } else {
            ms_nxt = ms_pf;
            mv_nxt = mv_pf;
        }
        v_pf = ld(Vo, s + 3);
        s_pf = ld(So, s + 1);
        if constexpr (PAL) {
            id_pf = ldm(s + 2);
        } else {
            ms_pf = ld(MS, s + 2);
            mv_pf = ld(MV, s);
        }

        // ---- stresses of row s
        const bool srow = s >= zA && s < zB;
        const bool zr1 = s >= 2 && s < nz - 1;
        const bool zr2 = s >= 1 && s < nz - 1;
        const int szs = (s >= 0 && s <= nz) ? A.zmap[s] : -1;
        T(*xv)[BW + 4] = lv[s & 1];
        xv[0][lane + 2] = vw[2].x;
        xv[1][lane + 2] = vw[2].z;
        xv[2][lane + 2] = vw[2].y;
        __syncthreads();

        const bool store_s = own && srow;
        T exx_n = so.x, eyy_n = so.y, zeta_n = so.z, sxy_n = so.w;
        T p = 0, sxx = 0, syy = 0;
        if (xr1 && zr1) {
            T dvx = d4(xv[0][lane + 3], vw[2].x, xv[0][lane + 4], xv[0][lane + 1], rdx);
            T dwx = d4(xv[1][lane + 3], vw[2].z, xv[1][lane + 4], xv[1][lane + 1], rdx);
            T dvz = d4(vw[2].y, vw[1].y, vw[3].y, vw[0].y, rdz);
            T dwz = d4(vw[2].w, vw[1].w, vw[3].w, vw[0].w, rdz);
            if (sx >= 0) {
                const long long r = (long long)s * A.wxs + sx;
                const T *po = psi_x(A, A.in, r);
                T *pn = psi_x(A, A.out, r);
                dvx = cpml(dvx, po + PX_DVX, pn + PX_DVX, store_s, axh, bxh, rkxh);
                dwx = cpml(dwx, po + PX_DWX, pn + PX_DWX, store_s, axh, bxh, rkxh);
            }
            if (szs >= 0) {
                const long long r = (long long)szs * P + c;
                const T *po = psi_z(A, A.in, r);
                T *pn = psi_z(A, A.out, r);
                const T az = A.tab[TB_AZ * A.tn + s], bz = A.tab[TB_BZ * A.tn + s], rkz = A.tab[TB_RKZ * A.tn + s];
                dvz = cpml(dvz, po + PZ_DVZ, pn + PZ_DVZ, store_s, az, bz, rkz);
                dwz = cpml(dwz, po + PZ_DWZ, pn + PZ_DWZ, store_s, az, bz, rkz);
            }
            exx_n = so.x + dt * dvx;
            eyy_n = so.y + dt * dvz;
            zeta_n = (so.z - dwx * dt) - dwz * dt;
            p = pore_pressure(exx_n, eyy_n, zeta_n, ms_cur.x, ms_cur.y);
            if (src_p && s == A.src_z) p += A.src_val;
            if (xr3) {
                const T lc = ms_cur.z, al = ms_cur.x;
                const T l2m = lc + T(2) * ms_cur.w;
                sxx = l2m * exx_n + lc * eyy_n - al * p;
                syy = lc * exx_n + l2m * eyy_n - al * p;
            }
        }
        if (xr2 && zr2) {
            T ddx = d4(vw[2].y, xv[2][lane + 1], xv[2][lane + 3], xv[2][lane], rdx);
            T ddz = d4(vw[3].x, vw[2].x, vw[4].x, vw[1].x, rdz);
            if (sx >= 0) {
                const long long r = (long long)s * A.wxs + sx;
                ddx = cpml(ddx, psi_x(A, A.in, r) + PX_DVY, psi_x(A, A.out, r) + PX_DVY, store_s, ax, bx, rkx);
            }
            if (szs >= 0) {
                // half-grid memory coefficients, full-grid K_z divisor (reference :224-226)
                const long long r = (long long)szs * P + c;
                ddz = cpml(ddz, psi_z(A, A.in, r) + PZ_DVX, psi_z(A, A.out, r) + PZ_DVX, store_s,
                           A.tab[TB_AZH * A.tn + s], A.tab[TB_BZH * A.tn + s], A.tab[TB_RKZ * A.tn + s]);
            }
            sxy_n = so.w + ms_nxt.w * (ddx + ddz) * dt;  // mu of point (x, z+1)
        }
        if (store_s) *(Vec4<T> *)(Sn + off(s)) = Vec4<T>{exx_n, eyy_n, zeta_n, sxy_n};
        ms_cur = ms_nxt;
        {
            T(*xs)[BW + 4] = ls[(s + 3) % 3];
            xs[0][lane + 2] = sxx;
            xs[1][lane + 2] = p;
            xs[2][lane + 2] = sxy_n;
        }
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

        // ---- velocities of row j = s - 2 (its stress row went into the LDS ring two steps ago)
        const int j = s - 2;
        const Vec4<T> mv_j = mv_cur;  // row j
        mv_cur = mv_nxt;              // row j+1, and row j of the next step
        if (j >= zA && j < zB && own) {
            const T(*xs)[BW + 4] = ls[(j + 3) % 3];
            const int szj = A.zmap[j];
            Vec4<T> v = vw[0];  // (vx, vy, wx, wy) of row j, step n
            if (xr4 && j >= 2 && j < nz) {
                T e1 = d4(xs[0][lane + 2], xs[0][lane + 1], xs[0][lane + 3], xs[0][lane], rdx);
                T e2 = d4(xs[1][lane + 2], xs[1][lane + 1], xs[1][lane + 3], xs[1][lane], rdx);
                T e3 = d4(sxyw[2], sxyw[1], sxyw[3], sxyw[0], rdz);
                if (sx >= 0) {
                    const long long r = (long long)j * A.wxs + sx;
                    const T *po = psi_x(A, A.in, r);
                    T *pn = psi_x(A, A.out, r);
                    e1 = cpml(e1, po + PX_DSXX, pn + PX_DSXX, true, ax, bx, rkx);
                    e2 = cpml(e2, po + PX_DPX, pn + PX_DPX, true, ax, bx, rkx);
                }
                if (szj >= 0) {
                    const long long r = (long long)szj * P + c;
                    e3 = cpml(e3, psi_z(A, A.in, r) + PZ_DSXY, psi_z(A, A.out, r) + PZ_DSXY, true,
                              A.tab[TB_AZ * A.tn + j], A.tab[TB_BZ * A.tn + j], A.tab[TB_RKZ * A.tn + j]);
                }
                biot_velocity(v.x, v.z, e1 + e3, e2, mv_j.x, mv_j.y, mv_j.z, mv_j.w, A.inv_dt);
            }
            if (xr5 && j >= 1 && j < nz - 1) {
                T e1 = d4(xs[2][lane + 3], xs[2][lane + 2], xs[2][lane + 4], xs[2][lane + 1], rdx);
                T e2 = d4(syyw[2], syyw[1], syyw[3], syyw[0], rdz);
                T e3 = d4(pw[2], pw[1], pw[3], pw[0], rdz);
                if (sx >= 0) {
                    const long long r = (long long)j * A.wxs + sx;
                    e1 = cpml(e1, psi_x(A, A.in, r) + PX_DSXY, psi_x(A, A.out, r) + PX_DSXY, true, axh, bxh, rkxh);
                }
                if (szj >= 0) {
                    const long long r = (long long)szj * P + c;
                    const T *po = psi_z(A, A.in, r);
                    T *pn = psi_z(A, A.out, r);
                    const T azh = A.tab[TB_AZH * A.tn + j], bzh = A.tab[TB_BZH * A.tn + j];
                    const T rkzh = A.tab[TB_RKZH * A.tn + j];
                    e2 = cpml(e2, po + PZ_DSYY, pn + PZ_DSYY, true, azh, bzh, rkzh);
                    e3 = cpml(e3, po + PZ_DPZ, pn + PZ_DPZ, true, azh, bzh, rkzh);
                }
                T div = e1 + e2;
                if (src_f && j == A.src_z) div += A.src_val;
                biot_velocity(v.y, v.w, div, e3, mv_cur.x, mv_cur.y, mv_cur.z, mv_cur.w, A.inv_dt);
            }
            *(Vec4<T> *)(Vn + off(j)) = v;
        }
    }
}

template <class T, bool PAL>
hipError_t launch_fused_bw(const StepArgs<T> &A, int bw, hipStream_t stream)
{
    const dim3 grid(8 * A.per_xcd);
    switch (bw) {
    case 128: fused_step<T, 128, PAL><<<grid, 128, 0, stream>>>(A); break;
    case 256: fused_step<T, 256, PAL><<<grid, 256, 0, stream>>>(A); break;
    case 512: fused_step<T, 512, PAL><<<grid, 512, 0, stream>>>(A); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class T>
hipError_t launch_fused(const StepArgs<T> &A, int bw, bool palette, hipStream_t stream)
{
    if (palette) {
        if (A.npal < 1 || A.npal > kMaxPalette || !A.mid || !A.pal) return hipErrorInvalidValue;
        return launch_fused_bw<T, true>(A, bw, stream);
    }
    return launch_fused_bw<T, false>(A, bw, stream);
}

template hipError_t launch_fused<float>(const StepArgs<float> &, int, bool, hipStream_t);
template hipError_t launch_fused<double>(const StepArgs<double> &, int, bool, hipStream_t);

}  // namespace poe2d
