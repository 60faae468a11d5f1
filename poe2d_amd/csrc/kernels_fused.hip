// Single-pass time step: one launch per step computes the stress update and the velocity update
// of a tile without writing the intermediate stresses to memory.
//
// Work decomposition.  A workgroup of BW threads owns a column strip of W = BW - 8 output columns
// and th output rows.  Thread `lane` holds column c = x0 - 4 + lane, so the block spans the tile
// plus a 4-column halo on each side.  The block marches down its rows; at row step s it
//   1. holds the old velocities of rows s-2 .. s+2 in registers (row s+3 is in flight),
//   2. forms the stresses of row s for every held column (strains, fluid content, shear stress,
//      pore pressure and normal stresses, with their C-PML memory variables), taking x
//      neighbours through LDS and z neighbours from registers,
//   3. forms the velocities of row j = s-2 from registers holding stress rows s-4 .. s and the x
//      neighbours of stress row j, which were put in an LDS ring two row steps earlier;
//   4. on recorded steps, gathers the receivers of row s-1 from the LDS rings, which still hold
//      that row's old velocities and new pore pressure.
// Each point's state vectors (V, S) are read with one 16-byte load each, issued one row step
// before use, and written with one 16-byte store each; every access is a uniform group base plus
// one 32-bit per-lane offset.  The coefficients come either from the MS and MV groups (one 16-byte
// load each) or, in palette mode, from a one-byte material index per point that selects an
// (MS, MV) entry of a palette held in LDS: 1 byte per point instead of 32.
//
// Register rings.  The row loop is unrolled by U = 6 and every per-row register ring (and LDS
// ring) has a size dividing U, so inside the unrolled body the slot of each row is a compile-time
// constant: a load lands directly in the register that later consumes it, with no register
// rotation at the loop latch (a rotation would force the load to complete within the row step
// that issued it).
//
// Stresses of the 2 halo columns and 2 halo rows on each side are recomputed from the unmodified
// old state (double buffering) instead of being exchanged.  One barrier per row step.
//
// Work list.  One launch advances every model of a batch (a single model is a batch of one): each
// block takes a (model, tile) item, reads that model's arguments from device memory and the step
// index from a device counter, so a captured sequence of launches replays for any step.  Items are
// placed so that consecutive items (neighbouring tiles of a row band) land on the same XCD (blocks
// are dispatched round-robin over the 8 XCDs), keeping halo re-reads inside one L2.
#include <cstdint>
#include <type_traits>

#include "step_common.h"

namespace poe2d {

template <class T, int BW, bool PAL, bool AVG>
__global__ void __launch_bounds__(BW)
    fused_step(const StepArgs<T> *__restrict__ models, const WorkItem *__restrict__ work, int nwork, int per_xcd,
               const long long *__restrict__ base, int k)
{
    constexpr int W = BW - 8;
    constexpr uint32_t VB = sizeof(Vec4<T>);
    typedef __attribute__((address_space(1))) nvec4<T> GN;
    typedef __attribute__((address_space(1))) T GT;
    __shared__ T lv[3][3][BW + 4];  // ring of 3 velocity rows: vx, wx, vy   (entry lane + 2)
    __shared__ T ls[3][3][BW + 4];  // ring of 3 stress rows:   sxx, p, sxy
    __shared__ Vec4<T> lpal[PAL ? 2 * kMaxPalette : 1];  // MS entries, then MV entries

    const int wi = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (wi >= nwork) return;
    const WorkItem item = work[wi];
    StepArgs<T> A = models[item.model];
    const long long n = *(const kcll *)base + k;
    at_step(A, n);
    const int tile = item.tile;
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
    gchar *const gb = (gchar *)A.grp;
    const gchar *const mb = (const gchar *)A.mat;
    const long long gsb = A.group_stride * (long long)VB;
    const gchar *MS = mb + G_MS * gsb;
    const gchar *MV = mb + G_MV * gsb;
    const gchar *Vo = gb + group_v(A.in) * gsb;
    const gchar *So = gb + group_s(A.in) * gsb;
    gchar *Vn = gb + group_v(A.out) * gsb;
    gchar *Sn = gb + group_s(A.out) * gsb;
    const GT *const tab = (const GT *)A.tab;
    GT *const psi = (GT *)A.psi;
    // memory-variable records of x-strip / z-strip record r in state buffer b
    auto px = [&](int b, long long r) -> GT * { return psi + ((long long)b * A.pxn + r) * NPSI_X; };
    auto pz = [&](int b, long long r) -> GT * {
        return psi + 2LL * NPSI_X * A.pxn + ((long long)b * A.pzn + r) * NPSI_Z;
    };
    const uint32_t colb = (uint32_t)(min(c, nx + 1) + kPadCols) * VB;
    const uint32_t rowb = (uint32_t)P * VB;
    auto off = [&](int row) -> uint32_t { return colb + (uint32_t)(row + kPadRows) * rowb; };
    auto ld = [&](const gchar *g, int row) -> Vec4<T> {
        const nvec4<T> v = *(const GN *)(g + off(row));
        return Vec4<T>{v.x, v.y, v.z, v.w};
    };
    const Vec4<T> zero4 = {T(0), T(0), T(0), T(0)};
    // palette mode: material index of a row (same pads as the groups, one byte per point)
    const uint32_t colm = (uint32_t)(min(c, nx + 1) + kPadCols);
    gcuchar *const mid = (gcuchar *)A.mid;
    auto ldm = [&](int row) -> int { return mid[colm + (uint32_t)(row + kPadRows) * (uint32_t)P]; };
    if constexpr (PAL) {
        const GN *const pal = (const GN *)A.pal;
        for (int i = lane; i < 2 * kMaxPalette; i += BW)
            if ((i & (kMaxPalette - 1)) < A.npal) {
                const nvec4<T> e = pal[i];
                lpal[i] = Vec4<T>{e.x, e.y, e.z, e.w};
            }
        __syncthreads();
    }

    // per-column C-PML profile (x strips only)
    const int sx = (c >= 0 && c <= nx) ? ((gcint *)A.xmap)[c] : -1;
    T ax = 0, bx = 1, rkx = 1, axh = 0, bxh = 1, rkxh = 1;
    if (sx >= 0) {
        ax = tab[TB_AX * A.tn + c]; bx = tab[TB_BX * A.tn + c]; rkx = tab[TB_RKX * A.tn + c];
        axh = tab[TB_AXH * A.tn + c]; bxh = tab[TB_BXH * A.tn + c]; rkxh = tab[TB_RKXH * A.tn + c];
    }
    const bool xr1 = c >= 1 && c < nx;      // stress1
    const bool xr2 = c >= 2 && c < nx;      // stress2
    const bool xr3 = c >= 1 && c < nx - 1;  // stress3
    const bool xr4 = c >= 2 && c < nx;      // vx
    const bool xr5 = c >= 1 && c < nx - 1;  // vz
    const bool src_p = A.src_kind == POE2D_SRC_PRESSURE_DGAUSS && c == A.src_x;
    // receivers of this step: slot and the tile's entries per row
    const int n32 = (int)n;
    kcint *const rec_ptr = (kcint *)A.rec_tile_ptr + (long long)tx * (nz + 1);  // offsets of this tile column
    const bool rec_on = A.rec_n > 0 && n32 % A.rec_dec == 0 && n32 / A.rec_dec < A.rec_nsamp &&
                        rec_ptr[zA] < rec_ptr[zB];  // receivers in this tile
    GT *const trace = (GT *)A.traces + (rec_on ? (long long)(n32 / A.rec_dec) * A.rec_ncomp * A.rec_n : 0);
    typedef int nint2 __attribute__((ext_vector_type(2)));
    const __attribute__((address_space(1))) nint2 *const rec_ent = (const __attribute__((address_space(1))) nint2 *)A.rec_entries;
    const bool src_f = A.src_kind == POE2D_SRC_FORCE_Y_RICKER && c == A.src_x;

    // Register rings.  Row r of a ring of size R sits in slot (r - s0) mod R, s0 = zA - 2 being the
    // first row step, so at phase PH of the unrolled group (s - s0 = PH mod 6) row s + k is in slot
    // (PH + k) mod R.
    const int s0 = zA - 2;
    const int send = zB + 2;
    Vec4<T> vr[6];              // velocities (vx, vy, wx, wy), rows s-2 .. s+3
    Vec4<T> sr[2];              // strains, fluid content, shear stress, rows s, s+1
    Vec4<T> msr[3];             // (alpha, M, lambda_c, mu), rows s, s+1 (and s+2 in flight)
    Vec4<T> mvr[3];             // (rho, rho_f, m, b), rows s-2, s-1 (and s in flight)
    int idr[6];                 // palette mode: material index, rows s-1 .. s+2
    T sxyr[6], syyr[6], pr[6];  // computed sxy rows s-4 .. s; syy and p rows s-3 .. s
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        sxyr[k] = T(0);
        syyr[k] = T(0);
        pr[k] = T(0);
        idr[k] = 0;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        msr[k] = zero4;
        mvr[k] = zero4;
    }
#pragma unroll
    for (int k = -2; k <= 2; ++k) vr[(k + 6) % 6] = ld(Vo, s0 + k);
    sr[0] = ld(So, s0);
    sr[1] = zero4;
    vr[3] = zero4;
    if constexpr (PAL) {
        idr[5] = ldm(s0 - 1);
        idr[0] = ldm(s0);
        idr[1] = ldm(s0 + 1);
        msr[0] = lpal[idr[0]];
    } else {
        msr[0] = ld(MS, s0);
        msr[1] = ld(MS, s0 + 1);
        mvr[2] = ld(MV, s0 - 1);
    }

    int s = s0;
    // One row step at phase PH of the unrolled group.
    auto row_step = [&](auto phase) __attribute__((always_inline)) {
        constexpr int PH = decltype(phase)::value;
        auto V = [&](int k) -> Vec4<T> & { return vr[(PH + k + 12) % 6]; };  // velocity row s+k
        const Vec4<T> so = sr[PH & 1];
        // ---- loads for later row steps
        V(3) = ld(Vo, s + 3);
        sr[(PH + 1) & 1] = ld(So, s + 1);
        const Vec4<T> &ms_cur = msr[PH % 3];      // row s
        Vec4<T> &ms_nxt = msr[(PH + 1) % 3];      // row s+1
        const Vec4<T> &mv_j = mvr[(PH + 1) % 3];  // row j = s-2
        Vec4<T> &mv_j1 = mvr[(PH + 2) % 3];       // row j+1 = s-1
        if constexpr (PAL) {
            ms_nxt = lpal[idr[(PH + 1) % 6]];
            mv_j1 = lpal[kMaxPalette + idr[(PH + 5) % 6]];
            idr[(PH + 2) % 6] = ldm(s + 2);
        } else {
            msr[(PH + 2) % 3] = ld(MS, s + 2);
            mvr[PH % 3] = ld(MV, s);
        }

        // ---- stresses of row s
        const bool srow = s >= zA && s < zB;
        const bool zr1 = s >= 2 && s < nz - 1;
        const bool zr2 = s >= 1 && s < nz - 1;
        const int szs = zstrip(A, s);
        T(*xv)[BW + 4] = lv[PH % 3];
        xv[0][lane + 2] = V(0).x;
        xv[1][lane + 2] = V(0).z;
        xv[2][lane + 2] = V(0).y;
        __syncthreads();

        // ---- receivers of row s-1: its old velocities and new pore pressure are in the ring
        // entries written one row step ago (visible after this step's barrier; the velocity ring
        // has 3 entries, so the next step's write goes to row s-2's entry, not this one)
        if (rec_on && s - 1 >= zA && s - 1 < zB) {
            const int lo = rec_ptr[s - 1], hi = rec_ptr[s];
            for (int e = lo + lane; e < hi; e += BW) {
                const nint2 en = rec_ent[e];
                const int li = en.x - tx * W + 6;  // LDS entry of column en.x
                int q = 0;
                if (A.rec_mask & POE2D_VX) trace[(long long)(q++) * A.rec_n + en.y] = lv[(PH + 2) % 3][0][li];
                if (A.rec_mask & POE2D_VY) trace[(long long)(q++) * A.rec_n + en.y] = lv[(PH + 2) % 3][2][li];
                if (A.rec_mask & POE2D_P) trace[(long long)q * A.rec_n + en.y] = ls[(PH + 2) % 3][1][li];
            }
        }

        const bool store_s = own && srow;
        T exx_n = so.x, eyy_n = so.y, zeta_n = so.z, sxy_n = so.w;
        T p = 0, sxx = 0, syy = 0;
        if (xr1 && zr1) {
            T dvx = d4(xv[0][lane + 3], V(0).x, xv[0][lane + 4], xv[0][lane + 1], rdx);
            T dwx = d4(xv[1][lane + 3], V(0).z, xv[1][lane + 4], xv[1][lane + 1], rdx);
            T dvz = d4(V(0).y, V(-1).y, V(1).y, V(-2).y, rdz);
            T dwz = d4(V(0).w, V(-1).w, V(1).w, V(-2).w, rdz);
            if (sx >= 0) {
                const long long r = (long long)s * A.wxs + sx;
                const GT *po = px(A.in, r);
                GT *pn = px(A.out, r);
                dvx = cpml(dvx, po + PX_DVX, pn + PX_DVX, store_s, axh, bxh, rkxh);
                dwx = cpml(dwx, po + PX_DWX, pn + PX_DWX, store_s, axh, bxh, rkxh);
            }
            if (szs >= 0) {
                const long long r = (long long)szs * P + c;
                const GT *po = pz(A.in, r);
                GT *pn = pz(A.out, r);
                const T az = tab[TB_AZ * A.tn + s], bz = tab[TB_BZ * A.tn + s], rkz = tab[TB_RKZ * A.tn + s];
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
            T ddx = d4(V(0).y, xv[2][lane + 1], xv[2][lane + 3], xv[2][lane], rdx);
            T ddz = d4(V(1).x, V(0).x, V(2).x, V(-1).x, rdz);
            if (sx >= 0) {
                const long long r = (long long)s * A.wxs + sx;
                ddx = cpml(ddx, px(A.in, r) + PX_DVY, px(A.out, r) + PX_DVY, store_s, ax, bx, rkx);
            }
            if (szs >= 0) {
                // half-grid memory coefficients, full-grid K_z divisor (reference :224-226)
                const long long r = (long long)szs * P + c;
                ddz = cpml(ddz, pz(A.in, r) + PZ_DVX, pz(A.out, r) + PZ_DVX, store_s,
                           tab[TB_AZH * A.tn + s], tab[TB_BZH * A.tn + s], tab[TB_RKZ * A.tn + s]);
            }
            // mu of point (x, z+1), or the harmonic mean of points (x, z) and (x, z+1)
            const T mu_xy = AVG ? harmonic_mu(ms_cur.w, ms_nxt.w) : ms_nxt.w;
            sxy_n = so.w + mu_xy * (ddx + ddz) * dt;
        }
        if (store_s) *(GN *)(Sn + off(s)) = nvec4<T>{exx_n, eyy_n, zeta_n, sxy_n};
        {
            T(*xs)[BW + 4] = ls[PH % 3];
            xs[0][lane + 2] = sxx;
            xs[1][lane + 2] = p;
            xs[2][lane + 2] = sxy_n;
        }
        sxyr[PH % 6] = sxy_n;
        syyr[PH % 6] = syy;
        pr[PH % 6] = p;
        auto SXY = [&](int k) -> T { return sxyr[(PH + k + 12) % 6]; };  // sxy of row s+k
        auto SYY = [&](int k) -> T { return syyr[(PH + k + 12) % 6]; };
        auto PP = [&](int k) -> T { return pr[(PH + k + 12) % 6]; };

        // ---- velocities of row j = s - 2 (its stress row went into the LDS ring two steps ago)
        const int j = s - 2;
        if (j >= zA && j < zB && own) {
            const T(*xs)[BW + 4] = ls[(PH + 1) % 3];
            const int szj = zstrip(A, j);
            Vec4<T> v = V(-2);  // (vx, vy, wx, wy) of row j, step n
            if (xr4 && j >= 2 && j < nz) {
                T e1 = d4(xs[0][lane + 2], xs[0][lane + 1], xs[0][lane + 3], xs[0][lane], rdx);
                T e2 = d4(xs[1][lane + 2], xs[1][lane + 1], xs[1][lane + 3], xs[1][lane], rdx);
                T e3 = d4(SXY(-2), SXY(-3), SXY(-1), SXY(-4), rdz);
                if (sx >= 0) {
                    const long long r = (long long)j * A.wxs + sx;
                    const GT *po = px(A.in, r);
                    GT *pn = px(A.out, r);
                    e1 = cpml(e1, po + PX_DSXX, pn + PX_DSXX, true, ax, bx, rkx);
                    e2 = cpml(e2, po + PX_DPX, pn + PX_DPX, true, ax, bx, rkx);
                }
                if (szj >= 0) {
                    const long long r = (long long)szj * P + c;
                    e3 = cpml(e3, pz(A.in, r) + PZ_DSXY, pz(A.out, r) + PZ_DSXY, true,
                              tab[TB_AZ * A.tn + j], tab[TB_BZ * A.tn + j], tab[TB_RKZ * A.tn + j]);
                }
                biot_velocity(v.x, v.z, e1 + e3, e2, mv_j.x, mv_j.y, mv_j.z, mv_j.w, A.inv_dt);
            }
            if (xr5 && j >= 1 && j < nz - 1) {
                T e1 = d4(xs[2][lane + 3], xs[2][lane + 2], xs[2][lane + 4], xs[2][lane + 1], rdx);
                T e2 = d4(SYY(-1), SYY(-2), SYY(0), SYY(-3), rdz);
                T e3 = d4(PP(-1), PP(-2), PP(0), PP(-3), rdz);
                if (sx >= 0) {
                    const long long r = (long long)j * A.wxs + sx;
                    e1 = cpml(e1, px(A.in, r) + PX_DSXY, px(A.out, r) + PX_DSXY, true, axh, bxh, rkxh);
                }
                if (szj >= 0) {
                    const long long r = (long long)szj * P + c;
                    const GT *po = pz(A.in, r);
                    GT *pn = pz(A.out, r);
                    const T azh = tab[TB_AZH * A.tn + j], bzh = tab[TB_BZH * A.tn + j];
                    const T rkzh = tab[TB_RKZH * A.tn + j];
                    e2 = cpml(e2, po + PZ_DSYY, pn + PZ_DSYY, true, azh, bzh, rkzh);
                    e3 = cpml(e3, po + PZ_DPZ, pn + PZ_DPZ, true, azh, bzh, rkzh);
                }
                T div = e1 + e2;
                if (src_f && j == A.src_z) div += A.src_val;
                // material of point (x, z+1), or the mean of points (x, z) and (x, z+1)
                const Vec4<T> mvy = AVG ? mean_rows(mv_j, mv_j1) : mv_j1;
                biot_velocity(v.y, v.w, div, e3, mvy.x, mvy.y, mvy.z, mvy.w, A.inv_dt);
            }
            *(GN *)(Vn + off(j)) = nvec4<T>{v.x, v.y, v.z, v.w};
        }
        ++s;
    };

    // Whole groups of 6 row steps: the last group may run up to 5 steps past zB + 2, over rows
    // that store nothing (every store is guarded by the tile's row range) and read only zero pad
    // rows (at most row nz + 10 < nz + 1 + kPadRows).
    for (int g = (send - s0 + 5) / 6; g > 0; --g) {
        row_step(std::integral_constant<int, 0>{});
        row_step(std::integral_constant<int, 1>{});
        row_step(std::integral_constant<int, 2>{});
        row_step(std::integral_constant<int, 3>{});
        row_step(std::integral_constant<int, 4>{});
        row_step(std::integral_constant<int, 5>{});
    }
}

template <class T, bool PAL, bool AVG>
hipError_t launch_fused_bw(const StepArgs<T> *models, const WorkItem *work, int nwork, int per_xcd, int bw,
                           const long long *base, int k, hipStream_t stream)
{
    const dim3 grid(8 * per_xcd);
    switch (bw) {
    case 128: fused_step<T, 128, PAL, AVG><<<grid, 128, 0, stream>>>(models, work, nwork, per_xcd, base, k); break;
    case 256: fused_step<T, 256, PAL, AVG><<<grid, 256, 0, stream>>>(models, work, nwork, per_xcd, base, k); break;
    case 512: fused_step<T, 512, PAL, AVG><<<grid, 512, 0, stream>>>(models, work, nwork, per_xcd, base, k); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class T>
hipError_t launch_fused(const StepArgs<T> *models, const WorkItem *work, int nwork, int per_xcd, int bw,
                        bool palette, bool avg, const long long *base, int k, hipStream_t stream)
{
    if (nwork < 1 || per_xcd < 1 || 8LL * per_xcd < nwork) return hipErrorInvalidValue;
    if (avg)
        return palette ? launch_fused_bw<T, true, true>(models, work, nwork, per_xcd, bw, base, k, stream)
                       : launch_fused_bw<T, false, true>(models, work, nwork, per_xcd, bw, base, k, stream);
    return palette ? launch_fused_bw<T, true, false>(models, work, nwork, per_xcd, bw, base, k, stream)
                   : launch_fused_bw<T, false, false>(models, work, nwork, per_xcd, bw, base, k, stream);
}

template hipError_t launch_fused<float>(const StepArgs<float> *, const WorkItem *, int, int, int, bool, bool,
                                        const long long *, int, hipStream_t);
template hipError_t launch_fused<double>(const StepArgs<double> *, const WorkItem *, int, int, int, bool, bool,
                                         const long long *, int, hipStream_t);

namespace {
__global__ void base_advance(long long *base, int delta) { base[0] += delta; }
}  // namespace

hipError_t launch_base_advance(long long *base, int delta, hipStream_t stream)
{
    base_advance<<<1, 1, 0, stream>>>(base, delta);
    return hipGetLastError();
}

}  // namespace poe2d
