// Two time steps per launch (temporal blocking of the single-pass step).
//
// A block owns BW - 12 output columns (6 halo columns on each side) and th rows of one model and
// marches down its rows as the single-pass kernel does, running two copies of that kernel's row
// step one after the other, each on its own time level:
//   stage 1 (step n)   at row step s: stresses of row s, velocities of row s - 2, from the state
//                      read from HBM (as the single-pass kernel);
//   stage 2 (step n+1) at row step s: stresses of row s - 4, velocities of row s - 6, from stage
//                      1's results, which never leave the chip.
// Stage 2 lags by 4 rows because its stresses of row r need the step-(n+1) velocities of rows
// r-2 .. r+2 (x velocity at r+2 for the shear-stress z derivative), the last of which stage 1
// produces at row step r+4.  Only stage 2 stores state (rows [zA, zB), the owned lanes); stage
// 1 computes every row and lane it can, and the rows and lanes whose inputs are incomplete hold
// values nothing valid depends on:
//   stage-1 stresses  valid on lanes [2, BW-3]  (x neighbours of the velocities read from HBM)
//   stage-1 velocities             [3, BW-4]
//   stage-2 stresses               [4, BW-6] (sxx, p), [5, BW-5] (sxy)
//   stage-2 velocities             [6, BW-7]   -> owned lanes [6, BW-6)
// and in rows, stage 2's rows [zA, zB) need stage-2 stresses of rows [zA-2, zB+1], step-(n+1)
// velocities of rows [zA-4, zB+3] and stage-1 stresses of rows [zA-6, zB+5]: the row steps run
// s = zA-6 .. zB+5.
//
// C-PML memory variables of step n+1 go through a third record set (psi_mid): stage 1 writes the
// records of the points whose stage-1 values are valid, stage 2 reads them back from the same
// lane four row steps later (a thread's store and its later load of the same address are
// ordered).  Neighbouring tiles compute the same stage-1 values for the halo points they share
// and store identical bits there.  Receivers record step n from stage 1 (row s-1 at row step s)
// and step n+1 from stage 2 (row s-5 at row step s), from the LDS rings.
//
// The state of step n is read from buffer `in` and that of step n+2 written to buffer in ^ 1, so
// the buffer holding the state flips once per launch: in = ((n - pair0) / 2 + pcur) & 1.
#include "step_common.h"

namespace poe2d {

// One tile of the launch.  EDGE = false: tiles none of whose points (halos included) lie in a
// C-PML strip and that hold no receivers; the C-PML and receiver code is then compiled out, and
// the registers this frees hold velocity loads further ahead.  Without the stage-2 receiver
// gather, the step-(n+1) velocity rows need an LDS ring of 3 instead of 6.  LDS rings, entry lane
// + 2 of each row: lv 3 step-n velocity rows (vx, wx, vy), ls 3 stage-1 stress rows (sxx, p, sxy),
// lv2 6 step-(n+1) velocity rows (the first R2 in use), ls2 3 stage-2 stress rows; lpal the MS,
// the MV and the velocity-constant entries (palette sections PAL_MS, PAL_MV, PAL_CN, kMaxPalette2
// entries each).  PAL = false: grids without a palette the LDS copy can hold read each point's
// coefficients from the MS and MV groups where they are used (32 B per point from HBM once per
// launch; the re-reads of the same rows by the other stage and the neighbouring row steps hit the
// caches) and form the velocity-update constants per point, which rounds as the palette's do.
template <class T, int BW, bool AVG, bool EDGE, bool PAL>
__device__ __forceinline__ void step2_tile(StepArgs<T> A, const int tile, const long long n, T (*lv)[3][BW + 4],
                                           T (*ls)[3][BW + 4], T (*lv2)[3][BW + 4], T (*ls2)[3][BW + 4],
                                           Vec4<T> *lpal)
{
    constexpr int W = fused2_tile_width(BW);
    constexpr uint32_t VB = sizeof(Vec4<T>);
    typedef __attribute__((address_space(1))) nvec4<T> GN;
    typedef __attribute__((address_space(1))) T GT;
    constexpr int R2 = EDGE ? 6 : 3;
    A.in = (int)((((n - A.pair0) >> 1) + A.pcur) & 1);
    A.out = A.in ^ 1;
    typedef __attribute__((address_space(1))) const T GCT;
    auto sample = [&](long long m) -> T {
        const long long q = m - A.src_n0;
        return (A.src_series && q >= 0 && q < A.src_len) ? ((GCT *)A.src_series)[q] : T(0);
    };
    const T src1 = sample(n), src2 = sample(n + 1);
    const int tx = tile % A.ntx;
    const int tz = tile / A.ntx;
    const int lane = threadIdx.x;
    const int nx = A.nx, nz = A.nz;
    const int c = tx * W - 6 + lane;
    const int zA = tz * A.th;
    const int zB = min(zA + A.th, nz + 1);
    const int P = A.pitch;
    // Interior tiles compute only columns [2, nx - 2] and rows [2, nz - 2] (the host makes every
    // other tile an edge tile), where each update's column and row guards hold.  The row guards
    // are compiled out of every interior tile and the column guards out of those without a
    // palette, so that a row step is a few large blocks of straight-line code (with the palette,
    // the longer live ranges of dropping the column guards too make the kernel spill).
    constexpr bool GUARD = EDGE || PAL;  // column guards
    constexpr bool ZGUARD = EDGE;        // row guards
    const bool own = lane >= 6 && lane < BW - 6 && (!GUARD || c <= nx);
    const bool ps1 = lane >= 2 && lane <= BW - 3;  // stage-1 stress records worth keeping
    const bool pv1 = lane >= 3 && lane <= BW - 4;  // stage-1 velocity records worth keeping
    const T rdx = A.rdx, rdz = A.rdz, dt = A.dt;

    gchar *const gb = (gchar *)A.grp;
    const long long gsb = A.group_stride * (long long)VB;
    const gchar *Vo = gb + group_v(A.in) * gsb;
    const gchar *So = gb + group_s(A.in) * gsb;
    gchar *Vn = gb + group_v(A.out) * gsb;
    gchar *Sn = gb + group_s(A.out) * gsb;
    const GT *const tab = (const GT *)A.tab;
    GT *const psi = (GT *)A.psi;
    GT *const psim = (GT *)A.psi_mid;
    // memory-variable records: x strip record r / z strip record r of state buffer b, and of the
    // step-(n+1) set
    auto px = [&](int b, long long r) -> GT * { return psi + ((long long)b * A.pxn + r) * NPSI_X; };
    auto pz = [&](int b, long long r) -> GT * {
        return psi + 2LL * NPSI_X * A.pxn + ((long long)b * A.pzn + r) * NPSI_Z;
    };
    auto pxm = [&](long long r) -> GT * { return psim + r * NPSI_X; };
    auto pzm = [&](long long r) -> GT * { return psim + A.pxn * NPSI_X + r * NPSI_Z; };
    // Lanes left of the grid's pad columns (c < -kPadCols at the first tile column) read pad column
    // -kPadCols, lanes right of the grid column nx+1: zero either way.  Rows past the bottom pad
    // rows (reached only by the over-run of the last group of row steps) read the last pad row.
    const int cc = min(max(c, -kPadCols), nx + 1);
    const uint32_t colb = (uint32_t)(cc + kPadCols) * VB;
    const uint32_t rowb = (uint32_t)P * VB;
    const int last_row = nz + kPadRows;
    auto off = [&](int row) -> uint32_t { return colb + (uint32_t)(min(row, last_row) + kPadRows) * rowb; };
    auto ld = [&](const gchar *g, int row) -> Vec4<T> {
        const nvec4<T> v = *(const GN *)(g + off(row));
        return Vec4<T>{v.x, v.y, v.z, v.w};
    };
    const Vec4<T> zero4 = {T(0), T(0), T(0), T(0)};
    const uint32_t colm = (uint32_t)(cc + kPadCols);
    gcuchar *const mid = (gcuchar *)A.mid;
    auto ldm = [&](int row) -> int { return mid[colm + (uint32_t)(min(row, last_row) + kPadRows) * (uint32_t)P]; };
    // coefficients of row `row` at this lane: from the palette entry `id` in LDS, or from the groups
    const gchar *const MSg = (const gchar *)A.mat + G_MS * gsb;
    const gchar *const MVg = (const gchar *)A.mat + G_MV * gsb;
    // (groups: the row's coefficients arrive through the register rings msr / mvr below)
    auto ms_at = [&](int id, const Vec4<T> &ring) -> Vec4<T> {
        if constexpr (PAL) return lpal[PAL_MS * kMaxPalette2 + id];
        else return ring;
    };
    auto mv_at = [&](int id, const Vec4<T> &ring) -> Vec4<T> {
        if constexpr (PAL) return lpal[PAL_MV * kMaxPalette2 + id];
        else return ring;
    };
    if constexpr (PAL) {
        const GN *const pal = (const GN *)A.pal;
        for (int i = lane; i < PAL_SECTIONS * kMaxPalette2; i += BW) {
            const int e = i % kMaxPalette2;
            if (e < A.npal) {
                const nvec4<T> v = pal[(i / kMaxPalette2) * kMaxPalette + e];
                lpal[i] = Vec4<T>{v.x, v.y, v.z, v.w};
            }
        }
    }

    // per-column C-PML profile (x strips only)
    const int sx = (EDGE && c >= 0 && c <= nx) ? ((gcint *)A.xmap)[c] : -1;
    T ax = 0, bx = 1, rkx = 1, axh = 0, bxh = 1, rkxh = 1;
    if (sx >= 0) {
        ax = tab[TB_AX * A.tn + c]; bx = tab[TB_BX * A.tn + c]; rkx = tab[TB_RKX * A.tn + c];
        axh = tab[TB_AXH * A.tn + c]; bxh = tab[TB_BXH * A.tn + c]; rkxh = tab[TB_RKXH * A.tn + c];
    }
    const bool xr1 = !GUARD || (c >= 1 && c < nx);      // stress1
    const bool xr2 = !GUARD || (c >= 2 && c < nx);      // stress2
    const bool xr3 = !GUARD || (c >= 1 && c < nx - 1);  // stress3
    const bool xr4 = !GUARD || (c >= 2 && c < nx);      // vx
    const bool xr5 = !GUARD || (c >= 1 && c < nx - 1);  // vz
    const bool src_p = A.src_kind == POE2D_SRC_PRESSURE_DGAUSS && c == A.src_x;
    const bool src_f = A.src_kind == POE2D_SRC_FORCE_Y_RICKER && c == A.src_x;
    // receivers of the two steps: slots and the tile's entries per row
    kcint *const rec_ptr = (kcint *)A.rec_tile_ptr + (long long)tx * (nz + 1);
    const bool rec_tile = EDGE && A.rec_n > 0 && rec_ptr[zA] < rec_ptr[zB];
    const int n1 = (int)n, n2 = n1 + 1;
    const bool rec1 = rec_tile && n1 % A.rec_dec == 0 && n1 / A.rec_dec < A.rec_nsamp;
    const bool rec2 = rec_tile && n2 % A.rec_dec == 0 && n2 / A.rec_dec < A.rec_nsamp;
    GT *const trace1 = (GT *)A.traces + (rec1 ? (long long)(n1 / A.rec_dec) * A.rec_ncomp * A.rec_n : 0);
    GT *const trace2 = (GT *)A.traces + (rec2 ? (long long)(n2 / A.rec_dec) * A.rec_ncomp * A.rec_n : 0);
    typedef int nint2 __attribute__((ext_vector_type(2)));
    const __attribute__((address_space(1))) nint2 *const rec_ent = (const __attribute__((address_space(1))) nint2 *)A.rec_entries;
    auto gather = [&](GT *trace, int row, const T (*xv)[BW + 4], const T (*xs)[BW + 4]) {
        const int lo = rec_ptr[row], hi = rec_ptr[row + 1];
        for (int e = lo + lane; e < hi; e += BW) {
            const nint2 en = rec_ent[e];
            const int li = en.x - tx * W + 8;  // LDS entry of column en.x
            int q = 0;
            if (A.rec_mask & POE2D_VX) trace[(long long)(q++) * A.rec_n + en.y] = xv[0][li];
            if (A.rec_mask & POE2D_VY) trace[(long long)(q++) * A.rec_n + en.y] = xv[2][li];
            if (A.rec_mask & POE2D_P) trace[(long long)q * A.rec_n + en.y] = xs[1][li];
        }
    };

    // One stress update of row r at this lane (the single-pass kernel's, with the velocity rows
    // r-2 .. r+2 given, the x neighbours of row r in xv, the memory records read through po_*
    // and written through pn_* when `keep`).  Returns the new (e_xx, e_yy, zeta, sigma_xy) and
    // sets sxx, p, syy.
    auto stress = [&](int r, const Vec4<T> &vm2, const Vec4<T> &vm1, const Vec4<T> &v0, const Vec4<T> &vp1,
                      const Vec4<T> &vp2, const Vec4<T> &so, const T (*xv)[BW + 4], int id_cur, int id_nxt,
                      const Vec4<T> &m_cur, const Vec4<T> &m_nxt, auto pxo, auto pxn_, auto pzo, auto pzn_,
                      bool keep, T srcv, T &sxx, T &p, T &syy) __attribute__((always_inline)) -> Vec4<T> {
        const bool zr1 = !ZGUARD || (r >= 2 && r < nz - 1);
        const bool zr2 = !ZGUARD || (r >= 1 && r < nz - 1);
        const bool u1 = xr1 && zr1, u2 = xr2 && zr2;
        const int szs = EDGE ? zstrip(A, r) : -1;
        // C-PML: the update's old records and row coefficients, all read before any record is written
        const long long qx = (long long)r * A.wxs + sx, qz = (long long)szs * P + c;
        T o_dvx = 0, o_dwx = 0, o_dvy = 0, o_dvz = 0, o_dwz = 0, o_dvxz = 0;
        T az = 0, bz = 0, rkz = 0, azh = 0, bzh = 0;
        if (sx >= 0) {
            if (u1) o_dvx = pxo(qx)[PX_DVX], o_dwx = pxo(qx)[PX_DWX];
            if (u2) o_dvy = pxo(qx)[PX_DVY];
        }
        if (szs >= 0) {
            if (u1) o_dvz = pzo(qz)[PZ_DVZ], o_dwz = pzo(qz)[PZ_DWZ];
            if (u2) o_dvxz = pzo(qz)[PZ_DVX];
            az = tab[TB_AZ * A.tn + r], bz = tab[TB_BZ * A.tn + r], rkz = tab[TB_RKZ * A.tn + r];
            azh = tab[TB_AZH * A.tn + r], bzh = tab[TB_BZH * A.tn + r];
        }
        T exx_n = so.x, eyy_n = so.y, zeta_n = so.z, sxy_n = so.w;
        p = 0;
        sxx = 0;
        syy = 0;
        if (u1) {
            T dvx = d4(xv[0][lane + 3], v0.x, xv[0][lane + 4], xv[0][lane + 1], rdx);
            T dwx = d4(xv[1][lane + 3], v0.z, xv[1][lane + 4], xv[1][lane + 1], rdx);
            T dvz = d4(v0.y, vm1.y, vp1.y, vm2.y, rdz);
            T dwz = d4(v0.w, vm1.w, vp1.w, vm2.w, rdz);
            if (sx >= 0) {
                dvx = cpml_from(dvx, o_dvx, pxn_(qx) + PX_DVX, keep, axh, bxh, rkxh);
                dwx = cpml_from(dwx, o_dwx, pxn_(qx) + PX_DWX, keep, axh, bxh, rkxh);
            }
            if (szs >= 0) {
                dvz = cpml_from(dvz, o_dvz, pzn_(qz) + PZ_DVZ, keep, az, bz, rkz);
                dwz = cpml_from(dwz, o_dwz, pzn_(qz) + PZ_DWZ, keep, az, bz, rkz);
            }
            exx_n = so.x + dt * dvx;
            eyy_n = so.y + dt * dvz;
            zeta_n = (so.z - dwx * dt) - dwz * dt;
            const Vec4<T> ms_cur = ms_at(id_cur, m_cur);
            p = pore_pressure(exx_n, eyy_n, zeta_n, ms_cur.x, ms_cur.y);
            if (src_p && r == A.src_z) p += srcv;
            if (xr3) {
                const T lc = ms_cur.z, al = ms_cur.x;
                const T l2m = lc + T(2) * ms_cur.w;
                sxx = l2m * exx_n + lc * eyy_n - al * p;
                syy = lc * exx_n + l2m * eyy_n - al * p;
            }
        }
        if (u2) {
            T ddx = d4(v0.y, xv[2][lane + 1], xv[2][lane + 3], xv[2][lane], rdx);
            T ddz = d4(vp1.x, v0.x, vp2.x, vm1.x, rdz);
            if (sx >= 0) ddx = cpml_from(ddx, o_dvy, pxn_(qx) + PX_DVY, keep, ax, bx, rkx);
            // half-grid memory coefficients, full-grid K_z divisor (reference :224-226)
            if (szs >= 0) ddz = cpml_from(ddz, o_dvxz, pzn_(qz) + PZ_DVX, keep, azh, bzh, rkz);
            const T mu_nxt = ms_at(id_nxt, m_nxt).w;
            const T mu_xy = AVG ? harmonic_mu(ms_at(id_cur, m_cur).w, mu_nxt) : mu_nxt;
            sxy_n = so.w + mu_xy * (ddx + ddz) * dt;
        }
        return Vec4<T>{exx_n, eyy_n, zeta_n, sxy_n};
    };

    // One velocity update of row j at this lane: v the step's old velocities of row j, xs the
    // new stress row j (x neighbours), sxy of rows j-2 .. j+1 and syy, p of rows j-1 .. j+2.
    auto velocity = [&](int j, Vec4<T> v, const T (*xs)[BW + 4], T sxy_m2, T sxy_m1, T sxy_0, T sxy_p1, T syy_m1,
                        T syy_0, T syy_p1, T syy_p2, T p_m1, T p_0, T p_p1, T p_p2, int id_j, int id_j1,
                        const Vec4<T> &m_j, const Vec4<T> &m_j1, auto pxo, auto pxn_, auto pzo, auto pzn_, bool keep,
                        T srcv) __attribute__((always_inline)) -> Vec4<T> {
        const int szj = EDGE ? zstrip(A, j) : -1;
        const bool u4 = xr4 && (!ZGUARD || (j >= 2 && j < nz)), u5 = xr5 && (!ZGUARD || (j >= 1 && j < nz - 1));
        // C-PML: the update's old records and row coefficients, all read before any record is written
        const long long qx = (long long)j * A.wxs + sx, qz = (long long)szj * P + c;
        T o_dsxx = 0, o_dpx = 0, o_dsxy = 0, o_dsxyz = 0, o_dsyy = 0, o_dpz = 0;
        T az = 0, bz = 0, rkz = 0, azh = 0, bzh = 0, rkzh = 0;
        if (sx >= 0) {
            if (u4) o_dsxx = pxo(qx)[PX_DSXX], o_dpx = pxo(qx)[PX_DPX];
            if (u5) o_dsxy = pxo(qx)[PX_DSXY];
        }
        if (szj >= 0) {
            if (u4) o_dsxyz = pzo(qz)[PZ_DSXY];
            if (u5) o_dsyy = pzo(qz)[PZ_DSYY], o_dpz = pzo(qz)[PZ_DPZ];
            az = tab[TB_AZ * A.tn + j], bz = tab[TB_BZ * A.tn + j], rkz = tab[TB_RKZ * A.tn + j];
            azh = tab[TB_AZH * A.tn + j], bzh = tab[TB_BZH * A.tn + j], rkzh = tab[TB_RKZH * A.tn + j];
        }
        if (u4) {
            T e1 = d4(xs[0][lane + 2], xs[0][lane + 1], xs[0][lane + 3], xs[0][lane], rdx);
            T e2 = d4(xs[1][lane + 2], xs[1][lane + 1], xs[1][lane + 3], xs[1][lane], rdx);
            T e3 = d4(sxy_0, sxy_m1, sxy_p1, sxy_m2, rdz);
            if (sx >= 0) {
                e1 = cpml_from(e1, o_dsxx, pxn_(qx) + PX_DSXX, keep, ax, bx, rkx);
                e2 = cpml_from(e2, o_dpx, pxn_(qx) + PX_DPX, keep, ax, bx, rkx);
            }
            if (szj >= 0) e3 = cpml_from(e3, o_dsxyz, pzn_(qz) + PZ_DSXY, keep, az, bz, rkz);
            const Vec4<T> mv_j = mv_at(id_j, m_j);
            if constexpr (PAL)
                biot_velocity_cn(v.x, v.z, e1 + e3, e2, mv_j.x, mv_j.y, mv_j.z, lpal[PAL_CN * kMaxPalette2 + id_j]);
            else
                biot_velocity(v.x, v.z, e1 + e3, e2, mv_j.x, mv_j.y, mv_j.z, mv_j.w, A.inv_dt);
        }
        if (u5) {
            T e1 = d4(xs[2][lane + 3], xs[2][lane + 2], xs[2][lane + 4], xs[2][lane + 1], rdx);
            T e2 = d4(syy_p1, syy_0, syy_p2, syy_m1, rdz);
            T e3 = d4(p_p1, p_0, p_p2, p_m1, rdz);
            if (sx >= 0) e1 = cpml_from(e1, o_dsxy, pxn_(qx) + PX_DSXY, keep, axh, bxh, rkxh);
            if (szj >= 0) {
                e2 = cpml_from(e2, o_dsyy, pzn_(qz) + PZ_DSYY, keep, azh, bzh, rkzh);
                e3 = cpml_from(e3, o_dpz, pzn_(qz) + PZ_DPZ, keep, azh, bzh, rkzh);
            }
            T div = e1 + e2;
            if (src_f && j == A.src_z) div += srcv;
            if constexpr (AVG) {
                const Vec4<T> mvy = mean_rows(mv_at(id_j, m_j), mv_at(id_j1, m_j1));
                biot_velocity(v.y, v.w, div, e3, mvy.x, mvy.y, mvy.z, mvy.w, A.inv_dt);
            } else if constexpr (PAL) {
                const Vec4<T> mv_j1 = lpal[PAL_MV * kMaxPalette2 + id_j1];
                biot_velocity_cn(v.y, v.w, div, e3, mv_j1.x, mv_j1.y, mv_j1.z, lpal[PAL_CN * kMaxPalette2 + id_j1]);
            } else {
                biot_velocity(v.y, v.w, div, e3, m_j1.x, m_j1.y, m_j1.z, m_j1.w, A.inv_dt);
            }
        }
        return v;
    };
    auto px_in = [&](long long q) { return px(A.in, q); };
    auto pz_in = [&](long long q) { return pz(A.in, q); };
    auto px_out = [&](long long q) { return px(A.out, q); };
    auto pz_out = [&](long long q) { return pz(A.out, q); };

    // Register rings; row r of a ring of size R sits in slot (r - s0) mod R, so at phase PH of the
    // unrolled group row s + k is in slot (PH + k) mod R.
    const int s0 = zA - 6;
    const int send = zB + 6;
    // The HBM loads run ahead: strains of row s+2 and material indices of row s+3 are requested at
    // the start of row step s, step-n velocities of row s+4 once row s-2's last use in row step s
    // is done (edge tiles) or of row s+5 at its start, two rows in flight in `pend` (interior).
    Vec4<T> vr[6];   // step-n velocities, rows s-2 .. s+3
    Vec4<T> pend[2]; // interior tiles: step-n velocities of rows s+3 (until moved to vr), s+4
    Vec4<T> sr[3];   // step-n strains, rows s .. s+2
    int idr[6];      // material index, rows s-1 .. s+3 (s-1 .. s+2 once row s-3 moved on)
    int idr2[6];     // material index, rows s-6 .. s-1 (the rows stage 2 and the v_x update read)
    T sxyr[6], syyr[6], pr[6];     // stage-1 stress rows s-4 .. s
    Vec4<T> v1r[6];  // step-(n+1) velocities, rows s-7 .. s-2
    Vec4<T> s1r[6];  // step-(n+1) strains, rows s-4 .. s
    T sxyr2[6], syyr2[6], pr2[6];  // stage-2 stress rows s-8 .. s-4
    // groups: coefficients of rows s-4 .. s+1 (MS) and s-6 .. s-1 (MV), each row's load issued one
    // row step before its first use, into the slot of the row six back, whose last use (stage 2)
    // has just passed
    Vec4<T> msr[6], mvr[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        msr[q] = mvr[q] = zero4;
        sxyr[q] = syyr[q] = pr[q] = T(0);
        sxyr2[q] = syyr2[q] = pr2[q] = T(0);
        idr[q] = idr2[q] = 0;
        v1r[q] = s1r[q] = zero4;
    }
#pragma unroll
    for (int q = -2; q <= 2; ++q) vr[(q + 6) % 6] = ld(Vo, s0 + q);
    if constexpr (EDGE) {
        vr[3] = ld(Vo, s0 + 3);
    } else {
        pend[0] = ld(Vo, s0 + 3);
        pend[1] = ld(Vo, s0 + 4);
    }
    sr[0] = ld(So, s0);
    sr[1] = ld(So, s0 + 1);
    sr[2] = zero4;
    if constexpr (PAL) {
        idr[5] = ldm(s0 - 1);
        idr[0] = ldm(s0);
        idr[1] = ldm(s0 + 1);
        idr[2] = ldm(s0 + 2);
    } else {
        msr[0] = ld(MSg, s0);
        msr[1] = ld(MSg, s0 + 1);
        mvr[4] = ld(MVg, s0 - 2);
        mvr[5] = ld(MVg, s0 - 1);
    }
    __syncthreads();  // palette in LDS

    int s = s0;
    auto row_step = [&](auto phase) __attribute__((always_inline)) {
        constexpr int PH = decltype(phase)::value;
        auto V = [&](int q) -> Vec4<T> & { return vr[(PH + q + 12) % 6]; };    // step-n velocity row s+q
        auto V1 = [&](int q) -> Vec4<T> & { return v1r[(PH + q + 12) % 6]; };  // step-(n+1) velocity row s+q
        const Vec4<T> so = sr[PH % 3];
        // ---- loads for later row steps and the material of this one
        if constexpr (!EDGE) {
            V(3) = pend[PH % 2];  // into the slot of row s-3, dead since row step s-1
            pend[PH % 2] = ld(Vo, s + 5);
        }
        sr[(PH + 2) % 3] = ld(So, s + 2);
        // material through the palette in LDS, looked up where used: stage 1 reads MS of rows s,
        // s+1 and MV of rows s-2, s-1; stage 2 MS of rows s-4, s-3 and MV of rows s-6, s-5
        if constexpr (PAL) {
            idr2[(PH + 5) % 6] = idr[(PH + 5) % 6];  // row s-1 moves to the older ring
            idr[(PH + 3) % 6] = ldm(s + 3);           // into the slot of row s-3, moved on at row step s-2
        }

        // ---- x neighbours through LDS: step-n velocities of row s, step-(n+1) velocities of row s-3
        {
            T(*xv)[BW + 4] = lv[PH % 3];
            xv[0][lane + 2] = V(0).x;
            xv[1][lane + 2] = V(0).z;
            xv[2][lane + 2] = V(0).y;
            T(*xv2)[BW + 4] = lv2[(PH + 3) % R2];
            xv2[0][lane + 2] = V1(-3).x;
            xv2[1][lane + 2] = V1(-3).z;
            xv2[2][lane + 2] = V1(-3).y;
        }
        __syncthreads();

        // ---- receivers: step n at row s-1 (stage 1), step n+1 at row s-5 (stage 2)
        if (rec1 && s - 1 >= zA && s - 1 < zB) gather(trace1, s - 1, lv[(PH + 2) % 3], ls[(PH + 2) % 3]);
        if (rec2 && s - 5 >= zA && s - 5 < zB) gather(trace2, s - 5, lv2[(PH + 1) % R2], ls2[(PH + 1) % 3]);

        // ---- stage 1: stresses of row s (step n), kept for stage 2 in the register ring
        T sxx, p, syy;
        const Vec4<T> s1 = stress(s, V(-2), V(-1), V(0), V(1), V(2), so, lv[PH % 3], idr[PH % 6], idr[(PH + 1) % 6],
                                  msr[PH % 6], msr[(PH + 1) % 6], px_in, pxm, pz_in, pzm, ps1, src1, sxx, p, syy);
        s1r[PH % 6] = s1;
        {
            T(*xs)[BW + 4] = ls[PH % 3];
            xs[0][lane + 2] = sxx;
            xs[1][lane + 2] = p;
            xs[2][lane + 2] = s1.w;
        }
        sxyr[PH % 6] = s1.w;
        syyr[PH % 6] = syy;
        pr[PH % 6] = p;
        auto SXY = [&](int q) -> T { return sxyr[(PH + q + 12) % 6]; };  // stage-1 row s+q
        auto SYY = [&](int q) -> T { return syyr[(PH + q + 12) % 6]; };
        auto PP = [&](int q) -> T { return pr[(PH + q + 12) % 6]; };

        // ---- stage 1: velocities of row s-2 (step n -> n+1); stage 2 reads rows [zA-4, zB+3]
        if (s - 2 >= zA - 4 && s - 2 <= zB + 3)
            V1(-2) = velocity(s - 2, V(-2), ls[(PH + 1) % 3], SXY(-4), SXY(-3), SXY(-2), SXY(-1), SYY(-3), SYY(-2),
                              SYY(-1), SYY(0), PP(-3), PP(-2), PP(-1), PP(0), idr2[(PH + 4) % 6], idr[(PH + 5) % 6],
                              mvr[(PH + 4) % 6], mvr[(PH + 5) % 6], px_in, pxm, pz_in, pzm, pv1, src1);
        else
            V1(-2) = V(-2);
        if constexpr (EDGE) V(4) = ld(Vo, s + 4);  // into the slot of row s-2, now dead

        // ---- stage 2: stresses of row s-4 (step n+1)
        const int r2 = s - 4;
        const bool srow = own && r2 >= zA && r2 < zB;
        // stage 2's velocities read the stresses of rows [zA-2, zB+1]; other rows are skipped
        T sxx2 = 0, p2 = 0, syy2 = 0;
        Vec4<T> s2 = zero4;
        if (r2 >= zA - 2 && r2 <= zB + 1) {
            s2 = stress(r2, V1(-6), V1(-5), V1(-4), V1(-3), V1(-2), s1r[(PH + 2) % 6], lv2[(PH + 2) % R2],
                        idr2[(PH + 2) % 6], idr2[(PH + 3) % 6], msr[(PH + 2) % 6], msr[(PH + 3) % 6], pxm, px_out,
                        pzm, pz_out, srow, src2, sxx2, p2, syy2);
            if (srow) *(GN *)(Sn + off(r2)) = nvec4<T>{s2.x, s2.y, s2.z, s2.w};
        }
        {
            T(*xs)[BW + 4] = ls2[(PH + 2) % 3];
            xs[0][lane + 2] = sxx2;
            xs[1][lane + 2] = p2;
            xs[2][lane + 2] = s2.w;
        }
        sxyr2[(PH + 2) % 6] = s2.w;
        syyr2[(PH + 2) % 6] = syy2;
        pr2[(PH + 2) % 6] = p2;
        auto SXY2 = [&](int q) -> T { return sxyr2[(PH + q + 12) % 6]; };  // stage-2 row s+q
        auto SYY2 = [&](int q) -> T { return syyr2[(PH + q + 12) % 6]; };
        auto PP2 = [&](int q) -> T { return pr2[(PH + q + 12) % 6]; };

        // ---- stage 2: velocities of row s-6 (step n+1 -> n+2)
        const int j2 = s - 6;
        if (j2 >= zA && j2 < zB) {
            const bool vrow = own;
            const Vec4<T> v2 = velocity(j2, V1(-6), ls2[PH % 3], SXY2(-8), SXY2(-7), SXY2(-6), SXY2(-5), SYY2(-7),
                                        SYY2(-6), SYY2(-5), SYY2(-4), PP2(-7), PP2(-6), PP2(-5), PP2(-4),
                                        idr2[PH % 6], idr2[(PH + 1) % 6], mvr[PH % 6], mvr[(PH + 1) % 6], pxm, px_out,
                                        pzm, pz_out, vrow, src2);
            if (vrow) *(GN *)(Vn + off(j2)) = nvec4<T>{v2.x, v2.y, v2.z, v2.w};
        }
        if constexpr (!PAL) {
            msr[(PH + 2) % 6] = ld(MSg, s + 2);  // the slot of row s-4, used by stage 2 above for the last time
            mvr[PH % 6] = ld(MVg, s);            // the slot of row s-6, likewise
        }
        ++s;
    };

    // Whole groups of 6 row steps: the last group may run up to 5 steps past zB + 5, over rows that
    // store nothing and read zero pad rows (loads clamp to the last pad row).
    for (int g = (send - s0 + 5) / 6; g > 0; --g) {
        row_step(std::integral_constant<int, 0>{});
        row_step(std::integral_constant<int, 1>{});
        row_step(std::integral_constant<int, 2>{});
        row_step(std::integral_constant<int, 3>{});
        row_step(std::integral_constant<int, 4>{});
        row_step(std::integral_constant<int, 5>{});
    }
}

// Three waves per SIMD: at most 168 VGPRs, and LDS for three blocks of 256 threads per CU (the
// palette copy is cut to kMaxPalette2 entries).  Each XCD's share of the work list starts with its
// edge tiles (tile index stored as ~tile), whose chains of C-PML record accesses make them
// latency-bound and about twice as long as interior tiles: started first, they overlap with the
// interior tiles that follow instead of finishing the launch on their own.  Items with a negative
// model index pad the shares to equal length.
template <class T, int BW, bool AVG, bool PAL>
__global__ void __launch_bounds__(BW, PAL ? 3 : 2)
    fused_step2(const StepArgs<T> *__restrict__ models, const WorkItem *__restrict__ work, int nwork, int per_xcd,
                const long long *__restrict__ base, int k)
{
    __shared__ T lv[3][3][BW + 4];
    __shared__ T ls[3][3][BW + 4];
    __shared__ T lv2[6][3][BW + 4];
    __shared__ T ls2[3][3][BW + 4];
    __shared__ Vec4<T> lpal[PAL_SECTIONS * kMaxPalette2];
    const int wi = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (wi >= nwork) return;
    const WorkItem item = work[wi];
    if (item.model < 0) return;
    const long long n = *(const kcll *)base + k;  // the launch advances steps n and n + 1
    if (item.tile < 0)
        step2_tile<T, BW, AVG, true, PAL>(models[item.model], ~item.tile, n, lv, ls, lv2, ls2, lpal);
    else
        step2_tile<T, BW, AVG, false, PAL>(models[item.model], item.tile, n, lv, ls, lv2, ls2, lpal);
}

template <class T, bool AVG, bool PAL>
hipError_t launch_fused2_bw(const StepArgs<T> *models, const WorkItem *work, int nwork, int per_xcd, int bw,
                            const long long *base, int k, hipStream_t stream)
{
    const dim3 grid(8 * per_xcd);
    switch (bw) {
    case 64: fused_step2<T, 64, AVG, PAL><<<grid, 64, 0, stream>>>(models, work, nwork, per_xcd, base, k); break;
    case 128: fused_step2<T, 128, AVG, PAL><<<grid, 128, 0, stream>>>(models, work, nwork, per_xcd, base, k); break;
    case 256: fused_step2<T, 256, AVG, PAL><<<grid, 256, 0, stream>>>(models, work, nwork, per_xcd, base, k); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class T>
hipError_t launch_fused2(const StepArgs<T> *models, const WorkItem *work, int nwork, int per_xcd, int bw, bool palette,
                         bool avg, const long long *base, int k, hipStream_t stream)
{
    if (nwork < 1 || per_xcd < 1 || 8LL * per_xcd < nwork) return hipErrorInvalidValue;
    if (palette)
        return avg ? launch_fused2_bw<T, true, true>(models, work, nwork, per_xcd, bw, base, k, stream)
                   : launch_fused2_bw<T, false, true>(models, work, nwork, per_xcd, bw, base, k, stream);
    return avg ? launch_fused2_bw<T, true, false>(models, work, nwork, per_xcd, bw, base, k, stream)
               : launch_fused2_bw<T, false, false>(models, work, nwork, per_xcd, bw, base, k, stream);
}

template hipError_t launch_fused2<float>(const StepArgs<float> *, const WorkItem *, int, int, int, bool, bool,
                                         const long long *, int, hipStream_t);

}  // namespace poe2d
