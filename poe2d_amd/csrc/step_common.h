// Device-side building blocks shared by the time-step kernels.
//
// Field layout.  Point data live in "groups": one 4-vector of reals per point, stored as an SoA
// array of 4-vectors with row pitch `pitch` (a multiple of 64 points).  Each group holds the four
// quantities one stage of the update consumes together, so a wave moves 64 x 16 B (FP32) of
// consecutive memory per load:
//   MS  (alpha, M, lambda_c, mu)       coefficients of the stress update     } material arena,
//   MV  (rho, rho_f, m, b)             coefficients of the velocity update   } shared by clones
//   V   (vx, vy, wx, wy)               velocities, double-buffered           } state arena,
//   S   (e_xx, e_yy, zeta, sigma_xy)   strains, fluid content, shear stress  } one per handle
//   I   (p, sigma_xx, sigma_yy, -)     intermediates of the two-pass schedule }
// The groups of an arena sit `group_stride` points apart; every group carries zero pad rows
// above/below and pad columns left/right of the (nz+1) x (nx+1) grid, and `origin` is the offset
// of point (0, 0).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/poe2d.h"

namespace poe2d {

template <class T>
struct alignas(4 * sizeof(T)) Vec4 {
    T x, y, z, w;
};

template <class T>
struct NativeVec4 {
    typedef T type __attribute__((ext_vector_type(4)));
};
template <class T>
using nvec4 = typename NativeVec4<T>::type;

// Global-address-space views of device pointers.  Pointers read from device memory (the
// work-list arguments) carry no address space, and loads through them would be flat loads
// (counted together with LDS traffic, no scalar base register); a kernel casts them once.
// Global loads and stores of a Vec4 go through the native 4-wide vector type nvec4<T>.
typedef __attribute__((address_space(1))) char gchar;
typedef __attribute__((address_space(1))) const unsigned char gcuchar;
typedef __attribute__((address_space(1))) const int gcint;
// The constant address space: loads through it of values uniform across a block (tables that no
// kernel of the launch writes) become scalar loads.
typedef __attribute__((address_space(4))) const int kcint;
typedef __attribute__((address_space(4))) const long long kcll;

enum { G_MS = 0, G_MV = 1, NMAT_GROUPS = 2 };                              // material arena
enum { G_V0 = 0, G_S0 = 1, G_V1 = 2, G_S1 = 3, G_I = 4, NSTATE_GROUPS = 5 };  // state arena
__host__ __device__ constexpr int group_v(int buf) { return G_V0 + 2 * buf; }
__host__ __device__ constexpr int group_s(int buf) { return G_S0 + 2 * buf; }

constexpr int kPadRows = 12;  // zero rows above and below the grid
constexpr int kPadCols = 4;  // zero columns left of the grid (>= 1 more on the right)

// C-PML memory variables.  Each strip point carries a record of 6 values per state buffer.
// x strips (columns with a nonzero x profile), record index z * wxs + strip column:
enum {
    PX_DVX = 0,   // d vx / dx in the strain update          (half-grid x)
    PX_DWX = 1,   // d wx / dx in the fluid-content update   (half-grid x)
    PX_DVY = 2,   // d vy / dx in the shear-stress update    (grid x)
    PX_DSXX = 3,  // d sxx / dx in the vx update             (grid x)
    PX_DPX = 4,   // d p / dx in the wx update               (grid x)
    PX_DSXY = 5,  // d sxy / dx in the vy update             (half-grid x)
    NPSI_X = 6
};
// z strips (rows with a nonzero z profile), record index strip row * pitch + x:
enum {
    PZ_DVZ = 0,   // d vy / dz in the strain update          (grid z)
    PZ_DWZ = 1,   // d wy / dz in the fluid-content update   (grid z)
    PZ_DVX = 2,   // d vx / dz in the shear-stress update    (half-grid z)
    PZ_DSXY = 3,  // d sxy / dz in the vx update             (grid z)
    PZ_DSYY = 4,  // d syy / dz in the vy update             (half-grid z)
    PZ_DPZ = 5,   // d p / dz in the wy update               (half-grid z)
    NPSI_Z = 6
};
constexpr int kMaxPalette = 256;  // distinct (MS, MV) tuples a one-byte material index can select
// palette sections of kMaxPalette Vec4 entries each: MS, MV, and (FP32 grids) the velocity
// update's constants of each MV entry (see biot_velocity_cn)
enum { PAL_MS = 0, PAL_MV = 1, PAL_CN = 2, PAL_SECTIONS = 3 };

// 1-D C-PML profiles, one table each, `tn` entries apart.  The K tables hold 1/K (K = 1 outside
// the layers, and wherever K_max = 1).
enum { TB_AX, TB_BX, TB_RKX, TB_AXH, TB_BXH, TB_RKXH, TB_AZ, TB_BZ, TB_RKZ, TB_AZH, TB_BZH, TB_RKZH, NTABLES };

// Everything a step kernel needs: the arena of groups, the memory-variable arena (x strips of
// both buffers, then z strips of both buffers), the profile tables and the strip maps.
template <class T>
struct StepArgs {
    int nx, nz, pitch, wxs;
    long long origin;        // offset of point (0,0) inside a group
    long long group_stride;  // points between consecutive groups of an arena
    long long pxn, pzn;      // x-strip / z-strip records per state buffer
    int tn;                  // entries per profile table
    int in, out;             // state buffer read / written by this step
    T rdx, rdz, dt, inv_dt;  // 1/(24 dx), 1/(24 dz), dt, 1/dt
    Vec4<T> *grp;            // state arena (V, S of both buffers, I)
    const Vec4<T> *mat;      // material arena (MS, MV)
    T *psi;                  // memory variables
    const T *tab;            // profile tables
    const int *xmap, *zmap;  // strip index of each column / row, -1 outside the strips
    int ztop, zbot;          // the z strips are rows [0, ztop) and [zbot, nz] (see zstrip)
    // material palette: when the grid holds at most kMaxPalette distinct coefficient tuples, a
    // one-byte index per point (laid out like a group, one byte per point) selects the MS entry
    // pal[i] and the MV entry pal[kMaxPalette + i]; npal = 0 when the grid has no palette
    const unsigned char *mid;
    const Vec4<T> *pal;
    int npal;
    int src_kind, src_x, src_z;
    T src_val;               // source sample of this step
    // Work-list launches take the step index n = base + k (base in device memory, k a launch
    // argument): the state buffer read is (n + par) & 1, and the source sample is
    // src_series[n - src_n0] for n in [src_n0, src_n0 + src_len), else 0.
    const T *src_series;
    long long src_n0;
    int src_len, par;
    // tiling of the single-pass kernel
    int th, ntx, ntiles, per_xcd;
    // receivers, gathered by the single-pass kernel: rec_entries holds (x, receiver index) sorted
    // by tile column, row and x; the entries of row z in tile column tx are
    // [rec_tile_ptr[tx * (nz+1) + z], rec_tile_ptr[tx * (nz+1) + z + 1]), so the rows of one tile
    // are contiguous.  Steps n with n % rec_dec == 0 and n / rec_dec < rec_nsamp write slot
    // n / rec_dec of traces ([slot][comp][receiver], components of rec_mask).
    const int *rec_tile_ptr;
    const int2 *rec_entries;
    T *traces;
    int rec_n, rec_ncomp, rec_dec, rec_nsamp;
    unsigned rec_mask;
    // the two-step kernel: memory variables of the intermediate step (x strips, then z strips,
    // one record set), and the buffer rule: the launch at step n reads buffer
    // ((n - pair0) / 2 + pcur) & 1 and writes the other
    T *psi_mid;
    long long pair0;
    int pcur;
};

// Step-dependent fields of a work-list launch at step n.
template <class T>
__device__ __forceinline__ void at_step(StepArgs<T> &A, long long n)
{
    A.in = (int)((n + A.par) & 1);
    A.out = A.in ^ 1;
    const long long k = n - A.src_n0;
    typedef __attribute__((address_space(1))) const T GCT;
    A.src_val = (A.src_series && k >= 0 && k < A.src_len) ? ((GCT *)A.src_series)[k] : T(0);
}

// Point (0, 0) of state group g / material group g.
template <class T>
__device__ __forceinline__ Vec4<T> *state_group(const StepArgs<T> &A, int g)
{
    return A.grp + A.origin + (long long)g * A.group_stride;
}
template <class T>
__device__ __forceinline__ const Vec4<T> *mat_group(const StepArgs<T> &A, int g)
{
    return A.mat + A.origin + (long long)g * A.group_stride;
}
// Memory-variable record of x-strip record `r` / z-strip record `r` in state buffer `buf`.
template <class T>
__device__ __forceinline__ T *psi_x(const StepArgs<T> &A, int buf, long long r)
{
    return A.psi + ((long long)buf * A.pxn + r) * NPSI_X;
}
template <class T>
__device__ __forceinline__ T *psi_z(const StepArgs<T> &A, int buf, long long r)
{
    return A.psi + 2LL * NPSI_X * A.pxn + ((long long)buf * A.pzn + r) * NPSI_Z;
}

// z-strip index of row z (uniform arithmetic, equal to zmap[z]), -1 outside the strips.
template <class T>
__device__ __forceinline__ int zstrip(const StepArgs<T> &A, int z)
{
    return z < 0 ? -1 : z < A.ztop ? z : z < A.zbot ? -1 : z <= A.nz ? A.ztop + (z - A.zbot) : -1;
}

// Fourth-order staggered difference (27 (f1 - f0) - (f2 - fm)) / (24 h) between f0 and f1, with
// f2 the far point beyond f1 and fm the far point before f0; rh = 1 / (24 h).
template <class T>
__device__ __forceinline__ T d4(T f1, T f0, T f2, T fm, T rh)
{
    return (T(27) * (f1 - f0) - (f2 - fm)) * rh;
}

// C-PML correction of a derivative d: psi_new = b psi_old + a d (stored when `store`), returns
// d / K + psi_new, with rk = 1/K.
template <class T, class PN>
__device__ __forceinline__ T cpml_from(T d, T psi_old, PN psi_n, bool store, T a, T b, T rk)
{
    const T s = b * psi_old + a * d;
    if (store) *psi_n = s;
    return d * rk + s;
}

// The same, reading psi_old from psi_o.  Kernels whose old and new records are different memory
// read every old record of an update first and call cpml_from: the compiler cannot tell the two
// record sets apart and would otherwise hold each read until the write before it is done, one
// memory latency per record.
template <class T, class PO, class PN>
__device__ __forceinline__ T cpml(T d, PO psi_o, PN psi_n, bool store, T a, T b, T rk)
{
    return cpml_from(d, T(*psi_o), psi_n, store, a, b, rk);
}

// Quotient used by the velocity update: IEEE division in FP64; in FP32 the hardware reciprocal
// (1 ulp) times the numerator, within 2 ulp of the IEEE quotient.
__device__ __forceinline__ float quot(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double quot(double a, double b) { return a / b; }

// Biot momentum update of one velocity point: solid velocity v and relative fluid velocity w
// from the stress divergence `div` and the pressure gradient `grad`,
//   [rho rho_f; rho_f m] d/dt [v; w] = [div; -grad - b w],
// with the viscous term b w taken Crank-Nicolson ((w_old + w_new) / 2), as the reference's vx and
// vz kernels do (seismic_CPML_2D_poroelastic_withCUDA.cu:293-311).
// The material's constants co = (rho m - rho_f^2) / dt, c1 = co + rho b / 2, c2 = co - rho b / 2,
// and the two numerators, shared by every form of the update so that all of them round alike.
template <class T>
__device__ __forceinline__ void cn_coeffs(T rho, T rhof, T m, T b, T inv_dt, T &co, T &c1, T &c2)
{
    co = (rho * m - rhof * rhof) * inv_dt;
    const T half_rb = rho * b * T(0.5);
    c1 = co + half_rb;
    c2 = co - half_rb;
}
template <class T>
__device__ __forceinline__ T cn_wnum(T c2, T wold, T div, T grad, T rho, T rhof)
{
    return c2 * wold + (-rhof * div - rho * grad);
}
template <class T>
__device__ __forceinline__ T cn_vnum(T div, T grad, T wavg, T m, T rhof, T rhofb)
{
    return m * div + rhof * grad + rhofb * wavg;
}
template <class T>
__device__ __forceinline__ void biot_velocity(T &v, T &w, T div, T grad, T rho, T rhof, T m, T b,
                                              T inv_dt)
{
    T co, c1, c2;
    cn_coeffs(rho, rhof, m, b, inv_dt, co, c1, c2);
    const T wold = w;
    const T wnew = quot(cn_wnum(c2, wold, div, grad, rho, rhof), c1);
    const T wavg = (wold + wnew) * T(0.5);
    v = v + quot(cn_vnum(div, grad, wavg, m, rhof, rhof * b), co);
    w = wnew;
}

// The FP32 update with a palette entry's constants dv = (c2, 1/c1, rho_f b, 1/co), as
// cn_palette_kernel computes them (cn_coeffs, the hardware reciprocal quot uses): equal bit for
// bit to biot_velocity on the entry's (rho, rho_f, m, b), without its per-point divisions.  The
// products with the reciprocals are rounded on their own (times_recip), as quot's are, so that
// no multiply-add forms across them.
__device__ __forceinline__ float times_recip(float a, float r) { return a * r; }
__device__ __forceinline__ void biot_velocity_cn(float &v, float &w, float div, float grad, float rho, float rhof,
                                                 float m, const Vec4<float> &dv)
{
    const float wold = w;
    const float wnew = times_recip(cn_wnum(dv.x, wold, div, grad, rho, rhof), dv.y);
    const float wavg = (wold + wnew) * 0.5f;
    v = v + times_recip(cn_vnum(div, grad, wavg, m, rhof, dv.z), dv.w);
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

// Stress-update index ranges of the reference kernels (seismic_CPML_2D_poroelastic_withCUDA.cu):
// stress1 x in [1,NX), z in [2,NZ-1) (:60-62); stress2 x in [2,NX), z in [1,NZ-1) (:212-214);
// stress3 x in [1,NX-1), z in [2,NZ) (:237-239); vx x in [2,NX), z in [2,NZ) (:289-291);
// vz x in [1,NX-1), z in [1,NZ-1) (:326-328).

// Material at the staggered points under POE2D_MATERIAL_AVERAGED: the shear modulus of sigma_xy
// at (x, z+1/2) as the harmonic mean of the two points, the velocity-row coefficients of v_y at
// (x, z+1/2) as the arithmetic means.
template <class T>
__device__ __forceinline__ T harmonic_mu(T a, T b)
{
    return T(2) / (T(1) / a + T(1) / b);
}

template <class T>
__device__ __forceinline__ Vec4<T> mean_rows(const Vec4<T> &a, const Vec4<T> &b)
{
    return Vec4<T>{T(0.5) * (a.x + b.x), T(0.5) * (a.y + b.y), T(0.5) * (a.z + b.z), T(0.5) * (a.w + b.w)};
}

// phase 0: stresses, 1: velocities; avg: POE2D_MATERIAL_AVERAGED.
template <class T>
void launch_twopass(const StepArgs<T> &A, hipStream_t stream, int phase, bool avg);

// One work item of the single-pass step: tile `tile` of model `model`.
struct WorkItem {
    int model, tile;
};

// Launches the single-pass step over a work list: block width bw (128, 256 or 512 threads), step
// index *base + k, model arguments from models[] (each with its own tiling and receivers; in, out
// and the source sample are set from the step).  With `palette` every model is read through its
// one-byte material index (needs npal > 0 for each).  nwork items, placed per_xcd to an XCD.
// avg: every model uses POE2D_MATERIAL_AVERAGED (else every model uses AS_WRITTEN).
template <class T>
hipError_t launch_fused(const StepArgs<T> *models, const WorkItem *work, int nwork, int per_xcd, int bw,
                        bool palette, bool avg, const long long *base, int k, hipStream_t stream);

// *base += delta (one thread).
hipError_t launch_base_advance(long long *base, int delta, hipStream_t stream);

// Output columns of one tile for block width bw (halo of 4 columns on each side).
inline int fused_tile_width(int bw) { return bw - 8; }

// The two-step kernel (FP32): output columns per tile (halo of 6 columns on each side), and the
// launch; n = *base + k is the first of the two steps.  With `palette` every model is read through
// its one-byte material index and a palette of at most kMaxPalette2 entries, else through the
// coefficient groups.
__host__ __device__ constexpr int fused2_tile_width(int bw) { return bw - 12; }
constexpr int kMaxPalette2 = 28;  // palette entries the two-step kernel keeps in LDS (blocks of 64 threads: 12 per CU)
// Items whose tile field is negative (~tile) are edge tiles: tiles touching a C-PML strip or
// holding receivers (the others run a variant without that code).
template <class T>
hipError_t launch_fused2(const StepArgs<T> *models, const WorkItem *work, int nwork, int per_xcd, int bw, bool palette,
                         bool avg, const long long *base, int k, hipStream_t stream);
// Rows and columns around a two-step tile whose points the kernel computes (halos of both steps and
// the over-run of its row loop): a tile with a C-PML strip point, or a point of the grid's two
// outer rows or columns on either side, in this range is an edge tile.
constexpr int kFused2RowsBefore = 10, kFused2RowsAfter = 12, kFused2ColsBefore = 6;

}  // namespace poe2d
