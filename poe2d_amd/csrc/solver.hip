// Solver handle behind the C ABI of include/poe2d.h: device memory, the time loop, gather,
// snapshots, checkpoint/restore and the stability check.
//
// Device memory of a handle (field layout described in step_common.h):
//   arena   NGROUPS groups of 4-vectors (coefficients MS, MV; state V, S for buffers 0 and 1;
//           two-pass intermediates I), each (nz+1+2*kPadRows) rows x pitch points, zero pads;
//   mid/pal when the grid holds at most kMaxPalette distinct coefficient tuples: a one-byte
//           material index per point (group layout, pads select entry 0) and the palette of
//           (MS, MV) entries, which the single-pass kernel reads instead of the MS/MV groups;
//   psi     C-PML memory variables: 6 x-strip and 6 z-strip arrays per state buffer;
//   tables  the 12 1-D C-PML profiles;
//   maps    strip index of each column and row.
// A single-pass step reads state buffer `cur` and writes buffer 1-cur.  The reference's stress1
// reads column nx+1 at x = nx-1, which in its flat array is column 0 of the next row and always
// zero; the zero pad column nx+1 gives the same value.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "aux_kernels.h"
#include "common.h"
#include "step_common.h"

using namespace poe2d;

namespace {

const float kRefPiF = 3.1415926535897932f;  // the reference's float PI (seismic_CPML_2d_poroelastic.cpp:9)
const double kStabilityThreshold = 1.0e25;  // STABILITY_THRESHOLD of the reference (declared, unused there)

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(POE2D_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

// single-pass (palette when the grid has one), two-pass, single-pass always on the MS/MV groups
enum { SCHED_FUSED = 0, SCHED_TWOPASS = 1, SCHED_FUSED_GROUPS = 2 };

}  // namespace

struct poe2d_solver {
    int precision = POE2D_F32;
    size_t rs = 4;  // sizeof(real)
    poe2d_grid grid{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    size_t pitch = 0, alloc_rows = 0;  // points per row, rows per group (with pads)
    int wxs = 0, wzs = 0;              // x-strip columns, z-strip rows
    int ztop = 0, zbot = 0;            // z strips: rows [0, ztop) and [zbot, nz]
    int tn = 0;                        // entries per profile table
    double cp_max = 0;
    int cur = 0;
    void *arena = nullptr;
    void *psi = nullptr;
    void *tab = nullptr;
    int *xmap = nullptr, *zmap = nullptr;
    unsigned char *mid = nullptr;  // material index per point (palette mode)
    void *pal = nullptr;           // 2 * kMaxPalette Vec4 entries
    int npal = 0;                  // palette entries, 0 when the grid has more than kMaxPalette
    float *red = nullptr;  // reduction scratch: [0] result, [1..] per-block partials
    int red_blocks = 1024;
    void *dense = nullptr;  // (nz+1) x (nx+1) scratch for snapshots, made on first use
    std::vector<double> htab[NTABLES];
    // schedule
    int schedule = SCHED_FUSED;
    int block_w = 256, tile_rows = 14;
    // source
    bool have_src = false;
    poe2d_source src{};
    double t0 = 0;
    std::vector<double> user_w;
    bool have_user_w = false;
    // receivers
    int nrec = 0, dec = 1, nsamp = 0, ncomp = 0;
    unsigned mask = 0;
    int *rix = nullptr, *riz = nullptr;
    void *traces = nullptr;  // [slot][comp][rec]
    int step = 0;
    double last_ms = 0;
    int last_steps = 0;

    size_t group_points() const { return pitch * alloc_rows; }
    size_t group_bytes() const { return group_points() * 4 * rs; }
    size_t origin() const { return (size_t)kPadRows * pitch + kPadCols; }
    size_t pxn() const { return (size_t)std::max(wxs, 1) * (grid.nz + 1); }
    size_t pzn() const { return (size_t)std::max(wzs, 1) * pitch; }
    size_t psi_half_elems() const { return NPSI_X * pxn() + NPSI_Z * pzn(); }  // one buffer's worth
    size_t trace_bytes() const { return (size_t)nsamp * ncomp * nrec * rs; }
    unsigned char *group_ptr(int g) const { return (unsigned char *)arena + (size_t)g * group_bytes(); }

    double source_value(int n) const
    {
        if (!have_src) return 0.0;
        const double t = n * grid.dt;
        if (src.t_cut > 0 && t >= src.t_cut) return 0.0;
        if (have_user_w) return n < (int)user_w.size() ? src.amp * user_w[n] : 0.0;
        return src.amp * wavelet_shape(src.kind, grid.f0, t0, t);
    }
};

namespace {

void free_all(poe2d_solver *s)
{
    auto f = [](void *&p) {
        if (p) (void)hipFree(p);
        p = nullptr;
    };
    f(s->arena);
    f(s->psi);
    f(s->tab);
    f(s->pal);
    void *q = s->mid; f(q); s->mid = nullptr;
    q = s->xmap; f(q); s->xmap = nullptr;
    q = s->zmap; f(q); s->zmap = nullptr;
    q = s->red; f(q); s->red = nullptr;
    q = s->rix; f(q); s->rix = nullptr;
    q = s->riz; f(q); s->riz = nullptr;
    f(s->traces);
    f(s->dense);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->ev0 = s->ev1 = nullptr;
    s->stream = nullptr;
}

poe2d_status alloc_zero(poe2d_solver *s, void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(e == hipErrorOutOfMemory ? POE2D_E_NOMEM : POE2D_E_HIP,
                    std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    HIP_TRY(hipMemsetAsync(*p, 0, bytes, s->stream));
    return POE2D_OK;
}

// Coefficients into the MS (alpha, M, lambda_c, mu) and MV (rho, rho_f, m, b) groups, and, when
// the grid holds at most kMaxPalette distinct tuples (bitwise, after rounding to T), into the
// palette and the one-byte material index.
template <class T>
poe2d_status upload_coeffs(poe2d_solver *s, const poe2d_coeffs *c)
{
    const poe2d_grid &g = s->grid;
    const size_t cpitch = c->pitch ? c->pitch : (size_t)g.nx + 1;
    const void *ms_src[4] = {c->alpha, c->M, c->lambdac, c->mu};
    const void *mv_src[4] = {c->rho, c->rhof, c->m, c->b};
    auto at = [&](const void *a, size_t i) {
        return c->dtype == POE2D_F64 ? ((const double *)a)[i] : (double)((const float *)a)[i];
    };
    std::vector<Vec4<T>> host(s->group_points(), Vec4<T>{T(0), T(0), T(0), T(0)});
    for (int grp = 0; grp < 2; ++grp) {
        const void **src = grp == 0 ? ms_src : mv_src;
        for (int z = 0; z <= g.nz; ++z)
            for (int x = 0; x <= g.nx; ++x) {
                const size_t i = (size_t)z * cpitch + x;
                double v[4];
                for (int k = 0; k < 4; ++k) {
                    v[k] = at(src[k], i);
                    if (!std::isfinite(v[k])) return fail(POE2D_E_INVALID, "create: non-finite coefficient");
                }
                if (grp == 1) {
                    // the velocity update divides by rho m - rho_f^2 (the 2x2 mass matrix)
                    const double det = v[0] * v[2] - v[1] * v[1];
                    if (!(det > 0) || !(v[3] >= 0) || !(v[0] > 0))
                        return fail(POE2D_E_INVALID, "create: coefficients need rho > 0, rho m > rho_f^2 and "
                                                     "b >= 0 (point " + std::to_string(x) + ", " + std::to_string(z) + ")");
                }
                host[s->origin() + (size_t)z * s->pitch + x] = Vec4<T>{(T)v[0], (T)v[1], (T)v[2], (T)v[3]};
            }
        HIP_TRY(hipMemcpy(s->group_ptr(grp == 0 ? G_MS : G_MV), host.data(), s->group_bytes(), hipMemcpyHostToDevice));
    }

    // palette: distinct 8-tuples in order of first appearance
    using Key = std::array<T, 8>;
    auto less = [](const Key &a, const Key &b) { return std::memcmp(a.data(), b.data(), sizeof(Key)) < 0; };
    std::map<Key, int, decltype(less)> seen(less);
    std::vector<Key> pal;
    std::vector<unsigned char> idx(s->group_points(), 0);
    Key last{};
    int last_id = -1;
    for (int z = 0; z <= g.nz && pal.size() <= (size_t)kMaxPalette; ++z)
        for (int x = 0; x <= g.nx; ++x) {
            const size_t i = (size_t)z * cpitch + x;
            Key k;
            for (int q = 0; q < 4; ++q) {
                k[q] = (T)at(ms_src[q], i);
                k[4 + q] = (T)at(mv_src[q], i);
            }
            if (last_id < 0 || std::memcmp(k.data(), last.data(), sizeof(Key)) != 0) {
                auto it = seen.find(k);
                if (it == seen.end()) {
                    it = seen.emplace(k, (int)pal.size()).first;
                    pal.push_back(k);
                    if (pal.size() > (size_t)kMaxPalette) break;
                }
                last = k;
                last_id = it->second;
            }
            idx[s->origin() + (size_t)z * s->pitch + x] = (unsigned char)last_id;
        }
    if (pal.size() > (size_t)kMaxPalette) return POE2D_OK;  // no palette: the groups serve every schedule
    std::vector<Vec4<T>> hp(2 * kMaxPalette, Vec4<T>{T(0), T(0), T(0), T(0)});
    for (size_t e = 0; e < pal.size(); ++e) {
        hp[e] = Vec4<T>{pal[e][0], pal[e][1], pal[e][2], pal[e][3]};
        hp[kMaxPalette + e] = Vec4<T>{pal[e][4], pal[e][5], pal[e][6], pal[e][7]};
    }
    void *q = nullptr;
    poe2d_status st = alloc_zero(s, &q, s->group_points());
    if (st) return st;
    s->mid = (unsigned char *)q;
    if ((st = alloc_zero(s, &s->pal, hp.size() * sizeof(Vec4<T>)))) return st;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(s->mid, idx.data(), idx.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->pal, hp.data(), hp.size() * sizeof(Vec4<T>), hipMemcpyHostToDevice));
    s->npal = (int)pal.size();
    return POE2D_OK;
}

template <class T>
poe2d_status upload_tables(poe2d_solver *s)
{
    // the K tables go to the device as 1/K, so the kernels multiply
    std::vector<T> h((size_t)NTABLES * s->tn, T(0));
    for (int k = 0; k < NTABLES; ++k) {
        const bool recip = k == TB_RKX || k == TB_RKXH || k == TB_RKZ || k == TB_RKZH;
        for (size_t i = 0; i < s->htab[k].size(); ++i)
            h[(size_t)k * s->tn + i] = (T)(recip ? 1.0 / s->htab[k][i] : s->htab[k][i]);
    }
    HIP_TRY(hipMemcpy(s->tab, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return POE2D_OK;
}

// Strip map of one axis: index into the strip arrays of every point inside the absorbing strips,
// else -1.  A point is active where any profile deviates from the no-absorption values (a = 0,
// K = 1); each strip is the contiguous hull of the active points of its half of the axis, so the
// map is [0, lo) -> 0..lo-1 and [hi, n) -> lo..lo+n-hi-1.  Points of the hull that are not active
// have a = 0 and K = 1: their memory variable stays zero and their derivative is returned
// unchanged, so the hull only changes which points carry a (zero) memory variable.
// t = {a, b, K, a_half, b_half, K_half}.
std::vector<int> strip_map(const std::vector<double> *t, int n, int *count, int *lo, int *hi)
{
    auto active = [&](int i) { return t[0][i] != 0.0 || t[3][i] != 0.0 || t[2][i] != 1.0 || t[5][i] != 1.0; };
    int l = 0, h = n;
    for (int i = 0; i < n / 2; ++i)
        if (active(i)) l = i + 1;
    for (int i = n - 1; i >= n / 2; --i)
        if (active(i)) h = i;
    std::vector<int> map(n, -1);
    for (int i = 0; i < l; ++i) map[i] = i;
    for (int i = h; i < n; ++i) map[i] = l + (i - h);
    *count = l + (n - h);
    *lo = l;
    *hi = h;
    return map;
}

template <class T>
StepArgs<T> step_args(poe2d_solver *s, int in, int out)
{
    StepArgs<T> A{};
    const poe2d_grid &g = s->grid;
    A.nx = g.nx;
    A.nz = g.nz;
    A.pitch = (int)s->pitch;
    A.wxs = s->wxs;
    A.origin = (long long)s->origin();
    A.group_stride = (long long)s->group_points();
    A.pxn = (long long)s->pxn();
    A.pzn = (long long)s->pzn();
    A.tn = s->tn;
    A.in = in;
    A.out = out;
    A.rdx = T(1.0 / (24.0 * g.dx));
    A.rdz = T(1.0 / (24.0 * g.dz));
    A.dt = T(g.dt);
    A.inv_dt = T(1.0 / g.dt);
    A.grp = (Vec4<T> *)s->arena;
    A.psi = (T *)s->psi;
    A.tab = (const T *)s->tab;
    A.xmap = s->xmap;
    A.zmap = s->zmap;
    A.ztop = s->ztop;
    A.zbot = s->zbot;
    A.mid = s->mid;
    A.pal = (const Vec4<T> *)s->pal;
    A.npal = s->npal;
    A.src_kind = s->have_src ? s->src.kind : -1;
    A.src_x = s->have_src ? s->src.ix : -1;
    A.src_z = s->have_src ? s->src.iz : -1;
    A.src_val = T(0);
    const int W = fused_tile_width(s->block_w);
    A.th = s->tile_rows;
    A.ntx = (g.nx + 1 + W - 1) / W;
    const int ntz = (g.nz + 1 + A.th - 1) / A.th;
    A.ntiles = A.ntx * ntz;
    A.per_xcd = (A.ntiles + 7) / 8;
    return A;
}

template <class T>
poe2d_status run_steps(poe2d_solver *s, int nsteps)
{
    Receivers R{s->nrec, s->mask, s->rix, s->riz};
    HIP_TRY(hipEventRecord(s->ev0, s->stream));
    if (s->schedule == SCHED_TWOPASS) {
        StepArgs<T> A = step_args<T>(s, s->cur, s->cur);
        for (int k = 0; k < nsteps; ++k) {
            const int n = s->step + k;
            A.src_val = (T)s->source_value(n);
            launch_twopass<T>(A, s->stream, 0);
            if (s->nrec > 0 && n % s->dec == 0 && n / s->dec < s->nsamp)
                // in place: v is still v^n, strains are already those of step n
                launch_gather<T>(A, s->cur, s->cur, R, (T *)s->traces + (size_t)(n / s->dec) * s->ncomp * s->nrec,
                                 s->stream);
            launch_twopass<T>(A, s->stream, 1);
        }
    } else {
        StepArgs<T> F[2] = {step_args<T>(s, 0, 1), step_args<T>(s, 1, 0)};
        for (int k = 0; k < nsteps; ++k) {
            const int n = s->step + k;
            StepArgs<T> &A = F[s->cur];
            A.src_val = (T)s->source_value(n);
            HIP_TRY(launch_fused<T>(A, s->block_w, s->schedule == SCHED_FUSED && s->npal > 0, s->stream));
            if (s->nrec > 0 && n % s->dec == 0 && n / s->dec < s->nsamp)
                launch_gather<T>(A, s->cur, 1 - s->cur, R, (T *)s->traces + (size_t)(n / s->dec) * s->ncomp * s->nrec,
                                 s->stream);
            s->cur = 1 - s->cur;
        }
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(s->ev1, s->stream));
    HIP_TRY(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms = ms;
    s->last_steps = nsteps;
    s->step += nsteps;
    // stability check: largest |v| over the grid
    launch_absmax<T>(step_args<T>(s, s->cur, s->cur), s->cur, s->red + 1, s->red_blocks, s->red, s->stream);
    float vmax = 0.f;
    HIP_TRY(hipMemcpyAsync(&vmax, s->red, sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (!(vmax <= kStabilityThreshold))
        return fail(POE2D_E_UNSTABLE, "run: wavefield became unstable (max |v| = " + std::to_string(vmax) +
                                          ") by step " + std::to_string(s->step));
    return POE2D_OK;
}

// One field of the current state (or p / sxx / syy recomputed from it) to the host.
template <class T>
poe2d_status field_to_host(poe2d_solver *s, int field, void *out)
{
    const size_t bytes = (size_t)(s->grid.nx + 1) * (s->grid.nz + 1) * sizeof(T);
    if (!s->dense) {
        poe2d_status st = alloc_zero(s, &s->dense, bytes);
        if (st) return st;
    }
    StepArgs<T> A = step_args<T>(s, s->cur, s->cur);
    T *dense = (T *)s->dense;
    if (field >= POE2D_FIELD_P) {
        A.src_val = (T)(s->step > 0 ? s->source_value(s->step - 1) : 0.0);
        launch_derived<T>(A, s->cur, field - POE2D_FIELD_P, dense, s->stream);
    } else {
        // V = (vx, vy, wx, wy), S = (e_xx, e_yy, zeta, sxy)
        static const int grp_of[8] = {0, 0, 0, 0, 1, 1, 1, 1};
        static const int comp_of[8] = {0, 1, 2, 3, 0, 1, 2, 3};
        const int g = grp_of[field] == 0 ? group_v(s->cur) : group_s(s->cur);
        launch_extract<T>(A, g, comp_of[field], dense, s->stream);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, dense, bytes, hipMemcpyDeviceToHost));
    return POE2D_OK;
}

poe2d_status check_handle(poe2d_solver *s)
{
    if (!s) return fail(POE2D_E_INVALID, "null solver handle");
    HIP_TRY(hipSetDevice(s->device));
    return POE2D_OK;
}

struct CkptHeader {
    char magic[8];
    int version, precision, nx, nz, step, nrec, ncomp, nsamp, wxs, wzs;
    unsigned long long pitch;
};

// The buffers that make up the state of step `step`: V and S groups of buffer b, its memory
// variables, and the traces.
template <class F>
void for_each_buffer(poe2d_solver *s, int b, F f)
{
    f(s->group_ptr(group_v(b)), s->group_bytes());
    f(s->group_ptr(group_s(b)), s->group_bytes());
    unsigned char *psi = (unsigned char *)s->psi;
    f(psi + (size_t)b * NPSI_X * s->pxn() * s->rs, NPSI_X * s->pxn() * s->rs);
    f(psi + 2 * NPSI_X * s->pxn() * s->rs + (size_t)b * NPSI_Z * s->pzn() * s->rs, NPSI_Z * s->pzn() * s->rs);
    if (s->traces) f(s->traces, s->trace_bytes());
}

}  // namespace

extern "C" {

poe2d_status poe2d_device_info(int *device_count, char *arch, size_t arch_len)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (device_count) *device_count = e == hipSuccess ? n : 0;
    if (e != hipSuccess || n == 0) return fail(POE2D_E_HIP, "no HIP device available");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, 0));
    if (arch && arch_len) {
        std::strncpy(arch, prop.gcnArchName, arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    return POE2D_OK;
}

poe2d_status poe2d_create(const poe2d_grid *g, const poe2d_coeffs *c, poe2d_solver **out)
{
    if (!out) return fail(POE2D_E_INVALID, "create: null output pointer");
    *out = nullptr;
    if (!g || !c) return fail(POE2D_E_INVALID, "create: null grid or coefficients");
    if (g->nx < 8 || g->nz < 8) return fail(POE2D_E_INVALID, "create: grid needs nx, nz >= 8");
    if (!(g->dx > 0) || !(g->dz > 0) || !(g->dt > 0)) return fail(POE2D_E_INVALID, "create: dx, dz, dt must be > 0");
    if (g->npml < 1 || 2 * g->npml + 8 > std::min(g->nx, g->nz))
        return fail(POE2D_E_INVALID, "create: npml must be >= 1 and leave an interior");
    if (g->npower < 1) return fail(POE2D_E_INVALID, "create: npower must be >= 1");
    if (!(g->rcoef > 0 && g->rcoef < 1)) return fail(POE2D_E_INVALID, "create: rcoef must be in (0, 1)");
    if (g->precision != POE2D_F32 && g->precision != POE2D_F64)
        return fail(POE2D_E_INVALID, "create: precision must be POE2D_F32 or POE2D_F64");
    if (c->dtype != POE2D_F32 && c->dtype != POE2D_F64)
        return fail(POE2D_E_INVALID, "create: coefficient dtype must be POE2D_F32 or POE2D_F64");
    if (!c->rho || !c->rhof || !c->m || !c->mu || !c->lambdac || !c->M || !c->alpha || !c->b)
        return fail(POE2D_E_INVALID, "create: null coefficient array");
    if (c->pitch && c->pitch < (size_t)g->nx + 1) return fail(POE2D_E_INVALID, "create: coefficient pitch < nx+1");
    if (!(g->f0 > 0)) return fail(POE2D_E_INVALID, "create: f0 must be > 0");

    poe2d_solver *s = new (std::nothrow) poe2d_solver;
    if (!s) return fail(POE2D_E_NOMEM, "create: out of host memory");
    s->grid = *g;
    s->precision = g->precision;
    s->rs = g->precision == POE2D_F32 ? 4 : 8;
    auto bail = [&](poe2d_status st) {
        free_all(s);
        delete s;
        return st;
    };
    hipError_t e = hipGetDevice(&s->device);
    if (e != hipSuccess) {
        fail(POE2D_E_HIP, std::string("create: no HIP device: ") + hipGetErrorString(e));
        return bail(POE2D_E_HIP);
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) {
        fail(POE2D_E_HIP, "create: stream/event creation failed");
        return bail(POE2D_E_HIP);
    }

    // stability and C-PML set-up
    s->cp_max = g->cp_max > 0 ? g->cp_max : cp_max_from_coeffs(*g, *c);
    if (!(s->cp_max > 0) || !std::isfinite(s->cp_max)) {
        fail(POE2D_E_INVALID, "create: coefficients give no positive P velocity");
        return bail(POE2D_E_INVALID);
    }
    const double courant = s->cp_max * g->dt / std::min(g->dx, g->dz);
    if (g->check_cfl && courant > 6.0 / 7.0 / std::sqrt(2.0)) {
        fail(POE2D_E_CFL, "create: time step is too large, simulation will be unstable (Courant " +
                              std::to_string(courant) + " > 6/7/sqrt(2))");
        return bail(POE2D_E_CFL);
    }
    double amax = g->alpha_max;
    if (amax <= 0)
        amax = g->precision == POE2D_F32 ? (double)(float)(2.e0 * kRefPiF * ((float)g->f0 / 2.e0))
                                         : 3.14159265358979323846 * g->f0;
    for (int axis = 0; axis < 2; ++axis) {
        const int n = (axis == 0 ? g->nx : g->nz) + 1;
        std::vector<double> *t = s->htab + (axis == 0 ? TB_AX : TB_AZ);
        for (int k = 0; k < 6; ++k) t[k].assign(n, 0.0);
        poe2d_status st = poe2d_cpml_profile(axis, n, axis == 0 ? g->dx : g->dz, g->dt, g->npml,
                                             g->npower, g->rcoef, g->kmax, amax, s->cp_max,
                                             g->precision, t[0].data(), t[1].data(), t[2].data(),
                                             t[3].data(), t[4].data(), t[5].data());
        if (st != POE2D_OK) return bail(st);
    }
    int xlo = 0, xhi = 0;
    std::vector<int> xm = strip_map(s->htab + TB_AX, g->nx + 1, &s->wxs, &xlo, &xhi);
    std::vector<int> zm = strip_map(s->htab + TB_AZ, g->nz + 1, &s->wzs, &s->ztop, &s->zbot);
    s->tn = std::max(g->nx, g->nz) + 1;

    s->pitch = (((size_t)kPadCols + g->nx + 2 + 63) / 64) * 64;
    s->alloc_rows = (size_t)g->nz + 1 + 2 * kPadRows;
    if (s->group_bytes() >= (1ull << 32)) {
        // the single-pass kernel addresses a group with 32-bit byte offsets
        fail(POE2D_E_INVALID, "create: grid too large (one field group exceeds 4 GiB)");
        return bail(POE2D_E_INVALID);
    }
    poe2d_status st = POE2D_OK;
    if ((st = alloc_zero(s, &s->arena, NGROUPS * s->group_bytes()))) return bail(st);
    if ((st = alloc_zero(s, &s->psi, 2 * s->psi_half_elems() * s->rs))) return bail(st);
    if ((st = alloc_zero(s, &s->tab, (size_t)NTABLES * s->tn * s->rs))) return bail(st);
    void *q = nullptr;
    if ((st = alloc_zero(s, &q, xm.size() * sizeof(int)))) return bail(st);
    s->xmap = (int *)q;
    if ((st = alloc_zero(s, &q, zm.size() * sizeof(int)))) return bail(st);
    s->zmap = (int *)q;
    if ((st = alloc_zero(s, &q, (s->red_blocks + 1) * sizeof(float)))) return bail(st);
    s->red = (float *)q;
    if (hipStreamSynchronize(s->stream) != hipSuccess ||
        hipMemcpy(s->xmap, xm.data(), xm.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->zmap, zm.data(), zm.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        fail(POE2D_E_HIP, "create: map upload failed");
        return bail(POE2D_E_HIP);
    }
    st = s->precision == POE2D_F32 ? upload_tables<float>(s) : upload_tables<double>(s);
    if (st) return bail(st);
    st = s->precision == POE2D_F32 ? upload_coeffs<float>(s, c) : upload_coeffs<double>(s, c);
    if (st) return bail(st);
    *out = s;
    return POE2D_OK;
}

poe2d_status poe2d_set_schedule(poe2d_solver *s, int schedule, int block_w, int tile_rows)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (schedule != SCHED_FUSED && schedule != SCHED_TWOPASS && schedule != SCHED_FUSED_GROUPS)
        return fail(POE2D_E_INVALID, "set_schedule: schedule must be 0 (single-pass), 1 (two-pass) or 2 "
                                     "(single-pass on the coefficient groups)");
    if (block_w != 128 && block_w != 256 && block_w != 512)
        return fail(POE2D_E_INVALID, "set_schedule: block width must be 128, 256 or 512");
    if (tile_rows < 4 || tile_rows > 1 << 20) return fail(POE2D_E_INVALID, "set_schedule: tile rows out of range");
    s->schedule = schedule;
    s->block_w = block_w;
    s->tile_rows = tile_rows;
    return POE2D_OK;
}

poe2d_status poe2d_material_palette(poe2d_solver *s, int *entries)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (!entries) return fail(POE2D_E_INVALID, "material_palette: null output");
    *entries = s->npal;
    return POE2D_OK;
}

poe2d_status poe2d_set_source(poe2d_solver *s, const poe2d_source *src)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (!src) return fail(POE2D_E_INVALID, "set_source: null source");
    const poe2d_grid &g = s->grid;
    if (src->kind == POE2D_SRC_PRESSURE_DGAUSS) {
        // p is only formed inside the strain-update region (x in [1,NX), z in [2,NZ-1))
        if (src->ix < 1 || src->ix >= g.nx || src->iz < 2 || src->iz >= g.nz - 1)
            return fail(POE2D_E_INVALID, "set_source: pressure source outside x in [1,NX), z in [2,NZ-1)");
    } else if (src->kind == POE2D_SRC_FORCE_Y_RICKER) {
        if (src->ix < 1 || src->ix >= g.nx - 1 || src->iz < 1 || src->iz >= g.nz - 1)
            return fail(POE2D_E_INVALID, "set_source: force source outside x in [1,NX-1), z in [1,NZ-1)");
    } else {
        return fail(POE2D_E_INVALID, "set_source: unknown source kind");
    }
    if (!std::isfinite(src->amp)) return fail(POE2D_E_INVALID, "set_source: non-finite amplitude");
    s->src = *src;
    s->t0 = src->t0 > 0 ? src->t0 : 1.0 / g.f0;
    s->have_src = true;
    return POE2D_OK;
}

poe2d_status poe2d_set_wavelet(poe2d_solver *s, const double *w, int n)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (n < 0 || (n > 0 && !w)) return fail(POE2D_E_INVALID, "set_wavelet: bad samples");
    s->user_w.assign(w, w + n);
    s->have_user_w = true;
    return POE2D_OK;
}

poe2d_status poe2d_set_receivers(poe2d_solver *s, const poe2d_receivers *r)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (!r || r->n < 0 || r->decimate < 1 || r->nsamp < 0 || (r->comp_mask & ~7u) ||
        (r->n > 0 && (!r->ix || !r->iz)))
        return fail(POE2D_E_INVALID, "set_receivers: bad receiver description");
    for (int k = 0; k < r->n; ++k)
        if (r->ix[k] < 0 || r->ix[k] > s->grid.nx || r->iz[k] < 0 || r->iz[k] > s->grid.nz)
            return fail(POE2D_E_INVALID, "set_receivers: receiver " + std::to_string(k) + " outside the grid");
    HIP_TRY(hipStreamSynchronize(s->stream));
    void *q = s->rix; if (q) (void)hipFree(q); s->rix = nullptr;
    q = s->riz; if (q) (void)hipFree(q); s->riz = nullptr;
    if (s->traces) (void)hipFree(s->traces);
    s->traces = nullptr;
    s->nrec = r->n;
    s->mask = r->comp_mask;
    s->ncomp = ((r->comp_mask & 1) != 0) + ((r->comp_mask & 2) != 0) + ((r->comp_mask & 4) != 0);
    s->dec = r->decimate;
    s->nsamp = r->nsamp;
    if (s->nrec == 0) return POE2D_OK;
    if ((st = alloc_zero(s, &q, (size_t)r->n * sizeof(int)))) return st;
    s->rix = (int *)q;
    if ((st = alloc_zero(s, &q, (size_t)r->n * sizeof(int)))) return st;
    s->riz = (int *)q;
    if ((st = alloc_zero(s, &s->traces, std::max<size_t>(s->trace_bytes(), 1)))) return st;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(s->rix, r->ix, (size_t)r->n * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(s->riz, r->iz, (size_t)r->n * sizeof(int), hipMemcpyHostToDevice));
    return POE2D_OK;
}

poe2d_status poe2d_run(poe2d_solver *s, int nsteps)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (nsteps < 0) return fail(POE2D_E_INVALID, "run: negative step count");
    if (nsteps == 0) return POE2D_OK;
    return s->precision == POE2D_F32 ? run_steps<float>(s, nsteps) : run_steps<double>(s, nsteps);
}

poe2d_status poe2d_last_run_ms(poe2d_solver *s, double *ms, int *nsteps)
{
    if (!s || !ms || !nsteps) return fail(POE2D_E_INVALID, "last_run_ms: null argument");
    *ms = s->last_ms;
    *nsteps = s->last_steps;
    return POE2D_OK;
}

poe2d_status poe2d_step_count(poe2d_solver *s, int *n)
{
    if (!s || !n) return fail(POE2D_E_INVALID, "step_count: null argument");
    *n = s->step;
    return POE2D_OK;
}

poe2d_status poe2d_seismograms(poe2d_solver *s, void *out, size_t bytes)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (!out) return fail(POE2D_E_INVALID, "seismograms: null output");
    const size_t need = s->trace_bytes();
    if (bytes < need) return fail(POE2D_E_INVALID, "seismograms: output buffer too small");
    if (need == 0) return POE2D_OK;
    std::vector<unsigned char> tmp(need);
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(tmp.data(), s->traces, need, hipMemcpyDeviceToHost));
    // device layout [slot][comp][rec] -> caller layout [comp][rec][slot]
    const size_t rs = s->rs;
    unsigned char *o = (unsigned char *)out;
    for (int k = 0; k < s->nsamp; ++k)
        for (int c = 0; c < s->ncomp; ++c)
            for (int r = 0; r < s->nrec; ++r)
                std::memcpy(o + (((size_t)c * s->nrec + r) * s->nsamp + k) * rs,
                            tmp.data() + (((size_t)k * s->ncomp + c) * s->nrec + r) * rs, rs);
    return POE2D_OK;
}

poe2d_status poe2d_snapshot(poe2d_solver *s, int field, void *out, size_t bytes)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (!out) return fail(POE2D_E_INVALID, "snapshot: null output");
    if (field < POE2D_FIELD_VX || field > POE2D_FIELD_SYY) return fail(POE2D_E_INVALID, "snapshot: unknown field");
    const size_t need = (size_t)(s->grid.nx + 1) * (s->grid.nz + 1) * s->rs;
    if (bytes < need) return fail(POE2D_E_INVALID, "snapshot: output buffer too small");
    return s->precision == POE2D_F32 ? field_to_host<float>(s, field, out) : field_to_host<double>(s, field, out);
}

poe2d_status poe2d_checkpoint_size(poe2d_solver *s, size_t *bytes)
{
    if (!s || !bytes) return fail(POE2D_E_INVALID, "checkpoint_size: null argument");
    size_t n = sizeof(CkptHeader);
    for_each_buffer(s, 0, [&](void *, size_t b) { n += b; });
    *bytes = n;
    return POE2D_OK;
}

poe2d_status poe2d_checkpoint(poe2d_solver *s, void *blob, size_t bytes)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    size_t need = 0;
    poe2d_checkpoint_size(s, &need);
    if (!blob || bytes < need) return fail(POE2D_E_INVALID, "checkpoint: buffer too small");
    CkptHeader h{};
    std::memcpy(h.magic, "POE2DCK", 8);
    h.version = 2; h.precision = s->precision; h.nx = s->grid.nx; h.nz = s->grid.nz;
    h.step = s->step; h.nrec = s->nrec; h.ncomp = s->ncomp; h.nsamp = s->nsamp;
    h.wxs = s->wxs; h.wzs = s->wzs; h.pitch = s->pitch;
    std::memcpy(blob, &h, sizeof h);
    HIP_TRY(hipStreamSynchronize(s->stream));
    unsigned char *p = (unsigned char *)blob + sizeof h;
    hipError_t err = hipSuccess;
    for_each_buffer(s, s->cur, [&](void *d, size_t b) {
        if (err == hipSuccess) err = hipMemcpy(p, d, b, hipMemcpyDeviceToHost);
        p += b;
    });
    HIP_TRY(err);
    return POE2D_OK;
}

poe2d_status poe2d_restore(poe2d_solver *s, const void *blob, size_t bytes)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    size_t need = 0;
    poe2d_checkpoint_size(s, &need);
    if (!blob || bytes < need) return fail(POE2D_E_INVALID, "restore: blob too small");
    CkptHeader h;
    std::memcpy(&h, blob, sizeof h);
    if (std::memcmp(h.magic, "POE2DCK", 8) || h.version != 2 || h.precision != s->precision ||
        h.nx != s->grid.nx || h.nz != s->grid.nz || h.nrec != s->nrec || h.ncomp != s->ncomp ||
        h.nsamp != s->nsamp || h.wxs != s->wxs || h.wzs != s->wzs || h.pitch != s->pitch)
        return fail(POE2D_E_INVALID, "restore: checkpoint does not match this solver's geometry");
    HIP_TRY(hipStreamSynchronize(s->stream));
    const unsigned char *p = (const unsigned char *)blob + sizeof h;
    hipError_t err = hipSuccess;
    for_each_buffer(s, s->cur, [&](void *d, size_t b) {
        if (err == hipSuccess) err = hipMemcpy(d, p, b, hipMemcpyHostToDevice);
        p += b;
    });
    HIP_TRY(err);
    s->step = h.step;
    return POE2D_OK;
}

poe2d_status poe2d_reset(poe2d_solver *s)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    // state groups V0, S0, V1, S1 and the intermediates are contiguous after MS, MV
    HIP_TRY(hipMemsetAsync(s->group_ptr(G_V0), 0, (NGROUPS - G_V0) * s->group_bytes(), s->stream));
    HIP_TRY(hipMemsetAsync(s->psi, 0, 2 * s->psi_half_elems() * s->rs, s->stream));
    if (s->traces) HIP_TRY(hipMemsetAsync(s->traces, 0, s->trace_bytes(), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->step = 0;
    s->cur = 0;
    return POE2D_OK;
}

poe2d_status poe2d_autotune(poe2d_solver *s, int nsteps, int *block_w, int *tile_rows)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (s->schedule == SCHED_TWOPASS) return fail(POE2D_E_INVALID, "autotune: the two-pass schedule has no tile shape");
    if (s->step != 0) return fail(POE2D_E_STATE, "autotune: needs a handle at step 0 (call poe2d_reset first)");
    if (nsteps <= 0) nsteps = 10;
    // Tile heights of 6k + 2 rows make whole groups of the kernel's 6-step row loop.
    static const int widths[] = {128, 256, 512};
    static const int heights[] = {8, 14, 20, 26, 32, 44, 62, 98};
    double best = 0;
    int best_w = s->block_w, best_h = s->tile_rows;
    for (int w : widths)
        for (int h : heights) {
            if (h > s->grid.nz + 1 && h != heights[0]) continue;
            s->block_w = w;
            s->tile_rows = h;
            if ((st = poe2d_run(s, 1)) || (st = poe2d_run(s, nsteps))) break;
            if (best == 0 || s->last_ms < best) {
                best = s->last_ms;
                best_w = w;
                best_h = h;
            }
        }
    s->block_w = best_w;
    s->tile_rows = best_h;
    const poe2d_status rst = poe2d_reset(s);
    if (st) return st;
    if (rst) return rst;
    if (block_w) *block_w = best_w;
    if (tile_rows) *tile_rows = best_h;
    return POE2D_OK;
}

void poe2d_destroy(poe2d_solver *s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
}

}  // extern "C"
