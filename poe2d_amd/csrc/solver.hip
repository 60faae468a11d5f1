// Solver handles and batches behind the C ABI of include/poe2d.h: device memory, the time loop,
// gather, snapshots, checkpoint/restore and the stability check.
//
// Device memory (field layout described in step_common.h):
//   model   (shared by a handle and its clones, read-only while stepping)
//           material arena: MS and MV groups; when the grid holds at most kMaxPalette distinct
//           coefficient tuples, a one-byte material index per point (group layout, pads select
//           entry 0) and the palette of (MS, MV) entries; the 12 C-PML profiles (1/K for the K
//           tables); the strip maps;
//   handle  state arena: V and S groups for state buffers 0 and 1 and the two-pass
//           intermediates I; C-PML memory variables (6 x-strip and 6 z-strip arrays per state
//           buffer); the source series of the current run; traces; the step counter and the
//           work list of its launches.
// A single-pass step reads state buffer `cur` and writes buffer 1-cur.  The reference's stress1
// reads column nx+1 at x = nx-1, which in its flat array is column 0 of the next row and always
// zero; the zero pad column nx+1 gives the same value.
//
// Time loop.  Single-pass steps are launched over a work list of (model, tile) items that reads
// each model's arguments from device memory and the step index from a device counter (a single
// handle is a batch of one); the kernel also gathers the receivers of recorded steps.  Launch k
// of a chunk computes step *base + k; a chunk of kGraphSteps launches and the advance of base is
// captured once in a HIP graph and replayed, the remainder is launched directly.  The two-pass schedule is launched
// directly with host-side arguments.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
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
const int kGraphSteps = 32;                 // steps per captured graph

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(POE2D_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

// single-pass (palette when the grid has one), two-pass, single-pass always on the MS/MV groups
enum { SCHED_FUSED = 0, SCHED_TWOPASS = 1, SCHED_FUSED_GROUPS = 2, SCHED_FUSED2 = 3 };

poe2d_status dev_alloc(void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 1));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(e == hipErrorOutOfMemory ? POE2D_E_NOMEM : POE2D_E_HIP,
                    std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    return POE2D_OK;
}

poe2d_status dev_alloc_zero(void **p, size_t bytes, hipStream_t stream)
{
    poe2d_status st = dev_alloc(p, bytes);
    if (st) return st;
    HIP_TRY(hipMemsetAsync(*p, 0, std::max<size_t>(bytes, 1), stream));
    return POE2D_OK;
}

void dev_free(void *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Grid- and material-derived device data: shared by a handle and its clones.
struct Model {
    int refs = 1;
    int device = 0;
    int precision = POE2D_F32;
    size_t rs = 4;
    poe2d_grid grid{};
    size_t pitch = 0, alloc_rows = 0;  // points per row, rows per group (with pads)
    int wxs = 0, wzs = 0;              // x-strip columns, z-strip rows
    int ztop = 0, zbot = 0;            // z strips: rows [0, ztop) and [zbot, nz]
    int xlo = 0, xhi = 0;              // x strips: columns [0, xlo) and [xhi, nx]
    int tn = 0;                        // entries per profile table
    double cp_max = 0;
    void *mat = nullptr;               // MS, MV groups
    unsigned char *mid = nullptr;      // material index per point (palette mode)
    void *pal = nullptr;               // PAL_SECTIONS * kMaxPalette Vec4 entries
    int npal = 0;                      // palette entries, 0 when the grid has more than kMaxPalette
    void *tab = nullptr;
    int *xmap = nullptr, *zmap = nullptr;
    std::vector<double> htab[NTABLES];

    size_t group_points() const { return pitch * alloc_rows; }
    size_t group_bytes() const { return group_points() * 4 * rs; }
    size_t origin() const { return (size_t)kPadRows * pitch + kPadCols; }
    size_t pxn() const { return (size_t)std::max(wxs, 1) * (grid.nz + 1); }
    size_t pzn() const { return (size_t)std::max(wzs, 1) * pitch; }
    size_t psi_half_elems() const { return NPSI_X * pxn() + NPSI_Z * pzn(); }  // one buffer's worth

    void release()
    {
        if (--refs > 0) return;
        (void)hipSetDevice(device);
        dev_free(mat);
        void *q = mid; dev_free(q); mid = nullptr;
        dev_free(pal);
        dev_free(tab);
        q = xmap; dev_free(q); xmap = nullptr;
        q = zmap; dev_free(q); zmap = nullptr;
        delete this;
    }
};

// Device arrays of one work-list launch sequence (a handle's own or a batch's) and its captured
// graph.  Arrays only grow, so their addresses stay put while the sizes allow, and a captured
// graph (which holds those addresses, the item count and the kernel choice) is reused across runs;
// the per-model arguments it points to are rewritten before every run.
struct StepList {
    void *models = nullptr;      // StepArgs<T>[nmodels]
    WorkItem *work = nullptr;    // [nwork]
    int *rec_ptr = nullptr;      // receiver offsets per (row, tile column), all models
    int2 *rec_ent = nullptr;     // receiver entries, all models
    long long *base = nullptr;   // step of launch 0 of the current chunk
    long long *h_base = nullptr; // pinned host staging of base
    size_t cap_models = 0, cap_work = 0, cap_ptr = 0, cap_ent = 0;
    int nmodels = 0, nwork = 0, per_xcd = 0, bw = 0;
    bool palette = false;
    bool avg = false;  // POE2D_MATERIAL_AVERAGED (all members alike)
    bool two = false;  // two-step kernel: each launch advances two steps
    int nint = 0;      // two-step kernel: interior tiles in the list
    int precision = POE2D_F32;
    // graph of kGraphSteps launches plus the base advance, and the launch it was captured for
    hipGraphExec_t exec = nullptr;
    struct Sig {
        void *models, *work, *base;
        int nwork, nint, per_xcd, bw, palette, avg, two, precision;
        bool operator==(const Sig &o) const
        {
            return models == o.models && work == o.work && base == o.base && nwork == o.nwork && nint == o.nint &&
                   per_xcd == o.per_xcd && bw == o.bw && palette == o.palette && avg == o.avg && two == o.two &&
                   precision == o.precision;
        }
    } sig{};

    Sig signature() const
    {
        Sig g{};
        g.models = models; g.work = work; g.base = base;
        g.nwork = nwork; g.nint = nint; g.per_xcd = per_xcd; g.bw = bw; g.palette = palette; g.avg = avg; g.two = two;
        g.precision = precision;
        return g;
    }

    void release()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        dev_free(models);
        void *q = work; dev_free(q); work = nullptr;
        q = rec_ptr; dev_free(q); rec_ptr = nullptr;
        q = rec_ent; dev_free(q); rec_ent = nullptr;
        q = base; dev_free(q); base = nullptr;
        if (h_base) (void)hipHostFree(h_base);
        h_base = nullptr;
        cap_models = cap_work = cap_ptr = cap_ent = 0;
    }
};

poe2d_status grow(void **p, size_t *cap, size_t bytes)
{
    if (bytes <= *cap && *p) return POE2D_OK;
    dev_free(*p);
    *cap = 0;
    poe2d_status st = dev_alloc(p, bytes);
    if (st) return st;
    *cap = bytes;
    return POE2D_OK;
}

}  // namespace

struct poe2d_solver {
    Model *m = nullptr;
    int device = 0;
    int precision = POE2D_F32;
    size_t rs = 4;
    poe2d_grid grid{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cur = 0;
    void *arena = nullptr;  // state groups
    void *psi = nullptr;
    void *psi_mid = nullptr;  // memory variables of the intermediate step (two-step kernel), on first use
    float *red = nullptr;   // reduction scratch: [0] result, [1..] per-block partials
    int red_blocks = 1024;
    void *dense = nullptr;  // (nz+1) x (nx+1) scratch for snapshots, made on first use
    void *src_dev = nullptr;  // source series of the current run
    size_t src_cap = 0;
    long long src_n0 = 0;
    int src_len = 0;
    // schedule
    int schedule = SCHED_FUSED;
    int block_w = 256, tile_rows = 14;
    bool graphs = true;
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
    StepList list;   // single-step launches
    StepList list2;  // two-step launches

    size_t group_bytes() const { return m->group_bytes(); }
    size_t trace_bytes() const { return (size_t)nsamp * ncomp * nrec * rs; }
    unsigned char *group_ptr(int g) const { return (unsigned char *)arena + (size_t)g * group_bytes(); }

    double source_value(long long n) const
    {
        if (!have_src) return 0.0;
        const double t = n * grid.dt;
        if (src.t_cut > 0 && t >= src.t_cut) return 0.0;
        if (have_user_w) return n < (long long)user_w.size() ? src.amp * user_w[n] : 0.0;
        return src.amp * wavelet_shape(src.kind, grid.f0, t0, t);
    }
};

struct poe2d_batch {
    std::vector<poe2d_solver *> members;
    int device = 0;
    int precision = POE2D_F32;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int block_w = 256, tile_rows = 14;
    bool graphs = true;
    double last_ms = 0;
    int last_steps = 0;
    StepList list;   // single-step launches
    StepList list2;  // two-step launches
};

namespace {

poe2d_status check_handle(poe2d_solver *s)
{
    if (!s) return fail(POE2D_E_INVALID, "null solver handle");
    HIP_TRY(hipSetDevice(s->device));
    return POE2D_OK;
}

void free_handle(poe2d_solver *s)
{
    s->list.release();
    s->list2.release();
    dev_free(s->arena);
    dev_free(s->psi);
    dev_free(s->psi_mid);
    void *q = s->red; dev_free(q); s->red = nullptr;
    q = s->rix; dev_free(q); s->rix = nullptr;
    q = s->riz; dev_free(q); s->riz = nullptr;
    dev_free(s->traces);
    dev_free(s->dense);
    dev_free(s->src_dev);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->ev0 = s->ev1 = nullptr;
    s->stream = nullptr;
    if (s->m) s->m->release();
    s->m = nullptr;
}

// Coefficients into the MS (alpha, M, lambda_c, mu) and MV (rho, rho_f, m, b) groups, and, when
// the grid holds at most kMaxPalette distinct tuples (bitwise, after rounding to T), into the
// palette and the one-byte material index.
template <class T>
poe2d_status upload_coeffs(Model *md, const poe2d_coeffs *c)
{
    const poe2d_grid &g = md->grid;
    const size_t cpitch = c->pitch ? c->pitch : (size_t)g.nx + 1;
    const void *ms_src[4] = {c->alpha, c->M, c->lambdac, c->mu};
    const void *mv_src[4] = {c->rho, c->rhof, c->m, c->b};
    auto raw = [&](const void *a, size_t i) {
        return c->dtype == POE2D_F64 ? ((const double *)a)[i] : (double)((const float *)a)[i];
    };
    // lambda (slot 2 of the MS group) follows the stress law
    auto at = [&](const void *a, size_t i) {
        return a == c->lambdac ? effective_lambda(g, raw(c->lambdac, i), raw(c->alpha, i), raw(c->M, i)) : raw(a, i);
    };
    // both groups are filled and checked before either is uploaded, so a refused set leaves the
    // model's coefficients as they were
    std::vector<Vec4<T>> hgrp[2];
    for (int grp = 0; grp < 2; ++grp) {
        std::vector<Vec4<T>> &host = hgrp[grp];
        host.assign(md->group_points(), Vec4<T>{T(0), T(0), T(0), T(0)});
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
                host[md->origin() + (size_t)z * md->pitch + x] = Vec4<T>{(T)v[0], (T)v[1], (T)v[2], (T)v[3]};
            }
    }
    for (int grp = 0; grp < 2; ++grp)
        HIP_TRY(hipMemcpy((unsigned char *)md->mat + (size_t)(grp == 0 ? G_MS : G_MV) * md->group_bytes(),
                          hgrp[grp].data(), md->group_bytes(), hipMemcpyHostToDevice));

    // palette: distinct 8-tuples in order of first appearance
    using Key = std::array<T, 8>;
    auto less = [](const Key &a, const Key &b) { return std::memcmp(a.data(), b.data(), sizeof(Key)) < 0; };
    std::map<Key, int, decltype(less)> seen(less);
    std::vector<Key> pal;
    std::vector<unsigned char> idx(md->group_points(), 0);
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
            idx[md->origin() + (size_t)z * md->pitch + x] = (unsigned char)last_id;
        }
    md->npal = 0;
    if (pal.size() > (size_t)kMaxPalette) return POE2D_OK;  // no palette: the groups serve every schedule
    std::vector<Vec4<T>> hp(PAL_SECTIONS * kMaxPalette, Vec4<T>{T(0), T(0), T(0), T(0)});
    for (size_t e = 0; e < pal.size(); ++e) {
        hp[PAL_MS * kMaxPalette + e] = Vec4<T>{pal[e][0], pal[e][1], pal[e][2], pal[e][3]};
        hp[PAL_MV * kMaxPalette + e] = Vec4<T>{pal[e][4], pal[e][5], pal[e][6], pal[e][7]};
    }
    // (a model whose coefficients are replaced keeps the index and palette arrays it has)
    poe2d_status st = POE2D_OK;
    if (!md->mid) {
        void *q = nullptr;
        if ((st = dev_alloc(&q, md->group_points()))) return st;
        md->mid = (unsigned char *)q;
    }
    if (!md->pal && (st = dev_alloc(&md->pal, hp.size() * sizeof(Vec4<T>)))) return st;
    HIP_TRY(hipMemcpy(md->mid, idx.data(), idx.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(md->pal, hp.data(), hp.size() * sizeof(Vec4<T>), hipMemcpyHostToDevice));
    md->npal = (int)pal.size();
    if constexpr (sizeof(T) == 4) {
        // the velocity update's constants of each entry, computed on the device as the kernels
        // would compute them per point (the same expressions and reciprocal)
        launch_cn_palette((Vec4<float> *)md->pal, md->npal, (float)(1.0 / g.dt), nullptr);
        HIP_TRY(hipGetLastError());
        // (the null stream only: other handles' streams are non-blocking and may be busy)
        HIP_TRY(hipStreamSynchronize(nullptr));
    }
    return POE2D_OK;
}

template <class T>
poe2d_status upload_tables(Model *md)
{
    // the K tables go to the device as 1/K, so the kernels multiply
    std::vector<T> h((size_t)NTABLES * md->tn, T(0));
    for (int k = 0; k < NTABLES; ++k) {
        const bool recip = k == TB_RKX || k == TB_RKXH || k == TB_RKZ || k == TB_RKZH;
        for (size_t i = 0; i < md->htab[k].size(); ++i)
            h[(size_t)k * md->tn + i] = (T)(recip ? 1.0 / md->htab[k][i] : md->htab[k][i]);
    }
    HIP_TRY(hipMemcpy(md->tab, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
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

// The P velocity of the model's coefficients (unless the grid fixes it), the Courant check and the
// C-PML profiles that follow from them, into md->cp_max and md->htab; the model is left unchanged
// when the check fails.
poe2d_status model_profiles(Model *md, const poe2d_coeffs *c, const char *what)
{
    const poe2d_grid *g = &md->grid;
    const double cp = g->cp_max > 0 ? g->cp_max : cp_max_from_coeffs(*g, *c);
    if (!(cp > 0) || !std::isfinite(cp))
        return fail(POE2D_E_INVALID, std::string(what) + ": coefficients give no positive P velocity");
    const double courant = cp * g->dt / std::min(g->dx, g->dz);
    if (g->check_cfl && courant > 6.0 / 7.0 / std::sqrt(2.0))
        return fail(POE2D_E_CFL, std::string(what) + ": time step is too large, simulation will be unstable (Courant " +
                                     std::to_string(courant) + " > 6/7/sqrt(2))");
    double amax = g->alpha_max;
    if (amax <= 0)
        amax = g->precision == POE2D_F32 ? (double)(float)(2.e0 * kRefPiF * ((float)g->f0 / 2.e0))
                                         : 3.14159265358979323846 * g->f0;
    std::vector<double> tab[NTABLES];
    for (int axis = 0; axis < 2; ++axis) {
        const int n = (axis == 0 ? g->nx : g->nz) + 1;
        std::vector<double> *t = tab + (axis == 0 ? TB_AX : TB_AZ);
        for (int k = 0; k < 6; ++k) t[k].assign(n, 0.0);
        poe2d_status st = poe2d_cpml_profile(axis, n, axis == 0 ? g->dx : g->dz, g->dt, g->npml,
                                             g->npower, g->rcoef, g->kmax, amax, cp,
                                             g->precision, t[0].data(), t[1].data(), t[2].data(),
                                             t[3].data(), t[4].data(), t[5].data());
        if (st != POE2D_OK) return st;
    }
    md->cp_max = cp;
    for (int k = 0; k < NTABLES; ++k) md->htab[k].swap(tab[k]);
    return POE2D_OK;
}

// Arguments of the step kernels for handle s, reading state buffer `in` and writing `out`, with
// tiles of bw - 8 columns (bw - 12 for the two-step kernel) by th rows.
template <class T>
StepArgs<T> step_args(const poe2d_solver *s, int in, int out, int bw, int th, bool two = false)
{
    StepArgs<T> A{};
    const Model *md = s->m;
    const poe2d_grid &g = md->grid;
    A.nx = g.nx;
    A.nz = g.nz;
    A.pitch = (int)md->pitch;
    A.wxs = md->wxs;
    A.origin = (long long)md->origin();
    A.group_stride = (long long)md->group_points();
    A.pxn = (long long)md->pxn();
    A.pzn = (long long)md->pzn();
    A.tn = md->tn;
    A.in = in;
    A.out = out;
    A.rdx = T(1.0 / (24.0 * g.dx));
    A.rdz = T(1.0 / (24.0 * g.dz));
    A.dt = T(g.dt);
    A.inv_dt = T(1.0 / g.dt);
    A.grp = (Vec4<T> *)s->arena;
    A.mat = (const Vec4<T> *)md->mat;
    A.psi = (T *)s->psi;
    A.tab = (const T *)md->tab;
    A.xmap = md->xmap;
    A.zmap = md->zmap;
    A.ztop = md->ztop;
    A.zbot = md->zbot;
    A.mid = md->mid;
    A.pal = (const Vec4<T> *)md->pal;
    A.npal = md->npal;
    A.src_kind = s->have_src ? s->src.kind : -1;
    A.src_x = s->have_src ? s->src.ix : -1;
    A.src_z = s->have_src ? s->src.iz : -1;
    A.src_val = T(0);
    A.src_series = (const T *)s->src_dev;
    A.src_n0 = s->src_n0;
    A.src_len = s->src_len;
    // step n = s->step reads state buffer cur
    A.par = (int)(((long long)s->cur - s->step) & 1);
    A.th = th;
    const int W = two ? fused2_tile_width(bw) : fused_tile_width(bw);
    A.ntx = (g.nx + 1 + W - 1) / W;
    A.psi_mid = (T *)s->psi_mid;
    A.pair0 = s->step;
    A.pcur = s->cur;
    A.ntiles = A.ntx * ((g.nz + 1 + th - 1) / th);
    A.per_xcd = (A.ntiles + 7) / 8;
    return A;
}

template <class T>
StepArgs<T> step_args(const poe2d_solver *s, int in, int out)
{
    return step_args<T>(s, in, out, s->block_w, s->tile_rows);
}

// Source samples of steps [n0, n0 + k) into the handle's device series.
template <class T>
poe2d_status upload_source(poe2d_solver *s, long long n0, int k)
{
    poe2d_status st = grow(&s->src_dev, &s->src_cap, (size_t)std::max(k, 1) * sizeof(T));
    if (st) return st;
    std::vector<T> h(std::max(k, 1), T(0));
    for (int i = 0; i < k; ++i) h[i] = (T)s->source_value(n0 + i);
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(s->src_dev, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    s->src_n0 = n0;
    s->src_len = k;
    return POE2D_OK;
}

// Fill `L` for the members (all at the same step, current source series uploaded): per-member
// arguments, the (member, tile) work list with block width bw and tile height th, and each
// member's receivers sorted by row and column with their offsets per (row, tile column).  The
// uploads wait for `stream` first.
template <class T>
poe2d_status build_list(StepList &L, const std::vector<poe2d_solver *> &members, int bw, int th, bool palette,
                        hipStream_t stream, bool two = false, int par_shift = 0)
{
    const int nm = (int)members.size();
    const int material = members[0]->grid.staggered_material;
    // a palette launch reads each member's material index and palette: refuse members without one
    // (or, for the two-step kernel, with more entries than its LDS copy holds)
    for (const poe2d_solver *m : members)
        if (palette && (m->m->npal < 1 || !m->m->mid || !m->m->pal || (two && m->m->npal > kMaxPalette2)))
            return fail(POE2D_E_INVALID, "run: a palette launch over a grid without a usable palette");
    for (const poe2d_solver *m : members)
        if (m->grid.staggered_material != material)
            return fail(POE2D_E_INVALID, "run: batch members differ in staggered material");
    std::vector<StepArgs<T>> args(nm);
    std::vector<WorkItem> work;
    std::vector<int> ptr;
    std::vector<int2> ent;
    std::vector<size_t> ptr_off(nm), ent_off(nm);
    const int W = two ? fused2_tile_width(bw) : fused_tile_width(bw);
    for (int i = 0; i < nm; ++i) {
        const poe2d_solver *s = members[i];
        args[i] = step_args<T>(s, 0, 1, bw, th, two);
        // a single step after par_shift step pairs: step n = step + 2 par_shift reads buffer
        // cur + par_shift, so (n + par) & 1 = (cur + par_shift) & 1
        if (par_shift) args[i].par = (int)(((long long)s->cur - s->step - par_shift) & 1);
        if ((long long)work.size() + args[i].ntiles > (1LL << 30))
            return fail(POE2D_E_INVALID, "run: too many tiles in one launch");
        if (!two)
            for (int t = 0; t < args[i].ntiles; ++t) work.push_back(WorkItem{i, t});
        // receivers sorted by (tile column, row, column), with offsets per (tile column, row)
        const int ntx = args[i].ntx, nrow = s->grid.nz + 1;
        std::vector<int> order(s->nrec);
        for (int r = 0; r < s->nrec; ++r) order[r] = r;
        std::vector<int> hx(s->nrec), hz(s->nrec);
        if (s->nrec > 0) {
            HIP_TRY(hipMemcpy(hx.data(), s->rix, s->nrec * sizeof(int), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(hz.data(), s->riz, s->nrec * sizeof(int), hipMemcpyDeviceToHost));
        }
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            if (hx[a] / W != hx[b] / W) return hx[a] / W < hx[b] / W;
            return hz[a] != hz[b] ? hz[a] < hz[b] : hx[a] < hx[b];
        });
        ptr_off[i] = ptr.size();
        ent_off[i] = ent.size();
        std::vector<int> cnt((size_t)ntx * nrow + 1, 0);
        for (int r : order) cnt[(size_t)(hx[r] / W) * nrow + hz[r] + 1]++;
        for (size_t q = 1; q < cnt.size(); ++q) cnt[q] += cnt[q - 1];
        ptr.insert(ptr.end(), cnt.begin(), cnt.end());
        for (int r : order) ent.push_back(int2{hx[r], r});
        if (two) {
            // edge tiles (stored as ~tile): a C-PML strip point, or a point of the grid's two outer
            // rows or columns on either side (whose updates only edge tiles guard), within the rows
            // and columns the kernel computes for the tile, or receivers in its rows
            const Model *md = s->m;
            const int nx = s->grid.nx, nz = s->grid.nz;
            for (int t = 0; t < args[i].ntiles; ++t) {
                const int tx = t % ntx, tz = t / ntx;
                const int zA = tz * th, zB = std::min(zA + th, nz + 1);
                const int c0 = tx * W - kFused2ColsBefore, c1 = c0 + bw - 1;
                const int r0 = zA - kFused2RowsBefore, r1 = zB + kFused2RowsAfter;
                const bool xs = c0 < std::max(md->xlo, 2) || c1 >= std::min(md->xhi, nx - 1);
                const bool zs = r0 < std::max(md->ztop, 2) || r1 >= std::min(md->zbot, nz - 1);
                const size_t p0 = (size_t)tx * nrow + zA + 1, p1 = (size_t)tx * nrow + zB + 1;
                const bool rec = s->nrec > 0 && cnt[p1 - 1] > cnt[p0 - 1];
                work.push_back(WorkItem{i, xs || zs || rec ? ~t : t});
            }
        }
        args[i].rec_n = s->nrec;
        args[i].rec_ncomp = s->ncomp;
        args[i].rec_dec = s->dec;
        args[i].rec_nsamp = s->nsamp;
        args[i].rec_mask = s->mask;
        args[i].traces = (T *)s->traces;
    }
    int nint = (int)work.size();
    int per_xcd = (int)((work.size() + 7) / 8);
    if (two) {
        // Edge tiles take about twice as long as interior ones and gather in the bands of the z
        // strips: each XCD's share of the launch gets every eighth edge tile, first, and then a
        // band of consecutive interior tiles, so that the XCDs finish together and the edge tiles
        // overlap with interior work.  Shares are padded to equal length with empty items.
        std::vector<WorkItem> edge, inter;
        for (const WorkItem &w : work) (w.tile < 0 ? edge : inter).push_back(w);
        nint = (int)inter.size();
        const int ne = (int)((edge.size() + 7) / 8), ni = (int)((inter.size() + 7) / 8);
        per_xcd = ne + ni;
        work.assign((size_t)8 * per_xcd, WorkItem{-1, 0});
        for (int x = 0; x < 8; ++x) {
            size_t o = (size_t)x * per_xcd;
            for (size_t e = x; e < edge.size(); e += 8) work[o++] = edge[e];
            for (size_t q = (size_t)x * ni; q < std::min(inter.size(), (size_t)(x + 1) * ni); ++q) work[o++] = inter[q];
        }
    }
    poe2d_status st;
    if ((st = grow(&L.models, &L.cap_models, args.size() * sizeof(StepArgs<T>)))) return st;
    void *w = L.work;
    st = grow(&w, &L.cap_work, work.size() * sizeof(WorkItem));
    L.work = (WorkItem *)w;
    if (st) return st;
    void *pp = L.rec_ptr;
    st = grow(&pp, &L.cap_ptr, ptr.size() * sizeof(int));
    L.rec_ptr = (int *)pp;
    if (st) return st;
    void *pe = L.rec_ent;
    st = grow(&pe, &L.cap_ent, std::max<size_t>(ent.size(), 1) * sizeof(int2));
    L.rec_ent = (int2 *)pe;
    if (st) return st;
    for (int i = 0; i < nm; ++i) {
        args[i].rec_tile_ptr = L.rec_ptr + ptr_off[i];
        args[i].rec_entries = L.rec_ent + ent_off[i];
    }
    if (!L.base) {
        void *q = nullptr;
        if ((st = dev_alloc(&q, sizeof(long long)))) return st;
        L.base = (long long *)q;
    }
    if (!L.h_base) HIP_TRY(hipHostMalloc((void **)&L.h_base, sizeof(long long), hipHostMallocDefault));
    HIP_TRY(hipStreamSynchronize(stream));
    HIP_TRY(hipMemcpy(L.models, args.data(), args.size() * sizeof(StepArgs<T>), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(L.work, work.data(), work.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(L.rec_ptr, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!ent.empty()) HIP_TRY(hipMemcpy(L.rec_ent, ent.data(), ent.size() * sizeof(int2), hipMemcpyHostToDevice));
    L.nmodels = nm;
    L.nint = two ? nint : (int)work.size();
    L.nwork = (int)work.size();
    L.per_xcd = per_xcd;
    L.bw = bw;
    L.palette = palette;
    L.avg = material == POE2D_MATERIAL_AVERAGED;
    L.two = two;
    L.precision = sizeof(T) == 4 ? POE2D_F32 : POE2D_F64;
    return POE2D_OK;
}

// Launches k = 0 .. count-1 of list L on `stream` (steps *base + k, or the step pairs *base + 2k
// for the two-step kernel), then advances *base past them.
template <class T>
hipError_t launch_list_chunk(const StepList &L, int count, hipStream_t stream)
{
    if (L.two) {
        if constexpr (sizeof(T) == 4) {
            for (int k = 0; k < count; ++k) {
                const hipError_t e = launch_fused2<T>((const StepArgs<T> *)L.models, L.work, L.nwork, L.per_xcd, L.bw,
                                                      L.palette, L.avg, L.base, 2 * k, stream);
                if (e != hipSuccess) return e;
            }
            return launch_base_advance(L.base, 2 * count, stream);
        }
        return hipErrorInvalidValue;
    }
    for (int k = 0; k < count; ++k) {
        hipError_t e = launch_fused<T>((const StepArgs<T> *)L.models, L.work, L.nwork, L.per_xcd, L.bw, L.palette,
                                       L.avg, L.base, k, stream);
        if (e != hipSuccess) return e;
    }
    return launch_base_advance(L.base, count, stream);
}

// k launches of list L from step n0 on `stream` (k steps, or k step pairs): graph replays of
// kGraphSteps steps, the remainder launched directly.
template <class T>
poe2d_status run_list(StepList &L, long long n0, int k, bool graphs, hipStream_t stream)
{
    *L.h_base = n0;
    HIP_TRY(hipMemcpyAsync(L.base, L.h_base, sizeof(long long), hipMemcpyHostToDevice, stream));
    int done = 0;
    const int chunk = L.two ? kGraphSteps / 2 : kGraphSteps;
    if (graphs && k >= chunk) {
        const StepList::Sig sig = L.signature();
        if (!L.exec || !(L.sig == sig)) {
            if (L.exec) (void)hipGraphExecDestroy(L.exec);
            L.exec = nullptr;
            hipGraph_t graph = nullptr;
            HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            const hipError_t e = launch_list_chunk<T>(L, chunk, stream);
            const hipError_t e2 = hipStreamEndCapture(stream, &graph);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return fail(POE2D_E_HIP, std::string("run: graph capture failed: ") +
                                             hipGetErrorString(e != hipSuccess ? e : e2));
            }
            const hipError_t e3 = hipGraphInstantiate(&L.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e3 != hipSuccess) {
                L.exec = nullptr;
                return fail(POE2D_E_HIP, std::string("run: graph instantiation failed: ") + hipGetErrorString(e3));
            }
            L.sig = sig;
        }
        for (; done + chunk <= k; done += chunk) HIP_TRY(hipGraphLaunch(L.exec, stream));
    }
    if (done < k) HIP_TRY(launch_list_chunk<T>(L, k - done, stream));
    return POE2D_OK;
}

// Largest |v| of the handle's current state; POE2D_E_UNSTABLE above the reference's threshold.
template <class T>
poe2d_status stability_check(poe2d_solver *s)
{
    StepArgs<T> A = step_args<T>(s, s->cur, s->cur);
    launch_absmax<T>(A, s->cur, s->red + 1, s->red_blocks, s->red, s->stream);
    float vmax = 0.f;
    HIP_TRY(hipMemcpyAsync(&vmax, s->red, sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (!(vmax <= kStabilityThreshold))
        return fail(POE2D_E_UNSTABLE, "run: wavefield became unstable (max |v| = " + std::to_string(vmax) +
                                          ") by step " + std::to_string(s->step));
    return POE2D_OK;
}

// The kernel a handle's steps run on: schedule 3 takes the two-step kernel on FP32 grids (through
// the palette where the grid's fits that kernel's LDS copy, else through the coefficient groups),
// and the single-step kernel (as schedule 0, 64-thread blocks widened to 128) on FP64 grids.
bool two_step_kernel(const poe2d_solver *s)
{
    return s->schedule == SCHED_FUSED2 && s->precision == POE2D_F32;
}
// Whether the two-step kernel reads handle s through its palette.
bool two_step_palette(const poe2d_solver *s) { return s->m->npal > 0 && s->m->npal <= kMaxPalette2; }
int effective_schedule(const poe2d_solver *s)
{
    return s->schedule == SCHED_FUSED2 && !two_step_kernel(s) ? SCHED_FUSED : s->schedule;
}

template <class T>
poe2d_status run_steps(poe2d_solver *s, int nsteps)
{
    poe2d_status st;
    if (s->schedule == SCHED_TWOPASS) {
        Receivers R{s->nrec, s->mask, s->rix, s->riz};
        StepArgs<T> A = step_args<T>(s, s->cur, s->cur);
        HIP_TRY(hipEventRecord(s->ev0, s->stream));
        for (int k = 0; k < nsteps; ++k) {
            const int n = s->step + k;
            A.src_val = (T)s->source_value(n);
            launch_twopass<T>(A, s->stream, 0, s->grid.staggered_material == POE2D_MATERIAL_AVERAGED);
            if (s->nrec > 0 && n % s->dec == 0 && n / s->dec < s->nsamp)
                // in place: v is still v^n, strains are already those of step n
                launch_gather<T>(A, s->cur, s->cur, R, (T *)s->traces + (size_t)(n / s->dec) * s->ncomp * s->nrec,
                                 s->stream);
            launch_twopass<T>(A, s->stream, 1, s->grid.staggered_material == POE2D_MATERIAL_AVERAGED);
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(s->ev1, s->stream));
    } else if (sizeof(T) == 4 && two_step_kernel(s)) {
        // step pairs on the two-step kernel, an odd last step on the single-step kernel
        const int pairs = nsteps / 2, rem = nsteps & 1;
        if (!s->psi_mid && (st = dev_alloc(&s->psi_mid, std::max<size_t>(s->m->psi_half_elems(), 1) * s->rs)))
            return st;
        if ((st = upload_source<T>(s, s->step, nsteps))) return st;
        if (pairs && (st = build_list<T>(s->list2, {s}, s->block_w, s->tile_rows, two_step_palette(s), s->stream, true)))
            return st;
        // (the single-step kernel has no 64-thread blocks; it reads the palette whenever the grid
        // has one, which may be larger than the two-step kernel's)
        if (rem && (st = build_list<T>(s->list, {s}, std::max(s->block_w, 128), s->tile_rows, s->m->npal > 0,
                                       s->stream, false, pairs)))
            return st;
        HIP_TRY(hipEventRecord(s->ev0, s->stream));
        if (pairs && (st = run_list<T>(s->list2, s->step, pairs, s->graphs, s->stream))) return st;
        if (rem && (st = run_list<T>(s->list, s->step + 2LL * pairs, 1, s->graphs, s->stream))) return st;
        HIP_TRY(hipEventRecord(s->ev1, s->stream));
        s->cur = (s->cur + pairs + rem) & 1;
    } else {
        if ((st = upload_source<T>(s, s->step, nsteps))) return st;
        const bool palette = (s->schedule == SCHED_FUSED || s->schedule == SCHED_FUSED2) && s->m->npal > 0;
        const int bw = s->schedule == SCHED_FUSED2 ? std::max(s->block_w, 128) : s->block_w;
        if ((st = build_list<T>(s->list, {s}, bw, s->tile_rows, palette, s->stream))) return st;
        HIP_TRY(hipEventRecord(s->ev0, s->stream));
        if ((st = run_list<T>(s->list, s->step, nsteps, s->graphs, s->stream))) return st;
        HIP_TRY(hipEventRecord(s->ev1, s->stream));
        s->cur = (s->cur + nsteps) & 1;
    }
    HIP_TRY(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->last_ms = ms;
    s->last_steps = nsteps;
    s->step += nsteps;
    return stability_check<T>(s);
}

// One field of the current state (or p / sxx / syy recomputed from it) to the host.
template <class T>
poe2d_status field_to_host(poe2d_solver *s, int field, void *out)
{
    const size_t bytes = (size_t)(s->grid.nx + 1) * (s->grid.nz + 1) * sizeof(T);
    if (!s->dense) {
        poe2d_status st = dev_alloc_zero(&s->dense, bytes, s->stream);
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

struct CkptHeader {
    char magic[8];
    int version, precision, nx, nz, step, nrec, ncomp, nsamp, wxs, wzs;
    unsigned long long pitch, alloc_rows;
};

// The buffers that make up the state of step `step`: V and S groups of buffer b, its memory
// variables, and the traces.
template <class F>
void for_each_buffer(poe2d_solver *s, int b, F f)
{
    const Model *md = s->m;
    f(s->group_ptr(group_v(b)), s->group_bytes());
    f(s->group_ptr(group_s(b)), s->group_bytes());
    unsigned char *psi = (unsigned char *)s->psi;
    f(psi + (size_t)b * NPSI_X * md->pxn() * s->rs, NPSI_X * md->pxn() * s->rs);
    f(psi + 2 * NPSI_X * md->pxn() * s->rs + (size_t)b * NPSI_Z * md->pzn() * s->rs, NPSI_Z * md->pzn() * s->rs);
    if (s->traces) f(s->traces, s->trace_bytes());
}

// A handle on model md (whose reference the caller hands over): stream, events, state arena,
// memory variables, reduction scratch.  On failure the reference is released.
poe2d_solver *new_handle(Model *md, poe2d_status *st)
{
    poe2d_solver *s = new (std::nothrow) poe2d_solver;
    if (!s) {
        md->release();
        *st = fail(POE2D_E_NOMEM, "create: out of host memory");
        return nullptr;
    }
    s->m = md;
    s->device = md->device;
    s->precision = md->precision;
    s->rs = md->rs;
    s->grid = md->grid;
    if (s->precision == POE2D_F32) {
        // FP32 handles start on schedule 3: the two-step kernel where the grid's palette fits it,
        // the single-step kernel elsewhere, the same results either way
        s->schedule = SCHED_FUSED2;
        s->tile_rows = 24;
    }
    auto bail = [&](poe2d_status e) -> poe2d_solver * {
        *st = e;
        free_handle(s);
        delete s;
        return nullptr;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
        return bail(fail(POE2D_E_HIP, "create: stream/event creation failed"));
    poe2d_status e;
    if ((e = dev_alloc_zero(&s->arena, NSTATE_GROUPS * md->group_bytes(), s->stream))) return bail(e);
    if ((e = dev_alloc_zero(&s->psi, 2 * md->psi_half_elems() * s->rs, s->stream))) return bail(e);
    void *q = nullptr;
    if ((e = dev_alloc_zero(&q, (s->red_blocks + 1) * sizeof(float), s->stream))) return bail(e);
    s->red = (float *)q;
    if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(POE2D_E_HIP, "create: synchronisation failed"));
    *st = POE2D_OK;
    return s;
}

poe2d_status check_batch(poe2d_batch *b)
{
    if (!b) return fail(POE2D_E_INVALID, "null batch handle");
    HIP_TRY(hipSetDevice(b->device));
    return POE2D_OK;
}

template <class T>
poe2d_status batch_run_steps(poe2d_batch *b, int nsteps)
{
    const long long n0 = b->members[0]->step;
    bool palette = true, two = sizeof(T) == 4, palette2 = true;
    for (poe2d_solver *s : b->members) {
        poe2d_status st = upload_source<T>(s, n0, nsteps);
        if (st) return st;
        palette = palette && (s->schedule == SCHED_FUSED || s->schedule == SCHED_FUSED2) && s->m->npal > 0;
        two = two && two_step_kernel(s);
        palette2 = palette2 && two_step_palette(s);
    }
    // (the single-step kernel widens 64-thread blocks to 128, below)
    if (two && b->block_w == 512)
        return fail(POE2D_E_INVALID, "batch_run: the two-step kernel takes block widths 64, 128 or 256");
    const int pairs = two ? nsteps / 2 : 0, rem = two ? nsteps & 1 : nsteps;
    poe2d_status st;
    if (pairs) {
        for (poe2d_solver *s : b->members)
            if (!s->psi_mid && (st = dev_alloc(&s->psi_mid, std::max<size_t>(s->m->psi_half_elems(), 1) * s->rs)))
                return st;
        if ((st = build_list<T>(b->list2, b->members, b->block_w, b->tile_rows, palette2, b->stream, true))) return st;
    }
    if (rem && (st = build_list<T>(b->list, b->members, std::max(b->block_w, 128), b->tile_rows, palette, b->stream,
                                   false, pairs)))
        return st;
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    if (pairs && (st = run_list<T>(b->list2, n0, pairs, b->graphs, b->stream))) return st;
    if (rem && (st = run_list<T>(b->list, n0 + 2LL * pairs, rem, b->graphs, b->stream))) return st;
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    HIP_TRY(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    b->last_steps = nsteps;
    poe2d_status first = POE2D_OK;
    for (poe2d_solver *s : b->members) {
        s->cur = (s->cur + pairs + rem) & 1;
        s->step += nsteps;
        s->last_ms = ms;
        s->last_steps = nsteps;
        const poe2d_status c = stability_check<T>(s);
        if (c && !first) first = c;
    }
    return first;
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
    if (g->stress_law != POE2D_STRESS_AS_WRITTEN && g->stress_law != POE2D_STRESS_DRAINED)
        return fail(POE2D_E_INVALID, "create: unknown stress law");
    if (g->staggered_material != POE2D_MATERIAL_AS_WRITTEN && g->staggered_material != POE2D_MATERIAL_AVERAGED)
        return fail(POE2D_E_INVALID, "create: unknown staggered material setting");

    Model *md = new (std::nothrow) Model;
    if (!md) return fail(POE2D_E_NOMEM, "create: out of host memory");
    md->grid = *g;
    md->precision = g->precision;
    md->rs = g->precision == POE2D_F32 ? 4 : 8;
    auto bail = [&](poe2d_status st) {
        md->release();
        return st;
    };
    hipError_t e = hipGetDevice(&md->device);
    if (e != hipSuccess) return bail(fail(POE2D_E_HIP, std::string("create: no HIP device: ") + hipGetErrorString(e)));

    // stability and C-PML set-up
    poe2d_status st = model_profiles(md, c, "create");
    if (st) return bail(st);
    std::vector<int> xm = strip_map(md->htab + TB_AX, g->nx + 1, &md->wxs, &md->xlo, &md->xhi);
    std::vector<int> zm = strip_map(md->htab + TB_AZ, g->nz + 1, &md->wzs, &md->ztop, &md->zbot);
    md->tn = std::max(g->nx, g->nz) + 1;

    md->pitch = (((size_t)kPadCols + g->nx + 2 + 63) / 64) * 64;
    md->alloc_rows = (size_t)g->nz + 1 + 2 * kPadRows;
    // the single-pass kernel addresses a group with 32-bit byte offsets
    if (md->group_bytes() >= (1ull << 32))
        return bail(fail(POE2D_E_INVALID, "create: grid too large (one field group exceeds 4 GiB)"));
    if ((st = dev_alloc_zero(&md->mat, NMAT_GROUPS * md->group_bytes(), nullptr))) return bail(st);
    if ((st = dev_alloc_zero(&md->tab, (size_t)NTABLES * md->tn * md->rs, nullptr))) return bail(st);
    void *q = nullptr;
    if ((st = dev_alloc(&q, xm.size() * sizeof(int)))) return bail(st);
    md->xmap = (int *)q;
    if ((st = dev_alloc(&q, zm.size() * sizeof(int)))) return bail(st);
    md->zmap = (int *)q;
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(md->xmap, xm.data(), xm.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(md->zmap, zm.data(), zm.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(POE2D_E_HIP, "create: map upload failed"));
    st = md->precision == POE2D_F32 ? upload_tables<float>(md) : upload_tables<double>(md);
    if (st) return bail(st);
    st = md->precision == POE2D_F32 ? upload_coeffs<float>(md, c) : upload_coeffs<double>(md, c);
    if (st) return bail(st);
    poe2d_solver *s = new_handle(md, &st);  // hands the model reference over
    if (!s) return st;
    *out = s;
    return POE2D_OK;
}

poe2d_status poe2d_clone(poe2d_solver *s, poe2d_solver **out)
{
    if (!out) return fail(POE2D_E_INVALID, "clone: null output pointer");
    *out = nullptr;
    poe2d_status st = check_handle(s);
    if (st) return st;
    s->m->refs++;
    poe2d_solver *c = new_handle(s->m, &st);
    if (!c) return st;
    c->schedule = s->schedule;
    c->block_w = s->block_w;
    c->tile_rows = s->tile_rows;
    c->graphs = s->graphs;
    *out = c;
    return POE2D_OK;
}

poe2d_status poe2d_set_coeffs(poe2d_solver *s, const poe2d_coeffs *c)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    Model *md = s->m;
    const poe2d_grid &g = md->grid;
    if (md->refs > 1) return fail(POE2D_E_STATE, "set_coeffs: the model is shared with clones");
    if (!c) return fail(POE2D_E_INVALID, "set_coeffs: null coefficients");
    if (c->dtype != POE2D_F32 && c->dtype != POE2D_F64)
        return fail(POE2D_E_INVALID, "set_coeffs: coefficient dtype must be POE2D_F32 or POE2D_F64");
    if (!c->rho || !c->rhof || !c->m || !c->mu || !c->lambdac || !c->M || !c->alpha || !c->b)
        return fail(POE2D_E_INVALID, "set_coeffs: null coefficient array");
    if (c->pitch && c->pitch < (size_t)g.nx + 1) return fail(POE2D_E_INVALID, "set_coeffs: coefficient pitch < nx+1");
    HIP_TRY(hipStreamSynchronize(s->stream));
    // The C-PML profiles scale with the P velocity; their strips (where a profile is active) do not,
    // so the memory variables keep their layout.
    // A refused set (Courant check, invalid coefficients) leaves the handle as it was.
    const int wxs = md->wxs, wzs = md->wzs, xlo = md->xlo, xhi = md->xhi, ztop = md->ztop, zbot = md->zbot;
    const double old_cp = md->cp_max;
    std::vector<double> old_tab[NTABLES];
    for (int k = 0; k < NTABLES; ++k) old_tab[k] = md->htab[k];
    auto restore = [&](poe2d_status e) {
        md->cp_max = old_cp;
        for (int k = 0; k < NTABLES; ++k) md->htab[k].swap(old_tab[k]);
        return e;
    };
    if ((st = model_profiles(md, c, "set_coeffs"))) return st;
    int nwx, nwz, lo, hi, top, bot;
    strip_map(md->htab + TB_AX, g.nx + 1, &nwx, &lo, &hi);
    strip_map(md->htab + TB_AZ, g.nz + 1, &nwz, &top, &bot);
    if (nwx != wxs || nwz != wzs || lo != xlo || hi != xhi || top != ztop || bot != zbot)
        return restore(fail(POE2D_E_INVALID, "set_coeffs: the C-PML strips would change (create a new handle)"));
    st = md->precision == POE2D_F32 ? upload_coeffs<float>(md, c) : upload_coeffs<double>(md, c);
    if (st) return restore(st);
    st = md->precision == POE2D_F32 ? upload_tables<float>(md) : upload_tables<double>(md);
    if (st) return st;
    return poe2d_reset(s);
}

poe2d_status poe2d_set_schedule(poe2d_solver *s, int schedule, int block_w, int tile_rows)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (schedule != SCHED_FUSED && schedule != SCHED_TWOPASS && schedule != SCHED_FUSED_GROUPS &&
        schedule != SCHED_FUSED2)
        return fail(POE2D_E_INVALID, "set_schedule: schedule must be 0 (single-pass), 1 (two-pass), 2 "
                                     "(single-pass on the coefficient groups) or 3 (two steps per launch)");
    if (schedule == SCHED_FUSED2 ? (block_w != 64 && block_w != 128 && block_w != 256)
                                 : (block_w != 128 && block_w != 256 && block_w != 512))
        return fail(POE2D_E_INVALID, "set_schedule: block width must be 128, 256 or 512 (64, 128 or 256 for the "
                                     "two-step schedule)");
    if (tile_rows < 4 || tile_rows > 1 << 20) return fail(POE2D_E_INVALID, "set_schedule: tile rows out of range");
    s->schedule = schedule;
    s->block_w = block_w;
    s->tile_rows = tile_rows;
    return POE2D_OK;
}

poe2d_status poe2d_set_graphs(poe2d_solver *s, int enable)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    s->graphs = enable != 0;
    return POE2D_OK;
}

poe2d_status poe2d_material_palette(poe2d_solver *s, int *entries)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (!entries) return fail(POE2D_E_INVALID, "material_palette: null output");
    *entries = s->m->npal;
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
    void *q = s->rix; dev_free(q); s->rix = nullptr;
    q = s->riz; dev_free(q); s->riz = nullptr;
    dev_free(s->traces);
    s->nrec = r->n;
    s->mask = r->comp_mask;
    s->ncomp = ((r->comp_mask & 1) != 0) + ((r->comp_mask & 2) != 0) + ((r->comp_mask & 4) != 0);
    s->dec = r->decimate;
    s->nsamp = r->nsamp;
    if (s->nrec == 0) return POE2D_OK;
    if ((st = dev_alloc(&q, (size_t)r->n * sizeof(int)))) return st;
    s->rix = (int *)q;
    if ((st = dev_alloc(&q, (size_t)r->n * sizeof(int)))) return st;
    s->riz = (int *)q;
    if ((st = dev_alloc_zero(&s->traces, s->trace_bytes(), s->stream))) return st;
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

poe2d_status poe2d_set_field(poe2d_solver *s, int field, const void *in, size_t bytes)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (!in) return fail(POE2D_E_INVALID, "set_field: null input");
    if (field < POE2D_FIELD_VX || field > POE2D_FIELD_SXY)
        return fail(POE2D_E_INVALID, "set_field: only the state fields VX .. SXY can be set");
    const size_t need = (size_t)(s->grid.nx + 1) * (s->grid.nz + 1) * s->rs;
    if (bytes != need) return fail(POE2D_E_INVALID, "set_field: input must hold (nx+1)(nz+1) values");
    if (!s->dense && (st = dev_alloc_zero(&s->dense, need, s->stream))) return st;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(s->dense, in, need, hipMemcpyHostToDevice));
    // V = (vx, vy, wx, wy), S = (e_xx, e_yy, zeta, sxy) of the current state buffer
    const int g = field < POE2D_FIELD_EXX ? group_v(s->cur) : group_s(s->cur);
    const int comp = field & 3;
    if (s->precision == POE2D_F32)
        launch_insert<float>(step_args<float>(s, s->cur, s->cur), g, comp, (const float *)s->dense, s->stream);
    else
        launch_insert<double>(step_args<double>(s, s->cur, s->cur), g, comp, (const double *)s->dense, s->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    return POE2D_OK;
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
    h.version = 3; h.precision = s->precision; h.nx = s->grid.nx; h.nz = s->grid.nz;
    h.step = s->step; h.nrec = s->nrec; h.ncomp = s->ncomp; h.nsamp = s->nsamp;
    h.wxs = s->m->wxs; h.wzs = s->m->wzs; h.pitch = s->m->pitch; h.alloc_rows = s->m->alloc_rows;
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
    if (std::memcmp(h.magic, "POE2DCK", 8) || h.version != 3 || h.precision != s->precision ||
        h.nx != s->grid.nx || h.nz != s->grid.nz || h.nrec != s->nrec || h.ncomp != s->ncomp ||
        h.nsamp != s->nsamp || h.wxs != s->m->wxs || h.wzs != s->m->wzs || h.pitch != s->m->pitch ||
        h.alloc_rows != s->m->alloc_rows)
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
    HIP_TRY(hipMemsetAsync(s->arena, 0, NSTATE_GROUPS * s->group_bytes(), s->stream));
    HIP_TRY(hipMemsetAsync(s->psi, 0, 2 * s->m->psi_half_elems() * s->rs, s->stream));
    if (s->traces) HIP_TRY(hipMemsetAsync(s->traces, 0, s->trace_bytes(), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->step = 0;
    s->cur = 0;
    return POE2D_OK;
}

}  // extern "C"

namespace {

// Tile shapes found by the autotuners, per grid class (device, precision, kernel, palette use and
// the grid shapes), for the life of the process.
std::mutex g_tune_mutex;
std::map<std::string, std::pair<int, int>> g_tune_cache;

std::string tune_key(const std::vector<poe2d_solver *> &members, int kernel)
{
    std::string k = std::to_string(members[0]->device) + "/" + std::to_string(members[0]->precision) + "/" +
                    std::to_string(kernel);
    std::vector<std::string> shapes;
    for (const poe2d_solver *s : members)
        shapes.push_back(std::to_string(s->grid.nx) + "x" + std::to_string(s->grid.nz) +
                         (s->m->npal == 0 ? "" : s->m->npal <= kMaxPalette2 ? "p" : "q"));
    std::sort(shapes.begin(), shapes.end());
    for (const std::string &q : shapes) k += "/" + q;
    return k;
}

bool tune_lookup(const std::string &key, int *w, int *h)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = g_tune_cache.find(key);
    if (it == g_tune_cache.end()) return false;
    *w = it->second.first;
    *h = it->second.second;
    return true;
}

void tune_store(const std::string &key, int w, int h)
{
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune_cache[key] = {w, h};
}

}  // namespace

extern "C" {

poe2d_status poe2d_autotune(poe2d_solver *s, int nsteps, int *block_w, int *tile_rows)
{
    poe2d_status st = check_handle(s);
    if (st) return st;
    if (s->schedule == SCHED_TWOPASS) return fail(POE2D_E_INVALID, "autotune: the two-pass schedule has no tile shape");
    if (s->step != 0) return fail(POE2D_E_STATE, "autotune: needs a handle at step 0 (call poe2d_reset first)");
    const std::string key = tune_key({s}, effective_schedule(s));
    if (nsteps >= 0) {
        int w, h;
        if (tune_lookup(key, &w, &h)) {
            s->block_w = w;
            s->tile_rows = h;
            if (block_w) *block_w = w;
            if (tile_rows) *tile_rows = h;
            return POE2D_OK;
        }
    }
    nsteps = nsteps == 0 ? 10 : std::abs(nsteps);
    // Tile heights of 6k + 2 rows (6k for the two-step kernel) make whole groups of the kernels'
    // 6-step row loops.
    const bool two = two_step_kernel(s);
    const std::vector<int> widths = two ? std::vector<int>{64, 128, 256} : std::vector<int>{128, 256, 512};
    const std::vector<int> heights = two ? std::vector<int>{6, 12, 18, 24, 30, 42, 60}
                                         : std::vector<int>{8, 14, 20, 26, 32, 44, 62, 98};
    // One timed run per shape, then the three fastest timed again over three times the steps, twice
    // each, the best of those kept: shapes within a few per cent of each other are told apart by
    // more than one sample.
    std::vector<std::pair<double, std::pair<int, int>>> cand;
    for (int w : widths)
        for (int h : heights) {
            if (h > s->grid.nz + 1 && h != heights[0]) continue;
            s->block_w = w;
            s->tile_rows = h;
            if ((st = poe2d_run(s, 1)) || (st = poe2d_run(s, nsteps))) break;
            cand.push_back({s->last_ms, {w, h}});
        }
    int best_w = s->block_w, best_h = s->tile_rows;
    if (!st && !cand.empty()) {
        std::sort(cand.begin(), cand.end());
        double best = 0;
        for (size_t c = 0; c < std::min<size_t>(3, cand.size()) && !st; ++c) {
            s->block_w = cand[c].second.first;
            s->tile_rows = cand[c].second.second;
            for (int r = 0; r < 2 && !st; ++r) {
                if ((st = poe2d_run(s, 1)) || (st = poe2d_run(s, 3 * nsteps))) break;
                if (best == 0 || s->last_ms < best) {
                    best = s->last_ms;
                    best_w = s->block_w;
                    best_h = s->tile_rows;
                }
            }
        }
    }
    s->block_w = best_w;
    s->tile_rows = best_h;
    const poe2d_status rst = poe2d_reset(s);
    if (st) return st;
    if (rst) return rst;
    tune_store(key, best_w, best_h);
    if (block_w) *block_w = best_w;
    if (tile_rows) *tile_rows = best_h;
    return POE2D_OK;
}

void poe2d_destroy(poe2d_solver *s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_handle(s);
    delete s;
}

// ---- batches

poe2d_status poe2d_batch_create(int n, poe2d_solver *const *members, poe2d_batch **out)
{
    if (!out) return fail(POE2D_E_INVALID, "batch_create: null output pointer");
    *out = nullptr;
    if (n < 1 || !members) return fail(POE2D_E_INVALID, "batch_create: needs at least one member");
    for (int i = 0; i < n; ++i) {
        poe2d_solver *s = members[i];
        if (!s) return fail(POE2D_E_INVALID, "batch_create: null member");
        if (s->device != members[0]->device || s->precision != members[0]->precision)
            return fail(POE2D_E_INVALID, "batch_create: members must share device and precision");
        if (s->schedule == SCHED_TWOPASS)
            return fail(POE2D_E_INVALID, "batch_create: members must use a single-pass schedule");
        for (int j = 0; j < i; ++j)
            if (members[j] == s) return fail(POE2D_E_INVALID, "batch_create: a handle appears twice");
    }
    poe2d_batch *b = new (std::nothrow) poe2d_batch;
    if (!b) return fail(POE2D_E_NOMEM, "batch_create: out of host memory");
    b->members.assign(members, members + n);
    b->device = members[0]->device;
    b->precision = members[0]->precision;
    b->block_w = members[0]->block_w;
    b->tile_rows = members[0]->tile_rows;
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        poe2d_batch_destroy(b);
        return fail(POE2D_E_HIP, "batch_create: stream/event creation failed");
    }
    *out = b;
    return POE2D_OK;
}

poe2d_status poe2d_batch_set_schedule(poe2d_batch *b, int block_w, int tile_rows)
{
    poe2d_status st = check_batch(b);
    if (st) return st;
    if (block_w != 64 && block_w != 128 && block_w != 256 && block_w != 512)
        return fail(POE2D_E_INVALID, "batch_set_schedule: block width must be 64 (two-step only), 128, 256 or 512 "
                                     "(single-step only)");
    if (tile_rows < 4 || tile_rows > 1 << 20) return fail(POE2D_E_INVALID, "batch_set_schedule: tile rows out of range");
    b->block_w = block_w;
    b->tile_rows = tile_rows;
    return POE2D_OK;
}

poe2d_status poe2d_batch_set_graphs(poe2d_batch *b, int enable)
{
    poe2d_status st = check_batch(b);
    if (st) return st;
    b->graphs = enable != 0;
    return POE2D_OK;
}

poe2d_status poe2d_batch_run(poe2d_batch *b, int nsteps)
{
    poe2d_status st = check_batch(b);
    if (st) return st;
    if (nsteps < 0) return fail(POE2D_E_INVALID, "batch_run: negative step count");
    for (poe2d_solver *s : b->members) {
        if (s->step != b->members[0]->step)
            return fail(POE2D_E_STATE, "batch_run: members are at different steps");
        if (s->schedule == SCHED_TWOPASS)
            return fail(POE2D_E_INVALID, "batch_run: members must use a single-pass schedule");
    }
    if (nsteps == 0) return POE2D_OK;
    return b->precision == POE2D_F32 ? batch_run_steps<float>(b, nsteps) : batch_run_steps<double>(b, nsteps);
}

poe2d_status poe2d_batch_last_run_ms(poe2d_batch *b, double *ms, int *nsteps)
{
    if (!b || !ms || !nsteps) return fail(POE2D_E_INVALID, "batch_last_run_ms: null argument");
    *ms = b->last_ms;
    *nsteps = b->last_steps;
    return POE2D_OK;
}

poe2d_status poe2d_batch_autotune(poe2d_batch *b, int nsteps, int *block_w, int *tile_rows)
{
    poe2d_status st = check_batch(b);
    if (st) return st;
    for (poe2d_solver *s : b->members)
        if (s->step != 0) return fail(POE2D_E_STATE, "batch_autotune: needs every member at step 0");
    int kinds = 0;
    for (poe2d_solver *s : b->members) kinds |= 1 << effective_schedule(s);
    const std::string key = "batch/" + tune_key(b->members, kinds);
    if (nsteps >= 0) {
        int w, h;
        if (tune_lookup(key, &w, &h)) {
            b->block_w = w;
            b->tile_rows = h;
            if (block_w) *block_w = w;
            if (tile_rows) *tile_rows = h;
            return POE2D_OK;
        }
    }
    nsteps = nsteps == 0 ? 10 : std::abs(nsteps);
    bool two = b->precision == POE2D_F32;
    for (poe2d_solver *s : b->members) two = two && two_step_kernel(s);
    const std::vector<int> widths = two ? std::vector<int>{64, 128, 256} : std::vector<int>{128, 256, 512};
    const std::vector<int> heights = two ? std::vector<int>{6, 12, 18, 24, 30, 42, 60}
                                         : std::vector<int>{8, 14, 20, 26, 32, 44, 62, 98};
    int nzmax = 0;
    for (poe2d_solver *s : b->members) nzmax = std::max(nzmax, s->grid.nz);
    // as poe2d_autotune: one timed run per shape, the three fastest timed again
    std::vector<std::pair<double, std::pair<int, int>>> cand;
    for (int w : widths)
        for (int h : heights) {
            if (h > nzmax + 1 && h != heights[0]) continue;
            b->block_w = w;
            b->tile_rows = h;
            if ((st = poe2d_batch_run(b, 1)) || (st = poe2d_batch_run(b, nsteps))) break;
            cand.push_back({b->last_ms, {w, h}});
        }
    int best_w = b->block_w, best_h = b->tile_rows;
    if (!st && !cand.empty()) {
        std::sort(cand.begin(), cand.end());
        double best = 0;
        for (size_t c = 0; c < std::min<size_t>(3, cand.size()) && !st; ++c) {
            b->block_w = cand[c].second.first;
            b->tile_rows = cand[c].second.second;
            for (int r = 0; r < 2 && !st; ++r) {
                if ((st = poe2d_batch_run(b, 1)) || (st = poe2d_batch_run(b, 3 * nsteps))) break;
                if (best == 0 || b->last_ms < best) {
                    best = b->last_ms;
                    best_w = b->block_w;
                    best_h = b->tile_rows;
                }
            }
        }
    }
    b->block_w = best_w;
    b->tile_rows = best_h;
    poe2d_status rst = POE2D_OK;
    for (poe2d_solver *s : b->members) {
        const poe2d_status r = poe2d_reset(s);
        if (r && !rst) rst = r;
    }
    if (st) return st;
    if (rst) return rst;
    tune_store(key, best_w, best_h);
    if (block_w) *block_w = best_w;
    if (tile_rows) *tile_rows = best_h;
    return POE2D_OK;
}

void poe2d_batch_destroy(poe2d_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->list.release();
    b->list2.release();
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
